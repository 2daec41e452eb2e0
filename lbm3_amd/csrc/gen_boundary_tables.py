#!/usr/bin/env python3
"""Generate the moment-based open-boundary coefficient tables.

The open boundary of the LBM3 solver (reference: include/boundaryConditions/
applyMBBC.h, restoreRho.h, restoreUxUyUz.h, produced by
tools/CppCodeGeneratorForMBBC/helperMain.py) reconstructs the populations that
enter the domain through a boundary cell from a set of prescribed equilibrium
moments.  For every one of the 26 outer normals (6 faces, 12 edges, 8 corners)
the method is fixed by linear algebra over the D3Q27 raw moments:

* known populations: every direction c with min(c * n) >= 0 (componentwise);
  unknown populations: the rest (9 on a face, 15 on an edge, 19 on a corner);
* MBBC: greedily pick, in the fixed raw-moment order below, moment rows that
  raise the rank of the unknown-population block until it is square, then
  invert it: f_u = (S M_u)^-1 (m_eq - S M_k f_k);
* restoreRho / restoreUxUyUz: the left null space C of M_u gives moment
  combinations that do not involve unknowns; sorted by total moment order, the
  first one determines rho (given u) and three independent ones determine u
  (given rho) on edges and corners.  Faces use a closed form for u.

This script re-derives those matrices with exact rational arithmetic (sympy)
and writes them as coefficient tables.  Evaluation order of every sum is the
increasing index order of the table, which is the order the generated reference
code uses, so a consumer that accumulates left to right without contraction
reproduces the reference arithmetic bit for bit.

Output: lbm3_amd/csrc/generated/boundary_tables.h (C99 and C++/HIP).
"""
import os
import sys

import sympy as sp

CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1]
CY = [0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1]
CZ = [0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1]

# Raw moments (exponents of cx, cy, cz) in the order the selection walks them.
MOMENTS = ["000", "100", "010", "001",
           "200", "020", "002", "011", "101", "110",
           "111", "210", "201", "120", "021", "102", "012",
           "022", "202", "220", "211", "121", "112"]
MOMENT_ORDER = [sum(int(ch) for ch in m) for m in MOMENTS]

# Equilibrium raw moments as coefficient columns over (rho, rho*ux, rho*uy, rho*uz):
# only the parts linear in rho or in rho*u are needed for the null-space systems.
THIRD, NINTH = sp.Rational(1, 3), sp.Rational(1, 9)


def _eq_linear_parts():
    q_all, q_u, q_rho = [], [], []
    for m in MOMENTS:
        a, b, c = (int(ch) for ch in m)
        row = [0, 0, 0, 0]
        # isotropic part: rho * (1/3)^(#even nonzero exponents == 2) for pure even moments
        if all(e % 2 == 0 for e in (a, b, c)):
            n2 = sum(1 for e in (a, b, c) if e == 2)
            row[0] = THIRD ** n2
        # first-order-in-u part for moments with exactly one odd exponent equal to 1
        odd = [ax for ax, e in enumerate((a, b, c)) if e % 2 == 1]
        if len(odd) == 1 and (a, b, c)[odd[0]] == 1:
            others = [e for ax, e in enumerate((a, b, c)) if ax != odd[0]]
            if all(e in (0, 2) for e in others):
                n2 = sum(1 for e in others if e == 2)
                row[1 + odd[0]] = THIRD ** n2
        q_all.append(row)
        q_u.append(row[1:])
        q_rho.append([row[0]])
    return sp.Matrix(q_all), sp.Matrix(q_u), sp.Matrix(q_rho)


Q_ALL, Q_U, Q_RHO = _eq_linear_parts()

# Normals in the order the reference generator enumerates them (applyMBBC.h branches).
NORMALS = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
           [0, 1, 1], [1, 0, 1], [1, 1, 0], [0, -1, -1], [-1, 0, -1], [-1, -1, 0],
           [0, 1, -1], [1, 0, -1], [1, -1, 0], [0, -1, 1], [-1, 0, 1], [-1, 1, 0],
           [1, 1, 1], [-1, -1, -1],
           [-1, 1, 1], [1, -1, 1], [1, 1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]


def moment_row(m, dirs):
    a, b, c = (int(ch) for ch in m)
    return [CX[q] ** a * CY[q] ** b * CZ[q] ** c for q in dirs]


def greedy_rows(mat, count, first=0):
    rows = [first]
    i = first + 1
    while len(rows) < count:
        if i >= mat.rows:
            raise RuntimeError("ran out of independent rows")
        if mat.extract(rows + [i], list(range(mat.cols))).rank() > mat.extract(rows, list(range(mat.cols))).rank():
            rows.append(i)
        i += 1
    return rows


def derive(normal):
    known = [q for q in range(27) if min(CX[q] * normal[0], CY[q] * normal[1], CZ[q] * normal[2]) >= 0]
    unknown = [q for q in range(27) if q not in known]
    mk = sp.Matrix([moment_row(m, known) for m in MOMENTS])
    mu = sp.Matrix([moment_row(m, unknown) for m in MOMENTS])

    # MBBC moment selection and inverse.
    sel = greedy_rows(mu, len(unknown))
    s_mu = mu.extract(sel, list(range(mu.cols)))
    inv = s_mu.inv()
    smk = mk.extract(sel, list(range(mk.cols)))

    # Left null space of M_u, sorted (stably) by summed order of its nonzero entries.
    null = mu.T.nullspace()

    def key(v):
        return sum(MOMENT_ORDER[i] for i, x in enumerate(v) if x != 0)

    null = sorted(null, key=key)
    c_mat = sp.Matrix.hstack(*null)
    ct = c_mat.T

    # rho from the first null combination.
    rho_s = (ct * Q_RHO)[0, 0]
    rho_u = list((ct * Q_U).row(0))
    rho_f = list((ct * mk).row(0))

    # u from three independent combinations (edges and corners).
    ctqu = ct * Q_U
    usel = greedy_rows(ctqu, 3)
    u_inv = ctqu.extract(usel, [0, 1, 2]).inv()
    u_f = [list((ct * mk).row(r)) for r in usel]
    u_q = [(ct * Q_RHO)[r, 0] for r in usel]
    return dict(known=known, unknown=unknown, sel=sel, inv=inv, smk=smk,
                rho_s=rho_s, rho_u=rho_u, rho_f=rho_f,
                u_inv=u_inv, u_f=u_f, u_q=u_q)


def nd(x):
    x = sp.Rational(x)
    return int(x.p), int(x.q)


def c_array(name, ctype, dims, values):
    def fmt(v, depth):
        if depth == len(dims):
            return str(v)
        return "{" + ",".join(fmt(e, depth + 1) for e in v) + "}"
    dimstr = "".join("[%d]" % d for d in dims)
    return "LBM3_BT_CONST %s %s%s = %s;\n" % (ctype, name, dimstr, fmt(values, 0))


def full_row(values, known):
    row = [0] * 27
    for v, q in zip(values, known):
        row[q] = v
    return row


def main(out_path):
    nu_max = 19
    mb_nu, mb_unknown, mb_moment, mb_smf, mb_inv_n, mb_inv_d = [], [], [], [], [], []
    rho_s_n, rho_s_d, rho_u_n, rho_u_d, rho_f_n, rho_f_d = [], [], [], [], [], []
    u_f_n, u_f_d, u_q_n, u_q_d, u_inv_n, u_inv_d = [], [], [], [], [], []
    for normal in NORMALS:
        d = derive(normal)
        n = len(d["unknown"])
        mb_nu.append(n)
        mb_unknown.append(d["unknown"] + [0] * (nu_max - n))
        mb_moment.append(d["sel"] + [0] * (nu_max - n))
        smf = [full_row([int(x) for x in d["smk"].row(r)], d["known"]) for r in range(n)]
        mb_smf.append(smf + [[0] * 27] * (nu_max - n))
        inv_n = [[0] * nu_max for _ in range(nu_max)]
        inv_d = [[1] * nu_max for _ in range(nu_max)]
        for r in range(n):
            for c in range(n):
                inv_n[r][c], inv_d[r][c] = nd(d["inv"][r, c])
        mb_inv_n.append(inv_n)
        mb_inv_d.append(inv_d)

        a, b = nd(d["rho_s"])
        rho_s_n.append(a)
        rho_s_d.append(b)
        rho_u_n.append([nd(x)[0] for x in d["rho_u"]])
        rho_u_d.append([nd(x)[1] for x in d["rho_u"]])
        rf = full_row(d["rho_f"], d["known"])
        rho_f_n.append([nd(x)[0] for x in rf])
        rho_f_d.append([nd(x)[1] for x in rf])

        ufn, ufd = [], []
        for row in d["u_f"]:
            fr = full_row(row, d["known"])
            ufn.append([nd(x)[0] for x in fr])
            ufd.append([nd(x)[1] for x in fr])
        u_f_n.append(ufn)
        u_f_d.append(ufd)
        u_q_n.append([nd(x)[0] for x in d["u_q"]])
        u_q_d.append([nd(x)[1] for x in d["u_q"]])
        u_inv_n.append([[nd(d["u_inv"][r, c])[0] for c in range(3)] for r in range(3)])
        u_inv_d.append([[nd(d["u_inv"][r, c])[1] for c in range(3)] for r in range(3)])

    codes = [(nx + 5) * 100 + (ny + 5) * 10 + (nz + 5) for nx, ny, nz in NORMALS]
    out = []
    out.append("/* Generated by lbm3_amd/csrc/gen_boundary_tables.py. Do not edit. */\n")
    out.append("#ifndef LBM3_BOUNDARY_TABLES_H\n#define LBM3_BOUNDARY_TABLES_H\n")
    out.append("#ifdef __cplusplus\n#define LBM3_BT_CONST static constexpr\n#else\n#define LBM3_BT_CONST static const\n#endif\n")
    out.append("#define LBM3_BT_NORMAL_COUNT 26\n#define LBM3_BT_UNKNOWN_MAX 19\n")
    out.append("/* Raw-moment ids: " + " ".join(MOMENTS) + " */\n")
    out.append(c_array("LBM3_BT_NORMAL", "signed char", [26, 3], NORMALS))
    out.append(c_array("LBM3_BT_NORMAL_CODE", "int", [26], codes))
    out.append("/* MBBC: unknown count, unknown directions, selected moments, known-population sign\n"
               "   rows (over all 27 directions, zero for unknowns), inverse (numerator/denominator). */\n")
    out.append(c_array("LBM3_BT_MBBC_NU", "unsigned char", [26], mb_nu))
    out.append(c_array("LBM3_BT_MBBC_UNKNOWN", "unsigned char", [26, nu_max], mb_unknown))
    out.append(c_array("LBM3_BT_MBBC_MOMENT", "unsigned char", [26, nu_max], mb_moment))
    out.append(c_array("LBM3_BT_MBBC_SMF", "signed char", [26, nu_max, 27], mb_smf))
    out.append(c_array("LBM3_BT_MBBC_INV_NUM", "short", [26, nu_max, nu_max], mb_inv_n))
    out.append(c_array("LBM3_BT_MBBC_INV_DEN", "short", [26, nu_max, nu_max], mb_inv_d))
    out.append("/* restoreRho: rho = (sum_q F[q] f_q) / (S + U . u). */\n")
    out.append(c_array("LBM3_BT_RHO_S_NUM", "short", [26], rho_s_n))
    out.append(c_array("LBM3_BT_RHO_S_DEN", "short", [26], rho_s_d))
    out.append(c_array("LBM3_BT_RHO_U_NUM", "short", [26, 3], rho_u_n))
    out.append(c_array("LBM3_BT_RHO_U_DEN", "short", [26, 3], rho_u_d))
    out.append(c_array("LBM3_BT_RHO_F_NUM", "short", [26, 27], rho_f_n))
    out.append(c_array("LBM3_BT_RHO_F_DEN", "short", [26, 27], rho_f_d))
    out.append("/* restoreUxUyUz on edges and corners: s_r = (sum_q F[r][q] f_q)/rho - Q[r]; u = INV s.\n"
               "   Faces (first six normals) use the closed-form face restoration instead. */\n")
    out.append(c_array("LBM3_BT_U_F_NUM", "short", [26, 3, 27], u_f_n))
    out.append(c_array("LBM3_BT_U_F_DEN", "short", [26, 3, 27], u_f_d))
    out.append(c_array("LBM3_BT_U_Q_NUM", "short", [26, 3], u_q_n))
    out.append(c_array("LBM3_BT_U_Q_DEN", "short", [26, 3], u_q_d))
    out.append(c_array("LBM3_BT_U_INV_NUM", "short", [26, 3, 3], u_inv_n))
    out.append(c_array("LBM3_BT_U_INV_DEN", "short", [26, 3, 3], u_inv_d))
    out.append("#endif\n")
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as fh:
        fh.write("".join(out))


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    target = sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "generated", "boundary_tables.h")
    main(target)
