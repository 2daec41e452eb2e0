"""Pin the open-boundary coefficient tables.

Two independent checks:

1. Where the reference tree is mounted, its generated boundary code
   (include/boundaryConditions/applyMBBC.h, restoreRho.h, restoreUxUyUz.h) is
   parsed as text and every coefficient, unknown list, selected moment and term
   order is compared with lbm3_amd/csrc/generated/boundary_tables.h.
2. Always: the MBBC solution reproduces the prescribed equilibrium moments for
   random known populations (the defining property of the method), and
   restoreRho / restoreUxUyUz invert the boundary relations for equilibrium
   populations.
"""
import os
import re
from fractions import Fraction

import numpy as np
import pytest

from tests.table_reader import load_tables

REF = "/root/reference/include/boundaryConditions"

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1])
W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
MOMENTS = ["000", "100", "010", "001", "200", "020", "002", "011", "101", "110",
           "111", "210", "201", "120", "021", "102", "012", "022", "202", "220", "211", "121", "112"]
# equilibrium raw-moment expressions exactly as the reference generator spells them
MEQ = ["rho", "rho * ux", "rho * uy", "rho * uz",
       "(1.f/3.f) * rho + rho * ux * ux", "(1.f/3.f) * rho + rho * uy * uy", "(1.f/3.f) * rho + rho * uz * uz",
       "rho * uy * uz", "rho * ux * uz", "rho * ux * uy", "0.f",
       "(1.f/3.f) * rho * uy + rho * ux * ux * uy", "(1.f/3.f) * rho * uz + rho * ux * ux * uz",
       "(1.f/3.f) * rho * ux + rho * ux * uy * uy", "(1.f/3.f) * rho * uz + rho * uy * uy * uz",
       "(1.f/3.f) * rho * ux + rho * ux * uz * uz", "(1.f/3.f) * rho * uy + rho * uy * uz * uz",
       "(1.f/9.f) * rho + (1.f/3.f) * rho * uy * uy + (1.f/3.f) * rho * uz * uz + rho * uy * uy * uz * uz",
       "(1.f/9.f) * rho + (1.f/3.f) * rho * ux * ux + (1.f/3.f) * rho * uz * uz + rho * ux * ux * uz * uz",
       "(1.f/9.f) * rho + (1.f/3.f) * rho * ux * ux + (1.f/3.f) * rho * uy * uy + rho * ux * ux * uy * uy",
       "(1.f/3.f) * rho * uy * uz", "(1.f/3.f) * rho * ux * uz", "(1.f/3.f) * rho * ux * uy"]


def frac(num, den):
    return Fraction(int(num), int(den))


def parse_coef(tok):
    tok = tok.strip()
    m = re.fullmatch(r"\((-?\d+)\.f/(\d+)\.f\)", tok)
    if m:
        return Fraction(int(m.group(1)), int(m.group(2)))
    m = re.fullmatch(r"\((-?\d+)\.f\)", tok)
    if m:
        return Fraction(int(m.group(1)))
    m = re.fullmatch(r"-?\d+", tok)
    if m:
        return Fraction(int(tok))
    raise ValueError(tok)


def parse_terms(expr, var):
    """Parse '+ a * var[i] - var[j] + var[k]' into an ordered list of (index, Fraction)."""
    terms = []
    pat = re.compile(r"([+-])\s*(?:(\([^()]*\))\s*\*\s*)?" + var + r"(\d+)\]?")
    pos = 0
    expr = expr.strip()
    while pos < len(expr):
        m = pat.match(expr, pos)
        if not m:
            raise ValueError("cannot parse %r at %d" % (expr, pos))
        sign = 1 if m.group(1) == "+" else -1
        coef = parse_coef(m.group(2)) if m.group(2) else Fraction(1)
        terms.append((int(m.group(3)), sign * coef))
        pos = m.end()
        while pos < len(expr) and expr[pos] == " ":
            pos += 1
    return terms


def branches(path):
    text = open(path).read()
    parts = re.split(r"if \( normalCode == (\d+) \)", text)
    out = {}
    for code, body in zip(parts[1::2], parts[2::2]):
        out[int(code)] = body
    return out


T = load_tables()


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_mbbc_table_matches_reference_generated_code():
    br = branches(os.path.join(REF, "applyMBBC.h"))
    assert len(br) == 26
    for n in range(26):
        body = br[T["NORMAL_CODE"][n]]
        nu = T["MBBC_NU"][n]
        # known-population sign rows
        smf = re.findall(r"const float smf(\d+) = (.*);", body)
        assert len(smf) == nu
        for r, (idx, expr) in enumerate(smf):
            terms = parse_terms(expr, r"f\[")
            row = np.zeros(27, int)
            for q, c in terms:
                row[q] = int(c)
            assert [q for q, _ in terms] == sorted(q for q, _ in terms)
            assert (row == np.array(T["MBBC_SMF"][n][r])).all(), (n, r)
        # selected moments
        ms = re.findall(r"const float m(\d+) = (.*);", body)
        assert [MEQ.index(e) for _, e in ms] == list(T["MBBC_MOMENT"][n][:nu])
        # unknowns and inverse
        rows = re.findall(r"\n\s*f\[(\d+)\] = (.*);", body)
        assert [int(u) for u, _ in rows] == list(T["MBBC_UNKNOWN"][n][:nu])
        for r, (_, expr) in enumerate(rows):
            got = {j: c for j, c in parse_terms(expr, "s")}
            for j in range(nu):
                want = frac(T["MBBC_INV_NUM"][n][r][j], T["MBBC_INV_DEN"][n][r][j])
                assert got.get(j, Fraction(0)) == want, (n, r, j)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_restore_rho_table_matches_reference_generated_code():
    br = branches(os.path.join(REF, "restoreRho.h"))
    for n in range(26):
        body = br[T["NORMAL_CODE"][n]]
        scqu = re.search(r"const float scqu = (.*);", body).group(1)
        terms = re.findall(r"\+ (\([^()]*\)) \* u([xyz])", scqu)
        assert [a for _, a in terms] == ["x", "y", "z"]
        for d, (c, _) in enumerate(terms):
            assert parse_coef(c) == frac(T["RHO_U_NUM"][n][d], T["RHO_U_DEN"][n][d])
        s = re.search(r"const float s = (.*) \+ scqu;", body).group(1)
        assert parse_coef(s) == frac(T["RHO_S_NUM"][n], T["RHO_S_DEN"][n])
        scmf = re.search(r"const float scmf = (.*);", body).group(1)
        got = parse_terms(scmf, r"f\[")
        want = [(q, frac(T["RHO_F_NUM"][n][q], T["RHO_F_DEN"][n][q])) for q in range(27)
                if T["RHO_F_NUM"][n][q] != 0]
        assert got == want, n


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_restore_u_table_matches_reference_generated_code():
    br = branches(os.path.join(REF, "restoreUxUyUz.h"))
    for n in range(6, 26):
        body = br[T["NORMAL_CODE"][n]]
        for r in range(3):
            expr = re.search(r"const float scmf%d = \(\s*(.*)\) / rho;" % r, body).group(1)
            got = parse_terms(expr, r"f\[")
            want = [(q, frac(T["U_F_NUM"][n][r][q], T["U_F_DEN"][n][r][q])) for q in range(27)
                    if T["U_F_NUM"][n][r][q] != 0]
            assert got == want, (n, r)
            q = re.search(r"const float s%d = scmf%d - (.*);" % (r, r), body).group(1)
            assert parse_coef(q) == frac(T["U_Q_NUM"][n][r], T["U_Q_DEN"][n][r])
        for d, ax in enumerate("xyz"):
            expr = re.search(r"\n\s*u%s = (.*);" % ax, body).group(1)
            got = {j: c for j, c in parse_terms(expr, "s")}
            for j in range(3):
                assert got.get(j, Fraction(0)) == frac(T["U_INV_NUM"][n][d][j], T["U_INV_DEN"][n][d][j])


def _meq(mid, rho, ux, uy, uz):
    a, b, c = (int(ch) for ch in MOMENTS[mid])
    # raw moment of the D3Q27 product-form equilibrium, exact in this basis
    def one(e, u):
        return {0: 1.0, 1: u, 2: 1.0 / 3.0 + u * u}[e]
    if (a, b, c) == (1, 1, 1):
        return 0.0  # the generator prescribes zero for m_111
    if sorted((a, b, c)) == [1, 1, 2]:
        # the generator keeps only the (1/3) rho u u part for 211-type moments
        us = [ux, uy, uz]
        odd = [i for i, e in enumerate((a, b, c)) if e == 1]
        return rho / 3.0 * us[odd[0]] * us[odd[1]]
    return rho * one(a, ux) * one(b, uy) * one(c, uz)


def test_mbbc_reproduces_prescribed_moments():
    rng = np.random.default_rng(1)
    for n in range(26):
        nu = T["MBBC_NU"][n]
        unk = T["MBBC_UNKNOWN"][n][:nu]
        f = W * (1 + 0.05 * rng.standard_normal(27))
        rho, ux, uy, uz = 1.02, 0.03, -0.02, 0.01
        s = []
        for r in range(nu):
            mid = T["MBBC_MOMENT"][n][r]
            s.append(_meq(mid, rho, ux, uy, uz) - np.dot(T["MBBC_SMF"][n][r], f))
        for r, q in enumerate(unk):
            f[q] = sum(frac(T["MBBC_INV_NUM"][n][r][j], T["MBBC_INV_DEN"][n][r][j]) * s[j] for j in range(nu))
            f[q] = float(f[q])
        for r in range(nu):
            mid = T["MBBC_MOMENT"][n][r]
            a, b, c = (int(ch) for ch in MOMENTS[mid])
            m = np.sum(CX ** a * CY ** b * CZ ** c * f)
            assert abs(m - _meq(mid, rho, ux, uy, uz)) < 1e-12, (n, MOMENTS[mid])


def _linear_moment_populations(rho, u):
    """Populations whose 23 raw moments equal the linear model the restoration
    assumes: m = Q [rho, rho u] (rho on even moments with 1/3 weights, rho u on
    the first-order-in-u moments).  Least-norm solution of M f = m."""
    mom = np.array([[CX[q] ** int(m[0]) * CY[q] ** int(m[1]) * CZ[q] ** int(m[2]) for q in range(27)]
                    for m in MOMENTS], float)
    target = []
    for m in MOMENTS:
        a, b, c = (int(ch) for ch in m)
        e = (a, b, c)
        val = 0.0
        if all(x % 2 == 0 for x in e):
            val = rho * (1.0 / 3.0) ** sum(1 for x in e if x == 2)
        odd = [i for i, x in enumerate(e) if x % 2 == 1]
        if len(odd) == 1 and e[odd[0]] == 1 and all(x in (0, 2) for i, x in enumerate(e) if i != odd[0]):
            val = rho * u[odd[0]] * (1.0 / 3.0) ** sum(1 for i, x in enumerate(e) if i != odd[0] and x == 2)
        target.append(val)
    f, *_ = np.linalg.lstsq(mom, np.array(target), rcond=None)
    assert np.allclose(mom @ f, target, atol=1e-13)
    return f


def test_restore_rho_inverts_linear_moment_relation():
    rho, u = 1.013, np.array([0.021, -0.017, 0.012])
    f = _linear_moment_populations(rho, u)
    for n in range(26):
        num = sum(T["RHO_F_NUM"][n][q] / T["RHO_F_DEN"][n][q] * f[q] for q in range(27))
        den = T["RHO_S_NUM"][n] / T["RHO_S_DEN"][n] + sum(
            T["RHO_U_NUM"][n][d] / T["RHO_U_DEN"][n][d] * u[d] for d in range(3))
        assert abs(num / den - rho) < 1e-12, n


def test_restore_u_inverts_linear_moment_relation():
    rho, u = 1.013, np.array([0.021, -0.017, 0.012])
    f = _linear_moment_populations(rho, u)
    for n in range(6, 26):
        s = []
        for r in range(3):
            acc = sum(T["U_F_NUM"][n][r][q] / T["U_F_DEN"][n][r][q] * f[q] for q in range(27)) / rho
            s.append(acc - T["U_Q_NUM"][n][r] / T["U_Q_DEN"][n][r])
        got = [sum(T["U_INV_NUM"][n][d][j] / T["U_INV_DEN"][n][d][j] * s[j] for j in range(3)) for d in range(3)]
        assert np.allclose(got, u, atol=1e-12), n


