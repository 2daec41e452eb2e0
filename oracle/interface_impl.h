/*
 * ORACLE (test infrastructure only): coarse <-> fine interface transfer of
 * updateInterface.h, instantiated for REAL=float and REAL=double by step.c.
 *
 *   NAME(if_reconstruct)   reconstructInterpolatedF      updateInterface.h:29-153
 *   NAME(if_cell)          stencil values of one cell    updateInterface.h:196-231, :404-438, :719-750
 *   NAME(if_coefficients)  interpolation coefficients    updateInterface.h:235-288 (F->C), :441-558 (C->F)
 *   NAME(if_fine_to_coarse)                              updateInterface.h:155-360
 *   NAME(if_coarse_to_fine)                              updateInterface.h:362-687
 *   NAME(if_leftovers)                                   updateInterface.h:689-884
 *
 * The coefficient sums keep the reference's grouping of the eight stencil
 * corners (corner n = self, i+, j+, ij+, k+, ik+, jk+, ijk+ for n = 0..7).
 */

#define RC(a, b) ((REAL)(a) / (REAL)(b))
#define F_AT(g, q, c) (((REAL *)(g)->f)[(size_t)(q) * (size_t)(g)->Info.cellCount + (size_t)(c)])

/* well-conditioned second-order moments -> shifted populations, third and
   higher cumulants at equilibrium (updateInterface.h:29-153) */
static void NAME(if_reconstruct)(REAL *f, REAL dRho, REAL ux, REAL uy, REAL uz, REAL k011, REAL k101, REAL k110,
                                 REAL k200, REAL k020, REAL k002)
{
    const REAL ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const REAL Kaa0 = RC(1, 36), Kab0 = RC(1, 9), Kac0 = RC(1, 36), Kba0 = RC(1, 9), Kbb0 = RC(4, 9), Kbc0 = RC(1, 9);
    const REAL Kca0 = RC(1, 36), Kcb0 = RC(1, 9), Kcc0 = RC(1, 36);
    const REAL Ka00 = RC(1, 6), Kb00 = RC(2, 3), Kc00 = RC(1, 6), Ka02 = RC(1, 18), Kb02 = RC(2, 9), Kc02 = RC(1, 18);
    const REAL k000 = dRho;
    const REAL k220 = dRho / (REAL)9, k022 = k220, k202 = k220;
    const REAL k222 = dRho / (REAL)27;
    const REAL h = (REAL)0.5, one = (REAL)1, two = (REAL)2, third = RC(1, 3), ninth = RC(1, 9);

    /* x chains */
    const REAL b00 = k000 * (one - ux2) - k200 - ux2;
    const REAL b01 = -two * ux * k101;
    const REAL b02 = k002 * (one - ux2) - k202 - ux2 / (REAL)3;
    const REAL b10 = -two * ux * k110;
    const REAL b11 = k011 * (one - ux2);
    const REAL b20 = k020 * (one - ux2) - k220 - ux2 / (REAL)3;
    const REAL b22 = k022 * (one - ux2) - k222 - ux2 / (REAL)9;
    const REAL a00 = ((k000 + one) * (ux2 - ux) + k200) * h;
    const REAL a01 = (k101 * (two * ux - one)) * h;
    const REAL a02 = ((k002 + third) * (ux2 - ux) + k202) * h;
    const REAL a10 = (k110 * (two * ux - one)) * h;
    const REAL a11 = (k011 * (ux2 - ux)) * h;
    const REAL a20 = ((k020 + third) * (ux2 - ux) + k220) * h;
    const REAL a22 = ((k022 + ninth) * (ux2 - ux) + k222) * h;
    const REAL c00 = ((k000 + one) * (ux2 + ux) + k200) * h;
    const REAL c01 = (k101 * (two * ux + one)) * h;
    const REAL c02 = ((k002 + third) * (ux2 + ux) + k202) * h;
    const REAL c10 = (k110 * (two * ux + one)) * h;
    const REAL c11 = (k011 * (ux2 + ux)) * h;
    const REAL c20 = ((k020 + third) * (ux2 + ux) + k220) * h;
    const REAL c22 = ((k022 + ninth) * (ux2 + ux) + k222) * h;

    /* y chains: centre (b = 0 in y), minus (a), plus (c) */
    const REAL ab0 = a00 * (one - uy2) - two * uy * a10 - a20 - Ka00 * uy2;
    const REAL ab1 = a01 * (one - uy2) - two * uy * a11 - (REAL)0;
    const REAL ab2 = a02 * (one - uy2) - two * uy * (REAL)0 - a22 - Ka02 * uy2;
    const REAL bb0 = b00 * (one - uy2) - two * uy * b10 - b20 - Kb00 * uy2;
    const REAL bb1 = b01 * (one - uy2) - two * uy * b11 - (REAL)0;
    const REAL bb2 = b02 * (one - uy2) - two * uy * (REAL)0 - b22 - Kb02 * uy2;
    const REAL cb0 = c00 * (one - uy2) - two * uy * c10 - c20 - Kc00 * uy2;
    const REAL cb1 = c01 * (one - uy2) - two * uy * c11 - (REAL)0;
    const REAL cb2 = c02 * (one - uy2) - two * uy * (REAL)0 - c22 - Kc02 * uy2;

    const REAL aa0 = ((a00 + Ka00) * (uy2 - uy) + a10 * (two * uy - one) + a20) * h;
    const REAL aa1 = (a01 * (uy2 - uy) + a11 * (two * uy - one) + (REAL)0) * h;
    const REAL aa2 = ((a02 + Ka02) * (uy2 - uy) + a22) * h;
    const REAL ba0 = ((b00 + Kb00) * (uy2 - uy) + b10 * (two * uy - one) + b20) * h;
    const REAL ba1 = (b01 * (uy2 - uy) + b11 * (two * uy - one)) * h;
    const REAL ba2 = ((b02 + Kb02) * (uy2 - uy) + b22) * h;
    const REAL ca0 = ((c00 + Kc00) * (uy2 - uy) + c10 * (two * uy - one) + c20) * h;
    const REAL ca1 = (c01 * (uy2 - uy) + c11 * (two * uy - one)) * h;
    const REAL ca2 = ((c02 + Kc02) * (uy2 - uy) + c22) * h;

    const REAL ac0 = ((a00 + Ka00) * (uy2 + uy) + a10 * (two * uy + one) + a20) * h;
    const REAL ac1 = (a01 * (uy2 + uy) + a11 * (two * uy + one)) * h;
    const REAL ac2 = ((a02 + Ka02) * (uy2 + uy) + a22) * h;
    const REAL bc0 = ((b00 + Kb00) * (uy2 + uy) + b10 * (two * uy + one) + b20) * h;
    const REAL bc1 = (b01 * (uy2 + uy) + b11 * (two * uy + one)) * h;
    const REAL bc2 = ((b02 + Kb02) * (uy2 + uy) + b22) * h;
    const REAL cc0 = ((c00 + Kc00) * (uy2 + uy) + c10 * (two * uy + one) + c20) * h;
    const REAL cc1 = (c01 * (uy2 + uy) + c11 * (two * uy + one)) * h;
    const REAL cc2 = ((c02 + Kc02) * (uy2 + uy) + c22) * h;

    /* z chains -> populations (zero z component, then minus, then plus) */
#define ZZ(k0, k1, k2, K) ((k0) * (one - uz2) - two * uz * (k1) - (k2) - (K) * uz2)
#define ZM(k0, k1, k2, K) ((((k0) + (K)) * (uz2 - uz) + (k1) * (two * uz - one) + (k2)) * h)
#define ZP(k0, k1, k2, K) ((((k0) + (K)) * (uz2 + uz) + (k1) * (two * uz + one) + (k2)) * h)
    f[MMO] = ZZ(aa0, aa1, aa2, Kaa0); f[MOO] = ZZ(ab0, ab1, ab2, Kab0); f[MPO] = ZZ(ac0, ac1, ac2, Kac0);
    f[OMO] = ZZ(ba0, ba1, ba2, Kba0); f[OOO] = ZZ(bb0, bb1, bb2, Kbb0); f[OPO] = ZZ(bc0, bc1, bc2, Kbc0);
    f[PMO] = ZZ(ca0, ca1, ca2, Kca0); f[POO] = ZZ(cb0, cb1, cb2, Kcb0); f[PPO] = ZZ(cc0, cc1, cc2, Kcc0);
    f[MMM] = ZM(aa0, aa1, aa2, Kaa0); f[MOM] = ZM(ab0, ab1, ab2, Kab0); f[MPM] = ZM(ac0, ac1, ac2, Kac0);
    f[OMM] = ZM(ba0, ba1, ba2, Kba0); f[OOM] = ZM(bb0, bb1, bb2, Kbb0); f[OPM] = ZM(bc0, bc1, bc2, Kbc0);
    f[PMM] = ZM(ca0, ca1, ca2, Kca0); f[POM] = ZM(cb0, cb1, cb2, Kcb0); f[PPM] = ZM(cc0, cc1, cc2, Kcc0);
    f[MMP] = ZP(aa0, aa1, aa2, Kaa0); f[MOP] = ZP(ab0, ab1, ab2, Kab0); f[MPP] = ZP(ac0, ac1, ac2, Kac0);
    f[OMP] = ZP(ba0, ba1, ba2, Kba0); f[OOP] = ZP(bb0, bb1, bb2, Kbb0); f[OPP] = ZP(bc0, bc1, bc2, Kbc0);
    f[PMP] = ZP(ca0, ca1, ca2, Kca0); f[POP] = ZP(cb0, cb1, cb2, Kcb0); f[PPP] = ZP(cc0, cc1, cc2, Kcc0);
#undef ZZ
#undef ZM
#undef ZP
}

/* the nine stencil values of one cell from its pre-collision populations:
   dRho, u and the rate-of-strain estimates from the second moments */
typedef struct { REAL d, x, y, z, xy, yz, xz, my, mz; } NAME(ifv);

static void NAME(if_values)(const REAL *f, REAL w1, NAME(ifv) *v)
{
    REAL m[4];
    NAME(orc_moments)(f, m);
    v->d = m[0]; v->x = m[1]; v->y = m[2]; v->z = m[3];
    const REAL rho = (REAL)1 + m[0];
    v->xy = -(REAL)3 * w1 * ((+f[11] + f[12] - f[15] - f[16] - f[19] - f[20] + f[21] + f[22] - f[23] - f[24] + f[25] + f[26]) / rho - v->x * v->y);
    v->yz = -(REAL)3 * w1 * ((-f[13] - f[14] + f[17] + f[18] - f[19] - f[20] - f[21] - f[22] + f[23] + f[24] + f[25] + f[26]) / rho - v->y * v->z);
    v->xz = -(REAL)3 * w1 * ((-f[7] - f[8] + f[9] + f[10] + f[19] + f[20] - f[21] - f[22] - f[23] - f[24] + f[25] + f[26]) / rho - v->x * v->z);
    v->my = -(REAL)1.5 * w1 * ((+f[1] + f[2] - f[5] - f[6] + f[7] + f[8] + f[9] + f[10] - f[13] - f[14] - f[17] - f[18]) / rho - (v->x * v->x - v->y * v->y));
    v->mz = -(REAL)1.5 * w1 * ((+f[1] + f[2] - f[3] - f[4] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]) / rho - (v->x * v->x - v->z * v->z));
}

static void NAME(if_cell)(const orc_grid *g, int cell, REAL w1, NAME(ifv) *v)
{
    int nb[8];
    NAME(decode)(g, cell, NULL, nb);
    REAL f[27];
    for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, g->Info.esotwistFlipper, q, &cc, &s); f[q] = F_AT(g, s, cc); }
    NAME(if_values)(f, w1, v);
}

static void NAME(if_write)(orc_grid *g, int cell, const REAL *f)
{
    int nb[8];
    NAME(decode)(g, cell, NULL, nb);
    for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, g->Info.esotwistFlipper, q, &cc, &s); F_AT(g, s, cc) = f[q]; }
}

/* the 2x2x2 stencil whose first (lowest i, j, k) cell is `c0`, corner order self, i+, j+, ij+, k+, ik+, jk+, ijk+ */
static void NAME(if_stencil)(const orc_grid *g, int c0, int *st)
{
    int nb[8];
    NAME(decode)(g, c0, NULL, nb);
    st[0] = nb[0]; st[1] = nb[1]; st[2] = nb[2]; st[3] = nb[4];
    st[4] = nb[3]; st[5] = nb[5]; st[6] = nb[6]; st[7] = nb[7];
}

/* interpolation coefficients (Schoenherr 2015, ch. 7; VirtualFluids form):
   out[0..7] d000 d100 d010 d001 d110 d101 d011 d111,
   out[8..18]  a000 a100 a010 a001 a110 a101 a011 a111 a200 a020 a002, the same for b (19..29) and c (30..40) */
static void NAME(if_coefficients)(const NAME(ifv) *s, REAL *out)
{
    REAL D[8], X[8], Y[8], Z[8], XY[8], YZ[8], XZ[8], MY[8], MZ[8];
    for (int n = 0; n < 8; ++n) {
        D[n] = s[n].d; X[n] = s[n].x; Y[n] = s[n].y; Z[n] = s[n].z;
        XY[n] = s[n].xy; YZ[n] = s[n].yz; XZ[n] = s[n].xz; MY[n] = s[n].my; MZ[n] = s[n].mz;
    }
    const REAL two = (REAL)2, eight = (REAL)8;
    REAL *d = out, *a = out + 8, *b = out + 19, *c = out + 30;
    /* :441-448 */
    d[0] = (REAL)0.125 * (((D[7] + D[0]) + (D[3] + D[4])) + ((D[1] + D[6]) + (D[5] + D[2])));
    d[1] = (REAL)0.25 * (((D[7] - D[0]) + (D[3] - D[4])) + ((D[1] - D[6]) + (D[5] - D[2])));
    d[2] = (REAL)0.25 * (((D[7] - D[0]) + (D[3] - D[4])) + ((D[6] - D[1]) + (D[2] - D[5])));
    d[3] = (REAL)0.25 * (((D[7] - D[0]) + (D[4] - D[3])) + ((D[6] - D[1]) + (D[5] - D[2])));
    d[4] = (REAL)0.5 * (((D[7] + D[0]) + (D[3] + D[4])) - ((D[1] + D[6]) + (D[5] + D[2])));
    d[5] = (REAL)0.5 * (((D[7] + D[0]) - (D[3] + D[4])) + ((D[5] + D[2]) - (D[1] + D[6])));
    d[6] = (REAL)0.5 * (((D[7] + D[0]) - (D[3] + D[4])) + ((D[1] + D[6]) - (D[5] + D[2])));
    d[7] = (((D[7] - D[0]) + (D[4] - D[3])) + ((D[1] - D[6]) + (D[2] - D[5])));
    /* :450-492 */
    a[0] = (REAL)0.015625 * (two * (((XY[0] - XY[7]) + (XY[4] - XY[3])) + ((XY[1] - XY[6]) + (XY[5] - XY[2])) +
                                    ((XZ[0] - XZ[7]) + (XZ[3] - XZ[4])) + ((XZ[1] - XZ[6]) + (XZ[2] - XZ[5])) +
                                    ((Y[7] + Y[0]) + (Y[3] + Y[4])) - ((Y[6] + Y[1]) + (Y[2] + Y[5])) +
                                    ((Z[7] + Z[0]) - (Z[3] + Z[4])) + ((Z[5] + Z[2]) - (Z[6] + Z[1]))) +
                             eight * (((X[7] + X[0]) + (X[3] + X[4])) + ((X[6] + X[1]) + (X[5] + X[2]))) +
                             ((MY[0] - MY[7]) + (MY[4] - MY[3])) + ((MY[6] - MY[1]) + (MY[2] - MY[5])) +
                             ((MZ[0] - MZ[7]) + (MZ[4] - MZ[3])) + ((MZ[6] - MZ[1]) + (MZ[2] - MZ[5])));
    b[0] = (REAL)0.015625 * (two * (((MY[7] - MY[0]) + (MY[3] - MY[4])) + ((MY[6] - MY[1]) + (MY[2] - MY[5])) +
                                    ((XY[0] - XY[7]) + (XY[4] - XY[3])) + ((XY[6] - XY[1]) + (XY[2] - XY[5])) +
                                    ((YZ[0] - YZ[7]) + (YZ[3] - YZ[4])) + ((YZ[1] - YZ[6]) + (YZ[2] - YZ[5])) +
                                    ((X[7] + X[0]) + (X[3] + X[4])) - ((X[2] + X[6]) + (X[1] + X[5])) +
                                    ((Z[7] + Z[0]) - (Z[3] + Z[4])) + ((Z[6] + Z[1]) - (Z[2] + Z[5]))) +
                             eight * (((Y[7] + Y[0]) + (Y[3] + Y[4])) + ((Y[6] + Y[1]) + (Y[2] + Y[5]))) +
                             ((MZ[0] - MZ[7]) + (MZ[4] - MZ[3])) + ((MZ[1] - MZ[6]) + (MZ[5] - MZ[2])));
    c[0] = (REAL)0.015625 * (two * (((MZ[7] - MZ[0]) + (MZ[4] - MZ[3])) + ((MZ[6] - MZ[1]) + (MZ[5] - MZ[2])) +
                                    ((XZ[0] - XZ[7]) + (XZ[4] - XZ[3])) + ((XZ[6] - XZ[1]) + (XZ[2] - XZ[5])) +
                                    ((YZ[0] - YZ[7]) + (YZ[4] - YZ[3])) + ((YZ[1] - YZ[6]) + (YZ[5] - YZ[2])) +
                                    ((X[7] + X[0]) - (X[4] + X[3])) + ((X[2] + X[5]) - (X[6] + X[1])) +
                                    ((Y[7] + Y[0]) - (Y[4] + Y[3])) + ((Y[6] + Y[1]) - (Y[2] + Y[5]))) +
                             eight * (((Z[7] + Z[0]) + (Z[3] + Z[4])) + ((Z[1] + Z[6]) + (Z[5] + Z[2]))) +
                             ((MY[0] - MY[7]) + (MY[3] - MY[4])) + ((MY[1] - MY[6]) + (MY[2] - MY[5])));
    /* first derivatives, :494-504 */
#define DX(V) ((REAL)0.25 * ((((V)[7] - (V)[0]) + ((V)[3] - (V)[4])) + (((V)[1] - (V)[6]) + ((V)[5] - (V)[2]))))
#define DY(V) ((REAL)0.25 * ((((V)[7] - (V)[0]) + ((V)[3] - (V)[4])) + (((V)[6] - (V)[1]) + ((V)[2] - (V)[5]))))
#define DZ(V) ((REAL)0.25 * ((((V)[7] - (V)[0]) + ((V)[4] - (V)[3])) + (((V)[6] - (V)[1]) + ((V)[5] - (V)[2]))))
    a[1] = DX(X); b[1] = DX(Y); c[1] = DX(Z);
    a[2] = DY(X); b[2] = DY(Y); c[2] = DY(Z);
    a[3] = DZ(X); b[3] = DZ(Y); c[3] = DZ(Z);
#undef DX
#undef DY
#undef DZ
    /* mixed derivatives, :542-558 */
#define DXY(V) ((REAL)0.5 * ((((V)[7] + (V)[0]) + ((V)[4] + (V)[3])) - (((V)[2] + (V)[5]) + ((V)[1] + (V)[6]))))
#define DXZ(V) ((REAL)0.5 * ((((V)[7] + (V)[0]) - ((V)[4] + (V)[3])) + (((V)[2] + (V)[5]) - ((V)[1] + (V)[6]))))
#define DYZ(V) ((REAL)0.5 * ((((V)[7] + (V)[0]) - ((V)[4] + (V)[3])) + (((V)[1] + (V)[6]) - ((V)[2] + (V)[5]))))
#define DXYZ(V) ((((V)[7] - (V)[0]) + ((V)[4] - (V)[3])) + (((V)[2] - (V)[5]) + ((V)[1] - (V)[6])))
    a[4] = DXY(X); b[4] = DXY(Y); c[4] = DXY(Z);
    a[5] = DXZ(X); b[5] = DXZ(Y); c[5] = DXZ(Z);
    a[6] = DYZ(X); b[6] = DYZ(Y); c[6] = DYZ(Z);
    a[7] = DXYZ(X); b[7] = DXYZ(Y); c[7] = DXYZ(Z);
#undef DXY
#undef DXZ
#undef DYZ
#undef DXYZ
    /* second derivatives, :506-540 */
    a[8] = (REAL)0.0625 * (two * (((Y[7] + Y[0]) + (Y[3] - Y[6])) + ((Y[4] - Y[1]) - (Y[2] + Y[5])) +
                                  ((Z[7] + Z[0]) - (Z[3] + Z[6])) + ((Z[2] + Z[5]) - (Z[4] + Z[1]))) +
                           ((MY[7] - MY[0]) + (MY[3] - MY[4])) + ((MY[1] - MY[6]) + (MY[5] - MY[2])) +
                           ((MZ[7] - MZ[0]) + (MZ[3] - MZ[4])) + ((MZ[1] - MZ[6]) + (MZ[5] - MZ[2])));
    b[8] = (REAL)0.125 * (two * (-((X[7] + X[0]) + (X[3] + X[4])) + ((X[6] + X[1]) + (X[2] + X[5]))) +
                          ((XY[7] - XY[0]) + (XY[3] - XY[4])) + ((XY[1] - XY[6]) + (XY[5] - XY[2])));
    c[8] = (REAL)0.125 * (two * (((X[3] + X[4]) - (X[7] + X[0])) + ((X[6] + X[1]) - (X[2] + X[5]))) +
                          ((XZ[7] - XZ[0]) + (XZ[3] - XZ[4])) + ((XZ[1] - XZ[6]) + (XZ[5] - XZ[2])));
    a[9] = (REAL)0.125 * (two * (-((Y[7] + Y[0]) + (Y[4] + Y[3])) + ((Y[6] + Y[1]) + (Y[2] + Y[5]))) +
                          ((XY[7] - XY[0]) + (XY[3] - XY[4])) + ((XY[6] - XY[1]) + (XY[2] - XY[5])));
    b[9] = (REAL)0.0625 * (two * (((MY[0] - MY[7]) + (MY[4] - MY[3])) + ((MY[1] - MY[6]) + (MY[5] - MY[2])) +
                                  ((X[7] + X[0]) + (X[3] + X[4])) - ((X[6] + X[1]) + (X[5] + X[2])) +
                                  ((Z[7] + Z[0]) - (Z[3] + Z[4])) + ((Z[6] + Z[1]) - (Z[2] + Z[5]))) +
                           ((MZ[7] - MZ[0]) + (MZ[3] - MZ[4])) + ((MZ[6] - MZ[1]) + (MZ[2] - MZ[5])));
    c[9] = (REAL)0.125 * (two * (((Y[4] + Y[3]) - (Y[7] + Y[0])) + ((Y[5] + Y[2]) - (Y[6] + Y[1]))) +
                          ((YZ[7] - YZ[0]) + (YZ[3] - YZ[4])) + ((YZ[6] - YZ[1]) + (YZ[2] - YZ[5])));
    a[10] = (REAL)0.125 * (two * (((Z[3] + Z[4]) - (Z[7] + Z[0])) + ((Z[6] + Z[1]) - (Z[5] + Z[2]))) +
                           ((XZ[7] - XZ[0]) + (XZ[4] - XZ[3])) + ((XZ[5] - XZ[2]) + (XZ[6] - XZ[1])));
    b[10] = (REAL)0.125 * (two * (((Z[3] + Z[4]) - (Z[7] + Z[0])) + ((Z[2] + Z[5]) - (Z[1] + Z[6]))) +
                           ((YZ[7] - YZ[0]) + (YZ[4] - YZ[3])) + ((YZ[5] - YZ[2]) + (YZ[6] - YZ[1])));
    c[10] = (REAL)0.0625 * (two * (((MZ[0] - MZ[7]) + (MZ[3] - MZ[4])) + ((MZ[2] - MZ[5]) + (MZ[1] - MZ[6])) +
                                   ((X[7] + X[0]) - (X[4] + X[3])) + ((X[2] + X[5]) - (X[1] + X[6])) +
                                   ((Y[7] + Y[0]) - (Y[4] + Y[3])) + ((Y[1] + Y[6]) - (Y[2] + Y[5]))) +
                            ((MY[7] - MY[0]) + (MY[4] - MY[3])) + ((MY[5] - MY[2]) + (MY[6] - MY[1])));
}

static inline REAL NAME(omega1)(const orc_grid *g)
{
    const REAL tau = (REAL)3 * (REAL)g->Info.nu + (REAL)0.5;
    return (REAL)1 / tau;
}

/* updateFineToCoarseInterface, updateInterface.h:155-360 */
static void NAME(if_fine_to_coarse)(orc_grid *gc, orc_grid *gf)
{
    const REAL w1c = NAME(omega1)(gc), w1f = NAME(omega1)(gf);
    const orc_interface *IF = &gc->F2C;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < IF->interfaceCount; ++x) {
        int st[8];
        NAME(if_stencil)(gf, IF->childMap[x], st);
        NAME(ifv) s[8];
        for (int n = 0; n < 8; ++n) NAME(if_cell)(gf, st[n], w1f, &s[n]);
        REAL d000 = 0;
        for (int n = 0; n < 8; ++n) d000 += s[n].d;
        d000 *= (REAL)0.125;
        REAL co[41];
        NAME(if_coefficients)(s, co);
        const REAL *a = co + 8, *b = co + 19, *c = co + 30;
        /* :290-344, the coarse cell sits at the stencil centre */
        const REAL dRho = d000, ux = a[0], uy = b[0], uz = c[0];
        const REAL rho = dRho + (REAL)1;
        const REAL sigma = (REAL)2;
        const REAL k011 = -(sigma * rho) / ((REAL)3 * w1c) * ((c[2] + b[3]) + (REAL)0);
        const REAL k101 = -(sigma * rho) / ((REAL)3 * w1c) * ((a[3] + c[1]) + (REAL)0);
        const REAL k110 = -(sigma * rho) / ((REAL)3 * w1c) * ((a[2] + b[1]) + (REAL)0);
        const REAL mxxMyy = -RC(2, 3) * ((a[1] - b[2]) + (REAL)0) * sigma / w1c * rho;
        const REAL mxxMzz = -RC(2, 3) * ((a[1] - c[3]) + (REAL)0) * sigma / w1c * rho;
        const REAL k200 = RC(1, 3) * (mxxMyy + mxxMzz + dRho);
        const REAL k020 = RC(1, 3) * (-(REAL)2 * mxxMyy + mxxMzz + dRho);
        const REAL k002 = RC(1, 3) * (mxxMyy - (REAL)2 * mxxMzz + dRho);
        REAL f[27];
        NAME(if_reconstruct)(f, dRho, ux, uy, uz, k011, k101, k110, k200, k020, k002);
        NAME(if_write)(gc, IF->index[x], f);
    }
}

/* updateCoarseToFineInterface, full blocks: updateInterface.h:362-687 */
static void NAME(if_coarse_to_fine)(orc_grid *gc, orc_grid *gf)
{
    const REAL w1c = NAME(omega1)(gc), w1f = NAME(omega1)(gf);
    const orc_interface *IF = &gc->C2F;
    static const REAL DXS[8] = {-0.25, 0.25, -0.25, 0.25, -0.25, 0.25, -0.25, 0.25};
    static const REAL DYS[8] = {-0.25, -0.25, 0.25, 0.25, -0.25, -0.25, 0.25, 0.25};
    static const REAL DZS[8] = {-0.25, -0.25, -0.25, -0.25, 0.25, 0.25, 0.25, 0.25};
#pragma omp parallel for schedule(static)
    for (int x = 0; x < IF->interfaceCount; ++x) {
        int st[8], tg[8];
        NAME(if_stencil)(gc, IF->index[x], st);
        NAME(ifv) s[8];
        for (int n = 0; n < 8; ++n) NAME(if_cell)(gc, st[n], w1c, &s[n]);
        REAL co[41];
        NAME(if_coefficients)(s, co);
        const REAL *d = co, *a = co + 8, *b = co + 19, *c = co + 30;
        NAME(if_stencil)(gf, IF->childMap[x], tg);
        for (int n = 0; n < 8; ++n) {
            const REAL dx = DXS[n], dy = DYS[n], dz = DZS[n];
            const REAL dRho = d[0] + d[1] * dx + d[2] * dy + d[3] * dz + d[4] * dx * dy + d[5] * dx * dz + d[6] * dy * dz +
                              d[7] * dx * dy * dz;
#define POLY(p) ((p)[0] + (p)[1] * dx + (p)[2] * dy + (p)[3] * dz + (p)[4] * dx * dy + (p)[5] * dx * dz + (p)[6] * dy * dz + \
                 (p)[7] * dx * dy * dz + (p)[8] * dx * dx + (p)[9] * dy * dy + (p)[10] * dz * dz)
            const REAL ux = POLY(a), uy = POLY(b), uz = POLY(c);
#undef POLY
            const REAL sigma = (REAL)0.5;
            const REAL rho = (REAL)1 + dRho;
            const REAL k011 = -RC(1, 3) * (b[3] + c[2] + (REAL)0 + b[5] * dx + c[4] * dx + b[6] * dy + (REAL)2 * c[9] * dy +
                                           b[7] * dx * dy + (REAL)2 * b[10] * dz + c[6] * dz + c[7] * dx * dz) * sigma / w1f * rho;
            const REAL k101 = -RC(1, 3) * (a[3] + c[1] + (REAL)0 + a[5] * dx + (REAL)2 * c[8] * dx + a[6] * dy + c[4] * dy +
                                           a[7] * dx * dy + (REAL)2 * a[10] * dz + c[5] * dz + c[7] * dy * dz) * sigma / w1f * rho;
            const REAL k110 = -RC(1, 3) * (a[2] + b[1] + (REAL)0 + a[4] * dx + (REAL)2 * b[8] * dx + (REAL)2 * a[9] * dy +
                                           b[4] * dy + a[6] * dz + b[5] * dz + a[7] * dx * dz + b[7] * dy * dz) * sigma / w1f * rho;
            const REAL mxxMyy = -RC(2, 3) * (a[1] - b[2] + (REAL)0 + (REAL)2 * a[8] * dx - b[4] * dx + a[4] * dy -
                                             (REAL)2 * b[9] * dy + a[5] * dz - b[6] * dz - b[7] * dx * dz + a[7] * dy * dz) * sigma / w1f * rho;
            const REAL mxxMzz = -RC(2, 3) * (a[1] - c[3] + (REAL)0 + (REAL)2 * a[8] * dx - c[5] * dx + a[4] * dy - c[6] * dy -
                                             c[7] * dx * dy + a[5] * dz - (REAL)2 * c[10] * dz + a[7] * dy * dz) * sigma / w1f * rho;
            const REAL k200 = RC(1, 3) * (mxxMyy + mxxMzz + dRho);
            const REAL k020 = RC(1, 3) * (-(REAL)2 * mxxMyy + mxxMzz + dRho);
            const REAL k002 = RC(1, 3) * (mxxMyy - (REAL)2 * mxxMzz + dRho);
            REAL f[27];
            NAME(if_reconstruct)(f, dRho, ux, uy, uz, k011, k101, k110, k200, k020, k002);
            NAME(if_write)(gf, tg[n], f);
        }
    }
}

/* ---- the linear interface, include/INTERFACE_VERSIONS_BACKUP/updateInterfaceLINEAR.h */

/* fitInterfaceLinear (:155-180): affine fit {c0, cx, cy, cz} of the eight corner values */
static void NAME(if_affine)(const REAL *v, REAL *p)
{
    p[0] = p[1] = p[2] = p[3] = 0;
    for (int i = 0; i < 8; ++i) {
        p[0] += (REAL)0.125 * v[i];
        p[1] += ((i & 1) ? (REAL)0.25 : -(REAL)0.25) * v[i];
        p[2] += ((i & 2) ? (REAL)0.25 : -(REAL)0.25) * v[i];
        p[3] += ((i & 4) ? (REAL)0.25 : -(REAL)0.25) * v[i];
    }
}

static inline REAL NAME(if_affine_at)(const REAL *p, REAL x, REAL y, REAL z) { return p[0] + p[1] * x + p[2] * y + p[3] * z; }

static void NAME(if_reconstruct_m)(REAL *f, REAL dRho, REAL ux, REAL uy, REAL uz, REAL k011, REAL k101, REAL k110,
                                   REAL mxxMyy, REAL mxxMzz)
{
    const REAL k200 = RC(1, 3) * (mxxMyy + mxxMzz + dRho);
    const REAL k020 = RC(1, 3) * (-(REAL)2 * mxxMyy + mxxMzz + dRho);
    const REAL k002 = RC(1, 3) * (mxxMyy - (REAL)2 * mxxMzz + dRho);
    NAME(if_reconstruct)(f, dRho, ux, uy, uz, k011, k101, k110, k200, k020, k002);
}

/* fine to coarse, :265-299 (the rest as updateInterface.h) */
static void NAME(if_fine_to_coarse_linear)(orc_grid *gc, orc_grid *gf)
{
    const REAL w1c = NAME(omega1)(gc), w1f = NAME(omega1)(gf);
    const orc_interface *IF = &gc->F2C;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < IF->interfaceCount; ++x) {
        int st[8];
        NAME(if_stencil)(gf, IF->childMap[x], st);
        NAME(ifv) s[8];
        for (int n = 0; n < 8; ++n) NAME(if_cell)(gf, st[n], w1f, &s[n]);
        REAL dRho = 0, jx = 0, jy = 0, jz = 0, qxy = 0, qyz = 0, qxz = 0, qxxMyy = 0, qxxMzz = 0;
        for (int i = 0; i < 8; ++i) {
            const REAL rhoSource = (REAL)1 + s[i].d;
            const REAL massWeight = (REAL)0.125 * rhoSource;
            dRho += (REAL)0.125 * s[i].d;
            jx += massWeight * s[i].x;
            jy += massWeight * s[i].y;
            jz += massWeight * s[i].z;
            qxy += massWeight * s[i].xy;
            qyz += massWeight * s[i].yz;
            qxz += massWeight * s[i].xz;
            qxxMyy += massWeight * s[i].my;
            qxxMzz += massWeight * s[i].mz;
        }
        const REAL rho = (REAL)1 + dRho;
        const REAL stressScale = (REAL)2 / w1c;
        REAL f[27];
        NAME(if_reconstruct_m)(f, dRho, jx / rho, jy / rho, jz / rho, -RC(1, 3) * stressScale * qyz,
                               -RC(1, 3) * stressScale * qxz, -RC(1, 3) * stressScale * qxy,
                               -RC(2, 3) * stressScale * qxxMyy, -RC(2, 3) * stressScale * qxxMzz);
        NAME(if_write)(gc, IF->index[x], f);
    }
}

/* coarse to fine, full blocks, :405-482 */
static void NAME(if_coarse_to_fine_linear)(orc_grid *gc, orc_grid *gf)
{
    const REAL w1c = NAME(omega1)(gc), w1f = NAME(omega1)(gf);
    const orc_interface *IF = &gc->C2F;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < IF->interfaceCount; ++x) {
        int st[8], tg[8];
        NAME(if_stencil)(gc, IF->index[x], st);
        NAME(ifv) s[8];
        for (int n = 0; n < 8; ++n) NAME(if_cell)(gc, st[n], w1c, &s[n]);
        REAL D[8], UX[8], UY[8], UZ[8], QXY[8], QYZ[8], QXZ[8], QMY[8], QMZ[8];
        REAL jx = 0, jy = 0, jz = 0;
        for (int i = 0; i < 8; ++i) {
            const REAL rhoSource = (REAL)1 + s[i].d;
            const REAL massWeight = (REAL)0.125 * rhoSource;
            D[i] = s[i].d; UX[i] = s[i].x; UY[i] = s[i].y; UZ[i] = s[i].z;
            jx += massWeight * s[i].x;
            jy += massWeight * s[i].y;
            jz += massWeight * s[i].z;
            QXY[i] = rhoSource * s[i].xy; QYZ[i] = rhoSource * s[i].yz; QXZ[i] = rhoSource * s[i].xz;
            QMY[i] = rhoSource * s[i].my; QMZ[i] = rhoSource * s[i].mz;
        }
        REAL d[4], a[4], b[4], c[4], qxy[4], qyz[4], qxz[4], qmy[4], qmz[4];
        NAME(if_affine)(D, d); NAME(if_affine)(UX, a); NAME(if_affine)(UY, b); NAME(if_affine)(UZ, c);
        NAME(if_affine)(QXY, qxy); NAME(if_affine)(QYZ, qyz); NAME(if_affine)(QXZ, qxz);
        NAME(if_affine)(QMY, qmy); NAME(if_affine)(QMZ, qmz);
        const REAL rhoMean = (REAL)1 + d[0];
        a[0] = (jx - (d[1] * a[1] + d[2] * a[2] + d[3] * a[3]) / (REAL)16) / rhoMean;
        b[0] = (jy - (d[1] * b[1] + d[2] * b[2] + d[3] * b[3]) / (REAL)16) / rhoMean;
        c[0] = (jz - (d[1] * c[1] + d[2] * c[2] + d[3] * c[3]) / (REAL)16) / rhoMean;
        NAME(if_stencil)(gf, IF->childMap[x], tg);
        const REAL stressScale = (REAL)0.5 / w1f;
        for (int n = 0; n < 8; ++n) {
            const REAL dx = (n & 1) ? (REAL)0.25 : -(REAL)0.25;
            const REAL dy = (n & 2) ? (REAL)0.25 : -(REAL)0.25;
            const REAL dz = (n & 4) ? (REAL)0.25 : -(REAL)0.25;
            REAL f[27];
            NAME(if_reconstruct_m)(f, NAME(if_affine_at)(d, dx, dy, dz), NAME(if_affine_at)(a, dx, dy, dz),
                                   NAME(if_affine_at)(b, dx, dy, dz), NAME(if_affine_at)(c, dx, dy, dz),
                                   -RC(1, 3) * stressScale * NAME(if_affine_at)(qyz, dx, dy, dz),
                                   -RC(1, 3) * stressScale * NAME(if_affine_at)(qxz, dx, dy, dz),
                                   -RC(1, 3) * stressScale * NAME(if_affine_at)(qxy, dx, dy, dz),
                                   -RC(2, 3) * stressScale * NAME(if_affine_at)(qmy, dx, dy, dz),
                                   -RC(2, 3) * stressScale * NAME(if_affine_at)(qmz, dx, dy, dz));
            NAME(if_write)(gf, tg[n], f);
        }
    }
}

/* leftover fine cells without a full coarse block: centre + three signed face
   neighbours, updateInterface.h:689-884 */
static void NAME(if_leftovers)(orc_sim *S, orc_grid *gc, orc_grid *gf)
{
    const REAL w1c = NAME(omega1)(gc), w1f = NAME(omega1)(gf);
    const orc_interface *IF = &gc->C2F;
    static const int DIRS[3] = {POO, OPO, OOP};
#pragma omp parallel for schedule(static)
    for (int x = 0; x < IF->leftoverCount; ++x) {
        const int cc = IF->leftoverParent[x];
        const int wm = gc->wallMap[cc];
        uint32_t wallData = 0u;
        int wallID = -1;
        if (wm >= 0) { wallData = gc->wallData[wm]; wallID = (int)((wallData >> 27) & 31u); }
        NAME(ifv) ctr;
        NAME(if_cell)(gc, cc, w1c, &ctr);
        int ijk[3], nbx[8];
        NAME(decode)(gc, cc, ijk, nbx);
        NAME(bcr) BC;
        NAME(bc_from)(&BC, &NAME(BC_DEFAULT));
        BC.wallID = wallID;
        BC.dRho = ctr.d; BC.ux = ctr.x; BC.uy = ctr.y; BC.uz = ctr.z;
        NAME(hook_local)(&S->C, &gc->Info, ijk[0], ijk[1], ijk[2], &BC);
        const REAL uxW = BC.ux, uyW = BC.uy, uzW = BC.uz;
        NAME(ifv) r = ctr;
        const int nbr[3] = {IF->leftoverNbrI[x], IF->leftoverNbrJ[x], IF->leftoverNbrK[x]};
        const REAL q = (REAL)0.25;
        for (int a = 0; a < 3; ++a) {
            int cell = nbr[a];
            const int opposite = cell < 0;
            if (opposite) cell = -cell - 1;
            NAME(ifv) v;
            NAME(if_cell)(gc, cell, w1c, &v);
            if (!opposite) {
                r.d += q * (v.d - ctr.d); r.x += q * (v.x - ctr.x); r.y += q * (v.y - ctr.y); r.z += q * (v.z - ctr.z);
                r.xy += q * (v.xy - ctr.xy); r.yz += q * (v.yz - ctr.yz); r.xz += q * (v.xz - ctr.xz);
                r.my += q * (v.my - ctr.my); r.mz += q * (v.mz - ctr.mz);
            } else {
                r.d -= q * (v.d - ctr.d);
                r.xy -= q * (v.xy - ctr.xy); r.yz -= q * (v.yz - ctr.yz); r.xz -= q * (v.xz - ctr.xz);
                r.my -= q * (v.my - ctr.my); r.mz -= q * (v.mz - ctr.mz);
                const int dir = DIRS[a];
                const int wall = (wallData & (1u << dir)) != 0u || (wallData & (1u << ORC_INV[dir])) != 0u;
                if (!wall) {
                    r.x -= q * (v.x - ctr.x); r.y -= q * (v.y - ctr.y); r.z -= q * (v.z - ctr.z);
                } else {
                    /* quadratic through the wall, the centre and the opposite cell */
                    r.x += RC(1, 24) * ((REAL)10 * uxW - (REAL)9 * ctr.x - v.x);
                    r.y += RC(1, 24) * ((REAL)10 * uyW - (REAL)9 * ctr.y - v.y);
                    r.z += RC(1, 24) * ((REAL)10 * uzW - (REAL)9 * ctr.z - v.z);
                }
            }
        }
        const REAL rho = r.d + (REAL)1;
        const REAL sigma = (REAL)0.5;
        const REAL k011 = -RC(1, 3) * (r.yz) * sigma / w1f * rho;
        const REAL k101 = -RC(1, 3) * (r.xz) * sigma / w1f * rho;
        const REAL k110 = -RC(1, 3) * (r.xy) * sigma / w1f * rho;
        const REAL mxxMyy = -RC(2, 3) * (r.my) * sigma / w1f * rho;
        const REAL mxxMzz = -RC(2, 3) * (r.mz) * sigma / w1f * rho;
        const REAL k200 = RC(1, 3) * (mxxMyy + mxxMzz + r.d);
        const REAL k020 = RC(1, 3) * (-(REAL)2 * mxxMyy + mxxMzz + r.d);
        const REAL k002 = RC(1, 3) * (mxxMyy - (REAL)2 * mxxMzz + r.d);
        REAL f[27];
        NAME(if_reconstruct)(f, r.d, r.x, r.y, r.z, k011, k101, k110, k200, k020, k002);
        NAME(if_write)(gf, IF->leftoverIndex[x], f);
    }
}

/* updateInterface, updateInterface.h:886-890 */
void NAME(orc_update_interface)(orc_sim *S, int coarse)
{
    orc_grid *gc = &S->g[coarse], *gf = &S->g[coarse + 1];
    if (S->C.interfaceScheme == LBM3_INTERFACE_LINEAR) {
        NAME(if_fine_to_coarse_linear)(gc, gf);
        NAME(if_coarse_to_fine_linear)(gc, gf);
    } else {
        NAME(if_fine_to_coarse)(gc, gf);
        NAME(if_coarse_to_fine)(gc, gf);
    }
    NAME(if_leftovers)(S, gc, gf);
}

/* test access: stencil values, coefficients and reconstruction of caller-given data */
void NAME(orc_if_values_batch)(const REAL *f, int n, REAL w1, REAL *out9)
{
    for (int c = 0; c < n; ++c) {
        NAME(ifv) v;
        NAME(if_values)(f + 27 * (size_t)c, w1, &v);
        REAL *o = out9 + 9 * (size_t)c;
        o[0] = v.d; o[1] = v.x; o[2] = v.y; o[3] = v.z; o[4] = v.xy; o[5] = v.yz; o[6] = v.xz; o[7] = v.my; o[8] = v.mz;
    }
}

void NAME(orc_if_coefficients)(const REAL *in9x8, REAL *out41)
{
    NAME(ifv) s[8];
    for (int n = 0; n < 8; ++n) {
        const REAL *i = in9x8 + 9 * n;
        s[n].d = i[0]; s[n].x = i[1]; s[n].y = i[2]; s[n].z = i[3]; s[n].xy = i[4]; s[n].yz = i[5]; s[n].xz = i[6];
        s[n].my = i[7]; s[n].mz = i[8];
    }
    NAME(if_coefficients)(s, out41);
}

void NAME(orc_if_reconstruct)(const REAL *in10, REAL *f)
{
    NAME(if_reconstruct)(f, in10[0], in10[1], in10[2], in10[3], in10[4], in10[5], in10[6], in10[7], in10[8], in10[9]);
}

#undef RC
#undef F_AT
