/*
 * ORACLE (test infrastructure only).  CPU restatement of the LBM3 per-cell
 * numerics, instantiated twice by numerics.c: REAL=float (the reference's FP32
 * path) and REAL=double (the FP64 build of the same arithmetic).
 *
 * Arithmetic order follows the reference line by line; build with
 * -ffp-contract=off so that no multiply-add is fused.
 */

#define RC(a, b) ((REAL)(a) / (REAL)(b))

/* A BCStruct whose floating fields are REAL (types.h:131-136). */
typedef struct {
    REAL dRho, ux, uy, uz, gx, gy, gz;
    int wallID, openBCID, useIBB, dirichletU, dirichletRho, periodic, nonReflective;
    REAL nonReflectiveBeta, collisionLimiter, nuMultiplier;
} NAME(bcr);

static inline void NAME(bc_from)(NAME(bcr) *o, const lbm3_bc *b)
{
    o->dRho = b->dRho; o->ux = b->ux; o->uy = b->uy; o->uz = b->uz;
    o->gx = b->gx; o->gy = b->gy; o->gz = b->gz;
    o->wallID = b->wallID; o->openBCID = b->openBCID; o->useIBB = b->useIBB;
    o->dirichletU = b->dirichletU; o->dirichletRho = b->dirichletRho;
    o->periodic = b->periodic; o->nonReflective = b->nonReflective;
    o->nonReflectiveBeta = b->nonReflectiveBeta;
    o->collisionLimiter = b->collisionLimiter; o->nuMultiplier = b->nuMultiplier;
}

static inline REAL NAME(wgt)(int q)
{
    /* DIRECTION_WEIGHTS, D3Q27Directions.h:40-42 */
    if (q == 0) return RC(8, 27);
    if (q < 7) return RC(2, 27);
    if (q < 19) return RC(1, 54);
    return RC(1, 216);
}

/* getFeq, cellFunctions.h:125-148 */
void NAME(orc_feq)(REAL dRho, REAL ux, REAL uy, REAL uz, REAL *feq)
{
    const REAL ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    for (int q = 0; q < 27; ++q) {
        const REAL cx = (REAL)ORC_CX[q], cy = (REAL)ORC_CY[q], cz = (REAL)ORC_CZ[q];
        const REAL dx = (REAL)3 * cx * ux + ((REAL)4.5 * cx * cx - (REAL)1.5) * ux2;
        const REAL dy = (REAL)3 * cy * uy + ((REAL)4.5 * cy * cy - (REAL)1.5) * uy2;
        const REAL dz = (REAL)3 * cz * uz + ((REAL)4.5 * cz * cz - (REAL)1.5) * uz2;
        const REAL xy = dx + dy + dx * dy;
        const REAL correction = xy + dz + xy * dz;
        feq[q] = NAME(wgt)(q) * ((dRho + correction) + dRho * correction);
    }
}

/* getFeqSingle, cellFunctions.h:150-168 */
static inline REAL NAME(feq_single)(REAL dRho, REAL ux, REAL uy, REAL uz, int q)
{
    const REAL ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const REAL cx = (REAL)ORC_CX[q], cy = (REAL)ORC_CY[q], cz = (REAL)ORC_CZ[q];
    const REAL dx = (REAL)3 * cx * ux + ((REAL)4.5 * cx * cx - (REAL)1.5) * ux2;
    const REAL dy = (REAL)3 * cy * uy + ((REAL)4.5 * cy * cy - (REAL)1.5) * uy2;
    const REAL dz = (REAL)3 * cz * uz + ((REAL)4.5 * cz * cz - (REAL)1.5) * uz2;
    const REAL xy = dx + dy + dx * dy;
    const REAL correction = xy + dz + xy * dz;
    return NAME(wgt)(q) * ((dRho + correction) + dRho * correction);
}

/* getDRhoUxUyUz, cellFunctions.h:257-279 */
void NAME(orc_moments)(const REAL *f, REAL *out4)
{
    const REAL dRho = (((f[PPP] + f[MMM]) + (f[PMP] + f[MPM])) + ((f[PPM] + f[MMP]) + (f[PMM] + f[MPP])))
        + (((f[OPP] + f[OMM]) + (f[OPM] + f[OMP])) + ((f[POP] + f[MOM]) + (f[POM] + f[MOP])) + ((f[PPO] + f[MMO]) + (f[PMO] + f[MPO])))
        + ((f[POO] + f[MOO]) + (f[OPO] + f[OMO]) + (f[OOP] + f[OOM])) + f[OOO];
    const REAL rhoInv = (REAL)1 / (dRho + (REAL)1);
    const REAL mx = ((((f[PPP] - f[MMM]) + (f[PMP] - f[MPM])) + ((f[PPM] - f[MMP]) + (f[PMM] - f[MPP])))
        + (((f[POP] - f[MOM]) + (f[POM] - f[MOP])) + ((f[PPO] - f[MMO]) + (f[PMO] - f[MPO])))
        + (f[POO] - f[MOO]));
    const REAL my = ((((f[PPP] - f[MMM]) - (f[PMP] - f[MPM])) + ((f[PPM] - f[MMP]) - (f[PMM] - f[MPP])))
        + (((f[OPP] - f[OMM]) + (f[OPM] - f[OMP])) + ((f[PPO] - f[MMO]) - (f[PMO] - f[MPO])))
        + (f[OPO] - f[OMO]));
    const REAL mz = ((((f[PPP] - f[MMM]) + (f[PMP] - f[MPM])) - ((f[PPM] - f[MMP]) + (f[PMM] - f[MPP])))
        + (((f[OPP] - f[OMM]) - (f[OPM] - f[OMP])) + ((f[POP] - f[MOM]) - (f[POM] - f[MOP])))
        + (f[OOP] - f[OOM]));
    out4[0] = dRho;
    out4[1] = mx * rhoInv;
    out4[2] = my * rhoInv;
    out4[3] = mz * rhoInv;
}

/* applyCollision, applyCollision.h:17-491 (well-conditioned D3Q27 cumulant, K17 / K15 AllOne) */
void NAME(orc_collide)(REAL *f, const NAME(bcr) *BC, REAL nu)
{
    const REAL gx = BC->gx, gy = BC->gy, gz = BC->gz;
    const REAL K_aa0 = RC(1, 36), K_ab0 = RC(1, 9), K_ac0 = RC(1, 36);
    const REAL K_ba0 = RC(1, 9), K_bb0 = RC(4, 9), K_bc0 = RC(1, 9);
    const REAL K_ca0 = RC(1, 36), K_cb0 = RC(1, 9), K_cc0 = RC(1, 36);
    const REAL K_a00 = RC(1, 6), K_b00 = RC(2, 3), K_c00 = RC(1, 6);
    const REAL K_a02 = RC(1, 18), K_b02 = RC(2, 9), K_c02 = RC(1, 18);
    const REAL one = 1, two = 2, half = RC(1, 2), third = RC(1, 3);

    const REAL k_aa0 = (f[MMP] + f[MMM]) + f[MMO];
    const REAL k_ab0 = (f[MOP] + f[MOM]) + f[MOO];
    const REAL k_ac0 = (f[MPP] + f[MPM]) + f[MPO];
    const REAL k_ba0 = (f[OMP] + f[OMM]) + f[OMO];
    const REAL k_bb0 = (f[OOP] + f[OOM]) + f[OOO];
    const REAL k_bc0 = (f[OPP] + f[OPM]) + f[OPO];
    const REAL k_ca0 = (f[PMP] + f[PMM]) + f[PMO];
    const REAL k_cb0 = (f[POP] + f[POM]) + f[POO];
    const REAL k_cc0 = (f[PPP] + f[PPM]) + f[PPO];

    const REAL dRho = ((k_aa0 + k_ab0) + k_ac0) + ((k_ba0 + k_bb0) + k_bc0) + ((k_ca0 + k_cb0) + k_cc0);
    const REAL rho = one + dRho;
    const REAL rhoInv = one / rho;

    const REAL momentumX = ((k_ca0 + k_cb0) + k_cc0) - ((k_aa0 + k_ab0) + k_ac0);
    const REAL momentumY = ((k_ac0 + k_bc0) + k_cc0) - ((k_aa0 + k_ba0) + k_ca0);
    const REAL momentumZ =
        (((f[MMP] - f[MMM]) + (f[MOP] - f[MOM])) + ((f[MPP] - f[MPM]) + (f[OMP] - f[OMM]))) +
        (((f[OOP] - f[OOM]) + (f[OPP] - f[OPM])) + ((f[PMP] - f[PMM]) + (f[POP] - f[POM]))) +
        (f[PPP] - f[PPM]);

    const REAL ux = (momentumX + half * gx) * rhoInv;
    const REAL uy = (momentumY + half * gy) * rhoInv;
    const REAL uz = (momentumZ + half * gz) * rhoInv;
    const REAL ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;

    const REAL k_aa1 = (f[MMP] - f[MMM]) - uz * (k_aa0 + K_aa0);
    const REAL k_ab1 = (f[MOP] - f[MOM]) - uz * (k_ab0 + K_ab0);
    const REAL k_ac1 = (f[MPP] - f[MPM]) - uz * (k_ac0 + K_ac0);
    const REAL k_ba1 = (f[OMP] - f[OMM]) - uz * (k_ba0 + K_ba0);
    const REAL k_bb1 = (f[OOP] - f[OOM]) - uz * (k_bb0 + K_bb0);
    const REAL k_bc1 = (f[OPP] - f[OPM]) - uz * (k_bc0 + K_bc0);
    const REAL k_ca1 = (f[PMP] - f[PMM]) - uz * (k_ca0 + K_ca0);
    const REAL k_cb1 = (f[POP] - f[POM]) - uz * (k_cb0 + K_cb0);
    const REAL k_cc1 = (f[PPP] - f[PPM]) - uz * (k_cc0 + K_cc0);

    const REAL k_aa2 = (f[MMP] + f[MMM]) - two * uz * (f[MMP] - f[MMM]) + uz2 * (k_aa0 + K_aa0);
    const REAL k_ab2 = (f[MOP] + f[MOM]) - two * uz * (f[MOP] - f[MOM]) + uz2 * (k_ab0 + K_ab0);
    const REAL k_ac2 = (f[MPP] + f[MPM]) - two * uz * (f[MPP] - f[MPM]) + uz2 * (k_ac0 + K_ac0);
    const REAL k_ba2 = (f[OMP] + f[OMM]) - two * uz * (f[OMP] - f[OMM]) + uz2 * (k_ba0 + K_ba0);
    const REAL k_bb2 = (f[OOP] + f[OOM]) - two * uz * (f[OOP] - f[OOM]) + uz2 * (k_bb0 + K_bb0);
    const REAL k_bc2 = (f[OPP] + f[OPM]) - two * uz * (f[OPP] - f[OPM]) + uz2 * (k_bc0 + K_bc0);
    const REAL k_ca2 = (f[PMP] + f[PMM]) - two * uz * (f[PMP] - f[PMM]) + uz2 * (k_ca0 + K_ca0);
    const REAL k_cb2 = (f[POP] + f[POM]) - two * uz * (f[POP] - f[POM]) + uz2 * (k_cb0 + K_cb0);
    const REAL k_cc2 = (f[PPP] + f[PPM]) - two * uz * (f[PPP] - f[PPM]) + uz2 * (k_cc0 + K_cc0);

    const REAL k_a00 = (k_ac0 + k_aa0) + k_ab0;
    const REAL k_b00 = (k_bc0 + k_ba0) + k_bb0;
    const REAL k_c00 = (k_cc0 + k_ca0) + k_cb0;
    const REAL k_a01 = (k_ac1 + k_aa1) + k_ab1;
    const REAL k_b01 = (k_bc1 + k_ba1) + k_bb1;
    const REAL k_c01 = (k_cc1 + k_ca1) + k_cb1;
    const REAL k_a02 = (k_ac2 + k_aa2) + k_ab2;
    const REAL k_b02 = (k_bc2 + k_ba2) + k_bb2;
    const REAL k_c02 = (k_cc2 + k_ca2) + k_cb2;

    const REAL k_a10 = (k_ac0 - k_aa0) - uy * (k_a00 + K_a00);
    const REAL k_b10 = (k_bc0 - k_ba0) - uy * (k_b00 + K_b00);
    const REAL k_c10 = (k_cc0 - k_ca0) - uy * (k_c00 + K_c00);
    const REAL k_a11 = (k_ac1 - k_aa1) - uy * k_a01;
    const REAL k_b11 = (k_bc1 - k_ba1) - uy * k_b01;
    const REAL k_c11 = (k_cc1 - k_ca1) - uy * k_c01;
    const REAL k_a12 = (k_ac2 - k_aa2) - uy * (k_a02 + K_a02);
    const REAL k_b12 = (k_bc2 - k_ba2) - uy * (k_b02 + K_b02);
    const REAL k_c12 = (k_cc2 - k_ca2) - uy * (k_c02 + K_c02);

    const REAL k_a20 = (k_ac0 + k_aa0) - two * uy * (k_ac0 - k_aa0) + uy2 * (k_a00 + K_a00);
    const REAL k_b20 = (k_bc0 + k_ba0) - two * uy * (k_bc0 - k_ba0) + uy2 * (k_b00 + K_b00);
    const REAL k_c20 = (k_cc0 + k_ca0) - two * uy * (k_cc0 - k_ca0) + uy2 * (k_c00 + K_c00);
    const REAL k_a21 = (k_ac1 + k_aa1) - two * uy * (k_ac1 - k_aa1) + uy2 * k_a01;
    const REAL k_b21 = (k_bc1 + k_ba1) - two * uy * (k_bc1 - k_ba1) + uy2 * k_b01;
    const REAL k_c21 = (k_cc1 + k_ca1) - two * uy * (k_cc1 - k_ca1) + uy2 * k_c01;
    const REAL k_a22 = (k_ac2 + k_aa2) - two * uy * (k_ac2 - k_aa2) + uy2 * (k_a02 + K_a02);
    const REAL k_b22 = (k_bc2 + k_ba2) - two * uy * (k_bc2 - k_ba2) + uy2 * (k_b02 + K_b02);
    const REAL k_c22 = (k_cc2 + k_ca2) - two * uy * (k_cc2 - k_ca2) + uy2 * (k_c02 + K_c02);

    const REAL k_000 = dRho;
    const REAL k_001 = (k_c01 + k_a01) + k_b01;
    const REAL k_002 = (k_c02 + k_a02) + k_b02;
    const REAL k_010 = (k_c10 + k_a10) + k_b10;
    const REAL k_011 = (k_c11 + k_a11) + k_b11;
    const REAL k_012 = (k_c12 + k_a12) + k_b12;
    const REAL k_020 = (k_c20 + k_a20) + k_b20;
    const REAL k_021 = (k_c21 + k_a21) + k_b21;
    const REAL k_022 = (k_c22 + k_a22) + k_b22;

    const REAL k_100 = (k_c00 - k_a00) - ux * (k_000 + one);
    const REAL k_101 = (k_c01 - k_a01) - ux * k_001;
    const REAL k_102 = (k_c02 - k_a02) - ux * (k_002 + third);
    const REAL k_110 = (k_c10 - k_a10) - ux * k_010;
    const REAL k_111 = (k_c11 - k_a11) - ux * k_011;
    const REAL k_112 = (k_c12 - k_a12) - ux * k_012;
    const REAL k_120 = (k_c20 - k_a20) - ux * (k_020 + third);
    const REAL k_121 = (k_c21 - k_a21) - ux * k_021;

    const REAL k_200 = (k_c00 + k_a00) - two * ux * (k_c00 - k_a00) + ux2 * (k_000 + one);
    const REAL k_201 = (k_c01 + k_a01) - two * ux * (k_c01 - k_a01) + ux2 * k_001;
    const REAL k_202 = (k_c02 + k_a02) - two * ux * (k_c02 - k_a02) + ux2 * (k_002 + third);
    const REAL k_210 = (k_c10 + k_a10) - two * ux * (k_c10 - k_a10) + ux2 * k_010;
    const REAL k_211 = (k_c11 + k_a11) - two * ux * (k_c11 - k_a11) + ux2 * k_011;
    const REAL k_220 = (k_c20 + k_a20) - two * ux * (k_c20 - k_a20) + ux2 * (k_020 + third);

    const REAL C_110 = k_110, C_101 = k_101, C_011 = k_011;
    const REAL C_200 = k_200, C_020 = k_020, C_002 = k_002;
    const REAL C_111 = k_111, C_120 = k_120, C_102 = k_102, C_210 = k_210;
    const REAL C_012 = k_012, C_201 = k_201, C_021 = k_021;

    const REAL C_211 = k_211 - ((k_200 + third) * k_011 + two * k_101 * k_110) * rhoInv;
    const REAL C_121 = k_121 - ((k_020 + third) * k_101 + two * k_110 * k_011) * rhoInv;
    const REAL C_112 = k_112 - ((k_002 + third) * k_110 + two * k_011 * k_101) * rhoInv;

    const REAL C_220 = k_220 - ((k_020 * k_200 + two * k_110 * k_110 + (k_200 + k_020) / (REAL)3) * rhoInv - dRho * rhoInv / (REAL)9);
    const REAL C_022 = k_022 - ((k_002 * k_020 + two * k_011 * k_011 + (k_002 + k_020) / (REAL)3) * rhoInv - dRho * rhoInv / (REAL)9);
    const REAL C_202 = k_202 - ((k_200 * k_002 + two * k_101 * k_101 + (k_200 + k_002) / (REAL)3) * rhoInv - dRho * rhoInv / (REAL)9);

    const REAL omega1 = one / ((REAL)3 * (nu * BC->nuMultiplier) + half);
    const REAL omega2 = one;

    const REAL Dxu = -omega1 * half * rhoInv * (two * C_200 - C_020 - C_002)
                   - omega2 * half * rhoInv * (C_200 + C_020 + C_002 - k_000);
    const REAL Dyv = Dxu + (REAL)1.5 * omega1 * rhoInv * (C_200 - C_020);
    const REAL Dzw = Dxu + (REAL)1.5 * omega1 * rhoInv * (C_200 - C_002);
    const REAL DxvDyu = -(REAL)3 * omega1 * rhoInv * C_110;
    const REAL DxwDzu = -(REAL)3 * omega1 * rhoInv * C_101;
    const REAL DywDzv = -(REAL)3 * omega1 * rhoInv * C_011;

    const REAL Cs_110 = (one - omega1) * C_110;
    const REAL Cs_101 = (one - omega1) * C_101;
    const REAL Cs_011 = (one - omega1) * C_011;

    const REAL Eq33RHS = (one - omega1) * (C_200 - C_020) - (REAL)3 * rho * (one - half * omega1) * (ux2 * Dxu - uy2 * Dyv);
    const REAL Eq34RHS = (one - omega1) * (C_200 - C_002) - (REAL)3 * rho * (one - half * omega1) * (ux2 * Dxu - uz2 * Dzw);
    const REAL Eq35RHS = k_000 * omega2 + (one - omega2) * (C_200 + C_020 + C_002)
                       - (REAL)3 * rho * (one - half * omega2) * (ux2 * Dxu + uy2 * Dyv + uz2 * Dzw);

    const REAL Cs_200 = (Eq33RHS + Eq34RHS + Eq35RHS) / (REAL)3;
    const REAL Cs_020 = (Eq34RHS - two * Eq33RHS + Eq35RHS) / (REAL)3;
    const REAL Cs_002 = (Eq33RHS - two * Eq34RHS + Eq35RHS) / (REAL)3;

    REAL Cs_120, Cs_102, Cs_210, Cs_012, Cs_201, Cs_021, Cs_111;
    REAL Cs_220, Cs_202, Cs_022, Cs_211, Cs_121, Cs_112;
    const REAL omega6 = one, omega7 = one, omega8 = one;
    const int useK17 = (BC->collisionLimiter > (REAL)0) && (BC->nuMultiplier <= one);

    if (useK17) {
        const REAL omega3 = (REAL)8 * (omega1 - two) * (omega2 * ((REAL)3 * omega1 - one) - (REAL)5 * omega1) /
            ((REAL)8 * ((REAL)5 - two * omega1) * omega1 + omega2 * ((REAL)8 + omega1 * ((REAL)9 * omega1 - (REAL)26)));
        const REAL omega4 = (REAL)8 * (omega1 - two) * (omega1 + omega2 * ((REAL)3 * omega1 - (REAL)7)) /
            (omega2 * ((REAL)56 - (REAL)42 * omega1 + (REAL)9 * omega1 * omega1) - (REAL)8 * omega1);
        const REAL omega5 = (REAL)24 * (omega1 - two) *
            ((REAL)4 * omega1 * omega1 + omega1 * omega2 * ((REAL)18 - (REAL)13 * omega1) +
             omega2 * omega2 * (two + omega1 * ((REAL)6 * omega1 - (REAL)11))) /
            ((REAL)16 * omega1 * omega1 * (omega1 - (REAL)6) -
             two * omega1 * omega2 * ((REAL)216 + (REAL)5 * omega1 * ((REAL)9 * omega1 - (REAL)46)) +
             omega2 * omega2 * (omega1 * ((REAL)3 * omega1 - (REAL)10) * ((REAL)15 * omega1 - (REAL)28) - (REAL)48));
        const REAL A = ((REAL)4 * omega1 * omega1 + two * omega1 * omega2 * (omega1 - (REAL)6) +
                        omega2 * omega2 * (omega1 * ((REAL)10 - (REAL)3 * omega1) - (REAL)4)) /
                       ((omega1 - omega2) * (omega2 * (two + (REAL)3 * omega1) - (REAL)8 * omega1));
        const REAL B = ((REAL)4 * omega1 * omega2 * ((REAL)9 * omega1 - (REAL)16) - (REAL)4 * omega1 * omega1 -
                        two * omega2 * omega2 * (two + (REAL)9 * omega1 * (omega1 - two))) /
                       ((REAL)3 * (omega1 - omega2) * (omega2 * (two + (REAL)3 * omega1) - (REAL)8 * omega1));
        const REAL lambda = BC->collisionLimiter;

        const REAL C_120p102 = C_120 + C_102, C_210p012 = C_210 + C_012, C_201p021 = C_201 + C_021;
        const REAL C_120m102 = C_120 - C_102, C_210m012 = C_210 - C_012, C_201m021 = C_201 - C_021;
        const REAL a1 = ORC_ABS(C_120p102), a2 = ORC_ABS(C_210p012), a3 = ORC_ABS(C_201p021);
        const REAL a4 = ORC_ABS(C_120m102), a5 = ORC_ABS(C_210m012), a6 = ORC_ABS(C_201m021);
        const REAL a7 = ORC_ABS(C_111);

        const REAL w1 = omega3 + (one - omega3) * a1 / (lambda + a1);
        const REAL w2 = omega3 + (one - omega3) * a2 / (lambda + a2);
        const REAL w3 = omega3 + (one - omega3) * a3 / (lambda + a3);
        const REAL w4 = omega4 + (one - omega4) * a4 / (lambda + a4);
        const REAL w5 = omega4 + (one - omega4) * a5 / (lambda + a5);
        const REAL w6 = omega4 + (one - omega4) * a6 / (lambda + a6);
        const REAL w7 = omega5 + (one - omega5) * a7 / (lambda + a7);

        const REAL Eq117 = (one - w1) * C_120p102;
        const REAL Eq118 = (one - w2) * C_210p012;
        const REAL Eq119 = (one - w3) * C_201p021;
        const REAL Eq120 = (one - w4) * C_120m102;
        const REAL Eq121 = (one - w5) * C_210m012;
        const REAL Eq122 = (one - w6) * C_201m021;

        Cs_120 = half * (Eq120 + Eq117);
        Cs_102 = half * (-Eq120 + Eq117);
        Cs_210 = half * (Eq121 + Eq118);
        Cs_012 = half * (-Eq121 + Eq118);
        Cs_201 = half * (Eq122 + Eq119);
        Cs_021 = half * (-Eq122 + Eq119);
        Cs_111 = (one - w7) * C_111;

        const REAL Eq43RHS = (two / (REAL)3) * (one / omega1 - half) * omega6 * A * rho * (Dxu - two * Dyv + Dzw)
                           + (one - omega6) * (C_220 - two * C_202 + C_022);
        const REAL Eq44RHS = (two / (REAL)3) * (one / omega1 - half) * omega6 * A * rho * (-two * Dxu + Dyv + Dzw)
                           + (one - omega6) * (C_220 + C_202 - two * C_022);
        const REAL Eq45RHS = -((REAL)4 / (REAL)3) * (one / omega1 - half) * omega7 * A * rho * (Dxu + Dyv + Dzw)
                           + (one - omega7) * (C_220 + C_202 + C_022);

        Cs_220 = (Eq43RHS + Eq44RHS + Eq45RHS) / (REAL)3;
        Cs_202 = (-Eq43RHS + Eq45RHS) / (REAL)3;
        Cs_022 = (-Eq44RHS + Eq45RHS) / (REAL)3;
        Cs_211 = -(one / (REAL)3) * (one / omega1 - half) * omega8 * B * rho * DywDzv + (one - omega8) * C_211;
        Cs_121 = -(one / (REAL)3) * (one / omega1 - half) * omega8 * B * rho * DxwDzu + (one - omega8) * C_121;
        Cs_112 = -(one / (REAL)3) * (one / omega1 - half) * omega8 * B * rho * DxvDyu + (one - omega8) * C_112;
    } else {
        Cs_120 = Cs_102 = Cs_210 = Cs_012 = Cs_201 = Cs_021 = Cs_111 = 0;
        Cs_220 = Cs_202 = Cs_022 = Cs_211 = Cs_121 = Cs_112 = 0;
    }

    const REAL ks_011 = Cs_011, ks_101 = Cs_101, ks_110 = Cs_110, ks_111 = Cs_111;
    const REAL ks_002 = Cs_002, ks_020 = Cs_020, ks_200 = Cs_200;
    const REAL ks_012 = Cs_012, ks_021 = Cs_021, ks_102 = Cs_102, ks_201 = Cs_201, ks_120 = Cs_120, ks_210 = Cs_210;

    const REAL ks_211 = Cs_211 + ((ks_200 + third) * ks_011 + two * ks_101 * ks_110) * rhoInv;
    const REAL ks_121 = Cs_121 + ((ks_020 + third) * ks_101 + two * ks_110 * ks_011) * rhoInv;
    const REAL ks_112 = Cs_112 + ((ks_002 + third) * ks_110 + two * ks_011 * ks_101) * rhoInv;

    const REAL ks_220 = Cs_220 + ((ks_020 * ks_200 + two * ks_110 * ks_110 + (ks_200 + ks_020) / (REAL)3) * rhoInv - dRho * rhoInv / (REAL)9);
    const REAL ks_022 = Cs_022 + ((ks_002 * ks_020 + two * ks_011 * ks_011 + (ks_002 + ks_020) / (REAL)3) * rhoInv - dRho * rhoInv / (REAL)9);
    const REAL ks_202 = Cs_202 + ((ks_200 * ks_002 + two * ks_101 * ks_101 + (ks_200 + ks_002) / (REAL)3) * rhoInv - dRho * rhoInv / (REAL)9);

    const REAL ks_122 = (ks_020 * ks_102 + ks_002 * ks_120 + (REAL)4 * ks_011 * ks_111 +
                         two * (ks_110 * ks_012 + ks_101 * ks_021) + (ks_120 + ks_102) / (REAL)3) * rhoInv;
    const REAL ks_212 = (ks_002 * ks_210 + ks_200 * ks_012 + (REAL)4 * ks_101 * ks_111 +
                         two * (ks_011 * ks_201 + ks_110 * ks_102) + (ks_210 + ks_012) / (REAL)3) * rhoInv;
    const REAL ks_221 = (ks_200 * ks_021 + ks_020 * ks_201 + (REAL)4 * ks_110 * ks_111 +
                         two * (ks_101 * ks_120 + ks_011 * ks_210) + (ks_021 + ks_201) / (REAL)3) * rhoInv;

    const REAL ks_222 = ((REAL)4 * ks_111 * ks_111 + ks_200 * ks_022 + ks_020 * ks_202 + ks_002 * ks_220 +
                         (REAL)4 * (ks_011 * ks_211 + ks_101 * ks_121 + ks_110 * ks_112) +
                         two * (ks_120 * ks_102 + ks_210 * ks_012 + ks_201 * ks_021)) * rhoInv
        - ((REAL)16 * ks_110 * ks_101 * ks_011 +
           (REAL)4 * (ks_101 * ks_101 * ks_020 + ks_011 * ks_011 * ks_200 + ks_110 * ks_110 * ks_002) +
           two * ks_200 * ks_020 * ks_002) * rhoInv * rhoInv
        + ((REAL)3 * (ks_022 + ks_202 + ks_220) + (ks_200 + ks_020 + ks_002)) * rhoInv / (REAL)9
        - (two / (REAL)3) *
          (two * (ks_101 * ks_101 + ks_011 * ks_011 + ks_110 * ks_110) +
           (ks_002 * ks_020 + ks_002 * ks_200 + ks_020 * ks_200) +
           (ks_002 + ks_020 + ks_200) / (REAL)3) * rhoInv * rhoInv
        - (dRho * dRho - dRho) * rhoInv * rhoInv / (REAL)27;

    const REAL ks_000 = k_000, ks_100 = -k_100, ks_010 = -k_010, ks_001 = -k_001;

    const REAL ks_b00 = ks_000 * (one - ux2) - two * ux * ks_100 - ks_200 - ux2;
    const REAL ks_b01 = ks_001 * (one - ux2) - two * ux * ks_101 - ks_201;
    const REAL ks_b02 = ks_002 * (one - ux2) - two * ux * ks_102 - ks_202 - ux2 / (REAL)3;
    const REAL ks_b10 = ks_010 * (one - ux2) - two * ux * ks_110 - ks_210;
    const REAL ks_b11 = ks_011 * (one - ux2) - two * ux * ks_111 - ks_211;
    const REAL ks_b12 = ks_012 * (one - ux2) - two * ux * ks_112 - ks_212;
    const REAL ks_b20 = ks_020 * (one - ux2) - two * ux * ks_120 - ks_220 - ux2 / (REAL)3;
    const REAL ks_b21 = ks_021 * (one - ux2) - two * ux * ks_121 - ks_221;
    const REAL ks_b22 = ks_022 * (one - ux2) - two * ux * ks_122 - ks_222 - ux2 / (REAL)9;

    const REAL ks_a00 = ((ks_000 + one) * (ux2 - ux) + ks_100 * (two * ux - one) + ks_200) * half;
    const REAL ks_a01 = (ks_001 * (ux2 - ux) + ks_101 * (two * ux - one) + ks_201) * half;
    const REAL ks_a02 = ((ks_002 + third) * (ux2 - ux) + ks_102 * (two * ux - one) + ks_202) * half;
    const REAL ks_a10 = (ks_010 * (ux2 - ux) + ks_110 * (two * ux - one) + ks_210) * half;
    const REAL ks_a11 = (ks_011 * (ux2 - ux) + ks_111 * (two * ux - one) + ks_211) * half;
    const REAL ks_a12 = (ks_012 * (ux2 - ux) + ks_112 * (two * ux - one) + ks_212) * half;
    const REAL ks_a20 = ((ks_020 + third) * (ux2 - ux) + ks_120 * (two * ux - one) + ks_220) * half;
    const REAL ks_a21 = (ks_021 * (ux2 - ux) + ks_121 * (two * ux - one) + ks_221) * half;
    const REAL ks_a22 = ((ks_022 + RC(1, 9)) * (ux2 - ux) + ks_122 * (two * ux - one) + ks_222) * half;

    const REAL ks_c00 = ((ks_000 + one) * (ux2 + ux) + ks_100 * (two * ux + one) + ks_200) * half;
    const REAL ks_c01 = (ks_001 * (ux2 + ux) + ks_101 * (two * ux + one) + ks_201) * half;
    const REAL ks_c02 = ((ks_002 + third) * (ux2 + ux) + ks_102 * (two * ux + one) + ks_202) * half;
    const REAL ks_c10 = (ks_010 * (ux2 + ux) + ks_110 * (two * ux + one) + ks_210) * half;
    const REAL ks_c11 = (ks_011 * (ux2 + ux) + ks_111 * (two * ux + one) + ks_211) * half;
    const REAL ks_c12 = (ks_012 * (ux2 + ux) + ks_112 * (two * ux + one) + ks_212) * half;
    const REAL ks_c20 = ((ks_020 + third) * (ux2 + ux) + ks_120 * (two * ux + one) + ks_220) * half;
    const REAL ks_c21 = (ks_021 * (ux2 + ux) + ks_121 * (two * ux + one) + ks_221) * half;
    const REAL ks_c22 = ((ks_022 + RC(1, 9)) * (ux2 + ux) + ks_122 * (two * ux + one) + ks_222) * half;

    const REAL ks_ab0 = ks_a00 * (one - uy2) - two * uy * ks_a10 - ks_a20 - K_a00 * uy2;
    const REAL ks_ab1 = ks_a01 * (one - uy2) - two * uy * ks_a11 - ks_a21;
    const REAL ks_ab2 = ks_a02 * (one - uy2) - two * uy * ks_a12 - ks_a22 - K_a02 * uy2;
    const REAL ks_bb0 = ks_b00 * (one - uy2) - two * uy * ks_b10 - ks_b20 - K_b00 * uy2;
    const REAL ks_bb1 = ks_b01 * (one - uy2) - two * uy * ks_b11 - ks_b21;
    const REAL ks_bb2 = ks_b02 * (one - uy2) - two * uy * ks_b12 - ks_b22 - K_b02 * uy2;
    const REAL ks_cb0 = ks_c00 * (one - uy2) - two * uy * ks_c10 - ks_c20 - K_c00 * uy2;
    const REAL ks_cb1 = ks_c01 * (one - uy2) - two * uy * ks_c11 - ks_c21;
    const REAL ks_cb2 = ks_c02 * (one - uy2) - two * uy * ks_c12 - ks_c22 - K_c02 * uy2;

    const REAL ks_aa0 = ((ks_a00 + K_a00) * (uy2 - uy) + ks_a10 * (two * uy - one) + ks_a20) * half;
    const REAL ks_aa1 = (ks_a01 * (uy2 - uy) + ks_a11 * (two * uy - one) + ks_a21) * half;
    const REAL ks_aa2 = ((ks_a02 + K_a02) * (uy2 - uy) + ks_a12 * (two * uy - one) + ks_a22) * half;
    const REAL ks_ba0 = ((ks_b00 + K_b00) * (uy2 - uy) + ks_b10 * (two * uy - one) + ks_b20) * half;
    const REAL ks_ba1 = (ks_b01 * (uy2 - uy) + ks_b11 * (two * uy - one) + ks_b21) * half;
    const REAL ks_ba2 = ((ks_b02 + K_b02) * (uy2 - uy) + ks_b12 * (two * uy - one) + ks_b22) * half;
    const REAL ks_ca0 = ((ks_c00 + K_c00) * (uy2 - uy) + ks_c10 * (two * uy - one) + ks_c20) * half;
    const REAL ks_ca1 = (ks_c01 * (uy2 - uy) + ks_c11 * (two * uy - one) + ks_c21) * half;
    const REAL ks_ca2 = ((ks_c02 + K_c02) * (uy2 - uy) + ks_c12 * (two * uy - one) + ks_c22) * half;

    const REAL ks_ac0 = ((ks_a00 + K_a00) * (uy2 + uy) + ks_a10 * (two * uy + one) + ks_a20) * half;
    const REAL ks_ac1 = (ks_a01 * (uy2 + uy) + ks_a11 * (two * uy + one) + ks_a21) * half;
    const REAL ks_ac2 = ((ks_a02 + K_a02) * (uy2 + uy) + ks_a12 * (two * uy + one) + ks_a22) * half;
    const REAL ks_bc0 = ((ks_b00 + K_b00) * (uy2 + uy) + ks_b10 * (two * uy + one) + ks_b20) * half;
    const REAL ks_bc1 = (ks_b01 * (uy2 + uy) + ks_b11 * (two * uy + one) + ks_b21) * half;
    const REAL ks_bc2 = ((ks_b02 + K_b02) * (uy2 + uy) + ks_b12 * (two * uy + one) + ks_b22) * half;
    const REAL ks_cc0 = ((ks_c00 + K_c00) * (uy2 + uy) + ks_c10 * (two * uy + one) + ks_c20) * half;
    const REAL ks_cc1 = (ks_c01 * (uy2 + uy) + ks_c11 * (two * uy + one) + ks_c21) * half;
    const REAL ks_cc2 = ((ks_c02 + K_c02) * (uy2 + uy) + ks_c12 * (two * uy + one) + ks_c22) * half;

    f[MMO] = ks_aa0 * (one - uz2) - two * uz * ks_aa1 - ks_aa2 - K_aa0 * uz2;
    f[MOO] = ks_ab0 * (one - uz2) - two * uz * ks_ab1 - ks_ab2 - K_ab0 * uz2;
    f[MPO] = ks_ac0 * (one - uz2) - two * uz * ks_ac1 - ks_ac2 - K_ac0 * uz2;
    f[OMO] = ks_ba0 * (one - uz2) - two * uz * ks_ba1 - ks_ba2 - K_ba0 * uz2;
    f[OOO] = ks_bb0 * (one - uz2) - two * uz * ks_bb1 - ks_bb2 - K_bb0 * uz2;
    f[OPO] = ks_bc0 * (one - uz2) - two * uz * ks_bc1 - ks_bc2 - K_bc0 * uz2;
    f[PMO] = ks_ca0 * (one - uz2) - two * uz * ks_ca1 - ks_ca2 - K_ca0 * uz2;
    f[POO] = ks_cb0 * (one - uz2) - two * uz * ks_cb1 - ks_cb2 - K_cb0 * uz2;
    f[PPO] = ks_cc0 * (one - uz2) - two * uz * ks_cc1 - ks_cc2 - K_cc0 * uz2;

    f[MMM] = ((ks_aa0 + K_aa0) * (uz2 - uz) + ks_aa1 * (two * uz - one) + ks_aa2) * half;
    f[MOM] = ((ks_ab0 + K_ab0) * (uz2 - uz) + ks_ab1 * (two * uz - one) + ks_ab2) * half;
    f[MPM] = ((ks_ac0 + K_ac0) * (uz2 - uz) + ks_ac1 * (two * uz - one) + ks_ac2) * half;
    f[OMM] = ((ks_ba0 + K_ba0) * (uz2 - uz) + ks_ba1 * (two * uz - one) + ks_ba2) * half;
    f[OOM] = ((ks_bb0 + K_bb0) * (uz2 - uz) + ks_bb1 * (two * uz - one) + ks_bb2) * half;
    f[OPM] = ((ks_bc0 + K_bc0) * (uz2 - uz) + ks_bc1 * (two * uz - one) + ks_bc2) * half;
    f[PMM] = ((ks_ca0 + K_ca0) * (uz2 - uz) + ks_ca1 * (two * uz - one) + ks_ca2) * half;
    f[POM] = ((ks_cb0 + K_cb0) * (uz2 - uz) + ks_cb1 * (two * uz - one) + ks_cb2) * half;
    f[PPM] = ((ks_cc0 + K_cc0) * (uz2 - uz) + ks_cc1 * (two * uz - one) + ks_cc2) * half;

    f[MMP] = ((ks_aa0 + K_aa0) * (uz2 + uz) + ks_aa1 * (two * uz + one) + ks_aa2) * half;
    f[MOP] = ((ks_ab0 + K_ab0) * (uz2 + uz) + ks_ab1 * (two * uz + one) + ks_ab2) * half;
    f[MPP] = ((ks_ac0 + K_ac0) * (uz2 + uz) + ks_ac1 * (two * uz + one) + ks_ac2) * half;
    f[OMP] = ((ks_ba0 + K_ba0) * (uz2 + uz) + ks_ba1 * (two * uz + one) + ks_ba2) * half;
    f[OOP] = ((ks_bb0 + K_bb0) * (uz2 + uz) + ks_bb1 * (two * uz + one) + ks_bb2) * half;
    f[OPP] = ((ks_bc0 + K_bc0) * (uz2 + uz) + ks_bc1 * (two * uz + one) + ks_bc2) * half;
    f[PMP] = ((ks_ca0 + K_ca0) * (uz2 + uz) + ks_ca1 * (two * uz + one) + ks_ca2) * half;
    f[POP] = ((ks_cb0 + K_cb0) * (uz2 + uz) + ks_cb1 * (two * uz + one) + ks_cb2) * half;
    f[PPP] = ((ks_cc0 + K_cc0) * (uz2 + uz) + ks_cc1 * (two * uz + one) + ks_cc2) * half;
}

/* ---------------------------------------------------------------------------
 * Open boundary: restoreRho (restoreRho.h:1-196), restoreUxUyUz
 * (restoreUxUyUz.h:1-409, faces by the closed form of
 * tools/CppCodeGeneratorForMBBC/restoreFaceUxUyUz.py), applyMBBC
 * (applyMBBC.h:1-1739).  Coefficients come from the generated table; sums run
 * in increasing index order as in the generated reference code.
 * ------------------------------------------------------------------------- */

static inline REAL NAME(rat)(int num, int den) { return (REAL)num / (REAL)den; }

/* accumulate  sum_j coef_j * x_j  over nonzero coefficients, first term exact */
#define ORC_ACC(acc, started, c, x)              \
    do {                                          \
        const REAL c_ = (c), x_ = (x);            \
        REAL t_;                                  \
        if (c_ == (REAL)1) t_ = x_;               \
        else if (c_ == (REAL)-1) t_ = -x_;        \
        else t_ = c_ * x_;                        \
        if (!(started)) { (acc) = t_; (started) = 1; } \
        else if (c_ == (REAL)1) (acc) = (acc) + x_;    \
        else if (c_ == (REAL)-1) (acc) = (acc) - x_;   \
        else (acc) = (acc) + t_;                  \
    } while (0)

void NAME(orc_restore_rho)(int n, NAME(bcr) *BC, const REAL *f)
{
    const REAL ux = BC->ux, uy = BC->uy, uz = BC->uz;
    /* scqu: all three terms are present with explicit coefficients */
    REAL scqu = NAME(rat)(LBM3_BT_RHO_U_NUM[n][0], LBM3_BT_RHO_U_DEN[n][0]) * ux;
    scqu = scqu + NAME(rat)(LBM3_BT_RHO_U_NUM[n][1], LBM3_BT_RHO_U_DEN[n][1]) * uy;
    scqu = scqu + NAME(rat)(LBM3_BT_RHO_U_NUM[n][2], LBM3_BT_RHO_U_DEN[n][2]) * uz;
    const REAL s = NAME(rat)(LBM3_BT_RHO_S_NUM[n], LBM3_BT_RHO_S_DEN[n]) + scqu;
    REAL scmf = 0; int st = 0;
    for (int q = 0; q < 27; ++q)
        if (LBM3_BT_RHO_F_NUM[n][q] != 0)
            ORC_ACC(scmf, st, NAME(rat)(LBM3_BT_RHO_F_NUM[n][q], LBM3_BT_RHO_F_DEN[n][q]), f[q]);
    const REAL rho = scmf / s;
    BC->dRho = rho - (REAL)1;
}

void NAME(orc_restore_u)(int n, NAME(bcr) *BC, const REAL *f)
{
    const REAL rho = BC->dRho + (REAL)1;
    const int nx = LBM3_BT_NORMAL[n][0], ny = LBM3_BT_NORMAL[n][1], nz = LBM3_BT_NORMAL[n][2];
    if (n < 6) {
        /* face: rhoZ = sum(c.n=0) f + 2 sum(c.n=+1) f, tangential k_nnt = 0 */
        const int axis = nx ? 0 : (ny ? 1 : 2);
        const int sgn = nx + ny + nz;
        REAL fZero = 0, fPlus = 0; int s0 = 0, s1 = 0;
        for (int q = 0; q < 27; ++q) {
            const int cn = nx * ORC_CX[q] + ny * ORC_CY[q] + nz * ORC_CZ[q];
            if (cn == 0) ORC_ACC(fZero, s0, (REAL)1, f[q]);
            else if (cn == 1) ORC_ACC(fPlus, s1, (REAL)1, f[q]);
        }
        const REAL uNormal = (fZero + (REAL)2 * fPlus) / rho - (REAL)1;
        REAL *uu[3] = {&BC->ux, &BC->uy, &BC->uz};
        *uu[axis] = sgn > 0 ? uNormal : -uNormal;
        const REAL tangentDenominator = fZero + (REAL)2 * (uNormal + (REAL)1) * rho * uNormal;
        for (int ax = 0; ax < 3; ++ax) {
            if (ax == axis) continue;
            const int *c = ax == 0 ? ORC_CX : (ax == 1 ? ORC_CY : ORC_CZ);
            REAL jZero = 0, jPlus = 0; int t0 = 0, t1 = 0;
            for (int q = 0; q < 27; ++q) {
                const int cn = nx * ORC_CX[q] + ny * ORC_CY[q] + nz * ORC_CZ[q];
                if (c[q] == 0) continue;
                if (cn == 0) ORC_ACC(jZero, t0, (REAL)c[q], f[q]);
                else if (cn == 1) ORC_ACC(jPlus, t1, (REAL)c[q], f[q]);
            }
            *uu[ax] = (((REAL)2 * uNormal + (REAL)1) * jZero + (REAL)4 * uNormal * jPlus) / tangentDenominator;
        }
        return;
    }
    REAL s[3];
    for (int r = 0; r < 3; ++r) {
        REAL acc = 0; int st = 0;
        for (int q = 0; q < 27; ++q)
            if (LBM3_BT_U_F_NUM[n][r][q] != 0)
                ORC_ACC(acc, st, NAME(rat)(LBM3_BT_U_F_NUM[n][r][q], LBM3_BT_U_F_DEN[n][r][q]), f[q]);
        const REAL scmf = acc / rho;
        s[r] = scmf - NAME(rat)(LBM3_BT_U_Q_NUM[n][r], LBM3_BT_U_Q_DEN[n][r]);
    }
    REAL u[3];
    for (int d = 0; d < 3; ++d) {
        REAL acc = 0; int st = 0;
        for (int j = 0; j < 3; ++j)
            if (LBM3_BT_U_INV_NUM[n][d][j] != 0)
                ORC_ACC(acc, st, NAME(rat)(LBM3_BT_U_INV_NUM[n][d][j], LBM3_BT_U_INV_DEN[n][d][j]), s[j]);
        u[d] = acc;
    }
    BC->ux = u[0]; BC->uy = u[1]; BC->uz = u[2];
}

/* equilibrium raw moments as the generator spells them (applyMBBC.h, m0..m8 lines) */
static inline REAL NAME(meq)(int mid, REAL rho, REAL ux, REAL uy, REAL uz)
{
    const REAL t = RC(1, 3), n9 = RC(1, 9);
    switch (mid) {
    case 0: return rho;
    case 1: return rho * ux;
    case 2: return rho * uy;
    case 3: return rho * uz;
    case 4: return t * rho + rho * ux * ux;
    case 5: return t * rho + rho * uy * uy;
    case 6: return t * rho + rho * uz * uz;
    case 7: return rho * uy * uz;
    case 8: return rho * ux * uz;
    case 9: return rho * ux * uy;
    case 10: return 0;
    case 11: return t * rho * uy + rho * ux * ux * uy;
    case 12: return t * rho * uz + rho * ux * ux * uz;
    case 13: return t * rho * ux + rho * ux * uy * uy;
    case 14: return t * rho * uz + rho * uy * uy * uz;
    case 15: return t * rho * ux + rho * ux * uz * uz;
    case 16: return t * rho * uy + rho * uy * uz * uz;
    case 17: return n9 * rho + t * rho * uy * uy + t * rho * uz * uz + rho * uy * uy * uz * uz;
    case 18: return n9 * rho + t * rho * ux * ux + t * rho * uz * uz + rho * ux * ux * uz * uz;
    case 19: return n9 * rho + t * rho * ux * ux + t * rho * uy * uy + rho * ux * ux * uy * uy;
    case 20: return t * rho * uy * uz;
    case 21: return t * rho * ux * uz;
    default: return t * rho * ux * uy;
    }
}

void NAME(orc_mbbc)(int n, const NAME(bcr) *BC, REAL *f)
{
    const REAL rho = BC->dRho + (REAL)1;
    const int nu = LBM3_BT_MBBC_NU[n];
    REAL s[19];
    for (int r = 0; r < nu; ++r) {
        REAL acc = 0; int st = 0;
        for (int q = 0; q < 27; ++q)
            if (LBM3_BT_MBBC_SMF[n][r][q] != 0) ORC_ACC(acc, st, (REAL)LBM3_BT_MBBC_SMF[n][r][q], f[q]);
        const REAL m = NAME(meq)(LBM3_BT_MBBC_MOMENT[n][r], rho, BC->ux, BC->uy, BC->uz);
        s[r] = m - acc;
    }
    for (int r = 0; r < nu; ++r) {
        REAL acc = 0; int st = 0;
        for (int j = 0; j < nu; ++j)
            if (LBM3_BT_MBBC_INV_NUM[n][r][j] != 0)
                ORC_ACC(acc, st, NAME(rat)(LBM3_BT_MBBC_INV_NUM[n][r][j], LBM3_BT_MBBC_INV_DEN[n][r][j]), s[j]);
        f[LBM3_BT_MBBC_UNKNOWN[n][r]] = acc;
    }
}

#undef ORC_ACC
#undef RC
