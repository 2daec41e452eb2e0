/*
 * ORACLE (test infrastructure only): per-step update of one grid level,
 * instantiated for REAL=float and REAL=double by step.c.  Sequential
 * restatement of updateGrid.h:16-396 and applyInitialCondition.h:8-84.
 * Esotwist makes every cell's read set private to it, so sequential order
 * gives the parallel result.
 */

#define F_AT(g, q, c) (((REAL *)(g)->f)[(size_t)(q) * (size_t)(g)->Info.cellCount + (size_t)(c)])

/* getCompressedIJKNBR, cellFunctions.h:21-50 + finishNBRPlus NBRFunctions.h:5-11 */
static void NAME(decode)(const orc_grid *g, int cell, int *ijk, int *nb)
{
    /* nb: self, iPlus, jPlus, kPlus, ijPlus, ikPlus, jkPlus, ijkPlus */
    const int N = g->Info.cellCount;
    const int shift = g->shifter[cell];
    int row, back = 0;
    if (shift >= 0) row = shift;
    else { row = g->shifter[cell + shift]; back = -shift; }
    if (ijk) { ijk[0] = g->ri[row] + back; ijk[1] = g->rj[row]; ijk[2] = g->rk[row]; }
    nb[0] = cell;
    nb[2] = g->rjp[row] + back;
    nb[3] = g->rkp[row] + back;
    nb[6] = g->rjkp[row] + back;
    nb[1] = cell + 1 >= N ? 0 : cell + 1;
    nb[4] = nb[2] + 1 >= N ? 0 : nb[2] + 1;
    nb[5] = nb[3] + 1 >= N ? 0 : nb[3] + 1;
    nb[7] = nb[6] + 1 >= N ? 0 : nb[6] + 1;
}

/* neighbour slot (index into nb[]) each direction reads from before collision,
   esotwistStreamingFunctions.h:17-44, and writes to after collision, :73-100 */
static const int NAME(PRE_NB)[27] = {0, 0, 1, 3, 0, 2, 0, 3, 1, 0, 5, 4, 0, 3, 2, 1, 2, 0, 6, 5, 2, 4, 3, 6, 1, 7, 0};
static const int NAME(POST_NB)[27] = {0, 1, 0, 0, 3, 0, 2, 1, 3, 5, 0, 0, 4, 2, 3, 2, 1, 6, 0, 2, 5, 3, 4, 1, 6, 0, 7};

static inline void NAME(pre_index)(const int *nb, int flip, int q, int *cell, int *slot)
{
    *cell = nb[NAME(PRE_NB)[q]];
    *slot = flip ? ORC_INV[q] : q;
}
static inline void NAME(post_index)(const int *nb, int flip, int q, int *cell, int *slot)
{
    *cell = nb[NAME(POST_NB)[q]];
    *slot = flip ? q : ORC_INV[q];
}

/* getLocalBC hook as rules */
static void NAME(hook_local)(const lbm3_case *C, const lbm3_info *I, int i, int j, int k, NAME(bcr) *BC)
{
    const int level = (int)I->gridID;
    for (int r = 0; r < C->nLocalRules; ++r) {
        const lbm3_local_rule *R = &C->localRules[r];
        if (R->levelMask && !(R->levelMask & (1 << level))) continue;
        if (R->wallID != -2 && R->wallID != BC->wallID) continue;
        if (R->set & LBM3_SET_VELOCITY) { BC->ux = R->ux; BC->uy = R->uy; BC->uz = R->uz; }
        if (R->set & LBM3_SET_ROTATION) {
            const float x = (float)i * I->res + I->ox, y = (float)j * I->res + I->oy, z = (float)k * I->res + I->oz;
            const float rx = x - R->origin[0], ry = y - R->origin[1], rz = z - R->origin[2];
            const float s = I->dtPhys / I->res;
            BC->ux = (R->omega[1] * rz - R->omega[2] * ry) * s;
            BC->uy = (R->omega[2] * rx - R->omega[0] * rz) * s;
            BC->uz = (R->omega[0] * ry - R->omega[1] * rx) * s;
        }
        if (R->set & LBM3_SET_FORCE) { BC->gx = R->gx; BC->gy = R->gy; BC->gz = R->gz; }
        if (R->set & LBM3_SET_LIMITER) BC->collisionLimiter = R->collisionLimiter;
        if (R->set & LBM3_SET_NU_MULTIPLIER) BC->nuMultiplier = R->nuMultiplier;
        if (R->set & LBM3_SET_USE_IBB) BC->useIBB = R->useIBB;
    }
}

/* getInitialCondition hook as data */
static void NAME(hook_init)(const lbm3_case *C, const lbm3_info *I, int i, int j, int k, NAME(bcr) *BC)
{
    const lbm3_init *In = &C->init;
    if (In->kind == LBM3_INIT_TAYLOR_GREEN) {
        const float x = (float)i * I->res + I->ox, y = (float)j * I->res + I->oy, z = (float)k * I->res + I->oz;
        const float X = x / In->length, Y = y / In->length, Z = z / In->length;
        const float u0 = In->u0;
        BC->ux = u0 * sinf(X) * cosf(Y) * cosf(Z);
        BC->uy = -u0 * cosf(X) * sinf(Y) * cosf(Z);
        BC->uz = 0.f;
        BC->dRho = In->pressure ? (3.f * u0 * u0 / 16.f) * (cosf(2.f * X) + cosf(2.f * Y)) * (2.f + cosf(2.f * Z)) : 0.f;
    } else {
        BC->dRho = In->dRho; BC->ux = In->ux; BC->uy = In->uy; BC->uz = In->uz;
        if (In->profile.radius > 0.f) {
            const float x = (float)i * I->res + I->ox, y = (float)j * I->res + I->oy, z = (float)k * I->res + I->oz;
            const float s = orc_profile_factor(&In->profile, x, y, z);
            BC->ux = (float)BC->ux * s; BC->uy = (float)BC->uy * s; BC->uz = (float)BC->uz * s;
        }
    }
}

static const lbm3_bc NAME(BC_DEFAULT) = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, -1, 0, 1, 0, 0, 0, 1, 7.f, 0.05f, 1.f};

/* applyInitialCondition, applyInitialCondition.h:8-84; fields4 (dRho, ux, uy,
   uz per cell, cell order) stands for the getInitialCondition hook when given */
void NAME(orc_init_level)(orc_sim *S, int L, const float *fields4)
{
    orc_grid *g = &S->g[L];
    const int N = g->Info.cellCount;
    const int flip = g->Info.esotwistFlipper;
    for (int c = 0; c < N; ++c) {
        int ijk[3], nb[8];
        NAME(decode)(g, c, ijk, nb);
        NAME(bcr) BC; NAME(bc_from)(&BC, &NAME(BC_DEFAULT));
        if (fields4) {
            BC.dRho = fields4[4 * (size_t)c]; BC.ux = fields4[4 * (size_t)c + 1];
            BC.uy = fields4[4 * (size_t)c + 2]; BC.uz = fields4[4 * (size_t)c + 3];
        } else {
            NAME(hook_init)(&S->C, &g->Info, ijk[0], ijk[1], ijk[2], &BC);
        }
        REAL f[27];
        NAME(orc_feq)(BC.dRho, BC.ux, BC.uy, BC.uz, f);
        for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, flip, q, &cc, &s); F_AT(g, s, cc) = f[q]; }
    }
    for (int b = 0; b < g->nOpen; ++b) {
        orc_openbc *O = &g->open[b];
        for (int x = 0; x < O->count; ++x) {
            const int c = O->index[x];
            int ijk[3], nb[8];
            NAME(decode)(g, c, ijk, nb);
            int nx = 0, ny = 0, nz = 0;
            if (ijk[0] == 0) nx = -1; else if (ijk[0] == g->Info.cellCountX - 1) nx = 1;
            if (ijk[1] == 0) ny = -1; else if (ijk[1] == g->Info.cellCountY - 1) ny = 1;
            if (ijk[2] == 0) nz = -1; else if (ijk[2] == g->Info.cellCountZ - 1) nz = 1;
            REAL f[27], m[4];
            for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, flip, q, &cc, &s); f[q] = F_AT(g, s, cc); }
            NAME(orc_moments)(f, m);
            ((REAL *)O->dRhoRef)[x] = m[0];
            ((REAL *)O->uNormalRef)[x] = (REAL)nx * m[1] + (REAL)ny * m[2] + (REAL)nz * m[3];
        }
    }
}

/* applyIBB, interpolatedBouncebackFunctions.h:43-157; returns force in g3 */
static void NAME(ibb)(orc_grid *g, const int *nb, int flip, REAL *fPost, const NAME(bcr) *BC,
                      uint32_t wallData, int wallMap, REAL *g3)
{
    const int W = g->W;
    int writeBuffer = 0;
    REAL fBuffer = 0;
    REAL m[4];
    NAME(orc_moments)(fPost, m);
    const REAL dRho = m[0];
    int linkIndex = 0;
    uint32_t code = 0u;
    for (int d = 1; d < 27; ++d) {
        const int inv = ORC_INV[d];
        const int exists = (wallData & (1u << d)) != 0u;
        if (!exists) {
            if (writeBuffer) fPost[inv] = fBuffer;
            writeBuffer = 0;
            continue;
        }
        const REAL eiDotFi = (REAL)ORC_CX[inv] * BC->ux + (REAL)ORC_CY[inv] * BC->uy + (REAL)ORC_CZ[inv] * BC->uz;
        const REAL wallTerm = (REAL)6 * NAME(wgt)(inv) * ((REAL)1 + dRho) * eiDotFi;
        REAL gamma = (REAL)0.5;
        if (BC->useIBB) {
            if (linkIndex % 4 == 0) code = g->linkLength[(linkIndex / 4) * W + wallMap];
            const uint32_t lc = (code >> (8 * (linkIndex % 4))) & 255u;
            gamma = (REAL)lc / (REAL)254;
            gamma = gamma < (REAL)0.01f ? (REAL)0.01f : (gamma > (REAL)1 ? (REAL)1 : gamma);
            ++linkIndex;
        }
        int pc, ps;
        NAME(pre_index)(nb, flip, d, &pc, &ps);
        const REAL fPre = F_AT(g, ps, pc);
        const REAL invDen = (REAL)1 / ((REAL)1 + gamma);
        const REAL fRes = (((REAL)1 - gamma) * invDen) * fPre + (gamma * invDen) * fPost[inv] + (gamma * invDen) * fPost[d]
                        + invDen * wallTerm;
        g3[0] += -(fPost[d]) * ((REAL)ORC_CX[d] - BC->ux) + (fRes) * ((REAL)ORC_CX[inv] - BC->ux);
        g3[1] += -(fPost[d]) * ((REAL)ORC_CY[d] - BC->uy) + (fRes) * ((REAL)ORC_CY[inv] - BC->uy);
        g3[2] += -(fPost[d]) * ((REAL)ORC_CZ[d] - BC->uz) + (fRes) * ((REAL)ORC_CZ[inv] - BC->uz);
        if (d % 2 == 0) {
            fPost[d] = fRes;
            if (writeBuffer) fPost[inv] = fBuffer;
            writeBuffer = 0;
        } else {
            fBuffer = fRes;
            writeBuffer = 1;
        }
    }
}

/* the open-boundary kernel body, updateGrid.h:192-354 */
static void NAME(open_cell)(orc_sim *S, orc_grid *g, orc_openbc *O, int x, int flip)
{
    const lbm3_info *I = &g->Info;
    const int c = O->index[x];
    const int wm = g->wallMap[c];
    int trackFlow = wm != -2;
    if (wm >= 0 && (g->wallData[wm] & 1u)) trackFlow = 0;
    int ijk[3], nb[8];
    NAME(decode)(g, c, ijk, nb);
    int nx = 0, ny = 0, nz = 0;
    if (ijk[0] == 0) nx = -1; else if (ijk[0] == I->cellCountX - 1) nx = 1;
    if (ijk[1] == 0) ny = -1; else if (ijk[1] == I->cellCountY - 1) ny = 1;
    if (ijk[2] == 0) nz = -1; else if (ijk[2] == I->cellCountZ - 1) nz = 1;
    REAL f[27];
    int cellIx[27], slotIx[27];
    REAL dRhoZ = 0;
    for (int q = 0; q < 27; ++q) {
        NAME(pre_index)(nb, flip, q, &cellIx[q], &slotIx[q]);
        const int unknown = nx * ORC_CX[q] < 0 || ny * ORC_CY[q] < 0 || nz * ORC_CZ[q] < 0;
        if (unknown) { f[q] = 0; continue; }
        f[q] = F_AT(g, slotIx[q], cellIx[q]);
        const int prod = nx * ORC_CX[q] + ny * ORC_CY[q] + nz * ORC_CZ[q];
        if (prod == 0) dRhoZ += f[q];
        else dRhoZ += (REAL)2 * f[q];
        f[q] += NAME(wgt)(q);
    }
    const REAL dRhoRef = ((REAL *)O->dRhoRef)[x];
    const REAL uNormalRef = ((REAL *)O->uNormalRef)[x];
    lbm3_bc hb = orc_hook_open_bc(&S->C, I, ijk[0], ijk[1], ijk[2]);
    NAME(bcr) BC; NAME(bc_from)(&BC, &hb);
    if (abs(nx) + abs(ny) + abs(nz) != 1) BC.nonReflective = 0;
    /* table index of this normal */
    int n = 0;
    const int code = (nx + 5) * 100 + (ny + 5) * 10 + (nz + 5);
    while (LBM3_BT_NORMAL_CODE[n] != code) ++n;
    const REAL cs = (REAL)ORC_INVSQRT3;
    if (S->C.openScheme == LBM3_OPEN_IMPEDANCE) {
        /* the impedance scheme, cases/nonReflectiveTest/
           updateGrid_SCHLAFFER_IMPEDANCE_SCHEME.h:250-311: dRhoRef and uNormalRef
           hold the previous completed state */
        const REAL tol = (REAL)S->C.rhoReflectionTolerance;
        const int useNR = BC.nonReflective && tol < (REAL)1;
        REAL dRhoNR = 0;
        if (useNR) {
            const REAL cs2 = cs * cs;
            const REAL bracket = uNormalRef + cs + (REAL)1;
            const REAL t = uNormalRef - dRhoZ;
            const REAL D = ORC_FMA(bracket, bracket, cs2);
            const REAL B = ORC_FMA(bracket, cs + t, -cs2 * dRhoRef);
            const REAL H = ORC_FMA(-t, (REAL)2 * cs + t, (REAL)2 * cs2 * dRhoRef);
            const REAL disc = ORC_FMA(D, H, B * B);
            dRhoNR = H / (B + ORC_SQRT(disc));
        }
        if (BC.dirichletRho) {
            if (useNR) {
                const REAL lo = dRhoNR - tol, hi = dRhoNR + tol;
                BC.dRho = BC.dRho < lo ? lo : (hi < BC.dRho ? hi : BC.dRho);
            }
            NAME(orc_restore_u)(n, &BC, f);
        } else if (BC.dirichletU) {
            if (useNR) {
                const REAL uMin = (dRhoZ - dRhoNR - tol) / (dRhoNR + (REAL)1 + tol);
                const REAL uMax = (dRhoZ - dRhoNR + tol) / (dRhoNR + (REAL)1 - tol);
                REAL *u = nx ? &BC.ux : (ny ? &BC.uy : &BC.uz);
                const int pos = nx + ny + nz > 0;
                const REAL lo = pos ? uMin : -uMax, hi = pos ? uMax : -uMin;
                *u = *u < lo ? lo : (hi < *u ? hi : *u);
            }
            NAME(orc_restore_rho)(n, &BC, f);
        }
    } else if (BC.dirichletRho) {
        if (BC.nonReflective) {
            const REAL dRhoTarget = BC.dRho;
            const REAL rhoTarget = dRhoTarget + (REAL)1;
            const REAL rhoTerm = ((dRhoZ - dRhoTarget) / rhoTarget);
            const REAL numerator = (REAL)2 * (uNormalRef + cs * rhoTerm);
            const REAL root = ((REAL)1 + cs - uNormalRef) * ((REAL)1 + cs - uNormalRef) + (REAL)4 * (uNormalRef + cs * rhoTerm);
            const REAL denominator = (REAL)1 + cs - uNormalRef + ORC_SQRT(root);
            const REAL uNormal = numerator / denominator;
            BC.dRho = (dRhoZ - uNormal) / ((REAL)1 + uNormal);
        }
        NAME(orc_restore_u)(n, &BC, f);
    } else if (BC.dirichletU) {
        if (BC.nonReflective) {
            REAL uT = 0;
            if (nx > 0) uT = BC.ux;
            if (nx < 0) uT = -BC.ux;
            if (ny > 0) uT = BC.uy;
            if (ny < 0) uT = -BC.uy;
            if (nz > 0) uT = BC.uz;
            if (nz < 0) uT = -BC.uz;
            const REAL rhoRef = dRhoRef + (REAL)1;
            const REAL rhoTerm = ((dRhoZ - dRhoRef) / rhoRef);
            const REAL numerator = (REAL)2 * (uT + cs * rhoTerm);
            const REAL root = ((REAL)1 + cs - uT) * ((REAL)1 + cs - uT) + (REAL)4 * (uT + cs * rhoTerm);
            const REAL denominator = (REAL)1 + cs - uT + ORC_SQRT(root);
            const REAL uNormal = numerator / denominator;
            if (nx > 0) BC.ux = uNormal;
            if (nx < 0) BC.ux = -uNormal;
            if (ny > 0) BC.uy = uNormal;
            if (ny < 0) BC.uy = -uNormal;
            if (nz > 0) BC.uz = uNormal;
            if (nz < 0) BC.uz = -uNormal;
        }
        NAME(orc_restore_rho)(n, &BC, f);
    }
    NAME(orc_mbbc)(n, &BC, f);
    for (int q = 0; q < 27; ++q) {
        const int unknown = nx * ORC_CX[q] < 0 || ny * ORC_CY[q] < 0 || nz * ORC_CZ[q] < 0;
        if (!unknown) continue;
        f[q] -= NAME(wgt)(q);
        F_AT(g, slotIx[q], cellIx[q]) = f[q];
    }
    const REAL uNormal = (REAL)nx * BC.ux + (REAL)ny * BC.uy + (REAL)nz * BC.uz;
    if (S->C.openScheme == LBM3_OPEN_IMPEDANCE) {
        /* the completed state for the next step (:326-331) */
        ((REAL *)O->dRhoRef)[x] = BC.dRho;
        ((REAL *)O->uNormalRef)[x] = uNormal;
    } else {
        const REAL beta = BC.nonReflectiveBeta;
        REAL alpha = beta / ((REAL)I->iterationsFinished + beta + (REAL)10);
        int every = 1;
        while (alpha < (REAL)1.e-5f) { alpha *= (REAL)3; every *= 3; }
        if (I->iterationsFinished % every == 0) {
            ((REAL *)O->dRhoRef)[x] = ORC_FMA(alpha, BC.dRho - dRhoRef, dRhoRef);
            ((REAL *)O->uNormalRef)[x] = ORC_FMA(alpha, uNormal - uNormalRef, uNormalRef);
        }
    }
    if (!S->C.trackOpenBoundaries || !trackFlow) return;
    ((REAL *)O->dRhoCum)[x] += BC.dRho;
    ((REAL *)O->uNormalCum)[x] += uNormal;
}

/* getRotorFraction, rotorFunctions.h:155-208 (projectXYZIntoRotorFrame :8-35,
   interpolateRotorCube :135-153): the cell's interpolation cell in the rotor
   frame and its solid fraction; 0 when it lies outside every marked block */
static int NAME(rotor_locate)(const orc_rotor *R, REAL c, REAL s, REAL x, REAL y, REAL z, int *rmOut, int *ijkOut,
                              REAL *wOut, REAL *frOut)
{
    REAL xr = x, yr = y, zr = z;
    const REAL dx = x - (REAL)R->ox, dy = y - (REAL)R->oy, dz = z - (REAL)R->oz;
    if (R->axis == LBM3_AXIS_X) { yr = (REAL)R->oy + c * dy + s * dz; zr = (REAL)R->oz - s * dy + c * dz; }
    else if (R->axis == LBM3_AXIS_Y) { xr = (REAL)R->ox + c * dx - s * dz; zr = (REAL)R->oz + s * dx + c * dz; }
    else if (R->axis == LBM3_AXIS_Z) { xr = (REAL)R->ox + c * dx + s * dy; yr = (REAL)R->oy - s * dx + c * dy; }
    if (xr < (REAL)R->B.xMin || xr >= (REAL)R->B.xMax || yr < (REAL)R->B.yMin || yr >= (REAL)R->B.yMax ||
        zr < (REAL)R->B.zMin || zr >= (REAL)R->B.zMax) return 0;
    const REAL xRel = xr - (REAL)R->B.xMin, yRel = yr - (REAL)R->B.yMin, zRel = zr - (REAL)R->B.zMin;
    const REAL res = (REAL)R->res;
    const int ii = (int)(xRel / res), ji = (int)(yRel / res), ki = (int)(zRel / res);
    if (ii >= R->cx || ji >= R->cy || ki >= R->cz) return 0;
    const int bx = R->cx / 4, by = R->cy / 4;
    const int block = (ki / 4) * bx * by + (ji / 4) * bx + ii / 4;
    const int rm = R->rotorMap[block];
    if (rm < 0) return 0;
    const uint32_t packed = R->interp[rm * 64 + (ki % 4) * 16 + (ji % 4) * 4 + (ii % 4)];
    const REAL xw = xRel / res - (REAL)ii, yw = yRel / res - (REAL)ji, zw = zRel / res - (REAL)ki;
    REAL v[8];
    for (unsigned q = 0; q < 8; ++q) v[q] = (REAL)((packed >> (4u * q)) & 0xFu);
    const REAL v00 = v[0] + xw * (v[1] - v[0]), v10 = v[2] + xw * (v[3] - v[2]);
    const REAL v01 = v[4] + xw * (v[5] - v[4]), v11 = v[6] + xw * (v[7] - v[6]);
    const REAL v0 = v00 + yw * (v10 - v00), v1 = v01 + yw * (v11 - v01);
    *frOut = (v0 + zw * (v1 - v0)) * (REAL)0.125f;
    *rmOut = rm;
    ijkOut[0] = ii; ijkOut[1] = ji; ijkOut[2] = ki;
    wOut[0] = xw; wOut[1] = yw; wOut[2] = zw;
    return 1;
}

/* processRotor, rotorFunctions.h:210-304 (getRotorForcing :63-100,
   projectForcingIntoRotorFrame :37-61): adds the rotor's share of the forcing
   toward rigid rotation to BC, in the compute precision with the rotor frame's
   cosine and sine of the step given */
static void NAME(process_rotor)(orc_rotor *R, const lbm3_info *I, REAL c, REAL s, REAL x, REAL y, REAL z, REAL rho,
                                REAL uxPre, REAL uyPre, REAL uzPre, REAL *cum, int track, NAME(bcr) *BC)
{
    int rm, ijk[3];
    REAL w[3], fr;
    if (!NAME(rotor_locate)(R, c, s, x, y, z, &rm, ijk, w, &fr)) return;
    const int ii = ijk[0], ji = ijk[1], ki = ijk[2];
    const REAL xw = w[0], yw = w[1], zw = w[2];
    const REAL dx = x - (REAL)R->ox, dy = y - (REAL)R->oy, dz = z - (REAL)R->oz;
    if (fr == (REAL)0) return;
    const REAL scale = (REAL)R->radiansPerSecond * (REAL)I->dtPhys / (REAL)I->res;
    REAL ut[3] = {0, 0, 0};
    if (R->axis == LBM3_AXIS_X) { ut[1] = -scale * dz; ut[2] = scale * dy; }
    else if (R->axis == LBM3_AXIS_Y) { ut[0] = scale * dz; ut[2] = -scale * dx; }
    else if (R->axis == LBM3_AXIS_Z) { ut[0] = -scale * dy; ut[1] = scale * dx; }
    REAL gx = rho * (ut[0] - uxPre), gy = rho * (ut[1] - uyPre), gz = rho * (ut[2] - uzPre);
    REAL ff = fr * fr * ((REAL)3 - (REAL)2 * fr);
    if (*cum + ff > (REAL)1) { ff = (REAL)1 - *cum; *cum = (REAL)1; }
    else *cum += ff;
    gx *= ff; gy *= ff; gz *= ff;
    BC->gx += gx; BC->gy += gy; BC->gz += gz;
    if (!track) return;
    if (R->axis == LBM3_AXIS_X) { const REAL t = gy; gy = c * t + s * gz; gz = -s * t + c * gz; }
    else if (R->axis == LBM3_AXIS_Y) { const REAL t = gx; gx = c * t - s * gz; gz = s * t + c * gz; }
    else if (R->axis == LBM3_AXIS_Z) { const REAL t = gx; gx = c * t + s * gy; gy = -s * t + c * gy; }
    int iF = (int)((REAL)0.75f * ((REAL)(ii % 4) + xw)), jF = (int)((REAL)0.75f * ((REAL)(ji % 4) + yw));
    int kF = (int)((REAL)0.75f * ((REAL)(ki % 4) + zw));
    iF = iF < 0 ? 0 : (iF > 2 ? 2 : iF); jF = jF < 0 ? 0 : (jF > 2 ? 2 : jF); kF = kF < 0 ? 0 : (kF > 2 ? 2 : kF);
    const size_t fi = (size_t)rm * 27 + (size_t)(kF * 9 + jF * 3 + iF), nf = (size_t)R->nonZero * 27;
    /* several cells share a force cell: FP64 sums, atomic under OpenMP */
#pragma omp atomic
    R->g[fi] += (double)gx;
#pragma omp atomic
    R->g[nf + fi] += (double)gy;
#pragma omp atomic
    R->g[2 * nf + fi] += (double)gz;
}

/* the rotor loop of the cell update, updateGrid.h:98-129 */
static void NAME(rotors_cell)(orc_grid *g, const REAL *cs, const int *ijk, REAL dRho, REAL ux, REAL uy, REAL uz,
                              int track, NAME(bcr) *BC)
{
    const lbm3_info *I = &g->Info;
    const REAL x = (REAL)((float)ijk[0] * I->res + I->ox), y = (REAL)((float)ijk[1] * I->res + I->oy),
               z = (REAL)((float)ijk[2] * I->res + I->oz);
    const REAL rho = dRho + (REAL)1, rhoInv = (REAL)1 / rho;
    const REAL uxPre = (ux * rho + BC->gx) * rhoInv, uyPre = (uy * rho + BC->gy) * rhoInv, uzPre = (uz * rho + BC->gz) * rhoInv;
    REAL cum = 0;
    for (int r = 0; r < g->nRotors; ++r) {
        NAME(process_rotor)(&g->rotors[r], I, cs[2 * r], cs[2 * r + 1], x, y, z, rho, uxPre, uyPre, uzPre, &cum, track, BC);
        if (cum >= (REAL)1) break;
    }
    if (cum > (REAL)0) BC->collisionLimiter *= (REAL)1 - (cum < (REAL)0 ? (REAL)0 : (cum > (REAL)1 ? (REAL)1 : cum));
}

/* updateSingleGrid, updateGrid.h:16-386 */
void NAME(orc_step_level)(orc_sim *S, int L)
{
    orc_grid *g = &S->g[L];
    lbm3_info *I = &g->Info;
    const int N = I->cellCount;
    const int flip = I->esotwistFlipper;
    const REAL nu = (REAL)I->nu;
    /* advance the rotors (updateGrid.h:36-49); the rotor frame's cosine and
       sine of this step, computed once per rotor in the compute precision */
    REAL cs[2 * LBM3_MAX_ROTORS];
    for (int r = 0; r < g->nRotors; ++r) {
        orc_rotor *R = &g->rotors[r];
        const double twoPi = 6.283185307179586476925286766559;
        R->radiansElapsed = fmod(R->radiansElapsed + (double)I->dtPhys * (double)R->radiansPerSecond, twoPi);
        if (R->radiansElapsed < 0.0) R->radiansElapsed += twoPi;
        if (R->radiansElapsed >= twoPi) R->radiansElapsed = 0.0;
        if (sizeof(REAL) == sizeof(float)) {
            cs[2 * r] = (REAL)cosf((float)R->radiansElapsed);
            cs[2 * r + 1] = (REAL)sinf((float)R->radiansElapsed);
        } else {
            cs[2 * r] = (REAL)cos(R->radiansElapsed);
            cs[2 * r + 1] = (REAL)sin(R->radiansElapsed);
        }
    }
    const int trackRotor = S->C.trackRotorForce;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < N; ++c) {
        const int wm = g->wallMap[c];
        if (wm == -3) continue;
        int trackForce = wm != -2;
        uint32_t wallData = 0u;
        int wallID = -1;
        if (wm >= 0) {
            wallData = g->wallData[wm];
            wallID = (int)((wallData >> 27) & 31u);
            if (wallData & 1u) trackForce = 0;
        }
        int ijk[3], nb[8];
        NAME(decode)(g, c, ijk, nb);
        REAL f[27];
        for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, flip, q, &cc, &s); f[q] = F_AT(g, s, cc); }
        REAL m[4];
        NAME(orc_moments)(f, m);
        NAME(bcr) BC; NAME(bc_from)(&BC, &NAME(BC_DEFAULT));
        BC.wallID = wallID;
        BC.dRho = m[0]; BC.ux = m[1]; BC.uy = m[2]; BC.uz = m[3];
        NAME(hook_local)(&S->C, I, ijk[0], ijk[1], ijk[2], &BC);
        if (g->nRotors > 0) NAME(rotors_cell)(g, cs, ijk, m[0], m[1], m[2], m[3], trackRotor && trackForce, &BC);
        NAME(orc_collide)(f, &BC, nu);
        if (wm < 0) {
            for (int q = 0; q < 27; ++q) { int cc, s; NAME(post_index)(nb, flip, q, &cc, &s); F_AT(g, s, cc) = f[q]; }
            continue;
        }
        REAL g3[3] = {0, 0, 0};
        NAME(ibb)(g, nb, flip, f, &BC, wallData, wm, g3);
        for (int d = 0; d < 27; ++d) {
            const int link = d != 0 && (wallData & (1u << d)) != 0u;
            int cc, s;
            if (link) NAME(pre_index)(nb, !flip, ORC_INV[d], &cc, &s);
            else NAME(post_index)(nb, flip, d, &cc, &s);
            F_AT(g, s, cc) = f[d];
        }
        if (!S->C.trackWallForce || !trackForce) continue;
        REAL *gw = (REAL *)g->gWall;
        gw[wm] += g3[0];
        gw[g->W + wm] += g3[1];
        gw[2 * g->W + wm] += g3[2];
    }
    I->esotwistFlipper = !I->esotwistFlipper;
    const int nflip = I->esotwistFlipper;
    for (int b = 0; b < g->nOpen; ++b) {
#pragma omp parallel for schedule(static)
        for (int x = 0; x < g->open[b].count; ++x) NAME(open_cell)(S, g, &g->open[b], x, nflip);
    }
#pragma omp parallel for schedule(static)
    for (int p = 0; p < g->nPeriodic; ++p) {
        int nb[8];
        REAL f[27];
        NAME(decode)(g, g->perSource[p], NULL, nb);
        for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, nflip, q, &cc, &s); f[q] = F_AT(g, s, cc); }
        NAME(decode)(g, g->perIndex[p], NULL, nb);
        for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, nflip, q, &cc, &s); F_AT(g, s, cc) = f[q]; }
    }
    I->iterationsSinceTrackerReset++;
    I->iterationsFinished++;
}

/* macroscopic fields of every cell from its pre-collision populations (cell order) */
void NAME(orc_macros_level)(const orc_sim *S, int L, REAL *out4)
{
    const orc_grid *g = &S->g[L];
    for (int c = 0; c < g->Info.cellCount; ++c) {
        int nb[8];
        REAL f[27];
        NAME(decode)(g, c, NULL, nb);
        for (int q = 0; q < 27; ++q) { int cc, s; NAME(pre_index)(nb, g->Info.esotwistFlipper, q, &cc, &s); f[q] = F_AT(g, s, cc); }
        NAME(orc_moments)(f, out4 + 4 * (size_t)c);
    }
}

#undef F_AT
