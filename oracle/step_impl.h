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
    }
}

static const lbm3_bc NAME(BC_DEFAULT) = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, -1, 0, 1, 0, 0, 0, 1, 7.f, 0.05f, 1.f};

/* applyInitialCondition, applyInitialCondition.h:8-84 */
void NAME(orc_init_level)(orc_sim *S, int L)
{
    orc_grid *g = &S->g[L];
    const int N = g->Info.cellCount;
    const int flip = g->Info.esotwistFlipper;
    for (int c = 0; c < N; ++c) {
        int ijk[3], nb[8];
        NAME(decode)(g, c, ijk, nb);
        NAME(bcr) BC; NAME(bc_from)(&BC, &NAME(BC_DEFAULT));
        NAME(hook_init)(&S->C, &g->Info, ijk[0], ijk[1], ijk[2], &BC);
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
    if (BC.dirichletRho) {
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
    const REAL beta = BC.nonReflectiveBeta;
    REAL alpha = beta / ((REAL)I->iterationsFinished + beta + (REAL)10);
    int every = 1;
    while (alpha < (REAL)1.e-5f) { alpha *= (REAL)3; every *= 3; }
    if (I->iterationsFinished % every == 0) {
        ((REAL *)O->dRhoRef)[x] = ORC_FMA(alpha, BC.dRho - dRhoRef, dRhoRef);
        ((REAL *)O->uNormalRef)[x] = ORC_FMA(alpha, uNormal - uNormalRef, uNormalRef);
    }
    if (!S->C.trackOpenBoundaries || !trackFlow) return;
    ((REAL *)O->dRhoCum)[x] += BC.dRho;
    ((REAL *)O->uNormalCum)[x] += uNormal;
}

/* updateSingleGrid, updateGrid.h:16-386 (no rotors) */
void NAME(orc_step_level)(orc_sim *S, int L)
{
    orc_grid *g = &S->g[L];
    lbm3_info *I = &g->Info;
    const int N = I->cellCount;
    const int flip = I->esotwistFlipper;
    const REAL nu = (REAL)I->nu;
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
