/*
 * ORACLE (test infrastructure only): CPU restatement of the LBM3 setup path --
 * STL ingest, z-ray voxelizer, markers, grid builder, IBB link building and the
 * compression into the per-step arrays.  Sequential; every function cites the
 * reference file:line it restates.  Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "oracle.h"

#define NEWI(n) ((int *)calloc((size_t)((n) > 0 ? (n) : 1), sizeof(int)))
#define NEWB(n) ((uint8_t *)calloc((size_t)((n) > 0 ? (n) : 1), 1))
#define NEWF(n) ((float *)calloc((size_t)((n) > 0 ? (n) : 1), sizeof(float)))
#define NEWL(n) ((long long *)calloc((size_t)((n) > 0 ? (n) : 1), sizeof(long long)))

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline float fmin3(float a, float b, float c) { float m = a < b ? a : b; return m < c ? m : c; }
static inline float fmax3(float a, float b, float c) { float m = a > b ? a : b; return m > c ? m : c; }

/* ------------------------------------------------------------------------
 * Hooks as data (case-file device functions; see include/lbm3mi_types.h)
 * ---------------------------------------------------------------------- */

/* getXYZFromIJKCellIndex, cellFunctions.h:103-108 */
static inline void xyz_of(const lbm3_info *I, int i, int j, int k, float *x, float *y, float *z)
{
    *x = (float)i * I->res + I->ox;
    *y = (float)j * I->res + I->oy;
    *z = (float)k * I->res + I->oz;
}

/* radial profile max(0, 1 - r^2 / radius^2) of a pipe case's hooks
   (poiseuilleFlowMain.cu:54-55,66-72); radius <= 0: 1 */
float orc_profile_factor(const lbm3_profile *P, float x, float y, float z)
{
    if (!(P->radius > 0.f)) return 1.f;
    const float dx = x - P->origin[0], dy = y - P->origin[1], dz = z - P->origin[2];
    float r2;
    if (P->axis == LBM3_AXIS_X) r2 = dy * dy + dz * dz;
    else if (P->axis == LBM3_AXIS_Y) r2 = dx * dx + dz * dz;
    else r2 = dx * dx + dy * dy;
    const float v = 1.f - r2 / (P->radius * P->radius);
    return 0.f < v ? v : 0.f;
}

/* the refinement cylinder of poiseuilleFlowMain.cu:38-45 (r^2 < radius^2) */
static int in_cylinder(const lbm3_profile *P, float x, float y, float z)
{
    if (!(P->radius > 0.f)) return 1;
    const float dx = x - P->origin[0], dy = y - P->origin[1], dz = z - P->origin[2];
    float r2;
    if (P->axis == LBM3_AXIS_X) r2 = dy * dy + dz * dz;
    else if (P->axis == LBM3_AXIS_Y) r2 = dx * dx + dz * dz;
    else r2 = dx * dx + dy * dy;
    return r2 < P->radius * P->radius;
}

static inline int face_class(int idx, int count)
{
    if (idx == 0) return LBM3_FACE_LOW;
    if (idx == count - 1) return LBM3_FACE_HIGH;
    return LBM3_FACE_MID;
}

/* getOpenBC hook: first matching rule wins */
lbm3_bc orc_hook_open_bc(const lbm3_case *C, const lbm3_info *I, int i, int j, int k)
{
    const int level = (int)I->gridID;
    const int fc[3] = {face_class(i, I->cellCountX), face_class(j, I->cellCountY), face_class(k, I->cellCountZ)};
    for (int r = 0; r < C->nOpenRules; ++r) {
        const lbm3_open_rule *R = &C->openRules[r];
        if (R->levelMask && !(R->levelMask & (1 << level))) continue;
        if ((R->faceMask[0] & fc[0]) && (R->faceMask[1] & fc[1]) && (R->faceMask[2] & fc[2])) {
            lbm3_bc b = R->bc;
            if (R->profile.radius > 0.f) {
                float x, y, z;
                xyz_of(I, i, j, k, &x, &y, &z);
                const float s = orc_profile_factor(&R->profile, x, y, z);
                b.ux *= s; b.uy *= s; b.uz *= s;
            }
            return b;
        }
    }
    return C->openDefault;
}

/* getRefinementModifier hook */
static void orc_hook_refine(const lbm3_case *C, const lbm3_info *I, int i, int j, int k, uint8_t *marker)
{
    const int level = (int)I->gridID;
    float x, y, z;
    xyz_of(I, i, j, k, &x, &y, &z);
    int haveSet = 0, inSet = 0;
    for (int r = 0; r < C->nRefineRules; ++r) {
        const lbm3_refine_rule *R = &C->refineRules[r];
        if (R->level != level) continue;
        const int in = x > R->xMin && x < R->xMax && y > R->yMin && y < R->yMax && z > R->zMin && z < R->zMax &&
                       in_cylinder(&R->cylinder, x, y, z);
        if (R->mode == LBM3_REFINE_SET) { haveSet = 1; inSet |= in; }
    }
    if (haveSet) *marker = (uint8_t)inSet;
    for (int r = 0; r < C->nRefineRules; ++r) {
        const lbm3_refine_rule *R = &C->refineRules[r];
        if (R->level != level) continue;
        const int in = x > R->xMin && x < R->xMax && y > R->yMin && y < R->yMax && z > R->zMin && z < R->zMax &&
                       in_cylinder(&R->cylinder, x, y, z);
        if (R->mode == LBM3_REFINE_ADD && in) *marker = 1;
        if (R->mode == LBM3_REFINE_REMOVE && in) *marker = 0;
    }
}

/* ------------------------------------------------------------------------
 * STL: readSTL STLFunctions.h:317-462, sortSTLToBins :5-128, checkSTLEdges :130-315
 * ---------------------------------------------------------------------- */

static int stl_bins(orc_stl *S, float RES)
{
    const lbm3_bounds *B = &S->Bounds;
    const int T = S->triangleCount;
    const float totalVolume = (B->xMax - B->xMin) * (B->yMax - B->yMin) * (B->zMax - B->zMin);
    const float targetBinVolume = totalVolume / ((float)T / (float)10);
    S->binSize = powf(targetBinVolume, 1.f / 3.f);
    const float bs = S->binSize;
    S->oxBin = B->xMin - bs - 2.f * RES;
    S->oyBin = B->yMin - bs - 2.f * RES;
    S->ozBin = B->zMin - bs - 2.f * RES;
    S->binCountX = (int)((B->xMax - B->xMin + 2.f * bs + 4.f * RES) / bs) + 1;
    S->binCountY = (int)((B->yMax - B->yMin + 2.f * bs + 4.f * RES) / bs) + 1;
    S->binCountZ = (int)((B->zMax - B->zMin + 2.f * bs + 4.f * RES) / bs) + 1;
    const int nb = S->binCountX * S->binCountY * S->binCountZ;
    const int bxy = S->binCountX * S->binCountY;
    S->firstInBin = NEWI(nb + 1);
    const float tol = bs * 0.2f + RES * 3.f;
    for (int pass = 0; pass < 2; ++pass) {
        int *cursor = pass ? NEWI(nb) : NULL;
        for (int t = 0; t < T; ++t) {
            const float ax = S->ax[t] - S->oxBin, ay = S->ay[t] - S->oyBin, az = S->az[t] - S->ozBin;
            const float bx = S->bx[t] - S->oxBin, by = S->by[t] - S->oyBin, bz = S->bz[t] - S->ozBin;
            const float cx = S->cx[t] - S->oxBin, cy = S->cy[t] - S->oyBin, cz = S->cz[t] - S->ozBin;
            const float xMin = fminf(ax, fminf(bx, cx)) - tol, yMin = fminf(ay, fminf(by, cy)) - tol;
            const float zMin = fminf(az, fminf(bz, cz)) - tol;
            const float xMax = fmaxf(ax, fmaxf(bx, cx)) + tol, yMax = fmaxf(ay, fmaxf(by, cy)) + tol;
            const float zMax = fmaxf(az, fmaxf(bz, cz)) + tol;
            const int i0 = imax(0, (int)(xMin / bs)), j0 = imax(0, (int)(yMin / bs)), k0 = imax(0, (int)(zMin / bs));
            const int i1 = imin(S->binCountX, (int)(xMax / bs) + 1);
            const int j1 = imin(S->binCountY, (int)(yMax / bs) + 1);
            const int k1 = imin(S->binCountZ, (int)(zMax / bs) + 1);
            for (int kb = k0; kb < k1; ++kb)
                for (int jb = j0; jb < j1; ++jb)
                    for (int ib = i0; ib < i1; ++ib) {
                        const int bin = bxy * kb + S->binCountX * jb + ib;
                        if (!pass) S->firstInBin[bin]++;
                        else S->binArray[S->firstInBin[bin] + cursor[bin]++] = t;
                    }
        }
        if (!pass) {
            long long run = 0;
            for (int b = 0; b <= nb; ++b) { const int c = S->firstInBin[b]; S->firstInBin[b] = (int)run; run += c; if (run > 0x7fffffffLL) return -1; }
            S->binArray = NEWI(run);
        } else free(cursor);
    }
    return 0;
}

static void stl_edges(orc_stl *S)
{
    const int T = S->triangleCount;
    const int bxy = S->binCountX * S->binCountY;
    S->edgeCount = NEWI(3 * T);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < T; ++t) {
        const float t1[9] = {S->ax[t], S->ay[t], S->az[t], S->bx[t], S->by[t], S->bz[t], S->cx[t], S->cy[t], S->cz[t]};
        for (int side = 0; side < 2; ++side) {
            /* side 0: bin of vertex B, edges AB and BC; side 1: bin of vertex A, edge CA */
            const float *p = side == 0 ? t1 + 3 : t1;
            const int ib = (int)((p[0] - S->oxBin) / S->binSize);
            const int jb = (int)((p[1] - S->oyBin) / S->binSize);
            const int kb = (int)((p[2] - S->ozBin) / S->binSize);
            const int bin = bxy * kb + S->binCountX * jb + ib;
            for (int r = S->firstInBin[bin]; r < S->firstInBin[bin + 1]; ++r) {
                const int u = S->binArray[r];
                const float t2[9] = {S->ax[u], S->ay[u], S->az[u], S->bx[u], S->by[u], S->bz[u], S->cx[u], S->cy[u], S->cz[u]};
                for (int e1 = side == 0 ? 0 : 2; e1 < (side == 0 ? 2 : 3); ++e1)
                    for (int e2 = 0; e2 < 3; ++e2) {
                        const float *a1 = t1 + 3 * e1, *b1 = t1 + 3 * ((e1 + 1) % 3);
                        const float *a2 = t2 + 3 * e2, *b2 = t2 + 3 * ((e2 + 1) % 3);
                        if (a1[0] == a2[0] && a1[1] == a2[1] && a1[2] == a2[2] && b1[0] == b2[0] && b1[1] == b2[1] && b1[2] == b2[2])
                            S->edgeCount[3 * t + e1]++;
                        if (a1[0] == b2[0] && a1[1] == b2[1] && a1[2] == b2[2] && b1[0] == a2[0] && b1[1] == a2[1] && b1[2] == a2[2])
                            S->edgeCount[3 * t + e1]++;
                    }
            }
        }
    }
}

int orc_stl_from_triangles(orc_stl *S, const float *tri, int n, float RES)
{
    memset(S, 0, sizeof(*S));
    float *a[9];
    for (int c = 0; c < 9; ++c) a[c] = NEWF(n);
    lbm3_bounds B = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    int w = 0;
    for (int t = 0; t < n; ++t) {
        const float *v = tri + 9 * t;
        const float ax = v[0], ay = v[1], az = v[2], bx = v[3], by = v[4], bz = v[5], cx = v[6], cy = v[7], cz = v[8];
        if ((ax == bx && ay == by && az == bz) || (cx == bx && cy == by && cz == bz) || (cx == ax && cy == ay && cz == az)) {
            S->excluded++;
            continue;
        }
        for (int c = 0; c < 9; ++c) a[c][w] = v[c];
        w++;
        B.xMin = fminf(B.xMin, fmin3(ax, bx, cx)); B.yMin = fminf(B.yMin, fmin3(ay, by, cy)); B.zMin = fminf(B.zMin, fmin3(az, bz, cz));
        B.xMax = fmaxf(B.xMax, fmax3(ax, bx, cx)); B.yMax = fmaxf(B.yMax, fmax3(ay, by, cy)); B.zMax = fmaxf(B.zMax, fmax3(az, bz, cz));
        const float rx = sqrtf(fmax3(ay * ay + az * az, by * by + bz * bz, cy * cy + cz * cz));
        const float ry = sqrtf(fmax3(ax * ax + az * az, bx * bx + bz * bz, cx * cx + cz * cz));
        const float rz = sqrtf(fmax3(ax * ax + ay * ay, bx * bx + by * by, cx * cx + cy * cy));
        B.rxMax = fmaxf(B.rxMax, rx); B.ryMax = fmaxf(B.ryMax, ry); B.rzMax = fmaxf(B.rzMax, rz);
    }
    S->triangleCount = w;
    S->ax = a[0]; S->ay = a[1]; S->az = a[2]; S->bx = a[3]; S->by = a[4]; S->bz = a[5];
    S->cx = a[6]; S->cy = a[7]; S->cz = a[8];
    S->Bounds = B;
    if (w == 0) return -1;
    if (stl_bins(S, RES)) return -1;   /* more than 2^31 bin entries */
    stl_edges(S);
    return 0;
}

int orc_stl_read(orc_stl *S, const char *path, float RES)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    char header[80];
    uint32_t n32 = 0;
    if (fread(header, 1, 80, fp) != 80 || fread(&n32, 4, 1, fp) != 1) { fclose(fp); return -2; }
    float *tri = (float *)malloc((size_t)n32 * 9 * sizeof(float) + 4);
    for (uint32_t t = 0; t < n32; ++t) {
        float rec[12]; uint16_t attr;
        if (fread(rec, 4, 12, fp) != 12 || fread(&attr, 2, 1, fp) != 1) { fclose(fp); free(tri); return -3; }
        memcpy(tri + 9 * t, rec + 3, 9 * sizeof(float));
    }
    fclose(fp);
    const int rc = orc_stl_from_triangles(S, tri, (int)n32, RES);
    free(tri);
    return rc;
}

/* shared-edge counts of an STL body (checkSTLEdges, STLFunctions.h:130-287):
 * out[3 t + e] for edge e (AB, BC, CA) of the t-th kept triangle; returns the
 * number of kept triangles (degenerate ones dropped as readSTL does), or a
 * negative code; out may be NULL to ask for the count */
int orc_stl_edge_counts(const float *tri9, int n, float resGlobal, int *out, int capacity)
{
    orc_stl S;
    if (orc_stl_from_triangles(&S, tri9, n, resGlobal) != 0) { orc_stl_free(&S); return -1; }
    const int T = S.triangleCount;
    if (out) {
        if (capacity < 3 * T) { orc_stl_free(&S); return -2; }
        memcpy(out, S.edgeCount, sizeof(int) * 3 * (size_t)T);
    }
    orc_stl_free(&S);
    return T;
}

void orc_stl_free(orc_stl *S)
{
    free(S->ax); free(S->ay); free(S->az); free(S->bx); free(S->by); free(S->bz);
    free(S->cx); free(S->cy); free(S->cz); free(S->binArray); free(S->firstInBin); free(S->edgeCount);
    memset(S, 0, sizeof(*S));
}

/* ------------------------------------------------------------------------
 * Voxelizer: getRayHitYesNo voxelizerFunctions.h:5-65, voxelizeSTL :67-535,
 * sumRayMaps :537-690, initializeVoxelizers :692-728
 * ---------------------------------------------------------------------- */

static int ray_hit(int i, int j, long long wab, long long wbc, long long wca,
                   int axI, int ayI, int bxI, int byI, int cxI, int cyI,
                   long long abx, long long bcx, long long cax)
{
    if (wab > 0 && wbc > 0 && wca > 0) return 1;
    if (wab < 0 || wbc < 0 || wca < 0) return 0;
    const int rx = i * 1000, ry = j * 1000;
    const long long ax0 = axI - rx, ay0 = ayI - ry, bx0 = bxI - rx, by0 = byI - ry, cx0 = cxI - rx, cy0 = cyI - ry;
    if (wab == 0) {
        if (abx == 0) return cx0 < 0;
        if (bx0 > 0 && cy0 * bx0 > by0 * cx0) return 1;
        if (bx0 < 0 && cy0 * bx0 < by0 * cx0) return 1;
    }
    if (wbc == 0) {
        if (bcx == 0) return ax0 < 0;
        if (cx0 > 0 && ay0 * cx0 > cy0 * ax0) return 1;
        if (cx0 < 0 && ay0 * cx0 < cy0 * ax0) return 1;
    }
    if (wca == 0) {
        if (cax == 0) return bx0 < 0;
        if (ax0 > 0 && by0 * ax0 > ay0 * bx0) return 1;
        if (ax0 < 0 && by0 * ax0 < ay0 * bx0) return 1;
    }
    return 0;
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static void voxelize(orc_raymap *R, const orc_stl *S, const lbm3_info *I)
{
    const int X = I->cellCountX, Y = I->cellCountY;
    R->nRays = X * Y;
    R->hitScan = NEWL(R->nRays + 1);
    long long *cnt = NEWL(R->nRays + 1);
    for (int pass = 0; pass < 2; ++pass) {
        for (int t = 0; t < S->triangleCount; ++t) {
            const float ax = S->ax[t] - I->ox, ay = S->ay[t] - I->oy, az = S->az[t] - I->oz;
            const float bx = S->bx[t] - I->ox, by = S->by[t] - I->oy, bz = S->bz[t] - I->oz;
            const float cx = S->cx[t] - I->ox, cy = S->cy[t] - I->oy, cz = S->cz[t] - I->oz;
            const float scale = 500.0f / I->res;
            const int axI = (int)roundf(ax * scale) * 2 + 1, ayI = (int)roundf(ay * scale) * 2 + 1;
            const int bxI = (int)roundf(bx * scale) * 2 + 1, byI = (int)roundf(by * scale) * 2 + 1;
            const int cxI = (int)roundf(cx * scale) * 2 + 1, cyI = (int)roundf(cy * scale) * 2 + 1;
            const int xMinI = imax(0, imin(imin(axI, bxI), imin(cxI, (X - 1) * 1000)));
            const int xMaxI = imin((X - 1) * 1000, imax(imax(axI, bxI), imax(cxI, 0)));
            const int yMinI = imax(0, imin(imin(ayI, byI), imin(cyI, (Y - 1) * 1000)));
            const int yMaxI = imin((Y - 1) * 1000, imax(imax(ayI, byI), imax(cyI, 0)));
            const int i0 = (xMinI + 999) / 1000, i1 = xMaxI / 1000 + 1;
            const int j0 = (yMinI + 999) / 1000, j1 = yMaxI / 1000 + 1;
            const long long xL0 = 1000LL * i0, yL0 = 1000LL * j0;
            const long long axL = axI - xL0, ayL = ayI - yL0, bxL = bxI - xL0, byL = byI - yL0, cxL = cxI - xL0, cyL = cyI - yL0;
            const long long abx = bxL - axL, aby = byL - ayL, bcx = cxL - bxL, bcy = cyL - byL, cax = axL - cxL, cay = ayL - cyL;
            const long long area = abx * bcy - aby * bcx;
            if (area == 0) continue;
            const long long qz = area < 0 ? -1 : 1;
            const long long wab0 = qz * (abx * (-byL) - aby * (-bxL));
            const long long wbc0 = qz * (bcx * (-cyL) - bcy * (-cxL));
            const long long wca0 = qz * (cax * (-ayL) - cay * (-axL));
            const long long dab_i = -qz * aby * 1000, dab_j = qz * abx * 1000;
            const long long dbc_i = -qz * bcy * 1000, dbc_j = qz * bcx * 1000;
            const long long dca_i = -qz * cay * 1000, dca_j = qz * cax * 1000;
            const float v1x = bx - ax, v1y = by - ay, v1z = bz - az, v2x = cx - ax, v2y = cy - ay, v2z = cz - az;
            const float nx = v1y * v2z - v1z * v2y, ny = v1z * v2x - v1x * v2z, nz = v1x * v2y - v1y * v2x;
            const float maxZ = fmaxf(fmaxf(az, bz), cz), minZ = fminf(fminf(az, bz), cz);
            const float midZ = 0.5f * (maxZ + minZ);
            if (i1 - i0 <= 0 || j1 - j0 <= 0) continue;
            for (int j = j0; j < j1; ++j)
                for (int i = i0; i < i1; ++i) {
                    const long long wab = wab0 + (long long)(i - i0) * dab_i + (long long)(j - j0) * dab_j;
                    const long long wbc = wbc0 + (long long)(i - i0) * dbc_i + (long long)(j - j0) * dbc_j;
                    const long long wca = wca0 + (long long)(i - i0) * dca_i + (long long)(j - j0) * dca_j;
                    if (!ray_hit(i, j, wab, wbc, wca, axI, ayI, bxI, byI, cxI, cyI, abx, bcx, cax)) continue;
                    const int ray = j * X + i;
                    if (!pass) { cnt[ray]++; continue; }
                    const float rayX = (float)i * I->res, rayY = (float)j * I->res;
                    float rayZ;
                    if (nz != 0.0f) {
                        rayZ = az - (nx * (rayX - ax) + ny * (rayY - ay)) / nz;
                        if (rayZ > maxZ) rayZ = maxZ;
                        else if (rayZ < minZ) rayZ = minZ;
                    } else rayZ = midZ;
                    R->rayMap[R->hitScan[ray] + cnt[ray]++] = (int)ceilf(rayZ / I->res);
                }
        }
        if (!pass) {
            long long run = 0;
            for (int r = 0; r < R->nRays; ++r) { R->hitScan[r] = run; run += cnt[r]; cnt[r] = 0; }
            R->hitScan[R->nRays] = run;
            R->totalHitCount = run;
            R->rayMap = NEWI(run);
        }
    }
    for (int r = 0; r < R->nRays; ++r)
        qsort(R->rayMap + R->hitScan[r], (size_t)(R->hitScan[r + 1] - R->hitScan[r]), sizeof(int), cmp_int);
    free(cnt);
}

static void raymap_union(orc_raymap *sum, const orc_raymap *bonus)
{
    const int n = sum->nRays;
    long long *scan = NEWL(n + 1);
    int *out = NULL;
    for (int pass = 0; pass < 2; ++pass) {
        for (int r = 0; r < n; ++r) {
            long long si = sum->hitScan[r], se = sum->hitScan[r + 1];
            long long bi = bonus->hitScan[r], be = bonus->hitScan[r + 1];
            long long w = pass ? scan[r] : 0;
            int active = 0;
            long long lastStart = 0, lastEnd = 0;
            while (si < se || bi < be) {
                long long cs, ce;
                if (bi >= be || (si < se && sum->rayMap[si] <= bonus->rayMap[bi])) {
                    cs = sum->rayMap[si]; ce = sum->rayMap[si + 1]; si += 2;
                } else {
                    cs = bonus->rayMap[bi]; ce = bonus->rayMap[bi + 1]; bi += 2;
                }
                if (!active) { lastStart = cs; lastEnd = ce; active = 1; }
                else if (cs <= lastEnd) { if (ce > lastEnd) lastEnd = ce; }
                else {
                    if (pass) { out[w] = (int)lastStart; out[w + 1] = (int)lastEnd; }
                    w += 2; lastStart = cs; lastEnd = ce;
                }
            }
            if (active) { if (pass) { out[w] = (int)lastStart; out[w + 1] = (int)lastEnd; } w += 2; }
            if (!pass) scan[r] = w;
        }
        if (!pass) {
            long long run = 0;
            for (int r = 0; r <= n; ++r) { const long long c = r < n ? scan[r] : 0; scan[r] = run; run += c; }
            out = NEWI(scan[n]);
        }
    }
    free(sum->hitScan); free(sum->rayMap);
    sum->hitScan = scan; sum->rayMap = out; sum->totalHitCount = scan[n];
}

typedef struct { lbm3_info Info; int nMaps; orc_raymap maps[LBM3_MAX_WALLS]; orc_raymap total; } voxelizer_t;

static void raymap_copy(orc_raymap *d, const orc_raymap *s)
{
    *d = *s;
    d->hitScan = NEWL(s->nRays + 1);
    memcpy(d->hitScan, s->hitScan, sizeof(long long) * (size_t)(s->nRays + 1));
    d->rayMap = NEWI(s->totalHitCount);
    memcpy(d->rayMap, s->rayMap, sizeof(int) * (size_t)s->totalHitCount);
}

/* ------------------------------------------------------------------------
 * Grid builder state (GridBuilderStruct, types.h:186-196)
 * ---------------------------------------------------------------------- */
typedef struct {
    lbm3_info Info;
    int *i, *j, *k, *jPlus, *kPlus, *jMinus, *kMinus;
    uint8_t *isGeo;
    int *parentMap;
    int *wallID, *wallIDFull, *wallList, W;
    uint8_t *linkExist, *linkPierce; /* [27][W] */
    float *linkLen;                  /* [27][W] */
    uint8_t *keep, *wall, *refine, *deep, *f2c, *c2f, *parentIf, *ifOverlap, *parentF2C, *needCoarse;
    lbm3_info Sk; uint8_t *skKeep;
} gb_t;

/* initializeGridInfo, gridBuilderFunctions.h:9-64 */
static void init_info(const lbm3_case *C, gb_t *G, int levels, lbm3_bounds B)
{
    for (int L = 0; L < levels; ++L) {
        lbm3_info *I = &G[L].Info;
        memset(I, 0, sizeof(*I));
        I->useRotors = 1;
        I->nu = 1.f; I->dtPhys = 1.f; I->res = 1.f;
        if (L == 0) {
            I->res = C->resGlobal;
            lbm3_info *S = &G[0].Sk;
            memset(S, 0, sizeof(*S));
            S->gridID = -1;
            S->res = I->res * 2.f;
            S->cellCountX = (int)((B.xMax - B.xMin) / S->res);
            S->cellCountY = (int)((B.yMax - B.yMin) / S->res);
            S->cellCountZ = (int)((B.zMax - B.zMin) / S->res);
            S->cellCount = S->cellCountX * S->cellCountY * S->cellCountZ;
            S->ox = B.xMin + ((B.xMax - B.xMin) - ((float)S->cellCountX * S->res) + S->res) / 2.0f;
            S->oy = B.yMin + ((B.yMax - B.yMin) - ((float)S->cellCountY * S->res) + S->res) / 2.0f;
            S->oz = B.zMin + ((B.zMax - B.zMin) - ((float)S->cellCountZ * S->res) + S->res) / 2.0f;
            I->cellCountX = S->cellCountX * 2; I->cellCountY = S->cellCountY * 2; I->cellCountZ = S->cellCountZ * 2;
            I->ox = S->ox - I->res * 0.5f; I->oy = S->oy - I->res * 0.5f; I->oz = S->oz - I->res * 0.5f;
            I->dtPhys = C->dtPhysGlobal;
        } else {
            const lbm3_info *P = &G[L - 1].Info;
            I->gridID = (float)L;
            I->res = P->res * 0.5f;
            I->cellCountX = P->cellCountX * 2; I->cellCountY = P->cellCountY * 2; I->cellCountZ = P->cellCountZ * 2;
            I->ox = P->ox - I->res * 0.5f; I->oy = P->oy - I->res * 0.5f; I->oz = P->oz - I->res * 0.5f;
            I->dtPhys = P->dtPhys * 0.5f;
        }
        I->nu = (I->dtPhys * C->nuPhys) / ((I->res / 1000.f) * (I->res / 1000.f));
    }
}

/* markWallCells, markerFunctions.h:18-52 */
static void mark_wall(uint8_t *m, const orc_raymap *R, const gb_t *G)
{
    const int N = G->Info.cellCount, X = G->Info.cellCountX;
    for (int c = 0; c < N; ++c) {
        m[c] = 0;
        const int ray = X * G->j[c] + G->i[c];
        for (long long s = R->hitScan[ray]; s < R->hitScan[ray + 1]; s += 2) {
            if (R->rayMap[s] > G->k[c]) break;
            if (R->rayMap[s + 1] > G->k[c]) { m[c] = 1; break; }
        }
    }
}

/* markSingleFinerFluid (skeleton and builder versions), markerFunctions.h:54-170 */
static int finer_fluid(const orc_raymap *R, int Xcoarse, int ic, int jc, int kc, int ds)
{
    const int i0 = ic * ds, j0 = jc * ds, k0 = kc * ds, kLast = k0 + ds - 1;
    for (int ja = 0; ja < ds; ++ja)
        for (int ia = 0; ia < ds; ++ia) {
            const int ray = (Xcoarse * ds) * (j0 + ja) + (i0 + ia);
            const long long s0 = R->hitScan[ray], e0 = R->hitScan[ray + 1];
            if (s0 == e0) return 1;
            for (long long s = s0; s < e0; s += 2) {
                const int kEnd = R->rayMap[s + 1];
                if (kEnd < k0) { if (s + 2 < e0) continue; else return 1; }
                else if (kEnd >= k0 && kEnd <= kLast) return 1;
                const int kStart = R->rayMap[s];
                if (kStart <= k0) break;
                else return 1;
            }
        }
    return 0;
}

/* markSingleFinerWall, markerFunctions.h:172-234 */
static int finer_wall(const orc_raymap *R, int Xcoarse, int ic, int jc, int kc, int ds)
{
    const int i0 = ic * ds, j0 = jc * ds, k0 = kc * ds, kLast = k0 + ds - 1;
    for (int ja = 0; ja < ds; ++ja)
        for (int ia = 0; ia < ds; ++ia) {
            const int ray = (Xcoarse * ds) * (j0 + ja) + (i0 + ia);
            for (long long s = R->hitScan[ray]; s < R->hitScan[ray + 1]; s += 2) {
                const int kStart = R->rayMap[s];
                if (kStart > kLast) break;
                else if (kStart >= k0) return 1;
                const int kEnd = R->rayMap[s + 1];
                if (kEnd <= k0) continue;
                return 1;
            }
        }
    return 0;
}

/* spreadMarkers on the dense skeleton box, markerFunctions.h:290-330 */
static void spread_skeleton(uint8_t *t, const uint8_t *s, const lbm3_info *I)
{
    const int X = I->cellCountX, Y = I->cellCountY, Z = I->cellCountZ, N = I->cellCount;
    memcpy(t, s, (size_t)N);
    for (int c = 0; c < N; ++c) {
        if (s[c]) continue;
        const int k = c / (X * Y), rem = c % (X * Y), j = rem / X, i = rem % X;
        for (int dk = -1; dk <= 1 && !t[c]; ++dk) {
            const int kn = k + dk;
            if (kn < 0 || kn >= Z) continue;
            for (int dj = -1; dj <= 1 && !t[c]; ++dj) {
                const int jn = j + dj;
                if (jn < 0 || jn >= Y) continue;
                for (int di = -1; di <= 1; ++di) {
                    if (!dk && !dj && !di) continue;
                    const int in = i + di;
                    if (in < 0 || in >= X) continue;
                    if (s[kn * (X * Y) + jn * X + in]) { t[c] = 1; break; }
                }
            }
        }
    }
}

/* spreadMarkers on a builder grid via geometric plus-links, markerFunctions.h:332-399 */
static void spread_grid(uint8_t *t, const uint8_t *s, const gb_t *G)
{
    const int N = G->Info.cellCount;
    memcpy(t, s, (size_t)N);
    for (int c = 0; c < N; ++c) {
        int nb[7];
        nb[0] = (c + 1 < N) ? c + 1 : 0;
        nb[1] = G->jPlus[c];
        nb[2] = (nb[1] + 1 < N) ? nb[1] + 1 : 0;
        nb[3] = G->kPlus[c];
        nb[4] = (nb[3] + 1 < N) ? nb[3] + 1 : 0;
        nb[5] = G->jPlus[nb[3]];
        nb[6] = (nb[5] + 1 < N) ? nb[5] + 1 : 0;
        const uint8_t geo = G->isGeo[c];
        int marker = s[c];
        if (!marker)
            for (int q = 0; q < 7; ++q)
                if (((geo >> q) & 1) && s[nb[q]]) { marker = 1; break; }
        if (marker) {
            t[c] = 1;
            for (int q = 0; q < 7; ++q)
                if ((geo >> q) & 1) t[nb[q]] = 1;
        }
    }
}

/* connectNBRInRows + buildNBRPlus, NBRFunctions.h:44-152 (rows sorted ascending) */
static void connect_rows(int *nbr, const gb_t *G, int axis)
{
    /* axis 1: rows of equal (i,k) ordered by cell index -> jPlus; axis 2: equal (i,j) -> kPlus */
    const int N = G->Info.cellCount, X = G->Info.cellCountX;
    const int R = axis == 1 ? X * G->Info.cellCountZ : X * G->Info.cellCountY;
    int *cnt = NEWI(R + 1), *start = NEWI(R + 1), *map = NEWI(N);
    for (int c = 0; c < N; ++c) cnt[(axis == 1 ? G->k[c] : G->j[c]) * X + G->i[c]]++;
    int run = 0;
    for (int r = 0; r < R; ++r) { start[r] = run; run += cnt[r]; cnt[r] = 0; }
    start[R] = run;
    for (int c = 0; c < N; ++c) { /* ascending c: rows come out sorted */
        const int r = (axis == 1 ? G->k[c] : G->j[c]) * X + G->i[c];
        map[start[r] + cnt[r]++] = c;
    }
    for (int r = 0; r < R; ++r) {
        const int s = start[r], n = start[r + 1] - s;
        if (!n) continue;
        for (int l = 0; l < n - 1; ++l) nbr[map[s + l]] = map[s + l + 1];
        nbr[map[s + n - 1]] = map[s];
    }
    free(cnt); free(start); free(map);
}

/* markGeometricNBRPlus, NBRFunctions.h:154-194 */
static void mark_geometric(gb_t *G)
{
    const int N = G->Info.cellCount;
    for (int c = 0; c < N; ++c) {
        const int ic = G->i[c], jc = G->j[c], kc = G->k[c];
        const int ip = (c + 1 < N) ? c + 1 : 0;
        const int jp = G->jPlus[c], ijp = (jp + 1 < N) ? jp + 1 : 0;
        const int kp = G->kPlus[c], ikp = (kp + 1 < N) ? kp + 1 : 0;
        const int jkp = G->jPlus[kp], ijkp = (jkp + 1 < N) ? jkp + 1 : 0;
        uint8_t b = 0;
#define GEO(bit, n, di, dj, dk) if (G->i[n] == ic + di && G->j[n] == jc + dj && G->k[n] == kc + dk) b |= (uint8_t)(1u << bit)
        GEO(0, ip, 1, 0, 0); GEO(1, jp, 0, 1, 0); GEO(2, ijp, 1, 1, 0); GEO(3, kp, 0, 0, 1);
        GEO(4, ikp, 1, 0, 1); GEO(5, jkp, 0, 1, 1); GEO(6, ijkp, 1, 1, 1);
#undef GEO
        G->isGeo[c] = b;
    }
}

static void build_nbr_plus(gb_t *G)
{
    free(G->jPlus); free(G->kPlus); free(G->isGeo);
    G->jPlus = NEWI(G->Info.cellCount); G->kPlus = NEWI(G->Info.cellCount); G->isGeo = NEWB(G->Info.cellCount);
    connect_rows(G->jPlus, G, 1);
    connect_rows(G->kPlus, G, 2);
    mark_geometric(G);
}

static long long g_sortX, g_sortXY;
static const int *g_si, *g_sj, *g_sk;
static int cmp_kji(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    const long long kx = (long long)g_sk[x] * g_sortXY + (long long)g_sj[x] * g_sortX + g_si[x];
    const long long ky = (long long)g_sk[y] * g_sortXY + (long long)g_sj[y] * g_sortX + g_si[y];
    return (kx > ky) - (kx < ky);
}

/* sortIJK, gridBuilderFunctions.h:204-234: ascending k, j, i */
static void sort_ijk(gb_t *G)
{
    const int N = G->Info.cellCount;
    int *perm = NEWI(N);
    for (int c = 0; c < N; ++c) perm[c] = c;
    g_sortX = G->Info.cellCountX; g_sortXY = (long long)G->Info.cellCountX * G->Info.cellCountY;
    g_si = G->i; g_sj = G->j; g_sk = G->k;
    qsort(perm, (size_t)N, sizeof(int), cmp_kji);
    int *ni = NEWI(N), *nj = NEWI(N), *nk = NEWI(N);
    for (int c = 0; c < N; ++c) { ni[c] = G->i[perm[c]]; nj[c] = G->j[perm[c]]; nk[c] = G->k[perm[c]]; }
    free(G->i); free(G->j); free(G->k); free(perm);
    G->i = ni; G->j = nj; G->k = nk;
}

/* binarySearchIJK, gridBuilderFunctions.h:238-297 (exact k,j,i lookup in a sorted grid) */
static int find_ijk(const gb_t *S, int iw, int jw, int kw)
{
    int start = 0, end = S->Info.cellCount, result = -1;
    while (end > start) {
        const int h = start + (end - start) / 2;
        const int kh = S->k[h], jh = S->j[h], ih = S->i[h];
        if (kh == kw && jh == jw && ih == iw) { result = h; break; }
        if (kh > kw || (kh == kw && jh > jw) || (kh == kw && jh == jw && ih > iw)) end = h;
        else start = h + 1;
    }
    if (result == -1 && start < S->Info.cellCount)
        if (S->k[start] == kw && S->j[start] == jw && S->i[start] == iw) result = start;
    return result;
}

static void parent_map(gb_t *G, const gb_t *P)
{
    for (int c = 0; c < G->Info.cellCount; ++c) G->parentMap[c] = find_ijk(P, G->i[c] / 2, G->j[c] / 2, G->k[c] / 2);
}

/* markKeepCells / markRefinementCells, markerFunctions.h:401-450 */
static void mark_keep_skeleton(gb_t *G0, const voxelizer_t *V, int levels)
{
    const lbm3_info *S = &G0->Sk;
    uint8_t *m = NEWB(S->cellCount);
    for (int c = 0; c < S->cellCount; ++c) {
        const int k = c / (S->cellCountX * S->cellCountY), rem = c % (S->cellCountX * S->cellCountY);
        const int j = rem / S->cellCountX, i = rem % S->cellCountX;
        int any = 0;
        for (int L = 0; L < levels && !any; ++L) any = finer_fluid(&V[L].total, S->cellCountX, i, j, k, 1 << (L + 1));
        m[c] = (uint8_t)any;
    }
    G0->skKeep = NEWB(S->cellCount);
    spread_skeleton(G0->skKeep, m, S);
    free(m);
}

static void mark_keep(gb_t *G, const voxelizer_t *V, int levels)
{
    const int N = G->Info.cellCount, L0 = (int)G->Info.gridID;
    uint8_t *m = NEWB(N);
    for (int c = 0; c < N; ++c) {
        int any = 0;
        for (int L = L0; L < levels && !any; ++L) any = finer_fluid(&V[L].total, G->Info.cellCountX, G->i[c], G->j[c], G->k[c], 1 << (L - L0));
        m[c] = (uint8_t)any;
    }
    spread_grid(G->keep, m, G);
    free(m);
    if (L0 > 0) for (int c = 0; c < N; ++c) G->keep[c] |= G->parentIf[c];
}

static void mark_refinement(const lbm3_case *C, gb_t *G, const voxelizer_t *V, int levels)
{
    const int N = G->Info.cellCount, L0 = (int)G->Info.gridID;
    mark_keep(G, V, levels);
    for (int c = 0; c < N; ++c) {
        int any = 0;
        for (int L = L0; L < levels && !any; ++L) any = finer_wall(&V[L].total, G->Info.cellCountX, G->i[c], G->j[c], G->k[c], 1 << (L - L0));
        G->deep[c] = (uint8_t)any;
    }
    uint8_t *buf = NEWB(N);
    for (int s = 0; s < C->wallRefinementCount; ++s) {
        memcpy(buf, G->deep, (size_t)N);
        spread_grid(G->deep, buf, G);
    }
    for (int c = 0; c < N; ++c) orc_hook_refine(C, &G->Info, G->i[c], G->j[c], G->k[c], &G->deep[c]);
    if (L0 > 0) for (int c = 0; c < N; ++c) G->deep[c] = G->deep[c] && !G->parentIf[c];
    for (int c = 0; c < N; ++c) G->deep[c] = G->deep[c] && G->keep[c];
    spread_grid(G->f2c, G->deep, G);
    for (int c = 0; c < N; ++c) G->f2c[c] = G->f2c[c] && G->keep[c] && !G->deep[c];
    spread_grid(buf, G->f2c, G);
    spread_grid(G->c2f, buf, G);
    for (int c = 0; c < N; ++c) G->c2f[c] = G->c2f[c] && G->keep[c] && !G->deep[c] && !G->f2c[c];
    for (int c = 0; c < N; ++c) G->refine[c] = G->deep[c] || G->f2c[c] || G->c2f[c];
    free(buf);
}

/* buildFinerGridBuilder, gridBuilderFunctions.h:66-202: 8 children per refined parent */
static void children_of(gb_t *F, const uint8_t *mark, int nParents, const int *pi, const int *pj, const int *pk,
                        int skeletonX, int skeletonXY)
{
    int cnt = 0;
    for (int p = 0; p < nParents; ++p) cnt += mark[p] ? 1 : 0;
    F->Info.cellCount = 8 * cnt;
    F->i = NEWI(F->Info.cellCount); F->j = NEWI(F->Info.cellCount); F->k = NEWI(F->Info.cellCount);
    int w = 0;
    for (int p = 0; p < nParents; ++p) {
        if (!mark[p]) continue;
        int ip, jp, kp;
        if (pi) { ip = pi[p]; jp = pj[p]; kp = pk[p]; }
        else { kp = p / skeletonXY; const int r = p % skeletonXY; jp = r / skeletonX; ip = r % skeletonX; }
        for (int ka = 0; ka < 2; ++ka)
            for (int ja = 0; ja < 2; ++ja)
                for (int ia = 0; ia < 2; ++ia) {
                    F->i[w] = 2 * ip + ia; F->j[w] = 2 * jp + ja; F->k[w] = 2 * kp + ka; w++;
                }
    }
}

static void alloc_markers(gb_t *G, int finest, int coarsest)
{
    const int N = G->Info.cellCount;
    G->keep = NEWB(N);
    if (!finest) { G->refine = NEWB(N); G->deep = NEWB(N); G->f2c = NEWB(N); G->c2f = NEWB(N); }
    if (!coarsest) { G->parentMap = NEWI(N); G->parentIf = NEWB(N); G->parentF2C = NEWB(N); G->needCoarse = NEWB(N); }
}

/* buildIJKFull, gridBuilderFunctions.h:299-400 */
static void build_ijk_full(const lbm3_case *C, gb_t *G, const voxelizer_t *V, int levels)
{
    for (int L = 0; L < levels; ++L) {
        gb_t *B = &G[L];
        const int finest = L == levels - 1, coarsest = L == 0;
        if (coarsest) {
            mark_keep_skeleton(B, V, levels);
            children_of(B, B->skKeep, B->Sk.cellCount, NULL, NULL, NULL, B->Sk.cellCountX, B->Sk.cellCountX * B->Sk.cellCountY);
        } else {
            mark_refinement(C, &G[L - 1], V, levels);
            children_of(B, G[L - 1].refine, G[L - 1].Info.cellCount, G[L - 1].i, G[L - 1].j, G[L - 1].k, 0, 0);
        }
        alloc_markers(B, finest, coarsest);
        sort_ijk(B);
        build_nbr_plus(B);
        if (!coarsest) {
            parent_map(B, &G[L - 1]);
            for (int c = 0; c < B->Info.cellCount; ++c) {
                const int p = B->parentMap[c];
                B->parentIf[c] = (uint8_t)((G[L - 1].c2f[p] + G[L - 1].f2c[p]) != 0);
                B->parentF2C[c] = G[L - 1].f2c[p];
            }
        }
    }
}

/* deleteExcessCells, gridBuilderFunctions.h:402-588 */
static void delete_excess(gb_t *G, const voxelizer_t *V, int levels)
{
    for (int L = 0; L < levels; ++L) {
        gb_t *B = &G[L];
        const int finest = L == levels - 1, coarsest = L == 0;
        int N = B->Info.cellCount;
        if (!coarsest) {
            parent_map(B, &G[L - 1]);
            for (int c = 0; c < N; ++c) B->parentIf[c] = B->parentMap[c] >= 0;
        }
        mark_wall(B->keep, &V[L].total, B);
        for (int c = 0; c < N; ++c) B->keep[c] = !B->keep[c];
        if (!coarsest) for (int c = 0; c < N; ++c) B->keep[c] |= B->parentIf[c];
        uint8_t *src = NEWB(N);
        memcpy(src, B->keep, (size_t)N);
        spread_grid(B->keep, src, B);
        if (!finest) for (int c = 0; c < N; ++c) B->keep[c] = B->keep[c] && !B->deep[c];
        if (!coarsest) {
            uint8_t *l3 = NEWB(N);
            memcpy(src, B->parentF2C, (size_t)N);
            spread_grid(l3, src, B); memcpy(src, l3, (size_t)N);
            spread_grid(l3, src, B); memcpy(src, l3, (size_t)N);
            spread_grid(l3, src, B);
            for (int c = 0; c < N; ++c) if (B->parentIf[c] && !l3[c]) B->keep[c] = 0;
            for (int c = 0; c < N; ++c) B->needCoarse[c] = B->parentIf[c] && B->keep[c];
            spread_grid(src, B->parentF2C, B);
            for (int c = 0; c < N; ++c) B->needCoarse[c] = B->needCoarse[c] && !src[c];
            free(l3);
        }
        free(src);
        /* compaction in cell order */
        int w = 0;
        for (int c = 0; c < N; ++c) {
            if (!B->keep[c]) continue;
            B->i[w] = B->i[c]; B->j[w] = B->j[c]; B->k[w] = B->k[c];
            if (!coarsest) { B->parentMap[w] = B->parentMap[c]; B->needCoarse[w] = B->needCoarse[c]; }
            if (!finest) { B->f2c[w] = B->f2c[c]; B->c2f[w] = B->c2f[c]; }
            w++;
        }
        B->Info.cellCount = N = w;
        B->wall = NEWB(N);
        B->ifOverlap = NEWB(N);
        free(B->skKeep); B->skKeep = NULL;
        build_nbr_plus(B);
        B->jMinus = NEWI(N); B->kMinus = NEWI(N);
        for (int c = 0; c < N; ++c) { B->jMinus[B->jPlus[c]] = c; B->kMinus[B->kPlus[c]] = c; }
        int imn = B->i[0], imx = B->i[0], jmn = B->j[0], jmx = B->j[0], kmn = B->k[0], kmx = B->k[0];
        for (int c = 1; c < N; ++c) {
            imn = imin(imn, B->i[c]); imx = imax(imx, B->i[c]); jmn = imin(jmn, B->j[c]); jmx = imax(jmx, B->j[c]);
            kmn = imin(kmn, B->k[c]); kmx = imax(kmx, B->k[c]);
        }
        lbm3_info *I = &B->Info;
        I->Bounds.xMin = I->ox + I->res * (float)imn - 0.5f * I->res;
        I->Bounds.xMax = I->ox + I->res * (float)imx + 0.5f * I->res;
        I->Bounds.yMin = I->oy + I->res * (float)jmn - 0.5f * I->res;
        I->Bounds.yMax = I->oy + I->res * (float)jmx + 0.5f * I->res;
        I->Bounds.zMin = I->oz + I->res * (float)kmn - 0.5f * I->res;
        I->Bounds.zMax = I->oz + I->res * (float)kmx + 0.5f * I->res;
    }
}

/* ------------------------------------------------------------------------
 * IBB links: buildLinkExistenceMarkerArray interpolatedBouncebackFunctions.h:181-479,
 * intersectRayTriangle :397-480, buildLinkLengthArray :482-660
 * ---------------------------------------------------------------------- */

static void link_existence(gb_t *G, int coarsest, int finest)
{
    const lbm3_info *I = &G->Info;
    const int N = I->cellCount, W = G->W;
    G->linkExist = NEWB(27 * W);
    G->linkPierce = NEWB(27 * W);
    for (int w = 0; w < W; ++w) {
        const int c = G->wallList[w];
        const int ic = G->i[c], jc = G->j[c], kc = G->k[c];
        int meIf = 0;
        if (!coarsest && G->parentMap[c] >= 0) meIf = 1;
        if (!finest && (G->f2c[c] || G->c2f[c])) meIf = 1;
        const int ip = (c + 1 >= N) ? 0 : c + 1, im = (c - 1 < 0) ? N - 1 : c - 1;
        const int jp = G->jPlus[c], kp = G->kPlus[c], jm = G->jMinus[c], km = G->kMinus[c];
        const int *JP = G->jPlus, *KP = G->kPlus, *JM = G->jMinus, *KM = G->kMinus;
        int nb[27];
        nb[OOO] = c; nb[POO] = ip; nb[MOO] = im; nb[OOM] = km; nb[OOP] = kp; nb[OMO] = jm; nb[OPO] = jp;
        nb[POM] = KM[ip]; nb[MOP] = KP[im]; nb[POP] = KP[ip]; nb[MOM] = KM[im];
        nb[MMO] = JM[im]; nb[PPO] = JP[ip]; nb[OPM] = KM[jp]; nb[OMP] = KP[jm];
        nb[MPO] = JP[im]; nb[PMO] = JM[ip]; nb[OPP] = JP[kp]; nb[OMM] = KM[jm];
        nb[MPM] = KM[JP[im]]; nb[PMP] = KP[JM[ip]]; nb[MMP] = KP[JM[im]]; nb[PPM] = KM[JP[ip]];
        nb[PMM] = KM[JM[ip]]; nb[MPP] = JP[KP[im]]; nb[MMM] = KM[JM[im]]; nb[PPP] = JP[KP[ip]];
        int geo[27];
        int all = 1, ipc = 1, jpc = 1, kpc = 1, imc = 1, jmc = 1, kmc = 1;
        for (int d = 1; d < 27; ++d) {
            geo[d] = 1;
            const int cx = ORC_CX[d], cy = ORC_CY[d], cz = ORC_CZ[d];
            const int ie = ic + cx, je = jc + cy, ke = kc + cz, n = nb[d];
            if (G->i[n] == ie && G->j[n] == je && G->k[n] == ke) continue;
            geo[d] = 0;
            if (ie < 0 || ie >= I->cellCountX || je < 0 || je >= I->cellCountY || ke < 0 || ke >= I->cellCountZ) continue;
            all = 0;
            if (cx > 0) ipc = 0;
            if (cy > 0) jpc = 0;
            if (cz > 0) kpc = 0;
            if (cx < 0) imc = 0;
            if (cy < 0) jmc = 0;
            if (cz < 0) kmc = 0;
        }
        const int interior = ic != 0 && ic != I->cellCountX - 1 && jc != 0 && jc != I->cellCountY - 1 && kc != 0 && kc != I->cellCountZ - 1;
        if (interior && !all) continue;
        for (int d = 1; d < 27; ++d) {
            const int n = nb[d];
            if (!interior) {
                if (!geo[d]) continue;
            }
            if (!G->wall[n]) continue;
            if (!interior) {
                const int cx = ORC_CX[d], cy = ORC_CY[d], cz = ORC_CZ[d];
                if (cx > 0 && !ipc) continue;
                if (cy > 0 && !jpc) continue;
                if (cz > 0 && !kpc) continue;
                if (cx < 0 && !imc) continue;
                if (cy < 0 && !jmc) continue;
                if (cz < 0 && !kmc) continue;
            }
            int nIf = 0;
            if (!coarsest && G->parentMap[n] >= 0) nIf = 1;
            if (!finest && (G->f2c[n] || G->c2f[n])) nIf = 1;
            G->linkExist[d * W + w] = 1;
            if (meIf || nIf) G->linkPierce[d * W + w] = 1;
        }
    }
}

static int intersect_ray_triangle(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz,
                                  float ex, float ey, float ez, float edgeTol, float *t)
{
    const float e1x = bx - ax, e1y = by - ay, e1z = bz - az;
    const float e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
    const float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const float area2 = sqrtf(nx * nx + ny * ny + nz * nz);
    if (area2 == 0.f) return 0;
    const float px = ey * e2z - ez * e2y, py = ez * e2x - ex * e2z, pz = ex * e2y - ey * e2x;
    const float det = e1x * px + e1y * py + e1z * pz;
    const float detEps = 1e-8f;
    if (det > -detEps && det < detEps) return 0;
    const float invDet = 1.f / det;
    const float lenAB = sqrtf(e1x * e1x + e1y * e1y + e1z * e1z);
    const float lenAC = sqrtf(e2x * e2x + e2y * e2y + e2z * e2z);
    const float bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
    const float lenBC = sqrtf(bcx * bcx + bcy * bcy + bcz * bcz);
    const float uTol = edgeTol * lenAC / area2, vTol = edgeTol * lenAB / area2, wTol = edgeTol * lenBC / area2;
    const float tx = -ax, ty = -ay, tz = -az;
    const float u = (tx * px + ty * py + tz * pz) * invDet;
    if (u < -uTol) return 0;
    const float qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    const float v = (ex * qx + ey * qy + ez * qz) * invDet;
    if (v < -vTol || u + v > 1.f + wTol) return 0;
    *t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
    return 1;
}

static void link_lengths(gb_t *G, const orc_stl *stls, int nStl)
{
    const lbm3_info *I = &G->Info;
    const int W = G->W;
    G->linkLen = NEWF(27 * W);
    for (int x = 0; x < 27 * W; ++x) G->linkLen[x] = 2.f;
    for (int s = 0; s < nStl; ++s) {
        const orc_stl *S = &stls[s];
        const int bxy = S->binCountX * S->binCountY;
        for (int w = 0; w < W; ++w) {
            const int c = G->wallList[w];
            float xc, yc, zc;
            xyz_of(I, G->i[c], G->j[c], G->k[c], &xc, &yc, &zc);
            const int ib = (int)((xc - S->oxBin) / S->binSize);
            const int jb = (int)((yc - S->oyBin) / S->binSize);
            const int kb = (int)((zc - S->ozBin) / S->binSize);
            if (ib < 0 || ib >= S->binCountX || jb < 0 || jb >= S->binCountY || kb < 0 || kb >= S->binCountZ) continue;
            const int bin = bxy * kb + S->binCountX * jb + ib;
            for (int r = S->firstInBin[bin]; r < S->firstInBin[bin + 1]; ++r) {
                const int t = S->binArray[r];
                const float ax = S->ax[t] - xc, ay = S->ay[t] - yc, az = S->az[t] - zc;
                const float bx = S->bx[t] - xc, by = S->by[t] - yc, bz = S->bz[t] - zc;
                const float cx = S->cx[t] - xc, cy = S->cy[t] - yc, cz = S->cz[t] - zc;
                for (int d = 1; d < 27; ++d) {
                    if (!G->linkExist[d * W + w]) continue;
                    if (G->linkPierce[d * W + w]) { G->linkLen[d * W + w] = 0.5f; continue; }
                    const float edgeTol = 0.1f * I->res;
                    const float qTol = 0.2f;
                    float tt;
                    if (!intersect_ray_triangle(ax, ay, az, bx, by, bz, cx, cy, cz, (float)ORC_CX[d], (float)ORC_CY[d],
                                                (float)ORC_CZ[d], edgeTol, &tt)) continue;
                    float q = tt / I->res;
                    float *slot = &G->linkLen[d * W + w];
                    if (q > 0 && q <= 1.f) {
                        q = q < 0.00001f ? 0.00001f : (q > 1.f ? 1.f : q);
                        const float qp = *slot;
                        if (qp < 0.f || q < qp) *slot = q;
                    } else if (q > -qTol && q <= 1.f + qTol) {
                        q = q < 0.00001f ? 0.00001f : (q > 1.f ? 1.f : q);
                        const float qp = *slot;
                        if (qp < 0.f) { if (q < -qp) *slot = -q; }
                        else if (qp > 1.f) *slot = -q;
                    }
                }
            }
        }
    }
    for (int w = 0; w < W; ++w)
        for (int d = 1; d < 27; ++d) {
            if (!G->linkExist[d * W + w]) continue;
            if (G->linkPierce[d * W + w]) { G->linkLen[d * W + w] = 0.5f; continue; }
            const float qp = G->linkLen[d * W + w];
            if (qp > 1.f) G->linkLen[d * W + w] = 0.5f;
            else if (qp < 0.f) G->linkLen[d * W + w] = -qp;
        }
}

/* buildWallMarkers, gridBuilderFunctions.h:590-739 */
static void build_wall_markers(gb_t *G, const voxelizer_t *V, const orc_stl *stls, int nStl, int levels)
{
    for (int L = 0; L < levels; ++L) {
        gb_t *B = &G[L];
        const int finest = L == levels - 1, coarsest = L == 0;
        const int N = B->Info.cellCount;
        mark_wall(B->wall, &V[L].total, B);
        if (!coarsest)
            for (int c = 0; c < N; ++c) if (B->parentMap[c] >= 0) B->wall[c] = G[L - 1].wall[B->parentMap[c]];
        B->wallID = NEWI(N);
        for (int c = 0; c < N; ++c) B->wallID[c] = -1;
        uint8_t *body = NEWB(N), *src = NEWB(N);
        for (int w = 0; w < V[L].nMaps; ++w) {
            mark_wall(body, &V[L].maps[w], B);
            if (!coarsest) {
                const gb_t *P = &G[L - 1];
                uint8_t *pb = NEWB(P->Info.cellCount);
                mark_wall(pb, &V[L - 1].maps[w], P);
                for (int c = 0; c < N; ++c) if (B->parentMap[c] >= 0) body[c] = pb[B->parentMap[c]];
                free(pb);
            }
            memcpy(src, body, (size_t)N);
            spread_grid(body, src, B);
            for (int c = 0; c < N; ++c) body[c] = body[c] && !B->wall[c];
            for (int c = 0; c < N; ++c) {
                if (!body[c]) continue;
                int id = w;
                if (!coarsest && B->parentMap[c] >= 0) {
                    const int pid = G[L - 1].wallIDFull[B->parentMap[c]];
                    if (pid >= 0) id = pid;
                }
                B->wallID[c] = id;
            }
        }
        free(body); free(src);
        int W = 0;
        for (int c = 0; c < N; ++c) W += B->wallID[c] >= 0;
        B->W = W;
        B->wallList = NEWI(W);
        B->wallIDFull = B->wallID;
        B->wallID = NEWI(W);
        W = 0;
        for (int c = 0; c < N; ++c)
            if (B->wallIDFull[c] >= 0) { B->wallList[W] = c; B->wallID[W] = B->wallIDFull[c]; W++; }
        link_existence(B, coarsest, finest);
        link_lengths(B, stls, nStl);
        if (!finest) for (int c = 0; c < N; ++c) B->ifOverlap[c] |= B->f2c[c];
        if (!coarsest)
            for (int c = 0; c < N; ++c) {
                const int p = B->parentMap[c];
                if (p >= 0 && G[L - 1].c2f[p]) B->ifOverlap[c] = 1;
            }
    }
}

/* packWallData, interpolatedBouncebackFunctions.h:662-695 */
static void pack_wall(uint32_t *wallData, uint32_t packed[7], int wallID, int overlap, const uint8_t ex[26], const float len[26])
{
    *wallData = (overlap ? 1u : 0u) | ((uint32_t)wallID << 27);
    for (int p = 0; p < 7; ++p) packed[p] = 0u;
    int li = 0;
    for (int i = 0; i < 26; ++i) {
        if (!ex[i]) continue;
        *wallData |= 1u << (i + 1);
        float q = len[i];
        q = q < 0.f ? 0.f : (q > 1.f ? 1.f : q);
        const uint32_t code = (uint32_t)(q * 254.f + 0.5f);
        packed[li / 4] |= code << (8 * (li % 4));
        li++;
    }
}

/* fillFineToCoarseInterface / fillCoarseToFineInterface, gridBuilderFunctions.h:741-1030 */
static void fill_f2c(orc_interface *IF, const uint8_t *mk, const int *childGlobal, int N)
{
    int n = 0;
    for (int c = 0; c < N; ++c) n += mk[c];
    IF->interfaceCount = n;
    IF->index = NEWI(n); IF->childMap = NEWI(n);
    n = 0;
    for (int c = 0; c < N; ++c) if (mk[c]) { IF->index[n] = c; IF->childMap[n] = childGlobal[c]; n++; }
}

static void nbr_plus(const int *jP, const int *kP, int N, int self, int *o /* self,ip,jp,ijp,kp,ikp,jkp,ijkp */)
{
    o[0] = self;
    o[2] = jP[self]; o[4] = kP[self]; o[6] = jP[o[4]];
    o[1] = self + 1 >= N ? 0 : self + 1;
    o[3] = o[2] + 1 >= N ? 0 : o[2] + 1;
    o[5] = o[4] + 1 >= N ? 0 : o[4] + 1;
    o[7] = o[6] + 1 >= N ? 0 : o[6] + 1;
}

static void fill_c2f(orc_interface *IF, const uint8_t *mk, const int *childGlobal, const gb_t *B, const gb_t *F)
{
    const int N = B->Info.cellCount, NF = F->Info.cellCount;
    uint8_t *full = NEWB(N);
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (int c = 0; c < N; ++c) {
        if (!mk[c]) continue;
        int nb[8];
        nbr_plus(B->jPlus, B->kPlus, N, c, nb);
        int ok = 1;
        for (int q = 1; q < 8 && ok; ++q) ok = mk[nb[q]];
        if (!ok) continue;
        const int ic = B->i[c], jc = B->j[c], kc = B->k[c];
        for (int q = 1; q < 8 && ok; ++q)
            ok = B->i[nb[q]] == ic + off[q][0] && B->j[nb[q]] == jc + off[q][1] && B->k[nb[q]] == kc + off[q][2];
        if (!ok) continue;
        const int top = childGlobal[nb[7]];
        if (top < 0) continue;
        int fn[8];
        fn[7] = top;
        fn[6] = top - 1 < 0 ? NF - 1 : top - 1;
        fn[4] = F->jMinus[fn[6]];
        fn[0] = F->kMinus[fn[4]];
        fn[2] = F->jPlus[fn[0]];
        fn[1] = fn[0] + 1 >= NF ? 0 : fn[0] + 1;
        fn[3] = fn[2] + 1 >= NF ? 0 : fn[2] + 1;
        fn[5] = fn[4] + 1 >= NF ? 0 : fn[4] + 1;
        for (int q = 0; q < 8 && ok; ++q)
            ok = F->i[fn[q]] == 2 * ic + 1 + off[q][0] && F->j[fn[q]] == 2 * jc + 1 + off[q][1] && F->k[fn[q]] == 2 * kc + 1 + off[q][2];
        if (ok) full[c] = 1;
    }
    int n = 0;
    for (int c = 0; c < N; ++c) n += full[c];
    IF->interfaceCount = n;
    IF->index = NEWI(n); IF->childMap = NEWI(n);
    uint8_t *recv = NEWB(NF);
    n = 0;
    for (int c = 0; c < N; ++c) {
        if (!full[c]) continue;
        int nb[8];
        nbr_plus(B->jPlus, B->kPlus, N, c, nb);
        const int top = childGlobal[nb[7]];
        const int jk = top - 1 < 0 ? NF - 1 : top - 1;
        const int self = F->kMinus[F->jMinus[jk]];
        IF->index[n] = c; IF->childMap[n] = self; n++;
        int fnb[8];
        nbr_plus(F->jPlus, F->kPlus, NF, self, fnb);
        for (int q = 0; q < 8; ++q) recv[fnb[q]] = 1;
    }
    int lo = 0;
    for (int c = 0; c < NF; ++c) lo += F->needCoarse[c] && !F->wall[c] && !recv[c];
    IF->leftoverCount = lo;
    IF->leftoverIndex = NEWI(lo); IF->leftoverParent = NEWI(lo);
    IF->leftoverNbrI = NEWI(lo); IF->leftoverNbrJ = NEWI(lo); IF->leftoverNbrK = NEWI(lo);
    lo = 0;
    for (int cf = 0; cf < NF; ++cf) {
        if (!(F->needCoarse[cf] && !F->wall[cf] && !recv[cf])) continue;
        const int cc = F->parentMap[cf];
        IF->leftoverIndex[lo] = cf; IF->leftoverParent[lo] = cc;
        const int ic = B->i[cc], jc = B->j[cc], kc = B->k[cc];
        int nI = cc, nJ = cc, nK = cc;
#define MATCH(x, di, dj, dk) (mk[x] && B->i[x] == ic + (di) && B->j[x] == jc + (dj) && B->k[x] == kc + (dk))
        {
            const int plus = cc + 1 >= N ? 0 : cc + 1, minus = cc - 1 < 0 ? N - 1 : cc - 1;
            if (F->i[cf] % 2 == 1) { if (MATCH(plus, 1, 0, 0)) nI = plus; else if (MATCH(minus, -1, 0, 0)) nI = -minus - 1; }
            else { if (MATCH(minus, -1, 0, 0)) nI = minus; else if (MATCH(plus, 1, 0, 0)) nI = -plus - 1; }
        }
        {
            const int plus = B->jPlus[cc], minus = B->jMinus[cc];
            if (F->j[cf] % 2 == 1) { if (MATCH(plus, 0, 1, 0)) nJ = plus; else if (MATCH(minus, 0, -1, 0)) nJ = -minus - 1; }
            else { if (MATCH(minus, 0, -1, 0)) nJ = minus; else if (MATCH(plus, 0, 1, 0)) nJ = -plus - 1; }
        }
        {
            const int plus = B->kPlus[cc], minus = B->kMinus[cc];
            if (F->k[cf] % 2 == 1) { if (MATCH(plus, 0, 0, 1)) nK = plus; else if (MATCH(minus, 0, 0, -1)) nK = -minus - 1; }
            else { if (MATCH(minus, 0, 0, -1)) nK = minus; else if (MATCH(plus, 0, 0, 1)) nK = -plus - 1; }
        }
#undef MATCH
        IF->leftoverNbrI[lo] = nI; IF->leftoverNbrJ[lo] = nJ; IF->leftoverNbrK[lo] = nK;
        lo++;
    }
    free(full); free(recv);
}

/* gridBuilderToGrid, gridBuilderFunctions.h:1032-1569 */
static int builder_to_grid(orc_sim *Sm, gb_t *G, int levels)
{
    const lbm3_case *C = &Sm->C;
    for (int L = 0; L < levels; ++L) {
        gb_t *B = &G[L];
        orc_grid *g = &Sm->g[L];
        const int finest = L == levels - 1;
        g->Info = B->Info;
        const lbm3_info *I = &g->Info;
        const int N = I->cellCount;
        /* compression, :1044-1126 */
        int *first = NEWI(N);
        int rows = 0;
        for (int c = 0; c < N; ++c) {
            int head = 1;
            if (c > 0) {
                const int jk = B->jPlus[B->kPlus[c]], jkp = B->jPlus[B->kPlus[c - 1]];
                head = B->i[c] != B->i[c - 1] + 1 || B->j[c] != B->j[c - 1] || B->k[c] != B->k[c - 1] ||
                       B->jPlus[c] != B->jPlus[c - 1] + 1 || B->kPlus[c] != B->kPlus[c - 1] + 1 || jk != jkp + 1;
            }
            first[c] = head ? c : first[c - 1];
            rows += head;
        }
        g->nRows = rows;
        g->shifter = NEWI(N);
        g->ri = NEWI(rows); g->rj = NEWI(rows); g->rk = NEWI(rows);
        g->rjp = NEWI(rows); g->rkp = NEWI(rows); g->rjkp = NEWI(rows);
        rows = 0;
        for (int c = 0; c < N; ++c) {
            if (first[c] == c) {
                g->shifter[c] = rows;
                g->ri[rows] = B->i[c]; g->rj[rows] = B->j[c]; g->rk[rows] = B->k[c];
                g->rjp[rows] = B->jPlus[c]; g->rkp[rows] = B->kPlus[c]; g->rjkp[rows] = B->jPlus[B->kPlus[c]];
                rows++;
            } else g->shifter[c] = first[c] - c;
        }
        free(first);
        /* wallMap, :1127-1146 */
        g->wallMap = NEWI(N);
        g->W = B->W;
        g->wallIndex = NEWI(g->W);
        for (int c = 0; c < N; ++c) g->wallMap[c] = B->wall[c] ? -3 : (B->ifOverlap[c] ? -2 : -1);
        for (int w = 0; w < g->W; ++w) { g->wallIndex[w] = B->wallList[w]; g->wallMap[B->wallList[w]] = w; }
        g->wallCount = 0;
        for (int c = 0; c < N; ++c) g->wallCount += B->wall[c];
        /* wallData + packed link codes, :1148-1180 */
        g->wallData = (uint32_t *)calloc((size_t)(g->W > 0 ? g->W : 1), 4);
        g->linkLength = (uint32_t *)calloc((size_t)(7 * (g->W > 0 ? g->W : 1)), 4);
        for (int w = 0; w < g->W; ++w) {
            uint8_t ex[26]; float len[26];
            for (int d = 1; d < 27; ++d) { ex[d - 1] = B->linkExist[d * g->W + w]; len[d - 1] = B->linkLen[d * g->W + w]; }
            uint32_t wd, packed[7];
            pack_wall(&wd, packed, B->wallID[w], B->ifOverlap[B->wallList[w]], ex, len);
            g->wallData[w] = wd;
            for (int p = 0; p < 7; ++p) g->linkLength[p * g->W + w] = packed[p];
        }
        /* interfaces, :1193-1230 */
        if (!finest) {
            gb_t *F = &G[L + 1];
            int *childGlobal = NEWI(N);
            for (int c = 0; c < N; ++c) childGlobal[c] = -1;
            for (int cf = 0; cf < F->Info.cellCount; ++cf) {
                const int cc = F->parentMap[cf];
                if (cc < 0) continue;
                if (F->i[cf] % 2 == 0 && F->j[cf] % 2 == 0 && F->k[cf] % 2 == 0) childGlobal[cc] = cf;
            }
            uint8_t *mk = NEWB(N);
            for (int c = 0; c < N; ++c) mk[c] = B->f2c[c] && !B->wall[c];
            fill_f2c(&g->F2C, mk, childGlobal, N);
            for (int c = 0; c < N; ++c) mk[c] = B->c2f[c] && !B->wall[c];
            fill_c2f(&g->C2F, mk, childGlobal, B, F);
            free(mk); free(childGlobal);
        }
        /* periodic overlap: no force tracking, :1232-1274 */
        const int layers = 1 << L;
        for (int c = 0; c < N; ++c) {
            if (B->wall[c]) continue;
            const int ic = B->i[c], jc = B->j[c], kc = B->k[c];
            int per = 0;
            if (ic < layers || ic >= I->cellCountX - layers) per |= orc_hook_open_bc(C, I, ic < layers ? 0 : I->cellCountX - 1, jc, kc).periodic;
            if (jc < layers || jc >= I->cellCountY - layers) per |= orc_hook_open_bc(C, I, ic, jc < layers ? 0 : I->cellCountY - 1, kc).periodic;
            if (kc < layers || kc >= I->cellCountZ - layers) per |= orc_hook_open_bc(C, I, ic, jc, kc < layers ? 0 : I->cellCountZ - 1).periodic;
            if (!per) continue;
            const int wm = g->wallMap[c];
            if (wm >= 0) g->wallData[wm] |= 1u;
            else if (wm == -1) g->wallMap[c] = -2;
        }
        /* open boundaries, :1276-1484 */
        int *bcid = NEWI(N);
        int bcMax = -1;
        for (int c = 0; c < N; ++c) {
            bcid[c] = -1;
            if (B->wall[c]) continue;
            const int ic = B->i[c], jc = B->j[c], kc = B->k[c];
            if (ic == 0 || ic == I->cellCountX - 1 || jc == 0 || jc == I->cellCountY - 1 || kc == 0 || kc == I->cellCountZ - 1) {
                lbm3_bc bc = orc_hook_open_bc(C, I, ic, jc, kc);
                bcid[c] = bc.periodic ? -1 : bc.openBCID;
            }
            if (bcid[c] > bcMax) bcMax = bcid[c];
        }
        g->nOpen = bcMax + 1;
        const size_t rsz = Sm->isDouble ? sizeof(double) : sizeof(float);
        for (int id = 0; id < g->nOpen; ++id) {
            orc_openbc *O = &g->open[id];
            O->openBCID = id;
            int n = 0;
            for (int c = 0; c < N; ++c) n += bcid[c] == id;
            O->count = n;
            O->index = NEWI(n); O->source = NEWI(n);
            O->dRhoRef = calloc((size_t)(n + 1), rsz); O->uNormalRef = calloc((size_t)(n + 1), rsz);
            O->dRhoCum = calloc((size_t)(n + 1), rsz); O->uNormalCum = calloc((size_t)(n + 1), rsz);
            n = 0;
            for (int c = 0; c < N; ++c) if (bcid[c] == id) O->index[n++] = c;
            O->trackFlowCount = 0;
            for (int x = 0; x < O->count; ++x) {
                const int c = O->index[x];
                const int wm = g->wallMap[c];
                int track = wm != -2;
                if (wm >= 0 && (g->wallData[wm] & 1u)) track = 0;
                O->trackFlowCount += track;
            }
            for (int x = 0; x < O->count; ++x) {
                const int c = O->index[x];
                const int ic = B->i[c], jc = B->j[c], kc = B->k[c];
                int nx = 0, ny = 0, nz = 0;
                if (ic == 0) nx = -1; else if (ic == I->cellCountX - 1) nx = 1;
                if (jc == 0) ny = -1; else if (jc == I->cellCountY - 1) ny = 1;
                if (kc == 0) nz = -1; else if (kc == I->cellCountZ - 1) nz = 1;
                const int faces = (nx != 0) + (ny != 0) + (nz != 0);
                const lbm3_bc bc = orc_hook_open_bc(C, I, ic, jc, kc);
                const int edgeFallback = bc.dirichletU && faces == 2;
                int pref = 0;
                if (edgeFallback) {
                    if (nx == 0) pref = ic <= I->cellCountX - 1 - ic ? 1 : -1;
                    else if (ny == 0) pref = jc <= I->cellCountY - 1 - jc ? 1 : -1;
                    else pref = kc <= I->cellCountZ - 1 - kc ? 1 : -1;
                }
                int src = -1;
                for (int cand = 0; cand < (edgeFallback ? 3 : 1); ++cand) {
                    int is = ic - nx, js = jc - ny, ks = kc - nz;
                    if (cand > 0) {
                        const int sh = cand == 1 ? pref : -pref;
                        if (nx == 0) is += sh; else if (ny == 0) js += sh; else ks += sh;
                    }
                    if (is <= 0 || is >= I->cellCountX - 1 || js <= 0 || js >= I->cellCountY - 1 || ks <= 0 || ks >= I->cellCountZ - 1) continue;
                    const int found = find_ijk(B, is, js, ks);
                    if (found < 0) continue;
                    if (bc.dirichletU && faces > 1 && B->wall[found]) continue;
                    src = found;
                    break;
                }
                O->source[x] = src;
                if (src < 0) {
                    snprintf(Sm->error, sizeof(Sm->error), "no valid open-BC source: grid=%d BCID=%d IJK=(%d,%d,%d)", L, id, ic, jc, kc);
                    free(bcid);
                    return -10;
                }
            }
        }
        /* periodic, :1486-1564 */
        int np = 0;
        for (int c = 0; c < N; ++c) {
            if (B->wall[c]) continue;
            const int ic = B->i[c], jc = B->j[c], kc = B->k[c];
            if (ic == 0 || ic == I->cellCountX - 1 || jc == 0 || jc == I->cellCountY - 1 || kc == 0 || kc == I->cellCountZ - 1)
                if (orc_hook_open_bc(C, I, ic, jc, kc).periodic) { bcid[c] = -2; np++; }
        }
        g->nPeriodic = np;
        g->perIndex = NEWI(np); g->perSource = NEWI(np);
        np = 0;
        const int cpc = (int)lround(pow(2.0, (double)I->gridID));
        for (int c = 0; c < N; ++c) {
            if (bcid[c] != -2) continue;
            int is = B->i[c], js = B->j[c], ks = B->k[c];
            if (is == 0) is = I->cellCountX - 2 * cpc; else if (is == I->cellCountX - 1) is = 2 * cpc - 1;
            if (js == 0) js = I->cellCountY - 2 * cpc; else if (js == I->cellCountY - 1) js = 2 * cpc - 1;
            if (ks == 0) ks = I->cellCountZ - 2 * cpc; else if (ks == I->cellCountZ - 1) ks = 2 * cpc - 1;
            int src = find_ijk(B, is, js, ks);
            if (src >= 0 && B->wall[src]) src = -1;
            g->perIndex[np] = c; g->perSource[np] = src; np++;
            if (src < 0) {
                snprintf(Sm->error, sizeof(Sm->error), "periodic source missing or wall at grid %d", L);
                free(bcid);
                return -11;
            }
        }
        free(bcid);
    }
    return 0;
}

/* buildGrids, gridBuilderFunctions.h:1577-1670 (without rotors) */
int orc_build_grids(orc_sim *Sm, const orc_stl *stls, int nStl, lbm3_bounds Domain)
{
    const lbm3_case *C = &Sm->C;
    const int levels = C->gridLevelCount;
    Sm->levels = levels;
    gb_t *G = (gb_t *)calloc((size_t)levels, sizeof(gb_t));
    voxelizer_t *V = (voxelizer_t *)calloc((size_t)levels, sizeof(voxelizer_t));
    init_info(C, G, levels, Domain);
    for (int L = 0; L < levels; ++L) {
        V[L].Info = G[L].Info;
        V[L].nMaps = nStl;
        for (int s = 0; s < nStl; ++s) {
            if (stls[s].triangleCount == 0) { snprintf(Sm->error, sizeof(Sm->error), "STL with 0 triangles"); return -1; }
            V[L].maps[s].gridID = L;
            voxelize(&V[L].maps[s], &stls[s], &V[L].Info);
        }
        if (nStl > 0) raymap_copy(&V[L].total, &V[L].maps[0]);
        else {
            V[L].total.nRays = V[L].Info.cellCountX * V[L].Info.cellCountY;
            V[L].total.hitScan = NEWL(V[L].total.nRays + 1);
            V[L].total.rayMap = NEWI(1);
        }
        V[L].total.gridID = L;
        for (int s = 1; s < nStl; ++s) raymap_union(&V[L].total, &V[L].maps[s]);
    }
    build_ijk_full(C, G, V, levels);
    delete_excess(G, V, levels);
    build_wall_markers(G, V, stls, nStl, levels);
    const int rc = builder_to_grid(Sm, G, levels);
    if (rc) return rc;
    Sm->fluidUpdatesPerIteration = 0;
    for (int L = 0; L < levels; ++L)
        Sm->fluidUpdatesPerIteration += (long long)(G[L].Info.cellCount - Sm->g[L].wallCount) * (1LL << L);
    /* keep the voxelizer result per level for flag tests */
    for (int L = 0; L < levels; ++L) raymap_copy(&Sm->rayTotal[L], &V[L].total);
    const size_t rsz = Sm->isDouble ? sizeof(double) : sizeof(float);
    for (int L = 0; L < levels; ++L) {
        orc_grid *g = &Sm->g[L];
        g->f = calloc((size_t)27 * (size_t)g->Info.cellCount, rsz);
        g->gWall = calloc((size_t)3 * (size_t)(g->W > 0 ? g->W : 1), rsz);
    }
    Sm->nStatic = nStl;
    /* builder temporaries go out of scope here in the reference; the oracle leaks
       them into process lifetime deliberately only when tests ask for them */
    for (int L = 0; L < levels; ++L) {
        gb_t *B = &G[L];
        free(B->i); free(B->j); free(B->k); free(B->jPlus); free(B->kPlus); free(B->jMinus); free(B->kMinus);
        free(B->isGeo); free(B->parentMap); free(B->wallID); free(B->wallIDFull); free(B->wallList);
        free(B->linkExist); free(B->linkPierce); free(B->linkLen); free(B->keep); free(B->wall); free(B->refine);
        free(B->deep); free(B->f2c); free(B->c2f); free(B->parentIf); free(B->ifOverlap); free(B->parentF2C);
        free(B->needCoarse); free(B->skKeep);
        for (int s = 0; s < V[L].nMaps; ++s) { free(V[L].maps[s].hitScan); free(V[L].maps[s].rayMap); }
        free(V[L].total.hitScan); free(V[L].total.rayMap);
    }
    free(G); free(V);
    return 0;
}

/* ------------------------------------------------------------------------
 * Rotors: buildRotors, rotorFunctions.h:315-585
 * ---------------------------------------------------------------------- */

/* is voxel (i, j, k) inside: a hit pair [kStart, kEnd) of ray (i, j) holds k
   (rotorFunctions.h:384-404; pairs are read only while both entries exist) */
static int rotor_voxel(const orc_raymap *R, int X, int i, int j, int k)
{
    const int ray = X * j + i;
    const long long end = R->hitScan[ray + 1];
    for (long long h = R->hitScan[ray]; h + 1 < end; h += 2) {
        if (R->rayMap[h] > k) break;
        if (R->rayMap[h + 1] > k) return 1;
    }
    return 0;
}

static void rotor_free(orc_rotor *Rt)
{
    free(Rt->rotorMap); free(Rt->index); free(Rt->interp); free(Rt->g);
    memset(Rt, 0, sizeof *Rt);
}

/* one rotor of one level, rotorFunctions.h:322-557 */
static void build_rotor(orc_rotor *Rt, const orc_stl *S, const lbm3_info *G)
{
    lbm3_info V;
    memset(&V, 0, sizeof V);
    V.ox = S->Bounds.xMin; V.oy = S->Bounds.yMin; V.oz = S->Bounds.zMin;
    V.res = G->res * 0.5f;
    V.cellCountX = (int)((S->Bounds.xMax - S->Bounds.xMin) / V.res) + 2;
    V.cellCountY = (int)((S->Bounds.yMax - S->Bounds.yMin) / V.res) + 2;
    V.cellCountZ = (int)((S->Bounds.zMax - S->Bounds.zMin) / V.res) + 2;
    orc_raymap R;
    memset(&R, 0, sizeof R);
    voxelize(&R, S, &V);
    memset(Rt, 0, sizeof *Rt);
    Rt->axis = LBM3_AXIS_NONE;
    Rt->res = V.res;
    Rt->B.xMin = V.ox - 1.5f * V.res;
    Rt->B.yMin = V.oy - 1.5f * V.res;
    Rt->B.zMin = V.oz - 1.5f * V.res;
    Rt->cx = ((V.cellCountX + 2 + 3) / 4) * 4;
    Rt->cy = ((V.cellCountY + 2 + 3) / 4) * 4;
    Rt->cz = ((V.cellCountZ + 2 + 3) / 4) * 4;
    Rt->B.xMax = Rt->B.xMin + Rt->res * (float)Rt->cx;
    Rt->B.yMax = Rt->B.yMin + Rt->res * (float)Rt->cy;
    Rt->B.zMax = Rt->B.zMin + Rt->res * (float)Rt->cz;
    const int bx = Rt->cx / 4, by = Rt->cy / 4, bz = Rt->cz / 4, bxy = bx * by;
    const int VX = V.cellCountX, VY = V.cellCountY;
    Rt->blockCount = bx * by * bz;
    /* 2) blocks holding any solid voxel of their 6^3 neighbourhood, :358-409 */
    uint8_t *mark = NEWB(Rt->blockCount);
    for (int b = 0; b < Rt->blockCount; ++b) {
        const int kb = b / bxy, jb = (b % bxy) / bx, ib = (b % bxy) % bx;
        const int i0 = imax(0, ib * 4 - 2), i1 = imin(VX, ib * 4 + 4);
        const int j0 = imax(0, jb * 4 - 2), j1 = imin(VY, jb * 4 + 4);
        for (int j = j0; j < j1 && !mark[b]; ++j)
            for (int i = i0; i < i1 && !mark[b]; ++i)
                for (int k = kb * 4 - 2; k < kb * 4 + 4; ++k)
                    if (rotor_voxel(&R, VX, i, j, k)) { mark[b] = 1; break; }
    }
    /* 3) compact index of the marked blocks, :411-435 */
    Rt->rotorMap = NEWI(Rt->blockCount);
    int nz = 0;
    for (int b = 0; b < Rt->blockCount; ++b) nz += mark[b];
    Rt->nonZero = nz;
    Rt->index = NEWI(nz);
    for (int b = 0, run = 0; b < Rt->blockCount; ++b) {
        if (mark[b]) { Rt->rotorMap[b] = run; Rt->index[run++] = b; }
        else Rt->rotorMap[b] = -1;
    }
    /* 4) corner counts of every interpolation cell, :437-547 */
    Rt->interp = (uint32_t *)calloc((size_t)(nz > 0 ? nz : 1) * 64, sizeof(uint32_t));
    for (int x = 0; x < nz; ++x) {
        const int b = Rt->index[x];
        const int kb = b / bxy, jb = (b % bxy) / bx, ib = (b % bxy) % bx;
        const int i0 = imax(0, ib * 4 - 2), i1 = imin(VX, ib * 4 + 4);
        const int j0 = imax(0, jb * 4 - 2), j1 = imin(VY, jb * 4 + 4);
        uint8_t vm[216];
        memset(vm, 0, sizeof vm);
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i)
                for (int k = kb * 4 - 2; k < kb * 4 + 4; ++k)
                    if (rotor_voxel(&R, VX, i, j, k))
                        vm[(k - (kb * 4 - 2)) * 36 + (j - (jb * 4 - 2)) * 6 + (i - (ib * 4 - 2))] = 1;
        uint8_t cnt[125];
        for (int kc = 0; kc < 5; ++kc)
            for (int jc = 0; jc < 5; ++jc)
                for (int ic = 0; ic < 5; ++ic) {
                    int n = 0;
                    for (int km = kc; km <= kc + 1; ++km)
                        for (int jm = jc; jm <= jc + 1; ++jm)
                            for (int im = ic; im <= ic + 1; ++im) n += vm[km * 36 + jm * 6 + im];
                    cnt[kc * 25 + jc * 5 + ic] = (uint8_t)n;
                }
        for (int ki = 0; ki < 4; ++ki)
            for (int ji = 0; ji < 4; ++ji)
                for (int ii = 0; ii < 4; ++ii) {
                    uint32_t packed = 0u;
                    int corner = 0;
                    for (int kc = ki; kc <= ki + 1; ++kc)
                        for (int jc = ji; jc <= ji + 1; ++jc)
                            for (int ic = ii; ic <= ii + 1; ++ic)
                                packed |= (uint32_t)cnt[kc * 25 + jc * 5 + ic] << (4u * (unsigned)corner++);
                    Rt->interp[x * 64 + ki * 16 + ji * 4 + ii] = packed;
                }
    }
    Rt->g = (double *)calloc((size_t)(nz > 0 ? nz : 1) * 27 * 3, sizeof(double));
    free(mark);
    free(R.hitScan);
    free(R.rayMap);
}

/* buildGrids' rotor pass, gridBuilderFunctions.h:1612-1619: every level whose
   bit is set in levelMask (the reference's Info.useRotors) gets all rotors */
int orc_build_rotors(orc_sim *S, const orc_stl *rotors, int nRotors, int levelMask)
{
    if (nRotors < 0 || nRotors > LBM3_MAX_ROTORS) return -1;
    for (int L = 0; L < S->levels; ++L) {
        orc_grid *g = &S->g[L];
        for (int r = 0; r < g->nRotors; ++r) rotor_free(&g->rotors[r]);
        free(g->rotors);
        g->rotors = NULL;
        g->nRotors = 0;
        g->Info.useRotors = (levelMask >> L) & 1;
        if (!g->Info.useRotors || nRotors == 0) continue;
        g->rotors = (orc_rotor *)calloc((size_t)nRotors, sizeof(orc_rotor));
        g->nRotors = nRotors;
        for (int r = 0; r < nRotors; ++r) build_rotor(&g->rotors[r], &rotors[r], &g->Info);
    }
    return 0;
}

/* the case's direct writes to grids[level].rotors[r].Info (BruteforcePump.cu:226-229) */
int orc_set_rotor(orc_sim *S, int level, int rotor, float radiansPerSecond, int axis, float ox, float oy, float oz)
{
    if (level < 0 || level >= S->levels || rotor < 0 || rotor >= S->g[level].nRotors) return -1;
    orc_rotor *R = &S->g[level].rotors[rotor];
    R->radiansPerSecond = radiansPerSecond;
    R->axis = axis;
    R->ox = ox; R->oy = oy; R->oz = oz;
    return 0;
}
