/*
 * ORACLE (test infrastructure only): per-cell numerics of the LBM3 hot path,
 * instantiated for float (_f32) and double (_f64).  See real_impl.h.
 */
#include <math.h>
#include "oracle.h"
#include "../lbm3_amd/csrc/generated/boundary_tables.h"

/* D3Q27Directions.h:36-41 */
const int ORC_CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1};
const int ORC_CY[27] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1};
const int ORC_CZ[27] = {0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1};
const int ORC_INV[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

#define REAL float
#define NAME(x) x##_f32
#define ORC_ABS fabsf
#include "real_impl.h"
#undef REAL
#undef NAME
#undef ORC_ABS

#define REAL double
#define NAME(x) x##_f64
#define ORC_ABS fabs
#include "real_impl.h"
#undef REAL
#undef NAME
#undef ORC_ABS

/* ctypes-friendly batch entry points: n cells, 27 populations each (cell-major) */
void orc_collide_batch_f32(float *f, int n, const lbm3_bc *bc, float nu)
{
    bcr_f32 b; bc_from_f32(&b, bc);
    for (int c = 0; c < n; ++c) orc_collide_f32(f + 27 * c, &b, nu);
}
void orc_collide_batch_f64(double *f, int n, const lbm3_bc *bc, double nu)
{
    bcr_f64 b; bc_from_f64(&b, bc);
    for (int c = 0; c < n; ++c) orc_collide_f64(f + 27 * c, &b, nu);
}
void orc_feq_batch_f32(const float *m4, int n, float *feq)
{
    for (int c = 0; c < n; ++c) orc_feq_f32(m4[4 * c], m4[4 * c + 1], m4[4 * c + 2], m4[4 * c + 3], feq + 27 * c);
}
void orc_feq_batch_f64(const double *m4, int n, double *feq)
{
    for (int c = 0; c < n; ++c) orc_feq_f64(m4[4 * c], m4[4 * c + 1], m4[4 * c + 2], m4[4 * c + 3], feq + 27 * c);
}
void orc_moments_batch_f32(const float *f, int n, float *m4)
{
    for (int c = 0; c < n; ++c) orc_moments_f32(f + 27 * c, m4 + 4 * c);
}
void orc_moments_batch_f64(const double *f, int n, double *m4)
{
    for (int c = 0; c < n; ++c) orc_moments_f64(f + 27 * c, m4 + 4 * c);
}
/* open-boundary pieces on one cell: normal index n (table order), bc in/out, f[27] in/out */
void orc_openbc_cell_f32(int n, int what, lbm3_bc *bc, float *f)
{
    bcr_f32 b; bc_from_f32(&b, bc);
    if (what == 0) orc_restore_rho_f32(n, &b, f);
    else if (what == 1) orc_restore_u_f32(n, &b, f);
    else orc_mbbc_f32(n, &b, f);
    bc->dRho = b.dRho; bc->ux = b.ux; bc->uy = b.uy; bc->uz = b.uz;
}
