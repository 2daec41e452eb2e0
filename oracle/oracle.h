/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C CPU restatement of the LBM3 hot path (reference stepanmuller/LBM3,
 * headers under include/), used only by tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py, as the checker.  Every function cites the
 * reference file:line it restates.  The product (lbm3_amd/, HIP for gfx950)
 * never links or calls anything here.
 *
 * The reference ships no tests or golden vectors for this path and cannot be
 * built here (TNL, nvcc), so the restatement is pinned by the mathematics and
 * the geometry (tests/test_oracle_pins.py, tests/test_boundary_tables.py):
 *   lattice tables      names -> vectors, inverse pairs
 *   feq / moments       closed-form zeroth, first and second moments
 *   collision           mass, momentum + g, feq fixed point, omega1 = 1 at rest,
 *                       cubic symmetry, shear decay (1 - omega1)
 *   step + periodic     brute-force collide + pull streaming + ghost copy
 *   viscosity           creeping Taylor-Green decay 6 nu k^2
 *   force + moment BC   plane Poiseuille bulk curvature, no-slip nodes
 *   MBBC / restore      tables equal the reference's generated code and
 *                       reproduce the prescribed moments; uniform flow kept
 *   voxelizer/builder   analytic box and sphere, link-code known answer
 *   interface transfer  reconstruction inverts the stencil values, the
 *                       interpolant reproduces its polynomial family, uniform
 *                       flow kept across levels, 2-level creeping TGV decay
 *                       (tests/test_oracle_interface_pins.py)
 *   rotors              corner-count packing and trilinear fraction, a box
 *                       rotor's blocks and counts by geometry, rigid-body
 *                       forcing inside a fully solid rotor, angle advance
 *                       (tests/test_oracle_rotor_pins.py)
 * Parity unpinned: the non-reflective characteristic update
 * (updateGrid.h:267-307) beyond the uniform-flow fixed point, the near-wall
 * layer of the moment-based walls and edge normals where ghost layers meet
 * wall faces.
 */
#ifndef LBM3_ORACLE_H
#define LBM3_ORACLE_H

#include <stdint.h>
#include "../include/lbm3mi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* D3Q27 direction ids, D3Q27Directions.h:8-34 */
enum { OOO = 0, POO, MOO, OOM, OOP, OMO, OPO, POM, MOP, POP, MOM, MMO, PPO, OPM, OMP,
       MPO, PMO, OPP, OMM, MPM, PMP, MMP, PPM, PMM, MPP, MMM, PPP };

extern const int ORC_CX[27], ORC_CY[27], ORC_CZ[27], ORC_INV[27];

/* ---- numerics (numerics.c / real_impl.h) ---- */
void orc_feq_f32(float dRho, float ux, float uy, float uz, float *feq);
void orc_feq_f64(double dRho, double ux, double uy, double uz, double *feq);
void orc_moments_f32(const float *f, float *out4);
void orc_moments_f64(const double *f, double *out4);

/* ---- STL (builder.c) ---- */
typedef struct orc_stl {
    int triangleCount;
    float *ax, *ay, *az, *bx, *by, *bz, *cx, *cy, *cz;
    lbm3_bounds Bounds;
    float binSize, oxBin, oyBin, ozBin;
    int binCountX, binCountY, binCountZ;
    int *binArray, *firstInBin;
    int *edgeCount; /* [triangleCount][3] shared-edge counts */
    int excluded;   /* triangles dropped for repeated vertices */
} orc_stl;

typedef struct orc_raymap {
    int gridID;
    long long totalHitCount;
    int nRays;
    long long *hitScan; /* nRays + 1 */
    int *rayMap;        /* totalHitCount */
} orc_raymap;

typedef struct orc_interface {
    int interfaceCount;
    int *index, *childMap;
    int leftoverCount;
    int *leftoverIndex, *leftoverParent, *leftoverNbrI, *leftoverNbrJ, *leftoverNbrK;
} orc_interface;

typedef struct orc_openbc {
    int openBCID, count, trackFlowCount;
    int *index, *source;
    void *dRhoRef, *uNormalRef, *dRhoCum, *uNormalCum; /* REAL[count] */
} orc_openbc;

/* RotorStruct / RotorInfoStruct, types.h:206-215 */
typedef struct orc_rotor {
    float res;                    /* half the level spacing */
    lbm3_bounds B;                /* interpolation-grid bounds */
    int cx, cy, cz;               /* interpolation cells, multiples of 4 */
    float ox, oy, oz;             /* a point of the rotation axis */
    int axis;                     /* LBM3_AXIS_* */
    float radiansPerSecond;
    double radiansElapsed;
    int blockCount, nonZero;
    int *rotorMap;                /* [blockCount] compact index or -1 */
    int *index;                   /* [nonZero] block of each compact index */
    uint32_t *interp;             /* [nonZero * 64] packed corner counts */
    double *g;                    /* [3][nonZero * 27] forcing accumulated per force cell */
} orc_rotor;

typedef struct orc_grid {
    lbm3_info Info;
    /* compressed IJK/NBR, types.h:162-164 */
    int *shifter; int nRows;
    int *ri, *rj, *rk, *rjp, *rkp, *rjkp;
    /* walls, types.h:202-204 */
    int *wallMap; int W; int *wallIndex; uint32_t *wallData; uint32_t *linkLength; /* [7][W] */
    void *gWall; /* REAL[3][W] force accumulators */
    orc_interface C2F, F2C;
    int nOpen; orc_openbc open[8];
    int nPeriodic; int *perIndex, *perSource;
    int wallCount; /* cells marked wall by the builder */
    void *f; /* REAL[27][N], index q*N + cell */
    int nRotors; struct orc_rotor *rotors;
} orc_grid;

typedef struct orc_sim {
    lbm3_case C;
    int isDouble;
    int levels;
    orc_grid g[LBM3_MAX_LEVELS];
    long long fluidUpdatesPerIteration;
    int nStatic; orc_stl *statics;
    /* voxelizer results per level (rayMapTotal), kept for flag parity tests */
    orc_raymap rayTotal[LBM3_MAX_LEVELS];
    char error[256];
} orc_sim;

/* STL */
int orc_stl_read(orc_stl *stl, const char *path, float resGlobal);
int orc_stl_from_triangles(orc_stl *stl, const float *tri9, int n, float resGlobal);
void orc_stl_free(orc_stl *stl);
int orc_stl_edge_counts(const float *tri9, int n, float resGlobal, int *out, int capacity);

/* build + run */
orc_sim *orc_sim_create(const lbm3_case *C, int isDouble);
int orc_build_grids(orc_sim *S, const orc_stl *statics, int nStatic, lbm3_bounds Domain);
void orc_apply_initial_condition(orc_sim *S, int level);
int orc_apply_initial_fields(orc_sim *S, int level, const float *fields4);
void orc_update_all_grids(orc_sim *S, int level);
void orc_update_single_grid(orc_sim *S, int level);
void orc_update_interface(orc_sim *S, int coarse);
void orc_sim_free(orc_sim *S);

/* rotors (rotorFunctions.h) */
int orc_build_rotors(orc_sim *S, const orc_stl *rotors, int nRotors, int levelMask);
int orc_set_rotor(orc_sim *S, int level, int rotor, float radiansPerSecond, int axis, float ox, float oy, float oz);
int orc_build_rotors_from_triangles(orc_sim *S, const float *tri, const int *triCounts, int nRotors, int levelMask);
float orc_rotor_fraction_f32(const orc_sim *S, int L, int r, float angle, float x, float y, float z);
int orc_rotor_info(const orc_sim *S, int L, int r, double *out12);
int orc_rotor_forces(const orc_sim *S, int L, int r, double *out6);
/* open-boundary flow statistics and their reset (trackerFunctions.h:127-223,476-486) */
int orc_open_boundary_stats(const orc_sim *S, int openBCID, double *out7);
void orc_reset_open_boundary_stats(orc_sim *S);

#ifdef __cplusplus
}
#endif
#endif
