/*
 * Plain C data types shared by the LBM3 MI355X library (liblbm3mi.so) and its
 * callers.  They restate the reference's case-file surface as data:
 *
 *  - lbm3_info       mirrors InfoStruct            (reference include/types.h:120-129)
 *  - lbm3_bounds     mirrors BoundsStruct          (include/types.h:117-118)
 *  - lbm3_bc         mirrors BCStruct              (include/types.h:131-136)
 *  - lbm3_case       the compile-time constants of a case file
 *                    (e.g. cases/taylorGreenVortex/taylorGreenVortexMain.cu:1-44)
 *                    plus its four device hooks expressed as rule tables:
 *      getRefinementModifier -> lbm3_refine_rule   (called at markerFunctions.h:273)
 *      getInitialCondition   -> lbm3_init          (applyInitialCondition.h:31)
 *      getOpenBC             -> lbm3_open_rule     (updateGrid.h:256, gridBuilderFunctions.h:1248-1500)
 *      getLocalBC            -> lbm3_local_rule    (updateGrid.h:95)
 *
 * No pointer in these structs refers to device memory.
 */
#ifndef LBM3MI_TYPES_H
#define LBM3MI_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM3_MAX_LEVELS 8
#define LBM3_MAX_RULES 16
#define LBM3_MAX_WALLS 31
#define LBM3_MAX_ROTORS 8        /* rotor STLs per level (rotorSTLs, buildGrids) */

/* rotation axis of a rotor (RotorInfoStruct rotateAlongX/Y/Z, types.h:206-211) */
#define LBM3_AXIS_NONE (-1)
#define LBM3_AXIS_X 0
#define LBM3_AXIS_Y 1
#define LBM3_AXIS_Z 2

typedef struct lbm3_bounds {
    float xMin, yMin, zMin, xMax, yMax, zMax;
    float rxMax, ryMax, rzMax;
} lbm3_bounds;

typedef struct lbm3_info {
    float gridID;
    unsigned long long gridMemoryBytes;
    int iterationsFinished;
    int iterationsSinceTrackerReset;
    float res, ox, oy, oz;
    lbm3_bounds Bounds;
    float nu, dtPhys;
    int cellCountX, cellCountY, cellCountZ;
    int cellCount;
    int esotwistFlipper;
    int useRotors;
    float iRegulatorInlet, iRegulatorOutlet;
} lbm3_info;

typedef struct lbm3_bc {
    float dRho, ux, uy, uz;
    float gx, gy, gz;
    int wallID;
    int openBCID;
    int useIBB;
    int dirichletU, dirichletRho, periodic, nonReflective;
    float nonReflectiveBeta;
    float collisionLimiter, nuMultiplier;
} lbm3_bc;

/* Face classes used by open-boundary rules: bit 0 = index 0 on that axis,
 * bit 1 = interior, bit 2 = index count-1. */
#define LBM3_FACE_LOW 1
#define LBM3_FACE_MID 2
#define LBM3_FACE_HIGH 4
#define LBM3_FACE_ANY 7

/* A radial profile about a line: the factor max(0, 1 - r^2 / radius^2), r the
 * distance of the cell centre from the line through `origin` along `axis`,
 * evaluated in FP32 as the pipe cases write it (r^2 = dx*dx + dy*dy,
 * poiseuilleFlowMain.cu:54-55,66-72).  radius <= 0 (a zeroed struct): no
 * profile, factor 1. */
typedef struct lbm3_profile {
    int axis;               /* LBM3_AXIS_X, _Y or _Z */
    float origin[3];        /* a point of the line [mm] */
    float radius;           /* [mm]; <= 0: no profile */
} lbm3_profile;

typedef struct lbm3_open_rule {
    int levelMask;          /* bit L: applies on level L; 0 means every level */
    int faceMask[3];        /* per axis x, y, z: LBM3_FACE_* bits a boundary cell must match */
    lbm3_bc bc;             /* the BCStruct the hook returns for matching cells */
    lbm3_profile profile;   /* bc.ux/uy/uz are scaled by the profile factor of the cell */
} lbm3_open_rule;

/* getLocalBC: rules are applied in order to the BCStruct that the kernel has
 * filled with wallID, dRho and velocity; later rules override earlier ones. */
#define LBM3_SET_VELOCITY 1       /* BC.ux/uy/uz := ux/uy/uz (wall velocity seen by IBB) */
#define LBM3_SET_ROTATION 2       /* BC.u := omega x (x - origin) at the cell centre */
#define LBM3_SET_FORCE 4          /* BC.gx/gy/gz := gx/gy/gz */
#define LBM3_SET_LIMITER 8        /* BC.collisionLimiter := collisionLimiter */
#define LBM3_SET_NU_MULTIPLIER 16 /* BC.nuMultiplier := nuMultiplier */
#define LBM3_SET_USE_IBB 32       /* BC.useIBB := useIBB */

typedef struct lbm3_local_rule {
    int levelMask;          /* 0: every level */
    int wallID;             /* -2: every cell; -1: cells not next to a wall; >=0: that wall body */
    int set;                /* LBM3_SET_* bits */
    float ux, uy, uz;
    float omega[3];         /* angular velocity [rad/s]; lattice velocity of the wall is
                               (omega x (x - origin)) * dtPhys / res on each level */
    float origin[3];        /* point on the rotation axis, physical [mm] */
    float gx, gy, gz;
    float collisionLimiter, nuMultiplier;
    int useIBB;
} lbm3_local_rule;

#define LBM3_INIT_UNIFORM 0
#define LBM3_INIT_TAYLOR_GREEN 1

typedef struct lbm3_init {
    int kind;
    float dRho, ux, uy, uz;  /* uniform state */
    float u0;                /* Taylor-Green amplitude, lattice units */
    float length;            /* Taylor-Green reference length [mm]: X = x / length */
    int pressure;            /* Taylor-Green: add the NASA pressure perturbation */
    lbm3_profile profile;    /* uniform state: ux/uy/uz scaled by the profile factor of the cell */
} lbm3_init;

#define LBM3_REFINE_SET 0    /* marker := inside (any SET box of the level) */
#define LBM3_REFINE_ADD 1    /* marker |= inside */
#define LBM3_REFINE_REMOVE 2 /* marker &= !inside */

typedef struct lbm3_refine_rule {
    int level;
    int mode;
    float xMin, yMin, zMin, xMax, yMax, zMax; /* open box, physical [mm] */
    lbm3_profile cylinder;  /* radius > 0: also r^2 < radius^2 about that line */
} lbm3_refine_rule;

typedef struct lbm3_case {
    int gridLevelCount;      /* GRID_LEVEL_COUNT */
    int wallRefinementCount; /* WALL_REFINEMENT_COUNT */
    float resGlobal;         /* RES_GLOBAL [mm] */
    float dtPhysGlobal;      /* DT_PHYS_GLOBAL [s] */
    float nuPhys;            /* NU_PHYS [m^2/s] */
    float rhoPhys;           /* RHO_PHYS [kg/m^3] */
    int trackWallForce, trackRotorForce, trackOpenBoundaries;
    lbm3_init init;
    int nOpenRules;
    lbm3_open_rule openRules[LBM3_MAX_RULES];
    lbm3_bc openDefault;     /* BCStruct for boundary cells no rule matches */
    int nLocalRules;
    lbm3_local_rule localRules[LBM3_MAX_RULES];
    int nRefineRules;
    lbm3_refine_rule refineRules[LBM3_MAX_RULES];
    /* Reference velocity of the 16-bit population encoding (no counterpart in the
     * reference): FP16/BF16 stores hold (f - feq(rho = 1, storeRef)) / w, so a flow
     * near that state keeps the full 16-bit precision for its deviation.  Zero gives
     * the plain shifted populations (f - w) / w.  Ignored by the FP32/FP64 stores. */
    float storeRef[3];
    /* Non-reflective scheme of the open boundaries (no counterpart field in the
     * reference, which swaps its updateGrid.h for another file):
     * LBM3_OPEN_CHARACTERISTIC (updateGrid.h:267-307: the target corrected along
     * the outgoing characteristic, reference state an exponential moving
     * average) or LBM3_OPEN_IMPEDANCE (nonReflectiveTest/
     * updateGrid_SCHLAFFER_IMPEDANCE_SCHEME.h:250-331: the target clamped to
     * within rhoReflectionTolerance of the non-reflecting density, reference
     * state the last completed one). */
    int openScheme;
    float rhoReflectionTolerance;   /* RHO_REFLECTION_TOLERANCE (nonReflectiveTestMain.cu:1); >= 1 disables */
    /* Coarse <-> fine interface transfer (no counterpart field in the reference,
     * which swaps its updateInterface.h for another file): LBM3_INTERFACE_QUADRATIC
     * (updateInterface.h, Geier's quadratic velocity interpolation) or
     * LBM3_INTERFACE_LINEAR (INTERFACE_VERSIONS_BACKUP/updateInterfaceLINEAR.h:
     * affine fits, volume-averaged density and momentum). */
    int interfaceScheme;
} lbm3_case;

/* section-cut planes (PlaneEnum of the reference's plotter) */
#define LBM3_PLANE_XY 0
#define LBM3_PLANE_ZY 1
#define LBM3_PLANE_ZX 2

#define LBM3_INTERFACE_QUADRATIC 0
#define LBM3_INTERFACE_LINEAR 1

#define LBM3_OPEN_CHARACTERISTIC 0
#define LBM3_OPEN_IMPEDANCE 1

/* Storage precisions of the population array (shifted f - w). */
#define LBM3_STORE_F32 0
#define LBM3_STORE_F16 1
#define LBM3_STORE_BF16 2
#define LBM3_STORE_F64 3
/* 16 bits: sign, 4 exponent bits, 11 mantissa bits -- one mantissa bit more
 * than F16 over the range the shifted encoding uses (|deviation| < 2, larger
 * values saturate); slower to encode and decode than F16's hardware
 * conversions.  No counterpart in the reference. */
#define LBM3_STORE_F16C 4

#ifdef __cplusplus
}
#endif
#endif
