/*
 * liblbm3mi -- C ABI of the LBM3 per-timestep hot path on MI355X (gfx950).
 *
 * The reference (stepanmuller/LBM3) has no FFI: a case file includes the
 * library headers and calls C++ functions on GridStruct vectors.  This header
 * is the thin C layer those calls map onto; every entry point names the
 * reference function it replaces.  Conventions:
 *   - an opaque lbm3_sim owns every device buffer of all grid levels and one
 *     HIP stream; calls on one sim are not thread safe;
 *   - host pointers are caller-owned and only read/written during the call;
 *   - every function returns 0 on success or a negative LBM3_E* code, with a
 *     message in lbm3_last_error(); no exception crosses the ABI.
 */
#ifndef LBM3MI_H
#define LBM3MI_H

#include <stdint.h>
#include "lbm3mi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LBM3_OK 0
#define LBM3_E_ARG -1      /* invalid argument */
#define LBM3_E_HIP -2      /* HIP runtime error */
#define LBM3_E_BUILD -3    /* grid builder failed (missing periodic/open-BC source, empty or open STL) */
#define LBM3_E_STATE -4    /* call out of order (e.g. step before build) */
#define LBM3_E_MEMORY -5   /* device allocation failed or exceeds the budget */
#define LBM3_E_DIVERGED -6 /* the watchdog found non-finite populations (lbm3_set_watchdog) */

typedef struct lbm3_sim lbm3_sim;

/* Host-side arrays of one built level, as produced by the reference builder
 * gridBuilderToGrid (gridBuilderFunctions.h:1032-1569).  Used by
 * lbm3_import_level; lbm3_download_array reads the same arrays back. */
typedef struct lbm3_level_arrays {
    lbm3_info info;
    int nRows;
    const int *shifter, *rowI, *rowJ, *rowK, *rowJPlus, *rowKPlus, *rowJKPlus;
    int W;                        /* wall-adjacent cells */
    const int *wallMap;           /* [cellCount] */
    const int *wallIndex;         /* [W] */
    const uint32_t *wallData;     /* [W] */
    const uint32_t *linkLength;   /* [7][W] */
    int nOpenIDs;                 /* open boundary ids 0..nOpenIDs-1 */
    const int *openCount;         /* [nOpenIDs] */
    const int *openIndex;         /* concatenated cell lists, BCID order */
    const int *openSource;        /* concatenated source cells */
    int nPeriodic;
    const int *periodicIndex, *periodicSource;
    int wallCount;                /* cells marked as wall */
    /* interface lists of this level towards the next finer one
     * (types.h InterfaceStruct; gridBuilderFunctions.h:741-1030); zero counts on the finest level */
    int f2cCount;
    const int *f2cIndex, *f2cChild;       /* coarse cell, first fine cell of its 2x2x2 block */
    int c2fCount;
    const int *c2fIndex, *c2fChild;       /* first coarse cell of the 2x2x2 block, first fine target */
    int leftoverCount;
    const int *leftoverIndex, *leftoverParent, *leftoverNbrI, *leftoverNbrJ, *leftoverNbrK;
} lbm3_level_arrays;

/* Library. */
const char *lbm3_last_error(void);
int lbm3_device_info(char *name, int nameLen, long long *hbmBytes, int *cuCount);

/* Simulation object: the case constants and hooks (lbm3_case) plus the
 * population storage precision (LBM3_STORE_*).  Replaces the
 * std::vector<GridStruct> grids of a case main(), e.g.
 * taylorGreenVortexMain.cu:185. */
int lbm3_sim_create(const lbm3_case *c, int storage, lbm3_sim **out);
int lbm3_sim_destroy(lbm3_sim *s);

/* buildGrids (gridBuilderFunctions.h:1577): voxelize the static STL bodies
 * (triangles as 9 floats ax,ay,az,bx,...,cz; triCounts[b] triangles per body),
 * build every level on the GPU and allocate the populations.  Returns the
 * reference's fluidUpdatesPerIteration (sum over levels of (cells - walls)*2^L).
 * Each body first passes readSTL's checks (STLFunctions.h:317-462): triangles
 * with a repeated vertex are dropped, and a body with an edge not shared by
 * exactly two triangles (checkSTLEdges, :130-315) is refused with LBM3_E_BUILD;
 * the counts stay readable as "stl<b>EdgeCount" (lbm3_download_array). */
int lbm3_build_grids(lbm3_sim *s, const float *triangles, const int *triCounts, int nBodies, lbm3_bounds domain,
                     long long *fluidUpdatesPerIteration);

/* Upload a level built elsewhere (same layout as gridBuilderToGrid output). */
int lbm3_import_level(lbm3_sim *s, int level, const lbm3_level_arrays *a);

/* The checks lbm3_import_level makes before any device call, on the host
 * only: sizes, null arrays with nonzero counts, rows numbered in the order of
 * their first cells and holding consecutive cells, every cell's j+/k+/jk+
 * neighbour (row neighbour + offset in the row, cellFunctions.h:41-46) inside
 * the level, wall, open-boundary, periodic and interface cells in range.
 * LBM3_E_ARG with a message naming the first failed check. */
int lbm3_validate_level(const lbm3_level_arrays *a);

/* applyInitialCondition (boundaryConditions/applyInitialCondition.h:8) on one level. */
int lbm3_apply_initial_condition(lbm3_sim *s, int level);
/* Per-cell initial state (dRho, ux, uy, uz) for every cell of the level, in cell order. */
int lbm3_apply_initial_fields(lbm3_sim *s, int level, const float *fields4);

/* updateSingleGrid (updateGrid.h:16): collide+stream, open BCs, periodic copy. */
int lbm3_update_single_grid(lbm3_sim *s, int level);
/* updateAllGrids (updateGrid.h:388): recursive 2:1 sub-cycling from `level`. */
int lbm3_update_all_grids(lbm3_sim *s, int level);
/* `steps` coarse time steps of updateAllGrids(grids, 0). */
int lbm3_run(lbm3_sim *s, int steps);
/* Divergence watchdog of lbm3_run (no counterpart in the reference, which has
 * no in-loop check): with period > 0, lbm3_run sums the mass over every level
 * (lbm3_mass) after each coarse step whose count is a multiple of `period`
 * and after its last step, and returns LBM3_E_DIVERGED, naming the step, when
 * the sum is not finite (a NaN or infinite population in a counted cell).
 * Each check reads the populations once and waits for the device.  0 (the
 * default) turns it off. */
int lbm3_set_watchdog(lbm3_sim *s, int period);
/* lbm3_run through a HIP graph (no counterpart in the reference): the launch
 * sequence of two coarse steps (every level's esotwist parity back where it
 * was) is captured once per call and replayed steps / 2 times, an odd last
 * step launched directly; returns when the work is done.  Results are those of
 * lbm3_run bit for bit.  The values that change from replay to replay -- the
 * iteration an open-boundary condition belongs to, whether a deferred one is
 * pending, the rotors' angles -- are read by the captured kernels from a small
 * device table refilled before each replay. */
int lbm3_run_graph(lbm3_sim *s, int steps);
int lbm3_synchronize(lbm3_sim *s);
/* Autotuner of the step's launch choices (no counterpart in the reference):
 * times `steps` whole coarse steps (after one untimed step) for every
 * combination of the bulk kernel's occupancy cap (dynamic LDS per block 0,
 * 24, 40 or 64 KiB), its block order (bit 0: round-robin over the XCDs or one
 * contiguous part per XCD; bit 1: the tiles in reverse order every other
 * step, so that a step starts on the lines the previous one left in the
 * last-level cache), the side kernels' placement (in order after the
 * bulk kernel: 0, on a second stream beside it: 1, and with several levels
 * also each finer level's sub-cycle on a stream of its own beside its coarser
 * level's step: 2) and, where a level completes its open boundary when it next
 * loads the open cells, the width of the aligned cell groups the open-cell
 * kernel takes (8 or 16 cells; 1: no groups, the bulk kernel completes and
 * steps the open cells itself) and what becomes of the cells on the x faces
 * (0: as the others; 1: stepped by the bulk kernel and completed by a
 * separate pass after streaming; 2: in groups while the bulk kernel takes the
 * other open cells; 3: in the bulk kernel with a block's lanes rotated by 32
 * cells, so that a row's two x-face cells share a wave), and, where a level
 * has rotors, whether the bulk kernel steps the rotor cells too instead of a
 * kernel of their own; times the six fastest again, three rounds in turn, keeps
 * the one with the best round and restores the simulation's state
 * as it was (the populations and every step counter; a
 * pending open-boundary condition stays pending unless the chosen open-cell
 * layout differs, which completes it first).  out[7 r + {LDS bytes,
 * block order, side streams, group cells, x-face choice, rotor cells in the
 * bulk kernel, ms per step in the sweep}] for the first `cap` candidates; *nOut the
 * candidates timed. */
int lbm3_autotune(lbm3_sim *s, int steps, double *out, int cap, int *nOut);
/* The launch choices in use: {bulk LDS bytes, block order, side streams, open
 * group cells, x-face choice, rotor cells in the bulk kernel}; lbm3_set_tune
 * sets them (LDS 0..65536, block order 0..3, side streams 0..3, group 1, 8 or
 * 16, x-face choice 0..3, rotor cells 0..1; a new layout of the open or rotor
 * cells completes a
 * pending open-boundary condition first).  Every choice gives the
 * same results bit for bit, except with 16-bit storage: that completion, and
 * every step's separate pass for the x faces, store the completed populations
 * rounded to the storage type. */
int lbm3_get_tune(lbm3_sim *s, int *out6);
int lbm3_set_tune(lbm3_sim *s, const int *in6);

/* Timed run: `steps` coarse steps between HIP events on the sim's stream.
 * out[0] = total ms; with kernelEvents != 0 also out[1] = summed ms of the
 * level-0 bulk collide-stream kernel (events around each launch), out[2] = the
 * number of those launches and out[3] = summed ms of the level-0 wall-adjacent
 * and open-cell kernels.  With kernelEvents the kernels run in order on one
 * stream (no overlap).  out holds 4 doubles.  No counterpart in the reference
 * (TNL::Timer). */
int lbm3_run_timed(lbm3_sim *s, int steps, int kernelEvents, double *out);

/* State access. */
int lbm3_get_info(lbm3_sim *s, int level, lbm3_info *out);
int lbm3_level_count(lbm3_sim *s);
/* (dRho, ux, uy, uz) of every cell from its pre-collision populations
 * (getDRhoUxUyUz, cellFunctions.h:257); out is float or double (FP64 store). */
int lbm3_download_macros(lbm3_sim *s, int level, void *out);
/* Populations [27][cellCount] in storage precision. */
int lbm3_download_populations(lbm3_sim *s, int level, void *out, long long bytes);
int lbm3_upload_populations(lbm3_sim *s, int level, const void *in, long long bytes);
/* Named built arrays (shifter, rowI.., wallMap, wallIndex, wallData, linkLength,
 * periodicIndex, periodicSource, openIndex, openSource, openMeta, rayHitScan,
 * rayMap, f2cIndex, ..., rotor<r>Map/Index/Interp/Force, stl<b>EdgeCount: the
 * shared-edge counts AB, BC, CA of every kept triangle of static body b of the
 * last build, any level).  "cellRow" is the per-step kernels' cell record, 8
 * bytes per cell: {kind, row} with kind >= 0 wall-adjacent entry, -1/-2 bulk,
 * -3 solid, -4 a cell the rotors can reach, <= -8 open cell (the last two
 * updated by kernels of their own).  With out == NULL only
 * *count is returned. */
int lbm3_download_array(lbm3_sim *s, int level, const char *name, void *out, long long capacity, long long *count);
/* The same arrays in device memory, not copied: *ptr points into the
 * simulation's own buffer (valid until the level is rebuilt or the sim is
 * destroyed), *count its element count.  "f" gives the populations (storage
 * type; tiled, see lbm3_population_layout), after any pending open-boundary
 * completion is applied.  Replaces the case files' direct view access
 * (Grid.fArray.getConstView(), Grid.IJKNBR.*, Grid.Wall.*:
 * taylorGreenVortexMain.cu:142-155). */
int lbm3_device_array(lbm3_sim *s, int level, const char *name, void **ptr, long long *count);
/* Population layout of "f": slot (q, cell) at (cell / tile) * 27 * tile +
 * q * tile + cell % tile; *bytes per value (4 FP32, 2 FP16/BF16, 8 FP64).
 * FP32 and FP64 hold the shifted populations f - w (applyCollision.h:8-10). */
int lbm3_population_layout(lbm3_sim *s, int *tile, int *bytes);

/* A case's four device hooks compiled with the case file and evaluated by it
 * (the reference inlines getRefinementModifier, getInitialCondition, getOpenBC
 * and getLocalBC into its kernels: markerFunctions.h:258-276,
 * applyInitialCondition.h:22-35, gridBuilderFunctions.h:1232-1300,1486-1503,
 * updateGrid.h:86-95,256).  Each callback evaluates its hook at n cells of the
 * level `info` describes, whose coordinates come as device arrays i, j, k;
 * writes n results into a device array; enqueues its work on `stream` (a
 * hipStream_t); returns 0 (anything else fails the calling entry point with
 * LBM3_E_ARG).  A null callback leaves that hook to the case's rule table.
 * Results are data: the library classifies them once (lbm3_build_grids,
 * lbm3_apply_initial_condition, lbm3_refresh_hooks), so the step kernels read
 * tables, not user code. */
typedef struct lbm3_hooks {
    void *user;
    /* out[c] bit 0: the marker the hook leaves for an incoming marker false,
     * bit 1: for an incoming marker true */
    int (*refinement)(void *user, const lbm3_info *info, const int *i, const int *j, const int *k, int n,
                      uint8_t *out, void *stream);
    /* out4[4c..4c+3] = BC.dRho, ux, uy, uz after the hook on a default BCStruct */
    int (*initial_condition)(void *user, const lbm3_info *info, const int *i, const int *j, const int *k, int n,
                             float *out4, void *stream);
    /* out[c] = the BCStruct after the hook on a default BCStruct */
    int (*open_bc)(void *user, const lbm3_info *info, const int *i, const int *j, const int *k, int n,
                   lbm3_bc *out, void *stream);
    /* in[c]: a default BCStruct with wallID and a state (dRho, ux, uy, uz) as the
     * step fills it; out[c] = the BCStruct after the hook.  The hook must not
     * depend on the state except by passing the velocity through (a wall that
     * moves with the fluid): the library evaluates it on two states and refuses
     * anything else, and refuses parameters (force, limiter, viscosity
     * multiplier, IBB) that differ between cells of one wall id or between
     * fluid cells of a level.  The wall velocity may differ from cell to cell. */
    int (*local_bc)(void *user, const lbm3_info *info, const int *i, const int *j, const int *k, const lbm3_bc *in,
                    int n, lbm3_bc *out, void *stream);
} lbm3_hooks;
/* Before lbm3_build_grids; hooks == NULL returns to the rule tables. */
int lbm3_set_hooks(lbm3_sim *s, const lbm3_hooks *hooks);
/* Host-side InfoStruct fields a case may change between steps (iRegulatorInlet,
 * iRegulatorOutlet, useRotors: BruteforcePump.cu:226-229,293-298), seen by the
 * hooks from then on. */
int lbm3_set_info_controls(lbm3_sim *s, int level, float iRegulatorInlet, float iRegulatorOutlet, int useRotors);
/* Evaluate the open-boundary and local hooks again on every built level (after
 * the case changed what they read) and replace the tables the step uses. */
int lbm3_refresh_hooks(lbm3_sim *s);

/* buildRotors (rotorFunctions.h:315-585, called by buildGrids for the levels
 * with Info.useRotors, gridBuilderFunctions.h:1612-1619): every level whose bit
 * is set in levelMask gets all nRotors rotor STLs (9 floats per triangle,
 * triCounts[r] triangles each), voxelized on the GPU at half the level spacing,
 * with 4x4x4 blocks of interpolation cells marked where solid voxels are near
 * and the 8 corner counts of each interpolation cell packed 4 bits each.
 * Replaces the level's rotors; a later build of the grids drops them. */
int lbm3_build_rotors(lbm3_sim *s, const float *triangles, const int *triCounts, int nRotors, int levelMask);
/* The case's writes to grids[level].rotors[r].Info after buildGrids
 * (BruteforcePump.cu:226-229): angular velocity [rad/s], axis LBM3_AXIS_X/Y/Z
 * (rotateAlongX/Y/Z) or LBM3_AXIS_NONE, and a point (ox, oy, oz) [mm] of the
 * axis.  Each updateSingleGrid advances the angle by dtPhys * radiansPerSecond
 * (updateGrid.h:36-49) and forces the cells inside the rotor toward rigid
 * rotation (processRotor, rotorFunctions.h:210-304). */
int lbm3_set_rotor(lbm3_sim *s, int level, int rotor, float radiansPerSecond, int axis, float ox, float oy, float oz);
/* RotorInfoStruct (types.h:206-211): out = {res, xMin, yMin, zMin, xMax, yMax,
 * zMax, cellCountX, cellCountY, cellCountZ, radiansElapsed, marked blocks}. */
int lbm3_rotor_info(lbm3_sim *s, int level, int rotor, double *out12);
/* Rotor force tracker sums (trackerFunctions.h:225-330) with FP64
 * accumulation, before the division by the iteration span and the unit
 * conversion: out = {Fx, Fy, Fz, Tx, Ty, Tz} in the rotor frame, the torque
 * about the rotation point with the force cells' centres as lever arms [mm].
 * Accumulated while the case tracks rotor force (trackRotorForce); reset by
 * lbm3_reset_wall_forces. */
int lbm3_rotor_forces(lbm3_sim *s, int level, int rotor, double *out6);

/* FP64 observables. */
/* kinetic energy sum over counted cells (taylorGreenVortexMain.cu:116-180, lattice units,
 * level L weighted by 8^-L) */
int lbm3_kinetic_energy(lbm3_sim *s, double *out);
/* Global mass balance in FP64 (no counterpart in the reference): over the same
 * counted cells as the kinetic energy, level L weighted by 8^-L,
 * out2[0] = sum of dRho and out2[1] = the number of counted cells, so the
 * mass in lattice units (rho0 = 1, one coarse cell = 1) is out2[1] + out2[0].
 * On a periodic box out2[0] stays constant to rounding. */
int lbm3_mass(lbm3_sim *s, double *out2);
/* per-body sums of the accumulated momentum-exchange force and its torque about
 * `origin` (trackerFunctions.h:356-474, FP64): out[body*6 + {Fx,Fy,Fz,Tx,Ty,Tz}] */
int lbm3_wall_forces(lbm3_sim *s, int level, const double *origin, int nBodies, double *out);
int lbm3_reset_wall_forces(lbm3_sim *s);   /* wall and rotor force accumulators */
/* lbm3_wall_forces without any allocation (no counterpart in the reference):
 * for callers that cannot allocate -- code under HIP graph capture, tight
 * tracker or serving loops that forbid hipMalloc.  Same arguments and result
 * (FP64, equal to lbm3_wall_forces up to the order of the partial sums).  The
 * partial sums live in a fixed device array of 2^20 doubles per GPU, held
 * until the process ends: wall-adjacent cells of the level, rounded up to a
 * multiple of 256, times 6 * nBodies / 256, plus 6 * nBodies, must fit, else
 * LBM3_E_ARG.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other host threads, other streams) race. */
int lbm3_wall_forces_noalloc(lbm3_sim *s, int level, const double *origin, int nBodies, double *out);
/* Open-boundary flow statistics of updateTracker (trackerFunctions.h:127-223)
 * for boundary openBCID over every level that has it, from the per-cell
 * accumulators time-averaged over the steps since the last
 * lbm3_reset_open_boundary_stats (the case must track open boundaries):
 * out7 = {tracked area [m^2], area-averaged normal velocity [m/s], mass flow
 * [kg/s], momentum thrust [N], area-averaged pressure [Pa], pressure power [W],
 * normal kinetic power [W]}.  Normal velocities are along the outer normal.
 * FP64 throughout (the reference reduces in FP32). */
int lbm3_open_boundary_stats(lbm3_sim *s, int openBCID, double *out7);
/* updateTracker (trackerFunctions.h:120-486) for one period: per wall body
 * walls[6 b + {Fx, Fy, Fz, Tx, Ty, Tz}] (N, N m about the origin, the force the
 * walls exert on the fluid), per rotor rotors[6 r + ...] (in the rotor frame,
 * torque about its rotation point), per open boundary open[7 b + ...] (as
 * lbm3_open_boundary_stats), each summed over the levels that have it and
 * time-averaged over the steps since the last reset; then every accumulator
 * and each level's iterationsSinceTrackerReset is reset.  FP64 (the
 * reference's reductions are FP32).  The families the case does not track
 * (trackWallForce, trackRotorForce, trackOpenBoundaries) must be asked for
 * with a count of 0, or come back as zeros. */
int lbm3_update_tracker(lbm3_sim *s, double *walls, int nWalls, double *rotors, int nRotors, double *open, int nOpen);
/* Section cut through every level (exportSectionCutPlot{XY,ZY,ZX} without
 * bounds or rotor frame, include/plotter/exportSectionCutPlot.h:11-330, 395-417):
 * the plane at `cut` [mm] along its normal (z for LBM3_PLANE_XY, x for
 * LBM3_PLANE_ZY, y for LBM3_PLANE_ZX), sampled at the finest level's cells
 * that fit 16 M pixels, coarser levels painted over finer pixel blocks and
 * blended between their two nearest planes.  dims = {vertical, horizontal
 * pixel counts}; with out non-null, out[(v * H + h) * 6 + k] holds the six
 * floats the reference writes to /dev/shm/sim_data.bin per pixel: pressure
 * [Pa], velocity along the horizontal, vertical and normal axes [m/s], solid
 * marker, cell size [mm].  bounds (null or all zero: the whole grid) crops the
 * image as the reference's bounded overloads do.  Rotor fractions are not
 * added to the marker. */
int lbm3_section_cut(lbm3_sim *s, int plane, float cut, const lbm3_bounds *bounds, float *out, long long capacity,
                     int *dims);
/* Checkpoint and restart (no counterpart in the reference, which keeps its
 * state in device memory only): lbm3_save_checkpoint writes the state a time
 * step changes -- populations, esotwist parity and iteration counters of every
 * level, open-boundary reference states and statistics, wall and rotor force
 * accumulators, rotor angles -- to `path`; lbm3_load_checkpoint reads it into
 * a simulation built from the same case and grid (checked: storage, level
 * count, cell, wall, open-boundary and rotor counts, grid sizes, origin,
 * resolution, time step and viscosity of every level), after which
 * its steps reproduce the saved simulation's bit for bit.  Any open boundary
 * condition still pending is saved as pending. */
int lbm3_save_checkpoint(lbm3_sim *s, const char *path);
int lbm3_load_checkpoint(lbm3_sim *s, const char *path);
/* exportVTK (include/exportVTK.h:17-158): every stored cell of every level as
 * one vertex of a legacy ASCII VTK polydata file at `path` (coordinates in
 * mm), with the point data wall, velocity_m_s, pressure_Pa, grid_id,
 * grid_resolution_mm and interface_overlap. */
int lbm3_export_vtk(lbm3_sim *s, const char *path);
/* getFlowReport{XY,ZY,ZX} (flowReportFunctions.h:11-372): the section image of
 * lbm3_section_cut reduced over its fluid pixels, with the plane's positive
 * normal: out7 = {area-averaged normal velocity [m/s], mass flow [kg/s],
 * momentum thrust [N], area-averaged pressure [Pa], pressure power [W],
 * normal kinetic power [W], fluid area [mm^2]} (FlowReportStruct order, FP64
 * sums; zeros where the section holds no fluid). */
int lbm3_flow_report(lbm3_sim *s, int plane, float cut, const lbm3_bounds *bounds, double *out7);
/* zero the open-boundary accumulators and every level's iterationsSinceTrackerReset
 * (the reset at the end of updateTracker, trackerFunctions.h:476-486) */
int lbm3_reset_open_boundary_stats(lbm3_sim *s);

/* applyCollision (applyCollision.h:17) on n independent cells, 27 shifted
 * populations each (cell-major), on the device. */
int lbm3_collide_cells_f32(float *f, int n, const lbm3_bc *bc, float nu);
int lbm3_collide_cells_f64(double *f, int n, const lbm3_bc *bc, double nu);

/* Continuous-batching server of single-level simulations.  No counterpart in
 * the reference (its batch runner, run_intakes_automated.py, solves cases one
 * after another); for serving many independent cases on one GPU.
 *
 * submit queues a request (its arrival time is the call time); each
 * lbm3_server_step admits queued requests in arrival order while at most
 * maxActive run and their estimated device bytes fit hbmBudgetBytes (building
 * their grids on the GPU), advances every active request by one time step in
 * one batched launch sequence, and retires those that reached their step
 * count: their ids are written to completed[0..cap) and *nCompleted counts
 * them.  Requests that fail to build complete immediately with an error status.
 * lbm3_server_result: out[0] queue wait s, out[1] time to solution s (submit to
 * finish), out[2] fluid updates, out[3] final kinetic energy, out[4..9] force
 * and torque of wall body 0 about the origin (FP64), out[10] status (0 or an
 * LBM3_E_* code), out[11] steps; n >= 12.  lbm3_server_macros: (dRho, ux, uy,
 * uz) per cell at the end, for requests submitted with keepMacros.
 * lbm3_server_stats: out[0] active, out[1] queued, out[2] device bytes held,
 * out[3] batched steps run, out[4] fluid updates run (5 doubles). */
typedef struct lbm3_server lbm3_server;
int lbm3_server_create(long long hbmBudgetBytes, int maxActive, lbm3_server **out);
int lbm3_server_destroy(lbm3_server *s);
int lbm3_server_submit(lbm3_server *s, const lbm3_case *c, int storage, const float *triangles, const int *triCounts,
                       int nBodies, lbm3_bounds domain, int steps, int keepMacros, long long *id);
int lbm3_server_step(lbm3_server *s, long long *completed, int cap, int *nCompleted);
int lbm3_server_result(lbm3_server *s, long long id, double *out, int n);
int lbm3_server_macros(lbm3_server *s, long long id, void *out, long long bytes);
int lbm3_server_stats(lbm3_server *s, double *out);

/* Plain device-to-device streaming copy of `bytes` (read + write), timed over
 * `reps` launches: the achievable-bandwidth probe and the reference kernel for
 * calibrating HBM byte counters.  No counterpart in the reference. */
int lbm3_copy_probe(long long bytes, int reps, double *msPerCopy);

#ifdef __cplusplus
}
#endif
#endif
