// Device-resident state of a simulation (all grid levels) behind the C ABI.
#pragma once
#include <string>
#include <array>
#include <vector>
#include "../../include/lbm3mi.h"
#include "step.hpp"
#include "hooks.hpp"

namespace lbm3 {

// a device buffer that frees itself
template <class T> struct DBuf {
    T *p = nullptr;
    long long n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(long long count)
    {
        release();
        if (count <= 0) return hipSuccess;
        const hipError_t e = hipMalloc((void **)&p, sizeof(T) * (size_t)count);
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
};

struct InterfaceLists {
    int interfaceCount = 0, leftoverCount = 0;
    DBuf<int> index, childMap, leftoverIndex, leftoverParent, leftoverNbrI, leftoverNbrJ, leftoverNbrK;
    // the eight neighbour cells (two int4) of every stencil corner the transfer
    // reads and of every cell it writes, derived from the two levels' topology
    // before the first transfer (nbReady false: to be derived again)
    DBuf<int4> srcNb, dstNb;
    bool nbReady = false;
};

struct Level {
    lbm3_info info{};
    LevelGeom geom{};
    int N = 0, nRows = 0, W = 0, wallCount = 0;
    DBuf<int> shifter, ri, rj, rk, rjp, rkp, rjkp;
    DBuf<int2> cellRow;              // {wallMap, row} per cell
    DBuf<int4> rowTab;               // {first cell, j+ row, k+ row, jk+ row} per row
    DBuf<int> wallMap, wallIndex;
    DBuf<uint32_t> wallData, linkLength;
    DBuf<double> gWall;
    // open boundaries, all BCIDs concatenated in BCID order
    int nOpenIDs = 0, nOpen = 0;
    std::vector<int> openCount;
    DBuf<int> openCell, openSource;
    DBuf<int4> openNb;               // [2][nOpen] neighbour cells of each open cell
    DBuf<uint32_t> openMeta;
    DBuf<uint64_t> bulkSkip;         // per 64 cells: bit set where the bulk kernel leaves the cell to others
    DBuf<int> tileRow;               // per 256 cells: row of the first cell
    DBuf<float> openScale;           // per open entry: velocity profile factor (null: no rule has a profile)
    DBuf<OpenBCRecord> openRecords;
    DBuf<unsigned char> openRefs;    // 2 arrays of nOpen compute-type values
    DBuf<double> openCum;            // 2 arrays of nOpen
    int nPeriodic = 0;
    DBuf<int> perIndex, perSource;
    InterfaceLists C2F, F2C;
    std::vector<int> fineRefs;       // imported lists: fine cells they name, checked against the finer level
    bool interfaceChecked = true;
    // open boundary condition applied by the bulk kernel when it loads the cell
    bool openFusionChecked = false;  // decided for the current lists
    bool fusedOpen = false;          // the level uses it
    bool openSettled = true;         // no boundary condition pending (after init, upload or settle)
    DBuf<int> openGroups;            // aligned cell groups holding open cells (k_open_groups)
    int nOpenGroups = 0;
    int openGroupCells = 8;          // cells per group of openGroups
    // group width 1: no groups, the bulk kernel completes and steps the open
    // cells itself (openBits marks them)
    bool openInBulk = false;
    DBuf<uint64_t> openBits;
    bool laneRot = false;            // Sim::Tune::openSplit 3: the bulk kernel's lanes rotated by 32 cells
    // Sim::Tune::rotorInBulk: the bulk kernel steps the rotor cells (rotorBits)
    bool rotorInBulk = false;
    DBuf<uint64_t> rotorBits;
    // with Sim::Tune::openSplit: the open-list entries the open-cell kernel
    // completes (fused) and those on the x faces, completed by a pass after
    // the step's streaming (separate)
    DBuf<int> fusedEntries, sepEntries;
    int nFusedEntries = 0, nSepEntries = 0;
    LocalBC bulk{};
    DBuf<LocalBC> wallBC;            // [32]
    DBuf<float> wallVel;             // [3][W] per-cell wall velocity of velMode 3 wall ids (else empty)
    // open-boundary condition records the open cells index (open_meta); empty:
    // the case's rule table and default (imported levels)
    std::vector<lbm3_bc> openBC;
    DBuf<unsigned char> f;           // [27][N] storage type
    // rotors of the level (buildRotors), device arrays and host-side motion
    struct Rotor {
        RotorDev dev{};              // arrays point into the buffers below
        double radiansElapsed = 0.0;
        int blockCount = 0;
        DBuf<int> rotorMap, index;
        DBuf<uint32_t> interp;
        DBuf<double> g;
    };
    std::vector<Rotor> rotors;
    DBuf<RotorDev> rotorDev;          // [rotors.size()]
    bool rotorChecked = true;         // candidate cells and records match the rotors' motion
    DBuf<int> rotorCells;             // cells the rotors can reach that have their own kernel
    int rotorCellCount = 0;
    DBuf<uint8_t> rotorCand;          // [N] 1 where a rotor can reach the cell
    // voxelizer output kept for flag checks
    DBuf<long long> rayHitScan;
    DBuf<int> rayMap;
    bool built = false;
};

struct Sim {
    lbm3_case C{};
    int storage = LBM3_STORE_F32;
    int levels = 0;
    Level L[LBM3_MAX_LEVELS];
    long long fluidUpdates = 0;
    Codec codec{};                   // 16-bit population encoding (case storeRef)
    hipStream_t stream = nullptr;    // every call's work is ordered on this stream
    hipStream_t aux = nullptr;       // wall-adjacent kernel beside the bulk kernel, joined back each step
    hipEvent_t fork = nullptr, join = nullptr;
    hipStream_t aux2 = nullptr;      // the rotor-cell kernel, beside the wall-adjacent one
    hipEvent_t join2 = nullptr;
    // Tune::overlap 2: level L >= 1 sub-cycles on levelStream[L], forked from
    // the coarser level's stream before that level's step (the sub-cycle
    // needs the last interface transfer, not the coarse step) and joined back
    // before the next transfer
    hipStream_t levelStream[LBM3_MAX_LEVELS] = {};
    hipEvent_t levelFork[LBM3_MAX_LEVELS] = {}, levelJoin[LBM3_MAX_LEVELS] = {};
    lbm3_hooks hooks{};              // case hooks evaluated by the case (null callbacks: rule tables)
    // launch choices of the step (lbm3_autotune picks them by timing)
    struct Tune {
        int bulkLds = LBM3_BULK_LDS;  // dynamic LDS per bulk block: caps resident bulk blocks per CU
        int xcdRemap = LBM3_XCD_REMAP;
        // 1: wall / open / rotor kernels on the second stream beside the bulk
        // kernel; 2: also each finer level's sub-cycle on a stream of its own,
        // beside its coarser level's step; 3: the wall-adjacent cells in
        // blocks ahead of the bulk blocks of one launch, the other side kernels
        // on the second stream; 0: everything in order
        int overlap = 1;
        int openGroup = 0;            // cells per open-boundary group (0: open_group_cells of the storage)
        // cells on the x faces: 0 as the others; 1 stepped by the bulk kernel
        // and completed by a pass after streaming; 2 (with openGroup 1) in
        // groups of the open-cell kernel while the bulk kernel takes the others;
        // 3 (with openGroup 1) in the bulk kernel with its lanes rotated so that
        // a row's two x-face cells share a wave
        int openSplit = 0;
        int rotorInBulk = 0;          // the rotor cells stepped by the bulk kernel instead of their own
    } tune;
    // graph capture: each level step takes the next slot of this device table
    // for the values that change between replays; a dry run only collects them
    StepParam *captureTable = nullptr;
    int captureSlot = 0;
    std::vector<StepParam> *dryRun = nullptr;
    DBuf<double> partial;            // reduction scratch
    DBuf<double> scalar;
    int watchdog = 0;                // lbm3_run checks for non-finite populations every this many coarse steps (0: off)
    std::vector<std::vector<int>> stlEdges;   // shared-edge counts [3 T] of each static body of the last build
    // timing of the level-0 collide-stream launches
    bool timing = false;
    std::vector<std::array<hipEvent_t, 3>> ev;   // before bulk, between, after wall kernel
    int evUsed = 0;
};

size_t storage_bytes(int storage);
size_t compute_bytes(int storage);
int finish_level_setup(Sim &s, int level);  // derived tables after a level's arrays exist
int set_error(int code, const std::string &msg);
int check_hip(hipError_t e, const char *what);

} // namespace lbm3

struct lbm3_sim {
    lbm3::Sim s;
};
