// Case-facing types of LBM3 (reference include/types.h:38-289) on the MI355X
// library: the same struct and field names a case file reads, with the grid
// arrays as views of the simulation's device buffers (liblbm3mi, include/
// lbm3mi.h).  A case file includes this header first, as in the reference
// (taylorGreenVortexMain.cu:46-49).
#pragma once

#include <iostream>
#include <sstream>
#include <cmath>
#include <fstream>
#include <cstdlib>
#include <cstdint>
#include <limits>
#include <climits>
#include <string>
#include <vector>
#include <algorithm>
#include <new>
#include <array>
#include <cstdio>
#include <stdexcept>
#include <unistd.h>
#include <initializer_list>
#include <utility>

#include "lbm3_tnl.h"
#include "lbm3mi.h"
#include "lattice.hpp"
#include "cumulant.hpp"

//------------------------------------------------------------------------------------
// arrays (types.h:38-112)
//------------------------------------------------------------------------------------

using BoolArrayType = TNL::Containers::Vector<bool, TNL::Devices::Cuda, size_t>;
using BoolArrayTypeCPU = TNL::Containers::Vector<bool, TNL::Devices::Host, size_t>;
using IntArrayType = TNL::Containers::Vector<int, TNL::Devices::Cuda, size_t>;
using IntConstViewType = IntArrayType::ConstViewType;
using IntArrayTypeCPU = TNL::Containers::Vector<int, TNL::Devices::Host, size_t>;
using IntPairType = TNL::Containers::StaticArray<2, int>;
using IntTripleType = TNL::Containers::StaticArray<3, int>;
using Uint8_tArrayType = TNL::Containers::Vector<uint8_t, TNL::Devices::Cuda, size_t>;
using Uint32_tArrayType = TNL::Containers::Vector<uint32_t, TNL::Devices::Cuda, size_t>;
using Uint32_tConstViewType = Uint32_tArrayType::ConstViewType;
using FloatArrayType = TNL::Containers::Vector<float, TNL::Devices::Cuda, size_t>;
using FloatViewType = FloatArrayType::ViewType;
using FloatArrayTypeCPU = TNL::Containers::Vector<float, TNL::Devices::Host, size_t>;
using LongLongArrayType = TNL::Containers::Vector<long long, TNL::Devices::Cuda, size_t>;
using FloatArray2DTypeCPU = TNL::Containers::HostArray2D<float>;
using IntArray2DTypeCPU = TNL::Containers::HostArray2D<int>;
using SixFloatArrayType = TNL::Containers::StaticArray<6, float>;

// The population array of a level: fView(q, cell) is the shifted population
// f_q - w_q of the cell's slot q, as in the reference's NDArray (27, cellCount)
// (types.h:90-93); the simulation stores it in tiles of `tile` cells
// (lbm3_population_layout), which the view's index follows.
template <class T> struct PopulationView {
    T *p = nullptr;
    int tileShift = 8;
    __host__ __device__ T &operator()(int q, int cell) const
    {
        const int tile = 1 << tileShift;
        return p[((size_t)(cell >> tileShift) * 27 + q) * (size_t)tile + (cell & (tile - 1))];
    }
};

// Taking a view settles the level first: a boundary condition the library
// applies when the next step loads the cell is applied now, so the view shows
// what the reference's array holds after updateAllGrids.
struct FloatArray2DType {
    using ViewType = PopulationView<float>;
    using ConstViewType = PopulationView<const float>;
    lbm3_sim *sim = nullptr;
    int level = 0;
    int tileShift = 8;
    float *settled() const
    {
        if (!sim) return nullptr;
        void *p = nullptr;
        long long n = 0;
        if (lbm3_device_array(sim, level, "f", &p, &n) != 0) throw std::runtime_error(lbm3_last_error());
        return static_cast<float *>(p);
    }
    ViewType getView() const { return ViewType{settled(), tileShift}; }
    ConstViewType getConstView() const { return ConstViewType{settled(), tileShift}; }
};

//------------------------------------------------------------------------------------
// structs (types.h:114-289)
//------------------------------------------------------------------------------------

struct BoundsStruct { float xMin = 0.f; float yMin = 0.f; float zMin = 0.f; float xMax = 0.f; float yMax = 0.f; float zMax = 0.f;
                      float rxMax = 0.f; float ryMax = 0.f; float rzMax = 0.f; };

struct InfoStruct { float gridID = 0; unsigned long long gridMemoryBytes = 0LL;
                    int iterationsFinished = 0; int iterationsSinceTrackerReset = 0;
                    float res = 1.f; float ox = 0.f; float oy = 0.f; float oz = 0.f;
                    BoundsStruct Bounds;
                    float nu = 1.f; float dtPhys = 1.f;
                    int cellCountX = 0; int cellCountY = 0; int cellCountZ = 0;
                    int cellCount = 0;
                    bool esotwistFlipper = 0;
                    bool useRotors = true;
                    float iRegulatorInlet = 0.f; float iRegulatorOutlet = 0.f; };

struct BCStruct { float dRho = 0.f; float ux = 0.f; float uy = 0.f; float uz = 0.f;
                  float gx = 0.f; float gy = 0.f; float gz = 0.f;
                  int wallID = -1; int openBCID = 0; bool useIBB = true;
                  bool dirichletU = false; bool dirichletRho = false; bool periodic = false;
                  bool nonReflective = true; float nonReflectiveBeta = 7.f;
                  float collisionLimiter = 0.05f; float nuMultiplier = 1.f; };

struct NBRStruct { int self;
                   int iPlus; int jPlus; int kPlus; int ijPlus; int ikPlus; int jkPlus; int ijkPlus;
                   int iMinus; int jMinus; int kMinus; };

struct CompressedIJKNBRStruct { IntArrayType shifterArray;
                                IntArrayType iArray; IntArrayType jArray; IntArrayType kArray;
                                IntArrayType jPlusArray; IntArrayType kPlusArray; IntArrayType jkPlusArray; };

struct InterfaceStruct { int interfaceCount = 0; IntArrayType indexArray; IntArrayType childMapArray;
                         int leftoverCount = 0; IntArrayType leftoverIndexArray; IntArrayType leftoverParentMapArray;
                         IntArrayType leftoverNbrIArray; IntArrayType leftoverNbrJArray; IntArrayType leftoverNbrKArray; };

// wallMapArray: -3 the cell is a wall, -2 free fluid under a parent interface
// or in a periodic overlap (no force tracking), -1 free fluid, >= 0 entry of a
// wall-adjacent cell (types.h:202-205).  The momentum-exchange forces are kept
// in FP64 by the library (lbm3_wall_forces), not in gx/gy/gzArray.
struct WallStruct { int wallAdjacentCount = 0; IntArrayType indexArray; IntArrayType wallMapArray; Uint32_tArrayType wallDataArray;
                    Uint32_tArrayType linkLengthArray;
                    FloatArrayType gxArray; FloatArrayType gyArray; FloatArrayType gzArray; };

struct RotorInfoStruct { float rotorID = 0; float radiansPerSecond = 0.f; double radiansElapsed = 0.f;
                         BoundsStruct Bounds; float res = 1.f; int cellCountX = 0; int cellCountY = 0; int cellCountZ = 0;
                         float ox = 0.f; float oy = 0.f; float oz = 0.f;
                         bool rotateAlongX = false; bool rotateAlongY = false; bool rotateAlongZ = false; };

struct RotorStruct { RotorInfoStruct Info;
                     IntArrayType indexArray; IntArrayType rotorMapArray; Uint32_tArrayType interpolationArray;
                     FloatArrayType gxArray; FloatArrayType gyArray; FloatArrayType gzArray;
                     // what the library was last told of the rotor's motion
                     float sentRadiansPerSecond = 0.f; int sentAxis = -2; };

struct OpenBCArrayStruct { int openBCID = 0; int openBCCount = 0; int trackFlowCount = 0;
                           IntArrayType indexArray; IntArrayType sourceArray;
                           FloatArrayType dRhoRefArray; FloatArrayType uNormalRefArray;
                           FloatArrayType dRhoCumulativeArray; FloatArrayType uNormalCumulativeArray; };

struct PeriodicBCStruct { int periodicBCCount = 0; IntArrayType indexArray; IntArrayType sourceArray; };

struct GridStruct { InfoStruct Info;
                    FloatArray2DType fArray;
                    CompressedIJKNBRStruct IJKNBR;
                    WallStruct Wall;
                    std::vector<RotorStruct> rotors;
                    InterfaceStruct CoarseToFineInterface; InterfaceStruct FineToCoarseInterface;
                    std::vector<OpenBCArrayStruct> openBCs; PeriodicBCStruct PeriodicBC;
                    // the simulation that owns every level's buffers, this level's index in
                    // it, and the host-side controls of Info last handed to it
                    lbm3_sim *sim = nullptr; int level = 0;
                    float sentIRegulatorInlet = 0.f; float sentIRegulatorOutlet = 0.f; bool sentUseRotors = true; };

// a binary STL read on the host (readSTL): triangles as 9 floats each
struct STLStructCPU { int triangleCount = 0;
                      FloatArrayTypeCPU axArray; FloatArrayTypeCPU ayArray; FloatArrayTypeCPU azArray;
                      FloatArrayTypeCPU bxArray; FloatArrayTypeCPU byArray; FloatArrayTypeCPU bzArray;
                      FloatArrayTypeCPU cxArray; FloatArrayTypeCPU cyArray; FloatArrayTypeCPU czArray;
                      BoundsStruct Bounds; };

struct STLStruct { int triangleCount = 0;
                   BoundsStruct Bounds;
                   std::vector<float> triangles;   // 9 floats per triangle: a, b, c vertices
                   STLStruct() = default; };

struct TrackerStruct { int iterationsFinished = 0; int openBCCount = 0; int wallCount = 0; int rotorCount = 0; float averagePercent = 30.f;
                       FloatArray2DTypeCPU normalVelocityArray; FloatArray2DTypeCPU massFlowArray; FloatArray2DTypeCPU momentumThrustArray;
                       FloatArray2DTypeCPU pressureArray; FloatArray2DTypeCPU pressurePowerArray; FloatArray2DTypeCPU normalKineticPowerArray;
                       FloatArray2DTypeCPU wallFxArray; FloatArray2DTypeCPU wallFyArray; FloatArray2DTypeCPU wallFzArray;
                       FloatArray2DTypeCPU wallTxArray; FloatArray2DTypeCPU wallTyArray; FloatArray2DTypeCPU wallTzArray;
                       FloatArray2DTypeCPU rotorFxArray; FloatArray2DTypeCPU rotorFyArray; FloatArray2DTypeCPU rotorFzArray;
                       FloatArray2DTypeCPU rotorTxArray; FloatArray2DTypeCPU rotorTyArray; FloatArray2DTypeCPU rotorTzArray;
                       bool TRACK_CUSTOM_VARIABLES = false; FloatArray2DTypeCPU customArray; std::vector<std::string> customNames;
                       std::vector<std::string> customUnits;
                       FloatArrayTypeCPU normalVelocity; FloatArrayTypeCPU massFlow; FloatArrayTypeCPU momentumThrust;
                       FloatArrayTypeCPU pressure; FloatArrayTypeCPU pressurePower; FloatArrayTypeCPU normalKineticPower;
                       FloatArrayTypeCPU wallFx; FloatArrayTypeCPU wallFy; FloatArrayTypeCPU wallFz;
                       FloatArrayTypeCPU wallTx; FloatArrayTypeCPU wallTy; FloatArrayTypeCPU wallTz;
                       FloatArrayTypeCPU rotorFx; FloatArrayTypeCPU rotorFy; FloatArrayTypeCPU rotorFz;
                       FloatArrayTypeCPU rotorTx; FloatArrayTypeCPU rotorTy; FloatArrayTypeCPU rotorTz; };

struct FlowReportStruct { float normalVelocity = 0.f; float massFlow = 0.f; float momentumThrust = 0.f;
                          float pressure = 0.f; float pressurePower = 0.f; float normalKineticPower = 0.f; float areamm2 = 0.f; };

//------------------------------------------------------------------------------------
// conversions to and from the C ABI's plain structs
//------------------------------------------------------------------------------------

namespace lbm3facade {

inline void check(int rc, const char *what)
{
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + lbm3_last_error());
}

__host__ __device__ inline InfoStruct info_from(const lbm3_info &a)
{
    InfoStruct I;
    I.gridID = a.gridID; I.gridMemoryBytes = a.gridMemoryBytes;
    I.iterationsFinished = a.iterationsFinished; I.iterationsSinceTrackerReset = a.iterationsSinceTrackerReset;
    I.res = a.res; I.ox = a.ox; I.oy = a.oy; I.oz = a.oz;
    I.Bounds.xMin = a.Bounds.xMin; I.Bounds.yMin = a.Bounds.yMin; I.Bounds.zMin = a.Bounds.zMin;
    I.Bounds.xMax = a.Bounds.xMax; I.Bounds.yMax = a.Bounds.yMax; I.Bounds.zMax = a.Bounds.zMax;
    I.Bounds.rxMax = a.Bounds.rxMax; I.Bounds.ryMax = a.Bounds.ryMax; I.Bounds.rzMax = a.Bounds.rzMax;
    I.nu = a.nu; I.dtPhys = a.dtPhys;
    I.cellCountX = a.cellCountX; I.cellCountY = a.cellCountY; I.cellCountZ = a.cellCountZ; I.cellCount = a.cellCount;
    I.esotwistFlipper = a.esotwistFlipper != 0; I.useRotors = a.useRotors != 0;
    I.iRegulatorInlet = a.iRegulatorInlet; I.iRegulatorOutlet = a.iRegulatorOutlet;
    return I;
}

__host__ __device__ inline lbm3_bc bc_to(const BCStruct &b)
{
    lbm3_bc r;
    r.dRho = b.dRho; r.ux = b.ux; r.uy = b.uy; r.uz = b.uz;
    r.gx = b.gx; r.gy = b.gy; r.gz = b.gz;
    r.wallID = b.wallID; r.openBCID = b.openBCID; r.useIBB = b.useIBB;
    r.dirichletU = b.dirichletU; r.dirichletRho = b.dirichletRho; r.periodic = b.periodic; r.nonReflective = b.nonReflective;
    r.nonReflectiveBeta = b.nonReflectiveBeta; r.collisionLimiter = b.collisionLimiter; r.nuMultiplier = b.nuMultiplier;
    return r;
}

__host__ __device__ inline BCStruct bc_from(const lbm3_bc &r)
{
    BCStruct b;
    b.dRho = r.dRho; b.ux = r.ux; b.uy = r.uy; b.uz = r.uz;
    b.gx = r.gx; b.gy = r.gy; b.gz = r.gz;
    b.wallID = r.wallID; b.openBCID = r.openBCID; b.useIBB = r.useIBB != 0;
    b.dirichletU = r.dirichletU != 0; b.dirichletRho = r.dirichletRho != 0; b.periodic = r.periodic != 0;
    b.nonReflective = r.nonReflective != 0; b.nonReflectiveBeta = r.nonReflectiveBeta;
    b.collisionLimiter = r.collisionLimiter; b.nuMultiplier = r.nuMultiplier;
    return b;
}

inline lbm3_bounds bounds_to(const BoundsStruct &b)
{
    lbm3_bounds r;
    r.xMin = b.xMin; r.yMin = b.yMin; r.zMin = b.zMin; r.xMax = b.xMax; r.yMax = b.yMax; r.zMax = b.zMax;
    r.rxMax = b.rxMax; r.ryMax = b.ryMax; r.rzMax = b.rzMax;
    return r;
}

} // namespace lbm3facade
