// Cell-level helpers a case file calls in its hooks and reductions
// (reference include/cellFunctions.h:21-301): compressed (i, j, k) and
// neighbour decode, cell centres, outer normal, equilibrium, macroscopic
// state and the conversions to physical units.
#pragma once
#include "types.h"
#include "D3Q27Directions.h"
#include "NBRFunctions.h"
#include "esotwistStreamingFunctions.h"

// Rows of x-consecutive cells: shifter >= 0 is the cell's row (the row's first
// cell), shifter < 0 the offset back to the row's first cell
// (cellFunctions.h:21-94, gridBuilderFunctions.h:1107-1125).
__host__ __device__ inline void lbm3_row_of(const int &cell, IntConstViewType &shifterView, int &row, int &back)
{
    const int shift = shifterView(cell);
    row = shift >= 0 ? shift : shifterView(cell + shift);
    back = shift >= 0 ? 0 : -shift;
}

__host__ __device__ inline void getCompressedIJKNBR(const int &cell, int &iCell, int &jCell, int &kCell, NBRStruct &NBR,
                                                    IntConstViewType &shifterView, IntConstViewType &iView,
                                                    IntConstViewType &jView, IntConstViewType &kView,
                                                    IntConstViewType &jPlusView, IntConstViewType &kPlusView,
                                                    IntConstViewType &jkPlusView, const InfoStruct &Info)
{
    int row, back;
    lbm3_row_of(cell, shifterView, row, back);
    iCell = iView(row) + back;
    jCell = jView(row);
    kCell = kView(row);
    NBR.self = cell;
    NBR.jPlus = jPlusView(row) + back;
    NBR.kPlus = kPlusView(row) + back;
    NBR.jkPlus = jkPlusView(row) + back;
    finishNBRPlus(NBR, Info);
}

__host__ __device__ inline void getCompressedIJK(const int &cell, int &iCell, int &jCell, int &kCell, IntConstViewType &shifterView,
                                                 IntConstViewType &iView, IntConstViewType &jView, IntConstViewType &kView,
                                                 const InfoStruct &Info)
{
    int row, back;
    lbm3_row_of(cell, shifterView, row, back);
    iCell = iView(row) + back;
    jCell = jView(row);
    kCell = kView(row);
}

__host__ __device__ inline void getCompressedNBR(const int &cell, NBRStruct &NBR, IntConstViewType &shifterView,
                                                 IntConstViewType &jPlusView, IntConstViewType &kPlusView,
                                                 IntConstViewType &jkPlusView, const InfoStruct &Info)
{
    int row, back;
    lbm3_row_of(cell, shifterView, row, back);
    NBR.self = cell;
    NBR.jPlus = jPlusView(row) + back;
    NBR.kPlus = kPlusView(row) + back;
    NBR.jkPlus = jkPlusView(row) + back;
    finishNBRPlus(NBR, Info);
}

// cell centre in mm, x = i * res + ox; never contracted into an FMA, so the
// host, the device and the library's own rule tables round alike
__host__ __device__ inline void getXYZFromIJKCellIndex(const int &iCell, const int &jCell, const int &kCell, float &x, float &y,
                                                       float &z, const InfoStruct &Info)
{
#pragma clang fp contract(off)
    x = float(iCell) * Info.res + Info.ox;
    y = float(jCell) * Info.res + Info.oy;
    z = float(kCell) * Info.res + Info.oz;
}

__host__ __device__ inline void getIJKCellIndexFromXYZ(int &iCell, int &jCell, int &kCell, const float &x, const float &y,
                                                       const float &z, const InfoStruct &Info)
{
#pragma clang fp contract(off)
    iCell = (int)((x - Info.ox) / Info.res + 0.5f);
    jCell = (int)((y - Info.oy) / Info.res + 0.5f);
    kCell = (int)((z - Info.oz) / Info.res + 0.5f);
}

// outer normal of a boundary cell: -1 / +1 on the first / last index of each axis
__host__ __device__ inline void getOuterNormal(const int &iCell, const int &jCell, const int &kCell, int &outerNormalX,
                                               int &outerNormalY, int &outerNormalZ, const InfoStruct &Info)
{
    auto n = [](int v, int count) { return v == 0 ? -1 : (v == count - 1 ? 1 : 0); };
    outerNormalX = n(iCell, Info.cellCountX);
    outerNormalY = n(jCell, Info.cellCountY);
    outerNormalZ = n(kCell, Info.cellCountZ);
}

// shifted equilibrium f_eq - w (cellFunctions.h:125-148), the library's form
__host__ __device__ inline void getFeq(const float &dRho, const float &ux, const float &uy, const float &uz, float (&feq)[27])
{
    for (int q = 0; q < 27; ++q) feq[q] = lbm3::feq_shifted<float>(q, dRho, ux, uy, uz);
}

__host__ __device__ inline float getFeqSingle(const float &dRho, const float &ux, const float &uy, const float &uz, const int &direction)
{
    return lbm3::feq_shifted<float>(direction, dRho, ux, uy, uz);
}

__host__ __device__ inline void getFneq(const float (&f)[27], const float (&feq)[27], float (&fneq)[27])
{
    for (int q = 0; q < 27; ++q) fneq[q] = f[q] - feq[q];
}

// dRho = sum of the shifted populations, u = momentum / (1 + dRho)
// (cellFunctions.h:257-279), the library's summation
__host__ __device__ inline void getDRhoUxUyUz(float &dRho, float &ux, float &uy, float &uz, const float (&f)[27])
{
    lbm3::moments<float>(f, dRho, ux, uy, uz);
}

// lattice to physical units (cellFunctions.h:282-301); RHO_PHYS is the case's
__host__ __device__ inline void convertToPhysicalVelocity(float &ux, float &uy, float &uz, const InfoStruct &Info)
{
    ux = ux * (Info.res / 1000.f) / Info.dtPhys;
    uy = uy * (Info.res / 1000.f) / Info.dtPhys;
    uz = uz * (Info.res / 1000.f) / Info.dtPhys;
}

__host__ __device__ inline void convertToPhysicalPressure(float &dRho, const InfoStruct &Info)
{
    const float soundSpeed = INVSQRT3 * (Info.res / 1000.f) / Info.dtPhys;
    dRho = dRho * RHO_PHYS * soundSpeed * soundSpeed;
}

__host__ __device__ inline void convertToPhysicalForce(float &gx, float &gy, float &gz, const InfoStruct &Info)
{
    const float l = Info.res / 1000.f;
    const float s = RHO_PHYS * l * l * l * l / (Info.dtPhys * Info.dtPhys);
    gx *= s;
    gy *= s;
    gz *= s;
}
