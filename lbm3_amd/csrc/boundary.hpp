// Moment-based open boundary (reference include/boundaryConditions/applyMBBC.h,
// restoreRho.h, restoreUxUyUz.h and the non-reflective characteristic update of
// include/updateGrid.h:260-307).  One template instance per outer normal: the
// coefficient tables (generated/boundary_tables.h) are compile-time constants,
// so after unrolling every instance is straight-line register code.
#pragma once
#include "lattice.hpp"
#include "generated/boundary_tables.h"

namespace lbm3 {

template <class T> __host__ __device__ constexpr T rat(int num, int den) { return T(num) / T(den); }

// raw equilibrium moment `mid` (order of the generator's moment list)
template <class T>
__host__ __device__ __forceinline__ T eq_moment(int mid, T rho, T ux, T uy, T uz)
{
    constexpr T t = T(1) / T(3), n9 = T(1) / T(9);
    switch (mid) {
    case 0: return rho;
    case 1: return rho * ux;
    case 2: return rho * uy;
    case 3: return rho * uz;
    case 4: return rho * (t + ux * ux);
    case 5: return rho * (t + uy * uy);
    case 6: return rho * (t + uz * uz);
    case 7: return rho * uy * uz;
    case 8: return rho * ux * uz;
    case 9: return rho * ux * uy;
    case 10: return T(0);
    case 11: return rho * uy * (t + ux * ux);
    case 12: return rho * uz * (t + ux * ux);
    case 13: return rho * ux * (t + uy * uy);
    case 14: return rho * uz * (t + uy * uy);
    case 15: return rho * ux * (t + uz * uz);
    case 16: return rho * uy * (t + uz * uz);
    case 17: return rho * (n9 + t * (uy * uy + uz * uz) + uy * uy * uz * uz);
    case 18: return rho * (n9 + t * (ux * ux + uz * uz) + ux * ux * uz * uz);
    case 19: return rho * (n9 + t * (ux * ux + uy * uy) + ux * ux * uy * uy);
    case 20: return t * rho * uy * uz;
    case 21: return t * rho * ux * uz;
    default: return t * rho * ux * uy;
    }
}

// restoreRho: density from the known populations and the prescribed velocity
template <int N, class T>
__host__ __device__ __forceinline__ T restore_rho(const T (&f)[27], T ux, T uy, T uz)
{
    T num = T(0);
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if (LBM3_BT_RHO_F_NUM[N][q] != 0) num += rat<T>(LBM3_BT_RHO_F_NUM[N][q], LBM3_BT_RHO_F_DEN[N][q]) * f[q];
    const T den = rat<T>(LBM3_BT_RHO_S_NUM[N], LBM3_BT_RHO_S_DEN[N]) + rat<T>(LBM3_BT_RHO_U_NUM[N][0], LBM3_BT_RHO_U_DEN[N][0]) * ux +
                  rat<T>(LBM3_BT_RHO_U_NUM[N][1], LBM3_BT_RHO_U_DEN[N][1]) * uy +
                  rat<T>(LBM3_BT_RHO_U_NUM[N][2], LBM3_BT_RHO_U_DEN[N][2]) * uz;
    return num / den;
}

// restoreUxUyUz: velocity from the known populations and the prescribed density
template <int N, class T>
__host__ __device__ __forceinline__ void restore_u(const T (&f)[27], T rho, T &ux, T &uy, T &uz)
{
    constexpr int nx = LBM3_BT_NORMAL[N][0], ny = LBM3_BT_NORMAL[N][1], nz = LBM3_BT_NORMAL[N][2];
    if constexpr (N < 6) {
        // face: closed form (zero tangential k_nnt)
        T fZero = T(0), fPlus = T(0), jz[3] = {T(0), T(0), T(0)}, jp[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            constexpr int dummy = 0; (void)dummy;
            const int cn = nx * cx(q) + ny * cy(q) + nz * cz(q);
            const int c[3] = {cx(q), cy(q), cz(q)};
            if (cn == 0) { fZero += f[q]; for (int a = 0; a < 3; ++a) jz[a] += T(c[a]) * f[q]; }
            else if (cn == 1) { fPlus += f[q]; for (int a = 0; a < 3; ++a) jp[a] += T(c[a]) * f[q]; }
        }
        const T un = (fZero + T(2) * fPlus) / rho - T(1);
        const T den = fZero + T(2) * (un + T(1)) * rho * un;
        T u[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) u[a] = ((T(2) * un + T(1)) * jz[a] + T(4) * un * jp[a]) / den;
        constexpr int axis = nx ? 0 : (ny ? 1 : 2);
        u[axis] = (nx + ny + nz) > 0 ? un : -un;
        ux = u[0]; uy = u[1]; uz = u[2];
    } else {
        T s[3];
        const T rinv = T(1) / rho;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < 27; ++q)
                if (LBM3_BT_U_F_NUM[N][r][q] != 0) acc += rat<T>(LBM3_BT_U_F_NUM[N][r][q], LBM3_BT_U_F_DEN[N][r][q]) * f[q];
            s[r] = acc * rinv - rat<T>(LBM3_BT_U_Q_NUM[N][r], LBM3_BT_U_Q_DEN[N][r]);
        }
        T u[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (LBM3_BT_U_INV_NUM[N][d][j] != 0) acc += rat<T>(LBM3_BT_U_INV_NUM[N][d][j], LBM3_BT_U_INV_DEN[N][d][j]) * s[j];
            u[d] = acc;
        }
        ux = u[0]; uy = u[1]; uz = u[2];
    }
}

// applyMBBC: unknown populations from the prescribed equilibrium moments
template <int N, class T>
__host__ __device__ __forceinline__ void mbbc(T (&f)[27], T rho, T ux, T uy, T uz)
{
    constexpr int nu = LBM3_BT_MBBC_NU[N];
    T s[nu];
#pragma unroll
    for (int r = 0; r < nu; ++r) {
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < 27; ++q)
            if (LBM3_BT_MBBC_SMF[N][r][q] != 0) acc += T(LBM3_BT_MBBC_SMF[N][r][q]) * f[q];
        s[r] = eq_moment<T>(LBM3_BT_MBBC_MOMENT[N][r], rho, ux, uy, uz) - acc;
    }
#pragma unroll
    for (int r = 0; r < nu; ++r) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < nu; ++j)
            if (LBM3_BT_MBBC_INV_NUM[N][r][j] != 0) acc += rat<T>(LBM3_BT_MBBC_INV_NUM[N][r][j], LBM3_BT_MBBC_INV_DEN[N][r][j]) * s[j];
        f[LBM3_BT_MBBC_UNKNOWN[N][r]] = acc;
    }
}

__host__ __device__ constexpr bool unknown_dir(int nx, int ny, int nz, int q)
{
    return nx * cx(q) < 0 || ny * cy(q) < 0 || nz * cz(q) < 0;
}

struct OpenBCRecord {      // a resolved getOpenBC result
    float dRho, ux, uy, uz;
    float beta;
    int dirichletU, dirichletRho, nonReflective;
    int impedance;         // the case's non-reflective scheme is LBM3_OPEN_IMPEDANCE
    float tolerance;       // its rhoReflectionTolerance
};

// std::clamp(v, lo, hi)
template <class T> __host__ __device__ __forceinline__ T clamp_to(T v, T lo, T hi)
{
    return v < lo ? lo : (hi < v ? hi : v);
}

// One open-boundary cell, normal table index N.  f holds the known
// pre-collision populations (shifted); unknown entries are overwritten.
// Returns the boundary state (dRho, normal velocity) for the reference update.
template <int N, class T>
__host__ __device__ __forceinline__ void open_cell(T (&f)[27], const OpenBCRecord &R, T dRhoRef, T uNormalRef,
                                                   T &dRhoOut, T &uNormalOut)
{
    constexpr int nx = LBM3_BT_NORMAL[N][0], ny = LBM3_BT_NORMAL[N][1], nz = LBM3_BT_NORMAL[N][2];
    constexpr bool face = N < 6;
    T dRhoZ = T(0);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        if (unknown_dir(nx, ny, nz, q)) continue;
        const int p = nx * cx(q) + ny * cy(q) + nz * cz(q);
        dRhoZ += p == 0 ? f[q] : T(2) * f[q];
        f[q] += weight<T>(q);   // the open boundary works on full populations
    }
    T dRho = T(R.dRho), ux = T(R.ux), uy = T(R.uy), uz = T(R.uz);
    const bool nonRefl = face && R.nonReflective;
    constexpr T cs = T(0.57735026918962576);
    constexpr int axis = nx ? 0 : (ny ? 1 : 2);
    constexpr T sgn = T(nx + ny + nz);
    if (R.impedance) {
        // impedance scheme (updateGrid_SCHLAFFER_IMPEDANCE_SCHEME.h:250-311): the
        // reference state is the previous completed one; the target is clamped
        // to within the tolerance of the non-reflecting density
        const T tol = T(R.tolerance);
        const bool useNR = nonRefl && tol < T(1);
        T dRhoNR = T(0);
        if (useNR) {
            const T cs2 = cs * cs;
            const T bracket = uNormalRef + cs + T(1);
            const T t = uNormalRef - dRhoZ;
            const T D = fma(bracket, bracket, cs2);
            const T B = fma(bracket, cs + t, -cs2 * dRhoRef);
            const T H = fma(-t, T(2) * cs + t, T(2) * cs2 * dRhoRef);
            const T disc = fma(D, H, B * B);
            dRhoNR = H / (B + sqrt(disc));
        }
        if (R.dirichletRho) {
            if (useNR) dRho = clamp_to(dRho, dRhoNR - tol, dRhoNR + tol);
            restore_u<N>(f, dRho + T(1), ux, uy, uz);
        } else if (R.dirichletU) {
            if (useNR) {
                const T uMin = (dRhoZ - dRhoNR - tol) / (dRhoNR + T(1) + tol);
                const T uMax = (dRhoZ - dRhoNR + tol) / (dRhoNR + T(1) - tol);
                T uvec[3] = {ux, uy, uz};
                uvec[axis] = sgn > T(0) ? clamp_to(uvec[axis], uMin, uMax) : clamp_to(uvec[axis], -uMax, -uMin);
                ux = uvec[0]; uy = uvec[1]; uz = uvec[2];
            }
            dRho = restore_rho<N>(f, ux, uy, uz) - T(1);
        }
    } else if (R.dirichletRho) {
        if (nonRefl) {
            const T rhoT = dRho + T(1);
            const T rt = (dRhoZ - dRho) / rhoT;
            const T a = T(1) + cs - uNormalRef;
            const T un = T(2) * (uNormalRef + cs * rt) / (a + sqrt(a * a + T(4) * (uNormalRef + cs * rt)));
            dRho = (dRhoZ - un) / (T(1) + un);
        }
        restore_u<N>(f, dRho + T(1), ux, uy, uz);
    } else if (R.dirichletU) {
        if (nonRefl) {
            T uvec[3] = {ux, uy, uz};
            const T target = sgn * uvec[axis];
            const T rt = (dRhoZ - dRhoRef) / (dRhoRef + T(1));
            const T a = T(1) + cs - target;
            const T un = T(2) * (target + cs * rt) / (a + sqrt(a * a + T(4) * (target + cs * rt)));
            uvec[axis] = sgn * un;
            ux = uvec[0]; uy = uvec[1]; uz = uvec[2];
        }
        dRho = restore_rho<N>(f, ux, uy, uz) - T(1);
    }
    mbbc<N>(f, dRho + T(1), ux, uy, uz);
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if (unknown_dir(nx, ny, nz, q)) f[q] -= weight<T>(q);
    dRhoOut = dRho;
    uNormalOut = T(nx) * ux + T(ny) * uy + T(nz) * uz;
}

// table index of an outer normal
__host__ __device__ __forceinline__ int normal_index(int nx, int ny, int nz)
{
    const int code = (nx + 5) * 100 + (ny + 5) * 10 + (nz + 5);
    int n = 0;
    while (n < 26 && LBM3_BT_NORMAL_CODE[n] != code) ++n;
    return n;
}

// per open cell: bits 0-4 the normal's table index, bit 5 set where the cell's
// flow is tracked, bits 6-31 the index of its boundary-condition record
constexpr uint32_t kOpenTrackBit = 1u << 5;
__host__ __device__ __forceinline__ uint32_t open_meta(int normal, int record, bool track)
{
    return (uint32_t)normal | (track ? kOpenTrackBit : 0u) | ((uint32_t)record << 6);
}
__host__ __device__ __forceinline__ uint32_t open_meta_record(uint32_t meta) { return meta >> 6; }

} // namespace lbm3
