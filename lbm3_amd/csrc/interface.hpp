// Coarse <-> fine transfer of the multilevel grid: quadratic velocity /
// trilinear density interpolation from a 2x2x2 stencil of cells, with the
// second-order (shear) moments carried across levels by the rate of strain.
//
// Method: Schoenherr 2015 (dissertation, ch. 7) as used by the reference
// (include/updateInterface.h:29-884).  Coordinates are in units of the
// stencil spacing with corner n of the stencil at
//   ((n & 1) - 1/2, ((n >> 1) & 1) - 1/2, ((n >> 2) & 1) - 1/2),
// i.e. corners self, i+, j+, ij+, k+, ik+, jk+, ijk+ for n = 0..7.  Each
// corner carries its density, velocity and rate-of-strain estimates
//   sxy = du/dy + dv/dx, syz, sxz,  nxy = du/dx - dv/dy,  nxz = du/dx - dw/dz
// (from its non-equilibrium second moments).  The interpolant is
//   rho: trilinear;  u_c: trilinear + sum_a q_ca x_a^2,
// whose trilinear part matches the corner values and whose second
// derivatives q_ca come from the corner strain rates:
//   c != a:  2 q_ca = d_a s_ac - d_a d_c u_a
//   c == a:  2 q_aa = d_a d_a u_a from d_a (n-combination) and the cross terms.
#pragma once
#include "lattice.hpp"
#include "cumulant.hpp"

namespace lbm3 {

// the nine stencil quantities of one cell
template <class T> struct CellValues {
    T d, u[3];
    T sxy, syz, sxz, nxy, nxz;
};

// dRho, u and the strain estimates of a cell from its shifted populations
// (updateInterface.h:196-231); w1 is omega1 of the cell's level
template <class T>
__host__ __device__ __forceinline__ CellValues<T> cell_values(const T (&f)[27], T w1)
{
    CellValues<T> v;
    moments(f, v.d, v.u[0], v.u[1], v.u[2]);
    const T rho = T(1) + v.d;
    T pxy = 0, pyz = 0, pxz = 0, pxxyy = 0, pxxzz = 0;
#pragma unroll
    for (int q = 1; q < 27; ++q) {
        const int x = cx(q), y = cy(q), z = cz(q);
        if (x * y > 0) pxy += f[q]; else if (x * y < 0) pxy -= f[q];
        if (y * z > 0) pyz += f[q]; else if (y * z < 0) pyz -= f[q];
        if (x * z > 0) pxz += f[q]; else if (x * z < 0) pxz -= f[q];
        const int dxy = x * x - y * y, dxz = x * x - z * z;   // the weight background cancels in these
        if (dxy > 0) pxxyy += f[q]; else if (dxy < 0) pxxyy -= f[q];
        if (dxz > 0) pxxzz += f[q]; else if (dxz < 0) pxxzz -= f[q];
    }
    const T ux = v.u[0], uy = v.u[1], uz = v.u[2];
    v.sxy = -T(3) * w1 * (pxy / rho - ux * uy);
    v.syz = -T(3) * w1 * (pyz / rho - uy * uz);
    v.sxz = -T(3) * w1 * (pxz / rho - ux * uz);
    v.nxy = -T(1.5) * w1 * (pxxyy / rho - (ux * ux - uy * uy));
    v.nxz = -T(1.5) * w1 * (pxxzz / rho - (ux * ux - uz * uz));
    return v;
}

// sum over the corners of sign(mask, n) * v[n], sign = product over the axes
// in mask (x = 1, y = 2, z = 4) of (+1 if the corner is on the + side else -1)
template <int MASK, class T> __host__ __device__ __forceinline__ T corner_sum(const T (&v)[8])
{
    T s = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int neg = __builtin_popcount(MASK & ~n) & 1;
        s = neg ? s - v[n] : s + v[n];
    }
    return s;
}

// interpolant of one stencil
template <class T> struct Interpolant {
    T d[8];          // rho: 1, x, y, z, xy, xz, yz, xyz
    T u[3][11];      // u_c: 1, x, y, z, xy, xz, yz, xyz, x^2, y^2, z^2
};

template <class T>
__host__ __device__ inline Interpolant<T> interpolant(const CellValues<T> (&cv)[8])
{
    Interpolant<T> P;
    T D[8], U[3][8], SXY[8], SYZ[8], SXZ[8], NXY[8], NXZ[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        D[n] = cv[n].d;
        U[0][n] = cv[n].u[0]; U[1][n] = cv[n].u[1]; U[2][n] = cv[n].u[2];
        SXY[n] = cv[n].sxy; SYZ[n] = cv[n].syz; SXZ[n] = cv[n].sxz; NXY[n] = cv[n].nxy; NXZ[n] = cv[n].nxz;
    }
    // trilinear parts: corners at +-1/2, so a corner sum with k axes is 8 (1/2)^k times the coefficient
    P.d[0] = T(0.125) * corner_sum<0>(D);
    P.d[1] = T(0.25) * corner_sum<1>(D); P.d[2] = T(0.25) * corner_sum<2>(D); P.d[3] = T(0.25) * corner_sum<4>(D);
    P.d[4] = T(0.5) * corner_sum<3>(D); P.d[5] = T(0.5) * corner_sum<5>(D); P.d[6] = T(0.5) * corner_sum<6>(D);
    P.d[7] = corner_sum<7>(D);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        P.u[c][1] = T(0.25) * corner_sum<1>(U[c]);
        P.u[c][2] = T(0.25) * corner_sum<2>(U[c]);
        P.u[c][3] = T(0.25) * corner_sum<4>(U[c]);
        P.u[c][4] = T(0.5) * corner_sum<3>(U[c]);
        P.u[c][5] = T(0.5) * corner_sum<5>(U[c]);
        P.u[c][6] = T(0.5) * corner_sum<6>(U[c]);
        P.u[c][7] = corner_sum<7>(U[c]);
    }
    // second derivatives from the strain rates: a corner sum over axis a of a
    // corner quantity is 4 times its a-derivative
    const T s8 = T(0.125), s16 = T(0.0625), q4 = T(0.25);
    // off-diagonal: u_c along a^2, c != a  ->  (d_a s_ac - d_a d_c u_a) / 2
    P.u[1][8] = s8 * corner_sum<1>(SXY) - q4 * corner_sum<3>(U[0]);    // v, x^2
    P.u[2][8] = s8 * corner_sum<1>(SXZ) - q4 * corner_sum<5>(U[0]);    // w, x^2
    P.u[0][9] = s8 * corner_sum<2>(SXY) - q4 * corner_sum<3>(U[1]);    // u, y^2
    P.u[2][9] = s8 * corner_sum<2>(SYZ) - q4 * corner_sum<6>(U[1]);    // w, y^2
    P.u[0][10] = s8 * corner_sum<4>(SXZ) - q4 * corner_sum<5>(U[2]);   // u, z^2
    P.u[1][10] = s8 * corner_sum<4>(SYZ) - q4 * corner_sum<6>(U[2]);   // v, z^2
    // diagonal: 2 d_a u_a - d_b u_b - d_c u_c is a combination of the n-rates; the
    // mixed derivatives supply the rest
    T NX[8], NY[8], NZ[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        NX[n] = NXY[n] + NXZ[n];              // 2 du/dx - dv/dy - dw/dz
        NY[n] = NXZ[n] - T(2) * NXY[n];       // 2 dv/dy - du/dx - dw/dz
        NZ[n] = NXY[n] - T(2) * NXZ[n];       // 2 dw/dz - du/dx - dv/dy
    }
    P.u[0][8] = s16 * corner_sum<1>(NX) + s8 * (corner_sum<3>(U[1]) + corner_sum<5>(U[2]));
    P.u[1][9] = s16 * corner_sum<2>(NY) + s8 * (corner_sum<3>(U[0]) + corner_sum<6>(U[2]));
    P.u[2][10] = s16 * corner_sum<4>(NZ) + s8 * (corner_sum<5>(U[0]) + corner_sum<6>(U[1]));
    // constant terms: the corner mean minus the quadratic part at the corners (x^2 = 1/4)
#pragma unroll
    for (int c = 0; c < 3; ++c)
        P.u[c][0] = T(0.125) * corner_sum<0>(U[c]) - q4 * (P.u[c][8] + P.u[c][9] + P.u[c][10]);
    return P;
}

// value and gradient of u at (x, y, z), density at (x, y, z)
template <class T>
__host__ __device__ __forceinline__ void evaluate(const Interpolant<T> &P, T x, T y, T z, T &d, T (&u)[3], T (&g)[3][3])
{
    d = P.d[0] + P.d[1] * x + P.d[2] * y + P.d[3] * z + P.d[4] * x * y + P.d[5] * x * z + P.d[6] * y * z + P.d[7] * x * y * z;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T *a = P.u[c];
        u[c] = a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * y + a[5] * x * z + a[6] * y * z + a[7] * x * y * z +
               a[8] * x * x + a[9] * y * y + a[10] * z * z;
        g[c][0] = a[1] + a[4] * y + a[5] * z + a[7] * y * z + T(2) * a[8] * x;
        g[c][1] = a[2] + a[4] * x + a[6] * z + a[7] * x * z + T(2) * a[9] * y;
        g[c][2] = a[3] + a[5] * x + a[6] * y + a[7] * x * y + T(2) * a[10] * z;
    }
}

// populations of a cell with density dRho, velocity u and rate of strain
// (sxy, syz, sxz, nxy, nxz) in the target level's units, scaled by
// sigma / omega1 (updateInterface.h:334-344, :645-661); third and higher
// cumulants at equilibrium (updateInterface.h:29-153)
template <class T>
__host__ __device__ __forceinline__ void reconstruct_k(T (&f)[27], T dRho, const T (&u)[3], T k011, T k101, T k110, T mxy,
                                                       T mxz)
{
    const T third = T(1) / T(3);
    const T k200 = third * (mxy + mxz + dRho);
    const T k020 = third * (-T(2) * mxy + mxz + dRho);
    const T k002 = third * (mxy - T(2) * mxz + dRho);
    const T k4 = dRho * (T(1) / T(9)), k6 = dRho * (T(1) / T(27));
    const CentralMoments<T> M{dRho, T(0), k200, T(0), k101, T(0), k002, T(0), k4,
                              T(0), k110, T(0), k011, T(0), T(0), T(0), T(0), T(0),
                              k020, T(0), k4, T(0), T(0), T(0), k4, T(0), k6};
    from_central_moments(f, M, u[0], u[1], u[2]);
}

template <class T>
__host__ __device__ __forceinline__ void reconstruct(T (&f)[27], T dRho, const T (&u)[3], T sxy, T syz, T sxz, T nxy, T nxz,
                                                     T sigmaOverOmega)
{
    const T rho = T(1) + dRho;
    const T third = T(1) / T(3);
    const T shear = -third * sigmaOverOmega * rho;
    const T normal = -T(2) / T(3) * sigmaOverOmega * rho;
    reconstruct_k(f, dRho, u, shear * syz, shear * sxz, shear * sxy, normal * nxy, normal * nxz);
}

// ---- the linear interface (INTERFACE_VERSIONS_BACKUP/updateInterfaceLINEAR.h)

// affine fit c0 + cx x + cy y + cz z of the eight corner values (:155-180)
template <class T> struct Affine {
    T c0, cx, cy, cz;
    __host__ __device__ T at(T x, T y, T z) const { return c0 + cx * x + cy * y + cz * z; }
};

template <class T> __host__ __device__ __forceinline__ Affine<T> affine_fit(const T (&v)[8])
{
    Affine<T> p{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        p.c0 += T(0.125) * v[n];
        p.cx += ((n & 1) ? T(0.25) : T(-0.25)) * v[n];
        p.cy += ((n & 2) ? T(0.25) : T(-0.25)) * v[n];
        p.cz += ((n & 4) ? T(0.25) : T(-0.25)) * v[n];
    }
    return p;
}

// F->C: density and momentum averaged over the eight fine cells, the shear
// moments from their density-weighted strain (:265-299); sigma = 2
template <class T> __host__ __device__ __forceinline__ void linear_fine_to_coarse(T (&f)[27], const CellValues<T> (&cv)[8],
                                                                                 T w1c)
{
    T dRho = 0, j[3] = {0, 0, 0}, qxy = 0, qyz = 0, qxz = 0, qxy2 = 0, qxz2 = 0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const T mw = T(0.125) * (T(1) + cv[n].d);
        dRho += T(0.125) * cv[n].d;
        j[0] += mw * cv[n].u[0]; j[1] += mw * cv[n].u[1]; j[2] += mw * cv[n].u[2];
        qxy += mw * cv[n].sxy; qyz += mw * cv[n].syz; qxz += mw * cv[n].sxz;
        qxy2 += mw * cv[n].nxy; qxz2 += mw * cv[n].nxz;
    }
    const T rho = T(1) + dRho;
    const T u[3] = {j[0] / rho, j[1] / rho, j[2] / rho};
    const T ss = T(2) / w1c;
    const T third = -(T(1) / T(3)) * ss, twoThirds = -(T(2) / T(3)) * ss;
    reconstruct_k(f, dRho, u, third * qyz, third * qxz, third * qxy, twoThirds * qxy2, twoThirds * qxz2);
}

// C->F at (x, y, z) = +-1/4: affine density, velocity and density-weighted
// strain, the velocity intercepts corrected so the eight targets keep the
// block's momentum (:405-482); sigma = 1/2
template <class T> __host__ __device__ __forceinline__ void linear_coarse_to_fine(T (&f)[27], const CellValues<T> (&cv)[8],
                                                                                 T w1f, T x, T y, T z)
{
    T D[8], U[3][8], QXY[8], QYZ[8], QXZ[8], QXY2[8], QXZ2[8];
    T jm[3] = {0, 0, 0};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const T rhoS = T(1) + cv[n].d;
        const T mw = T(0.125) * rhoS;
        D[n] = cv[n].d;
        U[0][n] = cv[n].u[0]; U[1][n] = cv[n].u[1]; U[2][n] = cv[n].u[2];
        jm[0] += mw * cv[n].u[0]; jm[1] += mw * cv[n].u[1]; jm[2] += mw * cv[n].u[2];
        QXY[n] = rhoS * cv[n].sxy; QYZ[n] = rhoS * cv[n].syz; QXZ[n] = rhoS * cv[n].sxz;
        QXY2[n] = rhoS * cv[n].nxy; QXZ2[n] = rhoS * cv[n].nxz;
    }
    const Affine<T> d = affine_fit(D);
    Affine<T> a[3] = {affine_fit(U[0]), affine_fit(U[1]), affine_fit(U[2])};
    const T rhoMean = T(1) + d.c0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        a[c].c0 = (jm[c] - (d.cx * a[c].cx + d.cy * a[c].cy + d.cz * a[c].cz) / T(16)) / rhoMean;
    const T dRho = d.at(x, y, z);
    const T u[3] = {a[0].at(x, y, z), a[1].at(x, y, z), a[2].at(x, y, z)};
    const T ss = T(0.5) / w1f;
    const T third = -(T(1) / T(3)) * ss, twoThirds = -(T(2) / T(3)) * ss;
    reconstruct_k(f, dRho, u, third * affine_fit(QYZ).at(x, y, z), third * affine_fit(QXZ).at(x, y, z),
                  third * affine_fit(QXY).at(x, y, z), twoThirds * affine_fit(QXY2).at(x, y, z),
                  twoThirds * affine_fit(QXZ2).at(x, y, z));
}

} // namespace lbm3
