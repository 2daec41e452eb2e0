// Register-resident D3Q27 cumulant collision on shifted populations f - w.
//
// Model: the well-conditioned cumulant LBM of Geier et al. 2015/2017 as used by
// the reference (include/applyCollision.h:17-491): K17 with limiter lambda on
// the seven third-order combinations and fourth order from velocity gradients,
// or the K15 "AllOne" variant when lambda <= 0 or the viscosity multiplier > 1;
// Guo-like half-force velocity shift with the full momentum added through the
// first central moment; omega2 = 1.
//
// Implementation: the central-moment transform is three passes (z, y, x) of
// one three-point kernel per chain of populations; the back transform is the
// inverse kernel in the reverse order.  Each chain carries the weight moment K
// of the well-conditioned form.  Everything stays in registers.
#pragma once
#include "lattice.hpp"

namespace lbm3 {

// Everything of the relaxation that depends only on the viscosity, the
// viscosity multiplier and the limiter, evaluated once per boundary-condition
// entry instead of once per cell (same expressions as the reference, in the
// compute precision).
template <class T> struct Relax {
    T gx, gy, gz;
    T w1;                 // omega1 = 1 / (3 nu nuMult + 1/2)
    T w3, w4, w5, A, B;   // K17 rates and fourth-order coefficients (omega2 = 1)
    T lambda;
    int k17;              // K17 if lambda > 0 and nuMult <= 1, else K15 AllOne
};

template <class T>
__host__ __device__ inline Relax<T> make_relax(T nu, T nuMult, T lambda, T gx, T gy, T gz)
{
    Relax<T> r;
    r.gx = gx; r.gy = gy; r.gz = gz;
    r.lambda = lambda;
    r.k17 = (lambda > T(0)) && (nuMult <= T(1));
    const T w1 = T(1) / (T(3) * (nu * nuMult) + T(0.5)), w2 = T(1);
    r.w1 = w1;
    r.w3 = T(8) * (w1 - T(2)) * (w2 * (T(3) * w1 - T(1)) - T(5) * w1) /
           (T(8) * (T(5) - T(2) * w1) * w1 + w2 * (T(8) + w1 * (T(9) * w1 - T(26))));
    r.w4 = T(8) * (w1 - T(2)) * (w1 + w2 * (T(3) * w1 - T(7))) / (w2 * (T(56) - T(42) * w1 + T(9) * w1 * w1) - T(8) * w1);
    r.w5 = T(24) * (w1 - T(2)) * (T(4) * w1 * w1 + w1 * w2 * (T(18) - T(13) * w1) + w2 * w2 * (T(2) + w1 * (T(6) * w1 - T(11)))) /
           (T(16) * w1 * w1 * (w1 - T(6)) - T(2) * w1 * w2 * (T(216) + T(5) * w1 * (T(9) * w1 - T(46))) +
            w2 * w2 * (w1 * (T(3) * w1 - T(10)) * (T(15) * w1 - T(28)) - T(48)));
    r.A = (T(4) * w1 * w1 + T(2) * w1 * w2 * (w1 - T(6)) + w2 * w2 * (w1 * (T(10) - T(3) * w1) - T(4))) /
          ((w1 - w2) * (w2 * (T(2) + T(3) * w1) - T(8) * w1));
    r.B = (T(4) * w1 * w2 * (T(9) * w1 - T(16)) - T(4) * w1 * w1 - T(2) * w2 * w2 * (T(2) + T(9) * w1 * (w1 - T(2)))) /
          (T(3) * (w1 - w2) * (w2 * (T(2) + T(3) * w1) - T(8) * w1));
    if (!r.k17) { r.w3 = r.w4 = r.w5 = r.A = r.B = T(0); }
    return r;
}

// reciprocal: the hardware estimate (about 1 ulp) refined by one Newton step
// on the residual e = 1 - x r (one FMA, no rounding next to 1), so the result
// is correctly rounded but for rare near-halfway cases and its error has no
// preferred sign.  The form
// r (2 - x r) rounds 2 - x r next to 1 and leaves a bias of a few 1e-8 that
// the collision turns into a loss of momentum every step (0.3% extra energy
// decay on a 128^3 Taylor-Green vortex).  Exact division in FP64.
__host__ __device__ __forceinline__ float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.f);
    return __builtin_fmaf(r, e, r);
#else
    return 1.f / x;
#endif
}
__host__ __device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }

// forward: (f-, f0, f+) about velocity u -> (k0, k1, k2); K is the weight moment of the chain
template <class T>
__host__ __device__ __forceinline__ void cm_forward(T m, T z, T p, T u, T u2, T K, T &k0, T &k1, T &k2)
{
    const T s = p + m, d = p - m;
    k0 = s + z;
    const T kk = k0 + K;
    k1 = d - u * kk;
    k2 = s - T(2) * u * d + u2 * kk;
}

// inverse of cm_forward
template <class T>
__host__ __device__ __forceinline__ void cm_backward(T k0, T k1, T k2, T u, T u2, T K, T &m, T &z, T &p)
{
    const T kk = k0 + K;
    const T S = kk * u2 + T(2) * u * k1 + k2;
    const T D = kk * u + k1;
    m = T(0.5) * (S - D);
    p = T(0.5) * (S + D);
    z = k0 - S;
}

// Well-conditioned central moments k_abc (x order a, y order b, z order c),
// stored x-chain by x-chain: {k0jk, k1jk, k2jk} for (j,k) = 00, 01, 02, 10, 11, 12, 20, 21, 22.
template <class T> struct CentralMoments {
    T m000, m100, m200, m001, m101, m201, m002, m102, m202;
    T m010, m110, m210, m011, m111, m211, m012, m112, m212;
    T m020, m120, m220, m021, m121, m221, m022, m122, m222;
};

// central moments about u -> shifted populations: the inverse three-point
// kernel along x, then y, then z, each chain carrying its weight moment
template <class T>
__host__ __device__ __forceinline__ void from_central_moments(T (&f)[27], const CentralMoments<T> &M, T ux, T uy, T uz)
{
    constexpr T W36 = T(1) / T(36), W9 = T(1) / T(9), W49 = T(4) / T(9);
    constexpr T W6 = T(1) / T(6), W23 = T(2) / T(3), W18 = T(1) / T(18), W29 = T(2) / T(9);
    constexpr T third = T(1) / T(3), ninth = T(1) / T(9);
    const T ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    T b_a00, b_b00, b_c00, b_a01, b_b01, b_c01, b_a02, b_b02, b_c02;
    T b_a10, b_b10, b_c10, b_a11, b_b11, b_c11, b_a12, b_b12, b_c12;
    T b_a20, b_b20, b_c20, b_a21, b_b21, b_c21, b_a22, b_b22, b_c22;
    cm_backward(M.m000, M.m100, M.m200, ux, ux2, T(1), b_a00, b_b00, b_c00);
    cm_backward(M.m001, M.m101, M.m201, ux, ux2, T(0), b_a01, b_b01, b_c01);
    cm_backward(M.m002, M.m102, M.m202, ux, ux2, third, b_a02, b_b02, b_c02);
    cm_backward(M.m010, M.m110, M.m210, ux, ux2, T(0), b_a10, b_b10, b_c10);
    cm_backward(M.m011, M.m111, M.m211, ux, ux2, T(0), b_a11, b_b11, b_c11);
    cm_backward(M.m012, M.m112, M.m212, ux, ux2, T(0), b_a12, b_b12, b_c12);
    cm_backward(M.m020, M.m120, M.m220, ux, ux2, third, b_a20, b_b20, b_c20);
    cm_backward(M.m021, M.m121, M.m221, ux, ux2, T(0), b_a21, b_b21, b_c21);
    cm_backward(M.m022, M.m122, M.m222, ux, ux2, ninth, b_a22, b_b22, b_c22);

    T a_aa0, a_ab0, a_ac0, a_ba0, a_bb0, a_bc0, a_ca0, a_cb0, a_cc0;
    T a_aa1, a_ab1, a_ac1, a_ba1, a_bb1, a_bc1, a_ca1, a_cb1, a_cc1;
    T a_aa2, a_ab2, a_ac2, a_ba2, a_bb2, a_bc2, a_ca2, a_cb2, a_cc2;
    cm_backward(b_a00, b_a10, b_a20, uy, uy2, W6, a_aa0, a_ab0, a_ac0);
    cm_backward(b_b00, b_b10, b_b20, uy, uy2, W23, a_ba0, a_bb0, a_bc0);
    cm_backward(b_c00, b_c10, b_c20, uy, uy2, W6, a_ca0, a_cb0, a_cc0);
    cm_backward(b_a01, b_a11, b_a21, uy, uy2, T(0), a_aa1, a_ab1, a_ac1);
    cm_backward(b_b01, b_b11, b_b21, uy, uy2, T(0), a_ba1, a_bb1, a_bc1);
    cm_backward(b_c01, b_c11, b_c21, uy, uy2, T(0), a_ca1, a_cb1, a_cc1);
    cm_backward(b_a02, b_a12, b_a22, uy, uy2, W18, a_aa2, a_ab2, a_ac2);
    cm_backward(b_b02, b_b12, b_b22, uy, uy2, W29, a_ba2, a_bb2, a_bc2);
    cm_backward(b_c02, b_c12, b_c22, uy, uy2, W18, a_ca2, a_cb2, a_cc2);

    cm_backward(a_aa0, a_aa1, a_aa2, uz, uz2, W36, f[MMM], f[MMO], f[MMP]);
    cm_backward(a_ab0, a_ab1, a_ab2, uz, uz2, W9, f[MOM], f[MOO], f[MOP]);
    cm_backward(a_ac0, a_ac1, a_ac2, uz, uz2, W36, f[MPM], f[MPO], f[MPP]);
    cm_backward(a_ba0, a_ba1, a_ba2, uz, uz2, W9, f[OMM], f[OMO], f[OMP]);
    cm_backward(a_bb0, a_bb1, a_bb2, uz, uz2, W49, f[OOM], f[OOO], f[OOP]);
    cm_backward(a_bc0, a_bc1, a_bc2, uz, uz2, W9, f[OPM], f[OPO], f[OPP]);
    cm_backward(a_ca0, a_ca1, a_ca2, uz, uz2, W36, f[PMM], f[PMO], f[PMP]);
    cm_backward(a_cb0, a_cb1, a_cb2, uz, uz2, W9, f[POM], f[POO], f[POP]);
    cm_backward(a_cc0, a_cc1, a_cc2, uz, uz2, W36, f[PPM], f[PPO], f[PPP]);
}

// f[27] in/out (shifted populations).  MODE 0: K17 or K15 from R.k17 at run
// time; 1: K17; 2: K15.
template <class T, int MODE = 0>
__host__ __device__ __forceinline__ void collide(T (&f)[27], const Relax<T> &R)
{
    // weight moments of the well-conditioned chains: z-pass (per x,y pair), y-pass (per x)
    constexpr T W36 = T(1) / T(36), W9 = T(1) / T(9), W49 = T(4) / T(9);
    constexpr T W6 = T(1) / T(6), W23 = T(2) / T(3), W18 = T(1) / T(18), W29 = T(2) / T(9);
    constexpr T third = T(1) / T(3), ninth = T(1) / T(9);

    // ---- density and momentum (exact for the stored f - w) ----
    const T s_aa = (f[MMP] + f[MMM]) + f[MMO], s_ab = (f[MOP] + f[MOM]) + f[MOO], s_ac = (f[MPP] + f[MPM]) + f[MPO];
    const T s_ba = (f[OMP] + f[OMM]) + f[OMO], s_bb = (f[OOP] + f[OOM]) + f[OOO], s_bc = (f[OPP] + f[OPM]) + f[OPO];
    const T s_ca = (f[PMP] + f[PMM]) + f[PMO], s_cb = (f[POP] + f[POM]) + f[POO], s_cc = (f[PPP] + f[PPM]) + f[PPO];
    const T xm = (s_aa + s_ab) + s_ac, x0 = (s_ba + s_bb) + s_bc, xp = (s_ca + s_cb) + s_cc;
    const T dRho = (xm + x0) + xp;
    const T rho = T(1) + dRho;
    const T rhoInv = fast_rcp(rho);
    const T jx = xp - xm;
    const T jy = ((s_ac + s_bc) + s_cc) - ((s_aa + s_ba) + s_ca);
    const T jz = (((f[MMP] - f[MMM]) + (f[MOP] - f[MOM])) + ((f[MPP] - f[MPM]) + (f[OMP] - f[OMM]))) +
                 (((f[OOP] - f[OOM]) + (f[OPP] - f[OPM])) + ((f[PMP] - f[PMM]) + (f[POP] - f[POM]))) + (f[PPP] - f[PPM]);
    const T ux = (jx + T(0.5) * R.gx) * rhoInv;
    const T uy = (jy + T(0.5) * R.gy) * rhoInv;
    const T uz = (jz + T(0.5) * R.gz) * rhoInv;
    const T ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;

    // ---- z pass: chains over (x,y) = (a,b,c) x (a,b,c) ----
    T z_aa0, z_aa1, z_aa2, z_ab0, z_ab1, z_ab2, z_ac0, z_ac1, z_ac2;
    T z_ba0, z_ba1, z_ba2, z_bb0, z_bb1, z_bb2, z_bc0, z_bc1, z_bc2;
    T z_ca0, z_ca1, z_ca2, z_cb0, z_cb1, z_cb2, z_cc0, z_cc1, z_cc2;
    cm_forward(f[MMM], f[MMO], f[MMP], uz, uz2, W36, z_aa0, z_aa1, z_aa2);
    cm_forward(f[MOM], f[MOO], f[MOP], uz, uz2, W9, z_ab0, z_ab1, z_ab2);
    cm_forward(f[MPM], f[MPO], f[MPP], uz, uz2, W36, z_ac0, z_ac1, z_ac2);
    cm_forward(f[OMM], f[OMO], f[OMP], uz, uz2, W9, z_ba0, z_ba1, z_ba2);
    cm_forward(f[OOM], f[OOO], f[OOP], uz, uz2, W49, z_bb0, z_bb1, z_bb2);
    cm_forward(f[OPM], f[OPO], f[OPP], uz, uz2, W9, z_bc0, z_bc1, z_bc2);
    cm_forward(f[PMM], f[PMO], f[PMP], uz, uz2, W36, z_ca0, z_ca1, z_ca2);
    cm_forward(f[POM], f[POO], f[POP], uz, uz2, W9, z_cb0, z_cb1, z_cb2);
    cm_forward(f[PPM], f[PPO], f[PPP], uz, uz2, W36, z_cc0, z_cc1, z_cc2);

    // ---- y pass: chains over (x, z-order) ----
    T y_a00, y_a10, y_a20, y_b00, y_b10, y_b20, y_c00, y_c10, y_c20;
    T y_a01, y_a11, y_a21, y_b01, y_b11, y_b21, y_c01, y_c11, y_c21;
    T y_a02, y_a12, y_a22, y_b02, y_b12, y_b22, y_c02, y_c12, y_c22;
    cm_forward(z_aa0, z_ab0, z_ac0, uy, uy2, W6, y_a00, y_a10, y_a20);
    cm_forward(z_ba0, z_bb0, z_bc0, uy, uy2, W23, y_b00, y_b10, y_b20);
    cm_forward(z_ca0, z_cb0, z_cc0, uy, uy2, W6, y_c00, y_c10, y_c20);
    cm_forward(z_aa1, z_ab1, z_ac1, uy, uy2, T(0), y_a01, y_a11, y_a21);
    cm_forward(z_ba1, z_bb1, z_bc1, uy, uy2, T(0), y_b01, y_b11, y_b21);
    cm_forward(z_ca1, z_cb1, z_cc1, uy, uy2, T(0), y_c01, y_c11, y_c21);
    cm_forward(z_aa2, z_ab2, z_ac2, uy, uy2, W18, y_a02, y_a12, y_a22);
    cm_forward(z_ba2, z_bb2, z_bc2, uy, uy2, W29, y_b02, y_b12, y_b22);
    cm_forward(z_ca2, z_cb2, z_cc2, uy, uy2, W18, y_c02, y_c12, y_c22);

    // ---- x pass: central moments k_abc (x order a, y order b, z order c) ----
    T k000, k100, k200, k001, k101, k201, k002, k102, k202;
    T k010, k110, k210, k011, k111, k211, k012, k112, k212;
    T k020, k120, k220, k021, k121, k221, k022, k122, k222;
    cm_forward(y_a00, y_b00, y_c00, ux, ux2, T(1), k000, k100, k200);
    // zeroth moment: use the stored-density sum so collision conserves exactly dRho
    k000 = dRho;
    {
        const T kk = dRho + T(1), s = y_c00 + y_a00, d = y_c00 - y_a00;
        k100 = d - ux * kk;
        k200 = s - T(2) * ux * d + ux2 * kk;
    }
    cm_forward(y_a01, y_b01, y_c01, ux, ux2, T(0), k001, k101, k201);
    cm_forward(y_a02, y_b02, y_c02, ux, ux2, third, k002, k102, k202);
    cm_forward(y_a10, y_b10, y_c10, ux, ux2, T(0), k010, k110, k210);
    cm_forward(y_a11, y_b11, y_c11, ux, ux2, T(0), k011, k111, k211);
    cm_forward(y_a12, y_b12, y_c12, ux, ux2, T(0), k012, k112, k212);
    cm_forward(y_a20, y_b20, y_c20, ux, ux2, third, k020, k120, k220);
    cm_forward(y_a21, y_b21, y_c21, ux, ux2, T(0), k021, k121, k221);
    cm_forward(y_a22, y_b22, y_c22, ux, ux2, ninth, k022, k122, k222);
    (void)k212; (void)k221; (void)k122; (void)k222;

    // fourth-order cumulants relax to their equilibrium (rate 1): their
    // pre-collision values, and the fourth-order moments, are not needed
    (void)k211; (void)k121; (void)k112; (void)k220; (void)k022; (void)k202;
    const T dRhoNinth = dRho * rhoInv * ninth;

    // ---- relaxation ----
    const T w1 = R.w1;
    const T hr = T(0.5) * rhoInv;
    const T Dxu = -w1 * hr * (T(2) * k200 - k020 - k002) - hr * (k200 + k020 + k002 - dRho);
    const T Dyv = Dxu + T(1.5) * w1 * rhoInv * (k200 - k020);
    const T Dzw = Dxu + T(1.5) * w1 * rhoInv * (k200 - k002);
    const T DxvDyu = -T(3) * w1 * rhoInv * k110;
    const T DxwDzu = -T(3) * w1 * rhoInv * k101;
    const T DywDzv = -T(3) * w1 * rhoInv * k011;

    const T om1 = T(1) - w1;
    const T s110 = om1 * k110, s101 = om1 * k101, s011 = om1 * k011;
    const T g1 = T(3) * rho * (T(1) - T(0.5) * w1);
    const T g2 = T(1.5) * rho;   // 3 rho (1 - omega2/2), omega2 = 1
    const T e33 = om1 * (k200 - k020) - g1 * (ux2 * Dxu - uy2 * Dyv);
    const T e34 = om1 * (k200 - k002) - g1 * (ux2 * Dxu - uz2 * Dzw);
    const T e35 = dRho - g2 * (ux2 * Dxu + uy2 * Dyv + uz2 * Dzw);
    const T s200 = (e33 + e34 + e35) * third;
    const T s020 = (e34 - T(2) * e33 + e35) * third;
    const T s002 = (e33 - T(2) * e34 + e35) * third;

    T s120 = 0, s102 = 0, s210 = 0, s012 = 0, s201 = 0, s021 = 0, s111 = 0;
    T c220 = 0, c202 = 0, c022 = 0, c211 = 0, c121 = 0, c112 = 0;
    const bool k17 = MODE == 1 || (MODE == 0 && R.k17);
    if (MODE != 2 && k17) {
        const T lambda = R.lambda;
        // limited relaxation: the rate tends to 1 where the cumulant is large compared to lambda
        auto limited = [&](T c, T w) {
            const T a = c < T(0) ? -c : c;
            const T wl = w + (T(1) - w) * a * fast_rcp(lambda + a);
            return (T(1) - wl) * c;
        };
        const T p1 = limited(k120 + k102, R.w3), p2 = limited(k210 + k012, R.w3), p3 = limited(k201 + k021, R.w3);
        const T m1 = limited(k120 - k102, R.w4), m2 = limited(k210 - k012, R.w4), m3 = limited(k201 - k021, R.w4);
        s120 = T(0.5) * (m1 + p1); s102 = T(0.5) * (p1 - m1);
        s210 = T(0.5) * (m2 + p2); s012 = T(0.5) * (p2 - m2);
        s201 = T(0.5) * (m3 + p3); s021 = T(0.5) * (p3 - m3);
        s111 = limited(k111, R.w5);
        const T visc = (T(1) / w1 - T(0.5)) * rho;
        const T a4 = T(2) / T(3) * visc * R.A;
        const T e43 = a4 * (Dxu - T(2) * Dyv + Dzw);
        const T e44 = a4 * (-T(2) * Dxu + Dyv + Dzw);
        const T e45 = -T(2) * a4 * (Dxu + Dyv + Dzw);
        c220 = (e43 + e44 + e45) * third;
        c202 = (e45 - e43) * third;
        c022 = (e45 - e44) * third;
        const T b4 = -third * visc * R.B;
        c211 = b4 * DywDzv;
        c121 = b4 * DxwDzu;
        c112 = b4 * DxvDyu;
    }

    // ---- post-collision central moments ----
    const T q211 = c211 + ((s200 + third) * s011 + T(2) * s101 * s110) * rhoInv;
    const T q121 = c121 + ((s020 + third) * s101 + T(2) * s110 * s011) * rhoInv;
    const T q112 = c112 + ((s002 + third) * s110 + T(2) * s011 * s101) * rhoInv;
    const T q220 = c220 + ((s020 * s200 + T(2) * s110 * s110 + (s200 + s020) * third) * rhoInv - dRhoNinth);
    const T q022 = c022 + ((s002 * s020 + T(2) * s011 * s011 + (s002 + s020) * third) * rhoInv - dRhoNinth);
    const T q202 = c202 + ((s200 * s002 + T(2) * s101 * s101 + (s200 + s002) * third) * rhoInv - dRhoNinth);
    const T q122 = (s020 * s102 + s002 * s120 + T(4) * s011 * s111 + T(2) * (s110 * s012 + s101 * s021) + (s120 + s102) * third) * rhoInv;
    const T q212 = (s002 * s210 + s200 * s012 + T(4) * s101 * s111 + T(2) * (s011 * s201 + s110 * s102) + (s210 + s012) * third) * rhoInv;
    const T q221 = (s200 * s021 + s020 * s201 + T(4) * s110 * s111 + T(2) * (s101 * s120 + s011 * s210) + (s021 + s201) * third) * rhoInv;
    const T r2 = rhoInv * rhoInv;
    const T q222 = (T(4) * s111 * s111 + s200 * q022 + s020 * q202 + s002 * q220 +
                    T(4) * (s011 * q211 + s101 * q121 + s110 * q112) + T(2) * (s120 * s102 + s210 * s012 + s201 * s021)) * rhoInv
                 - (T(16) * s110 * s101 * s011 + T(4) * (s101 * s101 * s020 + s011 * s011 * s200 + s110 * s110 * s002) +
                    T(2) * s200 * s020 * s002) * r2
                 + (T(3) * (q022 + q202 + q220) + (s200 + s020 + s002)) * rhoInv * ninth
                 - T(2) / T(3) * (T(2) * (s101 * s101 + s011 * s011 + s110 * s110) + (s002 * s020 + s002 * s200 + s020 * s200) +
                                  (s002 + s020 + s200) * third) * r2
                 - (dRho * dRho - dRho) * r2 * (T(1) / T(27));

    // ---- back transform (first moments flip sign: momentum gains the force) ----
    const CentralMoments<T> M{dRho, -k100, s200, -k001, s101, s201, s002, s102, q202,
                              -k010, s110, s210, s011, s111, q211, s012, q112, q212,
                              s020, s120, q220, s021, q121, q221, q022, q122, q222};
    from_central_moments(f, M, ux, uy, uz);
}

// getDRhoUxUyUz (cellFunctions.h:257-279): density deviation and velocity
template <class T>
__host__ __device__ __forceinline__ void moments(const T (&f)[27], T &dRho, T &ux, T &uy, T &uz)
{
    const T a = (f[PPP] + f[MMM]) + (f[PMP] + f[MPM]), b = (f[PPM] + f[MMP]) + (f[PMM] + f[MPP]);
    const T e = (f[OPP] + f[OMM]) + (f[OPM] + f[OMP]), g = (f[POP] + f[MOM]) + (f[POM] + f[MOP]);
    const T h = (f[PPO] + f[MMO]) + (f[PMO] + f[MPO]);
    dRho = (a + b) + (e + g + h) + ((f[POO] + f[MOO]) + (f[OPO] + f[OMO]) + (f[OOP] + f[OOM])) + f[OOO];
    const T rhoInv = T(1) / (dRho + T(1));
    const T d1 = f[PPP] - f[MMM], d2 = f[PMP] - f[MPM], d3 = f[PPM] - f[MMP], d4 = f[PMM] - f[MPP];
    const T mx = ((d1 + d2) + (d3 + d4)) + (((f[POP] - f[MOM]) + (f[POM] - f[MOP])) + ((f[PPO] - f[MMO]) + (f[PMO] - f[MPO]))) + (f[POO] - f[MOO]);
    const T my = ((d1 - d2) + (d3 - d4)) + (((f[OPP] - f[OMM]) + (f[OPM] - f[OMP])) + ((f[PPO] - f[MMO]) - (f[PMO] - f[MPO]))) + (f[OPO] - f[OMO]);
    const T mz = ((d1 + d2) - (d3 + d4)) + (((f[OPP] - f[OMM]) - (f[OPM] - f[OMP])) + ((f[POP] - f[MOM]) - (f[POM] - f[MOP]))) + (f[OOP] - f[OOM]);
    ux = mx * rhoInv;
    uy = my * rhoInv;
    uz = mz * rhoInv;
}

// getFeqSingle (cellFunctions.h:150-168): factorised equilibrium, shifted by the weight
template <class T>
__host__ __device__ __forceinline__ T feq_shifted(int q, T dRho, T ux, T uy, T uz)
{
    const T ax = T(3) * T(cx(q)) * ux + (T(4.5) * T(cx(q) * cx(q)) - T(1.5)) * ux * ux;
    const T ay = T(3) * T(cy(q)) * uy + (T(4.5) * T(cy(q) * cy(q)) - T(1.5)) * uy * uy;
    const T az = T(3) * T(cz(q)) * uz + (T(4.5) * T(cz(q) * cz(q)) - T(1.5)) * uz * uz;
    const T xy = ax + ay + ax * ay;
    const T corr = xy + az + xy * az;
    return weight<T>(q) * ((dRho + corr) + dRho * corr);
}

} // namespace lbm3
