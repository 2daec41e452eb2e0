// The four case-file hooks of the reference (getRefinementModifier,
// getInitialCondition, getOpenBC, getLocalBC) evaluated from the rule tables
// of lbm3_case.  Host and device.
#pragma once
#include <array>
#include <cstring>
#include <map>
#include <vector>
#include "../../include/lbm3mi_types.h"
#include "device.hpp"

namespace lbm3 {

__host__ __device__ inline lbm3_bc default_bc()
{
    lbm3_bc b{};
    b.wallID = -1; b.openBCID = 0; b.useIBB = 1; b.nonReflective = 1;
    b.nonReflectiveBeta = 7.f; b.collisionLimiter = 0.05f; b.nuMultiplier = 1.f;
    return b;
}

__host__ __device__ inline int face_class(int idx, int count)
{
    return idx == 0 ? LBM3_FACE_LOW : (idx == count - 1 ? LBM3_FACE_HIGH : LBM3_FACE_MID);
}

// getXYZFromIJKCellIndex (cellFunctions.h:103-108), never contracted
__host__ __device__ inline void cell_xyz(const LevelGeom &g, const int (&ijk)[3], float &x, float &y, float &z)
{
#pragma clang fp contract(off)
    x = float(ijk[0]) * g.res + g.ox;
    y = float(ijk[1]) * g.res + g.oy;
    z = float(ijk[2]) * g.res + g.oz;
}

// radial profile factor max(0, 1 - r^2 / radius^2) at (x, y, z), 1 without a
// profile; never contracted, so host, device and the oracle round alike
__host__ __device__ inline float profile_factor(const lbm3_profile &P, float x, float y, float z)
{
#pragma clang fp contract(off)
    if (!(P.radius > 0.f)) return 1.f;
    const float dx = x - P.origin[0], dy = y - P.origin[1], dz = z - P.origin[2];
    float r2;
    if (P.axis == LBM3_AXIS_X) r2 = dy * dy + dz * dz;
    else if (P.axis == LBM3_AXIS_Y) r2 = dx * dx + dz * dz;
    else r2 = dx * dx + dy * dy;
    const float v = 1.f - r2 / (P.radius * P.radius);
    return 0.f < v ? v : 0.f;
}

// inside a refinement rule's cylinder (always, without one)
__host__ __device__ inline bool in_cylinder(const lbm3_profile &P, float x, float y, float z)
{
#pragma clang fp contract(off)
    if (!(P.radius > 0.f)) return true;
    const float dx = x - P.origin[0], dy = y - P.origin[1], dz = z - P.origin[2];
    float r2;
    if (P.axis == LBM3_AXIS_X) r2 = dy * dy + dz * dz;
    else if (P.axis == LBM3_AXIS_Y) r2 = dx * dx + dz * dz;
    else r2 = dx * dx + dy * dy;
    return r2 < P.radius * P.radius;
}

// getOpenBC: index of the first matching rule, nOpenRules for the default record
__host__ __device__ inline int open_rule(const lbm3_case &c, int level, int X, int Y, int Z, int i, int j, int k)
{
    const int fx = face_class(i, X), fy = face_class(j, Y), fz = face_class(k, Z);
    for (int r = 0; r < c.nOpenRules; ++r) {
        const lbm3_open_rule &R = c.openRules[r];
        if (R.levelMask && !(R.levelMask & (1 << level))) continue;
        if ((R.faceMask[0] & fx) && (R.faceMask[1] & fy) && (R.faceMask[2] & fz)) return r;
    }
    return c.nOpenRules;
}

__host__ __device__ inline const lbm3_bc &open_bc(const lbm3_case &c, int rule)
{
    return rule < c.nOpenRules ? c.openRules[rule].bc : c.openDefault;
}

inline OpenBCRecord open_record(const lbm3_bc &b)
{
    OpenBCRecord r{};
    r.dRho = b.dRho; r.ux = b.ux; r.uy = b.uy; r.uz = b.uz;
    r.beta = b.nonReflectiveBeta;
    r.dirichletU = b.dirichletU; r.dirichletRho = b.dirichletRho; r.nonReflective = b.nonReflective;
    return r;
}

// getLocalBC for (level, wallID): rules in order, later ones override
inline LocalBC local_bc(const lbm3_case &c, int level, int wallID)
{
    const lbm3_bc d = default_bc();
    LocalBC b{};
    b.gx = d.gx; b.gy = d.gy; b.gz = d.gz; b.limiter = d.collisionLimiter; b.nuMult = d.nuMultiplier;
    b.useIBB = d.useIBB; b.velMode = 0;
    for (int r = 0; r < c.nLocalRules; ++r) {
        const lbm3_local_rule &R = c.localRules[r];
        if (R.levelMask && !(R.levelMask & (1 << level))) continue;
        if (R.wallID != -2 && R.wallID != wallID) continue;
        if (R.set & LBM3_SET_VELOCITY) { b.velMode = 1; b.ux = R.ux; b.uy = R.uy; b.uz = R.uz; }
        if (R.set & LBM3_SET_ROTATION) {
            b.velMode = 2;
            for (int a = 0; a < 3; ++a) { b.omega[a] = R.omega[a]; b.origin[a] = R.origin[a]; }
        }
        if (R.set & LBM3_SET_FORCE) { b.gx = R.gx; b.gy = R.gy; b.gz = R.gz; }
        if (R.set & LBM3_SET_LIMITER) b.limiter = R.collisionLimiter;
        if (R.set & LBM3_SET_NU_MULTIPLIER) b.nuMult = R.nuMultiplier;
        if (R.set & LBM3_SET_USE_IBB) b.useIBB = R.useIBB;
    }
    return b;
}

// getRefinementModifier
__host__ __device__ inline bool refine_hook(const lbm3_case &c, int level, float x, float y, float z, bool marker)
{
    bool haveSet = false, inSet = false;
    for (int r = 0; r < c.nRefineRules; ++r) {
        const lbm3_refine_rule &R = c.refineRules[r];
        if (R.level != level) continue;
        const bool in = x > R.xMin && x < R.xMax && y > R.yMin && y < R.yMax && z > R.zMin && z < R.zMax &&
                        in_cylinder(R.cylinder, x, y, z);
        if (R.mode == LBM3_REFINE_SET) { haveSet = true; inSet = inSet || in; }
    }
    if (haveSet) marker = inSet;
    for (int r = 0; r < c.nRefineRules; ++r) {
        const lbm3_refine_rule &R = c.refineRules[r];
        if (R.level != level) continue;
        const bool in = x > R.xMin && x < R.xMax && y > R.yMin && y < R.yMax && z > R.zMin && z < R.zMax &&
                        in_cylinder(R.cylinder, x, y, z);
        if (R.mode == LBM3_REFINE_ADD && in) marker = true;
        if (R.mode == LBM3_REFINE_REMOVE && in) marker = false;
    }
    return marker;
}

// distinct BCStructs of `calls` in order of first appearance (recs, appended)
// and the record of every call (idx); equal means bitwise equal
inline void dedupe_bc(const std::vector<lbm3_bc> &calls, std::vector<int> &idx, std::vector<lbm3_bc> &recs)
{
    using Key = std::array<uint32_t, sizeof(lbm3_bc) / 4>;
    static_assert(sizeof(lbm3_bc) % 4 == 0, "lbm3_bc holds 4-byte fields only");
    std::map<Key, int> seen;
    for (size_t r = 0; r < recs.size(); ++r) {
        Key k;
        std::memcpy(k.data(), &recs[r], sizeof(lbm3_bc));
        seen.emplace(k, (int)r);
    }
    idx.resize(calls.size());
    for (size_t x = 0; x < calls.size(); ++x) {
        Key k;
        std::memcpy(k.data(), &calls[x], sizeof(lbm3_bc));
        auto it = seen.find(k);
        if (it == seen.end()) {
            it = seen.emplace(k, (int)recs.size()).first;
            recs.push_back(calls[x]);
        }
        idx[x] = it->second;
    }
}

} // namespace lbm3
