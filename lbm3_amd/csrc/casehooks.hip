// Case hooks evaluated by the case's own code through lbm3_hooks callbacks
// (getRefinementModifier, getInitialCondition, getOpenBC, getLocalBC: the
// reference inlines them into its kernels, markerFunctions.h:258-276,
// applyInitialCondition.h:22-35, gridBuilderFunctions.h:1232-1300,1486-1503,
// updateGrid.h:86-95,256).  Their results become the same tables the rule
// tables produce: per-(level, wall id) local boundary conditions, open-boundary
// records, initial fields.  The builder's part (refinement, open-boundary
// classification) is in builder.hip.
#include <cstring>
#include "sim.hpp"
#include "levelargs.hpp"

namespace lbm3 {

#define HIPCHK(expr)                                  \
    do {                                              \
        const int rc_ = check_hip((expr), #expr);     \
        if (rc_) return rc_;                          \
    } while (0)

// (i, j, k) of every listed cell (or of cells 0..n-1 without a list)
__global__ __launch_bounds__(256) void k_cell_ijk(Topology t, const int *list, int n, int *i, int *j, int *k)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= n) return;
    int nb[8], ijk[3];
    decode(t, list ? list[x] : x, nb, ijk);
    i[x] = ijk[0];
    j[x] = ijk[1];
    k[x] = ijk[2];
}

static int cell_ijk(Level &l, const int *list, int n, DBuf<int> (&ijk)[3], hipStream_t s)
{
    for (DBuf<int> &a : ijk) HIPCHK(a.alloc(n > 0 ? n : 1));
    if (n > 0)
        hipLaunchKernelGGL(k_cell_ijk, dim3((n + 255) / 256), dim3(256), 0, s, topo_of(l), list, n, ijk[0].p, ijk[1].p,
                           ijk[2].p);
    return check_hip(hipGetLastError(), "cell coordinates");
}

// getInitialCondition of every cell of a level through the hook: (dRho, ux, uy, uz) per cell
int hook_initial_fields(Sim &s, int level, DBuf<float> &fields4)
{
    Level &l = s.L[level];
    DBuf<int> ijk[3];
    if (const int rc = cell_ijk(l, nullptr, l.N, ijk, s.stream)) return rc;
    HIPCHK(fields4.alloc(4LL * (l.N > 0 ? l.N : 1)));
    lbm3_info info = l.info;
    if (s.hooks.initial_condition(s.hooks.user, &info, ijk[0].p, ijk[1].p, ijk[2].p, l.N, fields4.p, (void *)s.stream) != 0)
        return set_error(LBM3_E_ARG, "the getInitialCondition hook callback failed");
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s.stream));
    return 0;
}

// ---------------------------------------------------------------------------
// getLocalBC: what the step reads of it is fixed per (level, wall id); the
// hook is evaluated at every cell on two different states and the answers
// are checked to be of that form
// ---------------------------------------------------------------------------
struct LocalIn {
    const int *wallMap;
    const uint32_t *wallData;
};

// the BCStruct the step hands to getLocalBC (updateGrid.h:86-94), state (d, u)
__global__ __launch_bounds__(256) void k_local_in(LocalIn a, int n, float d, float ux, float uy, float uz, lbm3_bc *in)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    lbm3_bc b = default_bc();
    const int wm = a.wallMap[c];
    if (wm >= 0) b.wallID = (int)((a.wallData[wm] >> 27) & 31u);
    b.dRho = d; b.ux = ux; b.uy = uy; b.uz = uz;
    in[c] = b;
}

// group of a cell: 0 not next to a wall, 1 + wall id next to one, -1 solid
__device__ __forceinline__ int local_group(const LocalIn &a, int c)
{
    const int wm = a.wallMap[c];
    if (wm == -3) return -1;
    return wm >= 0 ? 1 + (int)((a.wallData[wm] >> 27) & 31u) : 0;
}

__global__ __launch_bounds__(256) void k_local_first(LocalIn a, int n, int *first)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int g = local_group(a, c);
    if (g >= 0) atomicMin(&first[g], c);
}

__device__ __forceinline__ bool same_params(const lbm3_bc &x, const lbm3_bc &y, bool wall)
{
    bool s = x.gx == y.gx && x.gy == y.gy && x.gz == y.gz && x.collisionLimiter == y.collisionLimiter &&
             x.nuMultiplier == y.nuMultiplier;
    if (wall) s = s && x.useIBB == y.useIBB;
    return s;
}
__device__ __forceinline__ bool same_u(const lbm3_bc &x, const lbm3_bc &y) { return x.ux == y.ux && x.uy == y.uy && x.uz == y.uz; }

// velocity kind of a cell's answers: 0 passes the state's velocity through,
// 1 independent of the state, 2 neither
__device__ __forceinline__ int vel_kind(const lbm3_bc &inA, const lbm3_bc &outA, const lbm3_bc &inB, const lbm3_bc &outB)
{
    if (same_u(outA, inA) && same_u(outB, inB)) return 0;
    if (same_u(outA, outB)) return 1;
    return 2;
}

// flags per group: 1 state-dependent parameters, 2 parameters differ within the
// group, 4 velocity depends on the state otherwise than passing it through,
// 8 the velocity kind differs within the group, 16 fixed velocities differ
__global__ __launch_bounds__(256) void k_local_check(LocalIn a, int n, const int *first, const lbm3_bc *inA,
                                                     const lbm3_bc *outA, const lbm3_bc *inB, const lbm3_bc *outB, int *flags)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int g = local_group(a, c);
    if (g < 0) return;
    const bool wall = g > 0;
    int f = 0;
    if (!same_params(outA[c], outB[c], wall)) f |= 1;
    const int f0 = first[g];
    if (!same_params(outA[c], outA[f0], wall)) f |= 2;
    if (wall) {
        const int k = vel_kind(inA[c], outA[c], inB[c], outB[c]);
        const int k0 = vel_kind(inA[f0], outA[f0], inB[f0], outB[f0]);
        if (k == 2) f |= 4;
        if (k != k0) f |= 8;
        if (k == 1 && k0 == 1 && !same_u(outA[c], outA[f0])) f |= 16;
    }
    if (f) atomicOr(&flags[g], f);
}

// the wall velocity of every wall-adjacent cell, [3][W]
__global__ __launch_bounds__(256) void k_wall_velocity(const int *wallIndex, int W, const lbm3_bc *out, float *vel)
{
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= W) return;
    const lbm3_bc &b = out[wallIndex[w]];
    vel[w] = b.ux;
    vel[(size_t)W + w] = b.uy;
    vel[2 * (size_t)W + w] = b.uz;
}

static LocalBC local_from(const lbm3_bc &b, int velMode)
{
    LocalBC r{};
    r.gx = b.gx; r.gy = b.gy; r.gz = b.gz;
    r.limiter = b.collisionLimiter;
    r.nuMult = b.nuMultiplier;
    r.useIBB = b.useIBB;
    r.velMode = velMode;
    r.ux = b.ux; r.uy = b.uy; r.uz = b.uz;
    return r;
}

int hook_local_bc(Sim &s, int level, LocalBC &bulk, LocalBC (&wb)[32])
{
    Level &l = s.L[level];
    hipStream_t st = s.stream;
    const int n = l.N;
    const unsigned nb = (unsigned)((n + 255) / 256);
    DBuf<int> ijk[3];
    if (const int rc = cell_ijk(l, nullptr, n, ijk, st)) return rc;
    DBuf<lbm3_bc> inA, inB, outA, outB;
    HIPCHK(inA.alloc(n > 0 ? n : 1));
    HIPCHK(inB.alloc(n > 0 ? n : 1));
    HIPCHK(outA.alloc(n > 0 ? n : 1));
    HIPCHK(outB.alloc(n > 0 ? n : 1));
    const LocalIn li{l.wallMap.p, l.wallData.p};
    // two states no hook keys on by accident: at rest, and an arbitrary moving one
    if (n > 0) {
        hipLaunchKernelGGL(k_local_in, dim3(nb), dim3(256), 0, st, li, n, 0.f, 0.f, 0.f, 0.f, inA.p);
        hipLaunchKernelGGL(k_local_in, dim3(nb), dim3(256), 0, st, li, n, 0.0123f, 0.0171f, -0.0093f, 0.0211f, inB.p);
    }
    HIPCHK(hipGetLastError());
    lbm3_info info = l.info;
    if (s.hooks.local_bc(s.hooks.user, &info, ijk[0].p, ijk[1].p, ijk[2].p, inA.p, n, outA.p, (void *)st) != 0 ||
        s.hooks.local_bc(s.hooks.user, &info, ijk[0].p, ijk[1].p, ijk[2].p, inB.p, n, outB.p, (void *)st) != 0)
        return set_error(LBM3_E_ARG, "the getLocalBC hook callback failed");
    HIPCHK(hipGetLastError());
    DBuf<int> first, flags;
    HIPCHK(first.alloc(33));
    HIPCHK(flags.alloc(33));
    HIPCHK(hipMemsetAsync(first.p, 0x7f, sizeof(int) * 33, st));
    HIPCHK(hipMemsetAsync(flags.p, 0, sizeof(int) * 33, st));
    if (n > 0) {
        hipLaunchKernelGGL(k_local_first, dim3(nb), dim3(256), 0, st, li, n, first.p);
        hipLaunchKernelGGL(k_local_check, dim3(nb), dim3(256), 0, st, li, n, first.p, inA.p, outA.p, inB.p, outB.p, flags.p);
    }
    HIPCHK(hipGetLastError());
    int hf[33], hfl[33];
    HIPCHK(hipMemcpyAsync(hf, first.p, sizeof(hf), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hfl, flags.p, sizeof(hfl), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    bool perCellVelocity = false;
    lbm3_bc rep[33], repIn[33], repB[33], repInB[33];
    for (int g = 0; g < 33; ++g) {
        rep[g] = repIn[g] = repB[g] = repInB[g] = default_bc();
        const std::string who = g == 0 ? "fluid cells away from walls" : "cells next to wall id " + std::to_string(g - 1);
        if (hfl[g] & 1)
            return set_error(LBM3_E_ARG, "getLocalBC on level " + std::to_string(level) + ": force, limiter, viscosity "
                             "multiplier or IBB depend on the cell's state (" + who + "); not supported");
        if (hfl[g] & 2)
            return set_error(LBM3_E_ARG, "getLocalBC on level " + std::to_string(level) + ": force, limiter, viscosity "
                             "multiplier or IBB differ between " + who + "; not supported");
        if (hfl[g] & (4 | 8))
            return set_error(LBM3_E_ARG, "getLocalBC on level " + std::to_string(level) + ": the wall velocity of " + who +
                             " depends on the cell's state other than by taking it over; not supported");
        if (hf[g] >= 0 && hf[g] < n) {
            HIPCHK(hipMemcpy(&rep[g], outA.p + hf[g], sizeof(lbm3_bc), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&repIn[g], inA.p + hf[g], sizeof(lbm3_bc), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&repB[g], outB.p + hf[g], sizeof(lbm3_bc), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&repInB[g], inB.p + hf[g], sizeof(lbm3_bc), hipMemcpyDeviceToHost));
        }
    }
    bulk = local_from(rep[0], 0);
    for (int w = 0; w < 32; ++w) {
        const int g = w + 1;
        const lbm3_bc &a = rep[g];
        const bool keep = a.ux == repIn[g].ux && a.uy == repIn[g].uy && a.uz == repIn[g].uz &&
                          repB[g].ux == repInB[g].ux && repB[g].uy == repInB[g].uy && repB[g].uz == repInB[g].uz;
        int mode = keep ? 0 : 1;
        if (!keep && (hfl[g] & 16)) { mode = 3; perCellVelocity = true; }
        wb[w] = local_from(a, mode);
    }
    if (perCellVelocity && l.W > 0) {
        HIPCHK(l.wallVel.alloc(3LL * l.W));
        hipLaunchKernelGGL(k_wall_velocity, dim3((l.W + 255) / 256), dim3(256), 0, st, l.wallIndex.p, l.W, outA.p, l.wallVel.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// getOpenBC at the open cells again (after the case changed what it reads):
// new records, the cells' record indices replaced in their open_meta words
// ---------------------------------------------------------------------------
int upload_open_records_of(Sim &s, Level &l);
int settle_open_level(Sim &s, int level);

__global__ __launch_bounds__(256) void k_set_records(uint32_t *meta, const int *rec, int n)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= n) return;
    meta[x] = (meta[x] & 63u) | ((uint32_t)rec[x] << 6);
}

static int refresh_open(Sim &s, int level)
{
    Level &l = s.L[level];
    if (l.nOpen <= 0) return 0;
    if (const int rc = settle_open_level(s, level)) return rc;   // a pending completion keeps the values it was made with
    hipStream_t st = s.stream;
    DBuf<int> ijk[3];
    if (const int rc = cell_ijk(l, l.openCell.p, l.nOpen, ijk, st)) return rc;
    DBuf<lbm3_bc> out;
    HIPCHK(out.alloc(l.nOpen));
    lbm3_info info = l.info;
    if (s.hooks.open_bc(s.hooks.user, &info, ijk[0].p, ijk[1].p, ijk[2].p, l.nOpen, out.p, (void *)st) != 0)
        return set_error(LBM3_E_ARG, "the getOpenBC hook callback failed");
    HIPCHK(hipGetLastError());
    std::vector<lbm3_bc> calls((size_t)l.nOpen);
    HIPCHK(hipMemcpyAsync(calls.data(), out.p, sizeof(lbm3_bc) * calls.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // the cells' kinds are fixed by the build: an open cell stays one
    for (const lbm3_bc &b : calls)
        if (b.periodic || !(b.dirichletU || b.dirichletRho))
            return set_error(LBM3_E_ARG, "getOpenBC turned an open cell into a periodic or unconstrained one after the build");
    std::vector<int> idx;
    std::vector<lbm3_bc> recs;
    dedupe_bc(calls, idx, recs);
    DBuf<int> di;
    HIPCHK(di.alloc(l.nOpen));
    HIPCHK(hipMemcpyAsync(di.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_set_records, dim3((l.nOpen + 255) / 256), dim3(256), 0, st, l.openMeta.p, di.p, l.nOpen);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    l.openBC = recs;
    return upload_open_records_of(s, l);
}

int setup_local_bc_of(Sim &s, int level);

} // namespace lbm3

using namespace lbm3;

extern "C" {

int lbm3_set_hooks(lbm3_sim *h, const lbm3_hooks *hooks)
{
    if (!h) return set_error(LBM3_E_ARG, "null sim");
    h->s.hooks = hooks ? *hooks : lbm3_hooks{};
    return 0;
}

int lbm3_set_info_controls(lbm3_sim *h, int level, float iRegulatorInlet, float iRegulatorOutlet, int useRotors)
{
    if (!h || level < 0 || level >= h->s.levels) return set_error(LBM3_E_ARG, "bad level");
    lbm3_info &I = h->s.L[level].info;
    I.iRegulatorInlet = iRegulatorInlet;
    I.iRegulatorOutlet = iRegulatorOutlet;
    I.useRotors = useRotors;
    return 0;
}

int lbm3_refresh_hooks(lbm3_sim *h)
{
    if (!h) return set_error(LBM3_E_ARG, "null sim");
    Sim &s = h->s;
    for (int level = 0; level < s.levels; ++level) {
        if (!s.L[level].built) continue;
        if (s.hooks.open_bc)
            if (const int rc = refresh_open(s, level)) return rc;
        if (s.hooks.local_bc)
            if (const int rc = setup_local_bc_of(s, level)) return rc;
    }
    return 0;
}

} // extern "C"
