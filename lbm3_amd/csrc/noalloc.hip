// lbm3_wall_forces_noalloc: the per-body FP64 wall force and torque sums of
// lbm3_wall_forces without any device or host allocation, for callers that
// cannot allocate (code under HIP graph capture, tight tracker loops, serving
// loops that forbid hipMalloc).  The per-block partial sums and the final
// column sums live in one file-scope device array, `scratch`, of a fixed size;
// every call on a device shares that device's copy, so calls that overlap on
// one GPU race.
#include <mutex>
#include "sim.hpp"
#include "levelargs.hpp"

namespace lbm3 {

// 6 * nBodies partial sums per 256 wall cells, then the 6 * nBodies totals
constexpr long long kScratchDoubles = 1LL << 20;   // 8 MiB per device
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];
static std::once_flag g_once[kMaxDevices];
static hipError_t g_symErr[kMaxDevices];

// the device address of this device's `scratch`, looked up once per device
// with that device current
static int scratch_ptr(double **out)
{
    int dev = 0;
    if (const int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
    if (dev < 0 || dev >= kMaxDevices) return set_error(LBM3_E_ARG, "device index beyond the scratch table");
    std::call_once(g_once[dev], [dev]() {
        void *p = nullptr;
        g_symErr[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch[dev] = g_symErr[dev] == hipSuccess ? static_cast<double *>(p) : nullptr;
    });
    if (g_symErr[dev] != hipSuccess || !g_scratch[dev])
        return check_hip(g_symErr[dev] != hipSuccess ? g_symErr[dev] : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    *out = g_scratch[dev];
    return 0;
}

// out[c] = sum over rows r of part[r * cols + c], rows in order from 0 (the
// order lbm3_wall_forces adds the same partial sums on the host)
__global__ __launch_bounds__(256) void k_column_sums(const double *part, long long rows, int cols, double *out)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
    for (long long r = 0; r < rows; ++r) s += part[r * cols + c];
    out[c] = s;
}

} // namespace lbm3

using namespace lbm3;

extern "C" {

int lbm3_wall_forces_noalloc(lbm3_sim *h, int level, const double *origin, int nBodies, double *out)
{
    if (!h || !out || nBodies <= 0 || nBodies > LBM3_MAX_BODIES) return set_error(LBM3_E_ARG, "bad force arguments");
    Sim &s = h->s;
    if (level < 0 || level >= s.levels || !s.L[level].built) return set_error(LBM3_E_STATE, "level not built");
    Level &l = s.L[level];
    for (int i = 0; i < 6 * nBodies; ++i) out[i] = 0.0;
    if (l.W == 0) return 0;
    const long long nb = (l.W + 255) / 256;
    const int cols = 6 * nBodies;
    if (nb * cols + cols > kScratchDoubles)
        return set_error(LBM3_E_ARG, "wall-adjacent cells x bodies beyond the no-allocation scratch (" +
                                         std::to_string(kScratchDoubles) + " doubles)");
    double *sp = nullptr;
    if (const int rc = scratch_ptr(&sp)) return rc;
    const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
    // every block writes all of its cols partial sums, so nothing needs clearing
    launch_wall_force(l.gWall.p, l.W, l.wallData.p, l.wallIndex.p, topo_of(l), l.geom, o, nBodies, sp, s.stream);
    double *tot = sp + nb * cols;
    hipLaunchKernelGGL(k_column_sums, dim3((cols + 255) / 256), dim3(256), 0, s.stream, sp, nb, cols, tot);
    if (const int rc = check_hip(hipGetLastError(), "wall force launch")) return rc;
    if (const int rc = check_hip(hipMemcpyAsync(out, tot, sizeof(double) * (size_t)cols, hipMemcpyDeviceToHost, s.stream),
                                 "hipMemcpyAsync")) return rc;
    return check_hip(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
}

} // extern "C"
