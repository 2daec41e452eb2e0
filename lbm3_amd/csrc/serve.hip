// Continuous-batching server of single-level simulations (the serving loop;
// the reference has no counterpart: its batch runner solves cases one after
// another).
//
// Requests queue on submit; every lbm3_server_step admits queued requests
// while the HBM budget and the slot count allow (building their grids on the
// GPU), advances every active request by one time step with four launches for
// all of them together (bulk, wall-adjacent, open boundary, periodic copy:
// k_multi_*), and retires the requests that reached their step count.  A
// request admitted while others run joins the next batched step, so work of
// requests that arrived at different times overlaps on the device.  Each
// domain keeps its own esotwist parity, collision variant and boundary tables;
// the arithmetic per cell is that of the single-simulation path.
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <vector>
#include "sim.hpp"
#include "levelargs.hpp"

#define HIPCHK_RET(expr)                              \
    do {                                              \
        const int rc_ = check_hip((expr), #expr);     \
        if (rc_) return rc_;                          \
    } while (0)
#define HIPCHK(expr) HIPCHK_RET(expr)

namespace lbm3 {

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    long long id = 0;
    lbm3_case c{};
    int storage = LBM3_STORE_F32;
    std::vector<float> tri;
    std::vector<int> counts;
    lbm3_bounds dom{};
    int steps = 0;
    int keepMacros = 0;
    long long estBytes = 0;
    double tSubmit = 0, tStart = 0, tFinish = 0;
    lbm3_sim *sim = nullptr;
    int done = 0;
    bool finished = false;
    int status = 0;                      // 0 ok, < 0: the LBM3_E_* code the request failed with
    long long fluidUpdates = 0;          // per step
    double kineticEnergy = 0, force[6] = {0, 0, 0, 0, 0, 0};
    std::vector<unsigned char> macros;
};

} // namespace

// pinned host staging of the per-step domain descriptors: a ring of buffers,
// each reused only after the copy that last read it has completed
struct Stage {
    unsigned char *host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct Server {
    long long budget = 0, used = 0;
    int maxActive = 0;
    long long nextId = 1;
    std::deque<Request *> queue;
    std::vector<Request *> active;
    std::map<long long, std::unique_ptr<Request>> all;
    hipStream_t stream = nullptr;
    DBuf<unsigned char> desc[5];      // device descriptors and block prefixes, one per storage type
    Stage stage[8];
    int nextStage = 0;
    long long stepsRun = 0, updatesRun = 0;
};

static int stage_take(Server &sv, size_t bytes, Stage *&out)
{
    Stage &g = sv.stage[sv.nextStage];
    sv.nextStage = (sv.nextStage + 1) % 8;
    if (!g.done) HIPCHK_RET(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    if (g.pending) HIPCHK_RET(hipEventSynchronize(g.done));
    g.pending = false;
    if (g.cap < bytes) {
        if (g.host) HIPCHK_RET(hipHostFree(g.host));
        g.host = nullptr;
        g.cap = 0;
        HIPCHK_RET(hipHostMalloc((void **)&g.host, bytes, hipHostMallocDefault));
        g.cap = bytes;
    }
    out = &g;
    return 0;
}

// device bytes a request will hold while active, from its domain size: the
// populations and the per-cell topology and wall records
static long long estimate_bytes(const Request &r)
{
    const double res = r.c.resGlobal > 0 ? r.c.resGlobal : 1.0;
    const double nx = (r.dom.xMax - r.dom.xMin) / res + 2, ny = (r.dom.yMax - r.dom.yMin) / res + 2,
                 nz = (r.dom.zMax - r.dom.zMin) / res + 2;
    const double cells = nx * ny * nz;
    return (long long)(cells * (27.0 * (double)storage_bytes(r.storage) + 64.0)) + (64LL << 20);
}

static int admit(Server &sv, Request &r)
{
    int rc = lbm3_sim_create(&r.c, r.storage, &r.sim);
    long long updates = 0;
    if (!rc) rc = lbm3_build_grids(r.sim, r.tri.empty() ? nullptr : r.tri.data(), r.counts.empty() ? nullptr : r.counts.data(),
                                   (int)r.counts.size(), r.dom, &updates);
    if (!rc && wide_level<double>(r.sim->s.L[0].N)) rc = set_error(LBM3_E_MEMORY, "request too large for the batched kernels");
    if (!rc) rc = lbm3_apply_initial_condition(r.sim, 0);
    if (!rc) rc = check_hip(hipStreamSynchronize(r.sim->s.stream), "request build");
    if (rc) {
        if (r.sim) lbm3_sim_destroy(r.sim);
        r.sim = nullptr;
        r.status = rc;
        r.finished = true;
        r.tStart = r.tFinish = now_s();
        return rc;
    }
    // batched steps run the boundary condition as its own pass after streaming
    Level &l = r.sim->s.L[0];
    l.openFusionChecked = true;
    l.fusedOpen = false;
    l.openSettled = true;
    r.fluidUpdates = updates;
    r.tStart = now_s();
    sv.used += r.estBytes;
    return 0;
}

static void retire(Server &sv, Request &r)
{
    Sim &s = r.sim->s;
    (void)lbm3_kinetic_energy(r.sim, &r.kineticEnergy);
    if (s.C.trackWallForce && s.L[0].W > 0) {
        const double origin[3] = {0.0, 0.0, 0.0};
        (void)lbm3_wall_forces(r.sim, 0, origin, 1, r.force);
    }
    if (r.keepMacros) {
        r.macros.resize(4 * (size_t)s.L[0].N * compute_bytes(r.storage));
        (void)lbm3_download_macros(r.sim, 0, r.macros.data());
    }
    lbm3_sim_destroy(r.sim);
    r.sim = nullptr;
    r.finished = true;
    r.tFinish = now_s();
    sv.used -= r.estBytes;
}

template <class S> static int batched_step(Server &sv, int kind, const std::vector<Request *> &group)
{
    const int n = (int)group.size();
    if (!n) return 0;
    std::vector<MultiDesc<S>> D(n);
    std::vector<int> st(4 * (size_t)(n + 1));
    int blocks[4] = {0, 0, 0, 0};
    for (int d = 0; d < n; ++d) {
        Sim &s = group[d]->sim->s;
        Level &l = s.L[0];
        MultiDesc<S> &m = D[d];
        m.step = step_args<S>(s, l);
        m.step.open = FusedOpen<typename StoreTraits<S>::compute>{};
        m.open = open_args<S>(s, l, l.info.iterationsFinished);
        m.perIndex = l.perIndex.p;
        m.perSource = l.perSource.p;
        m.nPeriodic = l.nPeriodic;
        m.flip = l.info.esotwistFlipper;
        m.k17 = relax_of<typename StoreTraits<S>::compute>(l.bulk).k17;
        const int per[4] = {(l.N + 255) / 256, (l.W + 255) / 256, (l.nOpen + 255) / 256, (l.nPeriodic + 255) / 256};
        for (int k = 0; k < 4; ++k) {
            st[(size_t)k * (n + 1) + d] = blocks[k];
            blocks[k] += per[k];
        }
    }
    for (int k = 0; k < 4; ++k) st[(size_t)k * (n + 1) + n] = blocks[k];
    // one upload of [descriptors | block prefixes] into this storage type's device buffer
    const size_t db = (sizeof(MultiDesc<S>) * (size_t)n + 255) & ~size_t(255), sb = sizeof(int) * st.size();
    DBuf<unsigned char> &dev = sv.desc[kind];
    if ((size_t)dev.n < db + sb) {
        HIPCHK_RET(hipStreamSynchronize(sv.stream));   // earlier steps may still read the old buffer
        HIPCHK_RET(dev.alloc((long long)(2 * (db + sb))));
    }
    Stage *g = nullptr;
    if (const int rc = stage_take(sv, db + sb, g)) return rc;
    memcpy(g->host, D.data(), sizeof(MultiDesc<S>) * (size_t)n);
    memcpy(g->host + db, st.data(), sb);
    HIPCHK_RET(hipMemcpyAsync(dev.p, g->host, db + sb, hipMemcpyHostToDevice, sv.stream));
    HIPCHK_RET(hipEventRecord(g->done, sv.stream));
    g->pending = true;
    launch_multi_step<S>(reinterpret_cast<const MultiDesc<S> *>(dev.p), reinterpret_cast<const int *>(dev.p + db), n,
                         blocks, sv.stream);
    HIPCHK_RET(hipGetLastError());
    for (Request *r : group) {
        Level &l = r->sim->s.L[0];
        l.info.esotwistFlipper = l.info.esotwistFlipper ? 0 : 1;
        l.info.iterationsFinished++;
        l.info.iterationsSinceTrackerReset++;
        r->done++;
        sv.updatesRun += r->fluidUpdates;
    }
    return 0;
}

} // namespace lbm3

struct lbm3_server {
    lbm3::Server s;
};

using namespace lbm3;

extern "C" {

int lbm3_server_create(long long hbmBudgetBytes, int maxActive, lbm3_server **out)
{
    if (!out || hbmBudgetBytes <= 0 || maxActive <= 0) return set_error(LBM3_E_ARG, "bad server arguments");
    lbm3_server *h = new (std::nothrow) lbm3_server();
    if (!h) return set_error(LBM3_E_MEMORY, "host allocation");
    h->s.budget = hbmBudgetBytes;
    h->s.maxActive = maxActive;
    const int rc = check_hip(hipStreamCreateWithFlags(&h->s.stream, hipStreamNonBlocking), "server stream");
    if (rc) { delete h; return rc; }
    *out = h;
    return 0;
}

int lbm3_server_destroy(lbm3_server *h)
{
    if (!h) return 0;
    (void)hipStreamSynchronize(h->s.stream);
    for (Request *r : h->s.active)
        if (r->sim) lbm3_sim_destroy(r->sim);
    for (Stage &g : h->s.stage) {
        if (g.host) (void)hipHostFree(g.host);
        if (g.done) (void)hipEventDestroy(g.done);
    }
    (void)hipStreamDestroy(h->s.stream);
    delete h;
    return 0;
}

int lbm3_server_submit(lbm3_server *h, const lbm3_case *c, int storage, const float *triangles, const int *triCounts,
                       int nBodies, lbm3_bounds domain, int steps, int keepMacros, long long *id)
{
    if (!h || !c || !id || steps <= 0 || nBodies < 0 || (nBodies > 0 && (!triangles || !triCounts)))
        return set_error(LBM3_E_ARG, "bad request");
    if (c->gridLevelCount != 1) return set_error(LBM3_E_ARG, "the server runs single-level cases");
    if (storage < LBM3_STORE_F32 || storage > LBM3_STORE_F16C) return set_error(LBM3_E_ARG, "unknown storage");
    Server &sv = h->s;
    auto r = std::make_unique<Request>();
    r->id = sv.nextId++;
    r->c = *c;
    r->storage = storage;
    long long total = 0;
    for (int b = 0; b < nBodies; ++b) {
        if (triCounts[b] < 0) return set_error(LBM3_E_ARG, "negative triangle count");
        r->counts.push_back(triCounts[b]);
        total += triCounts[b];
    }
    if (total > 0) r->tri.assign(triangles, triangles + 9 * total);
    r->dom = domain;
    r->steps = steps;
    r->keepMacros = keepMacros;
    r->estBytes = estimate_bytes(*r);
    if (r->estBytes > sv.budget) return set_error(LBM3_E_MEMORY, "request exceeds the server's HBM budget");
    r->tSubmit = now_s();
    *id = r->id;
    sv.queue.push_back(r.get());
    sv.all[r->id] = std::move(r);
    return 0;
}

int lbm3_server_step(lbm3_server *h, long long *completed, int cap, int *nCompleted)
{
    if (!h || !nCompleted || (cap > 0 && !completed)) return set_error(LBM3_E_ARG, "bad server step arguments");
    Server &sv = h->s;
    *nCompleted = 0;
    auto report = [&](Request *r) {
        if (*nCompleted < cap) completed[*nCompleted] = r->id;
        (*nCompleted)++;
    };
    // admit in arrival order while memory and slots allow
    while (!sv.queue.empty() && (int)sv.active.size() < sv.maxActive && sv.used + sv.queue.front()->estBytes <= sv.budget) {
        Request *r = sv.queue.front();
        sv.queue.pop_front();
        if (admit(sv, *r)) { report(r); continue; }
        sv.active.push_back(r);
    }
    if (sv.active.empty()) return 0;
    std::vector<Request *> groups[5];
    for (Request *r : sv.active) groups[r->storage].push_back(r);
    int rc = batched_step<float>(sv, LBM3_STORE_F32, groups[LBM3_STORE_F32]);
    if (!rc) rc = batched_step<__half>(sv, LBM3_STORE_F16, groups[LBM3_STORE_F16]);
    if (!rc) rc = batched_step<__hip_bfloat16>(sv, LBM3_STORE_BF16, groups[LBM3_STORE_BF16]);
    if (!rc) rc = batched_step<double>(sv, LBM3_STORE_F64, groups[LBM3_STORE_F64]);
    if (!rc) rc = batched_step<f16c>(sv, LBM3_STORE_F16C, groups[LBM3_STORE_F16C]);
    if (rc) return rc;
    sv.stepsRun++;
    bool any = false;
    for (Request *r : sv.active) any = any || r->done >= r->steps;
    if (!any) return 0;
    HIPCHK(hipStreamSynchronize(sv.stream));
    std::vector<Request *> still;
    for (Request *r : sv.active) {
        if (r->done >= r->steps) {
            retire(sv, *r);
            report(r);
        } else {
            still.push_back(r);
        }
    }
    sv.active.swap(still);
    return 0;
}

int lbm3_server_result(lbm3_server *h, long long id, double *out, int n)
{
    if (!h || !out || n < 12) return set_error(LBM3_E_ARG, "bad result arguments");
    auto it = h->s.all.find(id);
    if (it == h->s.all.end()) return set_error(LBM3_E_ARG, "unknown request");
    const Request &r = *it->second;
    if (!r.finished) return set_error(LBM3_E_STATE, "request not finished");
    out[0] = r.tStart - r.tSubmit;
    out[1] = r.tFinish - r.tSubmit;
    out[2] = (double)r.fluidUpdates * (double)r.steps;
    out[3] = r.kineticEnergy;
    for (int k = 0; k < 6; ++k) out[4 + k] = r.force[k];
    out[10] = (double)r.status;
    out[11] = (double)r.steps;
    return 0;
}

int lbm3_server_macros(lbm3_server *h, long long id, void *out, long long bytes)
{
    if (!h || !out) return set_error(LBM3_E_ARG, "null argument");
    auto it = h->s.all.find(id);
    if (it == h->s.all.end()) return set_error(LBM3_E_ARG, "unknown request");
    const Request &r = *it->second;
    if (!r.finished || !r.keepMacros || r.status) return set_error(LBM3_E_STATE, "no macros kept for this request");
    if (bytes < (long long)r.macros.size()) return set_error(LBM3_E_ARG, "buffer too small");
    memcpy(out, r.macros.data(), r.macros.size());
    return 0;
}

int lbm3_server_stats(lbm3_server *h, double *out)
{
    if (!h || !out) return set_error(LBM3_E_ARG, "null argument");
    const Server &sv = h->s;
    out[0] = (double)sv.active.size();
    out[1] = (double)sv.queue.size();
    out[2] = (double)sv.used;
    out[3] = (double)sv.stepsRun;
    out[4] = (double)sv.updatesRun;
    return 0;
}

} // extern "C"
