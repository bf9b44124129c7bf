// swd_plan.hip -- plans, device memory and the operators of libswd's C-ABI.
//
// A plan owns one HIP stream (default priority) on one device, the model in
// the kernels' form (dt*buoyancy at the two half points, dt*rho*vp^2 and
// rho*vp^2, each rounded exactly as the reference rounds them per cell), the
// CPML vectors, and a growing workspace from which every call carves its
// wavefields, strips, gathers and scratch.  Operators run a batch of shots;
// on SWD_PATH_BATCHED every wavefield of the batch shares each launch, on
// SWD_PATH_PLAIN shots run one after another with one kernel per sub-step,
// which is the single-shot baseline.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "swd.h"
#include "swd_internal.h"
#include "swd_kernels.h"
#include "swd_launch.h"
#include "swd_loops.h"

// ------------------------------------------------------------------ errors
static thread_local char g_last_error[1024] = "";

int swd_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" int swd_last_error(char *buf, size_t n)
{
    if (!buf || n == 0) return SWD_E_ARG;
    std::snprintf(buf, n, "%s", g_last_error);
    return SWD_OK;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return swd_fail(e_ == hipErrorOutOfMemory ? SWD_E_NOMEM : SWD_E_DEVICE, "%s failed: %s (%s:%d)", \
                            #expr, hipGetErrorString(e_), __FILE__, __LINE__);                            \
    } while (0)

#define SWD_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != SWD_OK) return r_; \
    } while (0)

extern "C" int swd_get_device_count(int *n)
{
    if (!n) return swd_fail(SWD_E_ARG, "swd_get_device_count: null argument");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *n = 0;
        return swd_fail(SWD_E_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *n = c;
    return SWD_OK;
}

// ------------------------------------------------------------------ plan
namespace {

struct Plan {
    int dev = 0;
    hipStream_t stream = nullptr;
    swd_grid g{};
    int pmlc = 0, ntx = 0, ntz = 0, ldz = 0, nframe = 0;
    size_t nfield = 0;   // ntx*ldz
    int path = SWD_PATH_FUSED;
    bool have_coefs = false, have_model = false;
    float *d_coef = nullptr;    // 4*ntx + 4*ntz
    float *d_model = nullptr;   // 4*nfield
    char *ws = nullptr;
    size_t ws_cap = 0;
    hipEvent_t ev[4] = {};
    hipEvent_t tev[4] = {};     // autotuner
    swd_stats stats{};
    KGrid kg{};
    int tile_req = 0;           // fused tile width asked for (0 = tune)
    int group_fwd = 0, group_bwd = 0;     // RTM shots per fused forward / backward loop (0 = sized to the cache)
    std::map<int, int> tuned_fwd, tuned_bwd;   // tuned tile width by instances per launch
    std::map<int, int> tuned_cig;              // the same for the backward with angle gathers
    std::map<int, int> tuned_born;             // the same for the Born forward pairs
    int streams = 2;            // RTM shot groups run on this many streams at once (1 .. kMaxStreams)
    hipStream_t side[3] = {};   // the other streams (default priority)
    hipEvent_t fork = nullptr, join[3] = {};
    bool graphs = true;         // replay fused time loops from captured hipGraphs
    std::map<std::vector<long long>, hipGraphExec_t> graph_cache;
    char *aux = nullptr;        // multi-grid launch tables
    size_t aux_cap = 0;
    size_t mem_limit = 0;       // device bytes a call may use (0 = the free device memory)
    std::vector<unsigned char> updown_last;   // what swd_updown_image_dev last uploaded (see there)
};

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Sequential carving of one workspace block.
struct Carver {
    size_t off = 0;
    template <class T>
    size_t add(size_t count)
    {
        const size_t o = align_up(off, 256);
        off = o + count * sizeof(T);
        return o;
    }
};

void drop_graphs(Plan &P)
{
    for (auto &kv : P.graph_cache) (void)hipGraphExecDestroy(kv.second);
    P.graph_cache.clear();
}

// Every entry point lays its own tables out in the workspace: what
// swd_updown_image_dev uploaded is stale after any other call (keep = false).
int ensure_ws(Plan &P, size_t bytes, bool keep_updown = false)
{
    if (!keep_updown) P.updown_last.clear();
    if (bytes <= P.ws_cap) return SWD_OK;
    P.updown_last.clear();
    drop_graphs(P);   // captured launches refer to the old workspace
    if (P.ws) {
        HIP_TRY(hipStreamSynchronize(P.stream));
        HIP_TRY(hipFree(P.ws));
        P.ws = nullptr;
        P.ws_cap = 0;
    }
    const size_t cap = align_up(bytes + bytes / 8, 1 << 20);
    HIP_TRY(hipMalloc(&P.ws, cap));
    P.ws_cap = cap;
    return SWD_OK;
}

template <class T>
T *at(Plan &P, size_t off) { return reinterpret_cast<T *>(P.ws + off); }

int ensure_aux(Plan &P, size_t bytes)
{
    if (bytes <= P.aux_cap) return SWD_OK;
    drop_graphs(P);
    if (P.aux) {
        HIP_TRY(hipStreamSynchronize(P.stream));
        HIP_TRY(hipFree(P.aux));
        P.aux = nullptr;
        P.aux_cap = 0;
    }
    const size_t cap = align_up(bytes + bytes / 4, 1 << 16);
    HIP_TRY(hipMalloc(&P.aux, cap));
    P.aux_cap = cap;
    return SWD_OK;
}

// Runs a launch sequence through a captured hipGraph: the first call with a
// key records the sequence (record() issues the launches on P.stream under
// capture) and instantiates it; later calls with the same key (same kernels,
// arguments and workspace addresses) replay it.  Only kernel launches are
// recorded.
template <class F>
int graph_run(Plan &P, const std::vector<long long> &key, F record, hipStream_t st)
{
    auto it = P.graph_cache.find(key);
    if (it == P.graph_cache.end()) {
        if (P.graph_cache.size() >= 16) drop_graphs(P);
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = record();
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc != SWD_OK) {
            if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIP_TRY(e);
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ei);
        it = P.graph_cache.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, st));
    return SWD_OK;
}

template <class F>
int graph_run(Plan &P, const std::vector<long long> &key, F record)
{
    return graph_run(P, key, record, P.stream);
}

int check_plan(void *plan, Plan *&P)
{
    if (!plan) return swd_fail(SWD_E_ARG, "null plan");
    P = static_cast<Plan *>(plan);
    if (!P->have_coefs || !P->have_model)
        return swd_fail(SWD_E_STATE, "plan needs swd_plan_set_coefs and swd_plan_set_model before running");
    HIP_TRY(hipSetDevice(P->dev));
    return SWD_OK;
}

// Receiver lookup of one shot: colmap[ntx] -> row map offset, rowmap -> trace
// of the node (the last trace on a node wins, as the reference's receiver
// loops do), dup = traces that share a node with a later trace.
struct RecvMap {
    std::vector<int> colmap, rowmap;
    std::vector<int> prim;   // primary trace of each trace
};

int build_recv(const Plan &P, const swd_shot &s, RecvMap &m)
{
    if (s.s_ix < 0 || s.s_ix >= P.ntx || s.s_iz < 0 || s.s_iz >= P.ntz)
        return swd_fail(SWD_E_ARG, "source (%d,%d) outside the padded grid %dx%d", s.s_ix, s.s_iz, P.ntx, P.ntz);
    if (s.nrec < 0 || (s.nrec > 0 && (!s.rec_ix || !s.rec_iz))) return swd_fail(SWD_E_ARG, "bad receiver list");
    m.colmap.assign(P.ntx, -1);
    m.rowmap.clear();
    for (int tr = 0; tr < s.nrec; tr++) {
        const int ix = s.rec_ix[tr], iz = s.rec_iz[tr];
        if (ix < 0 || ix >= P.ntx || iz < 0 || iz >= P.ntz)
            return swd_fail(SWD_E_ARG, "receiver %d at (%d,%d) outside the padded grid", tr, ix, iz);
        if (m.colmap[ix] < 0) {
            m.colmap[ix] = static_cast<int>(m.rowmap.size());
            m.rowmap.resize(m.rowmap.size() + P.ntz, -1);
        }
        m.rowmap[m.colmap[ix] + iz] = tr;
    }
    m.prim.resize(s.nrec);
    for (int tr = 0; tr < s.nrec; tr++) m.prim[tr] = m.rowmap[m.colmap[s.rec_ix[tr]] + s.rec_iz[tr]];
    return SWD_OK;
}

// Host-side description of a batch of shots, uploaded in one block.
struct Batch {
    int n = 0;
    std::vector<RecvMap> maps;
    std::vector<int> toff;         // first gather row of each shot
    int ntr = 0;
    std::vector<int> ints;         // all colmaps and rowmaps, concatenated
    std::vector<size_t> col_off, row_off;
    std::vector<int2> dup_pairs;   // (dup row, primary row), global rows
};

int build_batch(const Plan &P, int nshots, const swd_shot *shots, Batch &b)
{
    if (nshots < 1 || !shots) return swd_fail(SWD_E_ARG, "need at least one shot");
    b.n = nshots;
    b.maps.resize(nshots);
    b.toff.resize(nshots);
    b.col_off.resize(nshots);
    b.row_off.resize(nshots);
    b.ntr = 0;
    b.ints.clear();
    b.dup_pairs.clear();
    for (int k = 0; k < nshots; k++) {
        SWD_TRY(build_recv(P, shots[k], b.maps[k]));
        b.toff[k] = b.ntr;
        b.col_off[k] = b.ints.size();
        b.ints.insert(b.ints.end(), b.maps[k].colmap.begin(), b.maps[k].colmap.end());
        b.row_off[k] = b.ints.size();
        b.ints.insert(b.ints.end(), b.maps[k].rowmap.begin(), b.maps[k].rowmap.end());
        for (int tr = 0; tr < shots[k].nrec; tr++)
            if (b.maps[k].prim[tr] != tr) b.dup_pairs.push_back(make_int2(b.ntr + tr, b.ntr + b.maps[k].prim[tr]));
        b.ntr += shots[k].nrec;
    }
    if (b.ints.empty()) b.ints.push_back(-1);
    return SWD_OK;
}

void stats_begin(Plan &P, int which) { (void)hipEventRecord(P.ev[2 * which], P.stream); }
double stats_end(Plan &P, int which)
{
    (void)hipEventRecord(P.ev[2 * which + 1], P.stream);
    (void)hipEventSynchronize(P.ev[2 * which + 1]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, P.ev[2 * which], P.ev[2 * which + 1]);
    return ms;
}

}  // namespace

extern "C" int swd_plan_create(const swd_grid *g, int device, void **out)
{
    if (!g || !out) return swd_fail(SWD_E_ARG, "swd_plan_create: null argument");
    *out = nullptr;
    if (g->Lc < 1 || g->Lc > SWD_MAX_LC) return swd_fail(SWD_E_ARG, "Lc=%d outside 1..%d", g->Lc, SWD_MAX_LC);
    if (g->nx < 2 * g->Lc + 1 || g->nz < 2 * g->Lc + 1 || g->pml < 0 || g->nt < 2 || !(g->dx > 0) ||
        !(g->dz > 0) || !(g->dt > 0))
        return swd_fail(SWD_E_ARG, "invalid grid (nx=%d nz=%d pml=%d Lc=%d nt=%d)", g->nx, g->nz, g->pml, g->Lc, g->nt);
    {
        // The kernels form field offsets (ix*ldz + iz, ix*nz + iz over the interior) in 32 bits
        // and address them as bytes, so a field must stay below 2^30 floats.
        const long long pmlc = static_cast<long long>(g->pml) + g->Lc;
        const long long ntx = g->nx + 2 * pmlc, ldz = static_cast<long long>(align_up(g->nz + 2 * pmlc, 32));
        if (ntx * ldz >= (1LL << 30))
            return swd_fail(SWD_E_RANGE, "grid of %lld x %lld padded cells: a field exceeds 2^30 floats", ntx, ldz);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return swd_fail(SWD_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return swd_fail(SWD_E_ARG, "device %d outside 0..%d", device, ndev - 1);
    Plan *P = new Plan();
    P->dev = device;
    P->g = *g;
    P->pmlc = g->pml + g->Lc;
    P->ntx = g->nx + 2 * P->pmlc;
    P->ntz = g->nz + 2 * P->pmlc;
    P->ldz = static_cast<int>(align_up(P->ntz, 32));
    P->nfield = static_cast<size_t>(P->ntx) * P->ldz;
    P->nframe = 2 * g->Lc * (g->nx + g->nz - 2 * g->Lc);
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&P->stream);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&P->ev[i]);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&P->tev[i]);
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipStreamCreate(&P->side[i]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&P->fork, hipEventDisableTiming);
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&P->join[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&P->d_coef, sizeof(float) * 4 * (P->ntx + P->ntz));
    if (e == hipSuccess) e = hipMalloc(&P->d_model, sizeof(float) * 4 * P->nfield);
    if (e == hipSuccess) e = hipMemset(P->d_model, 0, sizeof(float) * 4 * P->nfield);
    if (e != hipSuccess) {
        swd_plan_destroy(P);
        return swd_fail(SWD_E_DEVICE, "swd_plan_create: %s", hipGetErrorString(e));
    }
    KGrid &k = P->kg;
    k.ntx = P->ntx; k.ntz = P->ntz; k.ldz = P->ldz;
    k.Lc = g->Lc; k.pmlc = P->pmlc; k.nx = g->nx; k.nz = g->nz; k.nt = g->nt; k.nframe = P->nframe;
    k.idx = static_cast<float>(1.0 / g->dx);   // float one_over_dx = 1.0/dx (RK:47)
    k.idz = static_cast<float>(1.0 / g->dz);
    k.fx0 = k.fz0 = 1;   // no fast tiles until the coefficients are known
    k.fx1 = k.fz1 = 0;
    float *c = P->d_coef;
    k.ax = c; k.axh = c + P->ntx; k.bx = c + 2 * P->ntx; k.bxh = c + 3 * P->ntx;
    float *cz = c + 4 * P->ntx;
    k.az = cz; k.azh = cz + P->ntz; k.bz = cz + 2 * P->ntz; k.bzh = cz + 3 * P->ntz;
    k.Bx = P->d_model; k.Bz = P->d_model + P->nfield; k.K = P->d_model + 2 * P->nfield; k.K0 = P->d_model + 3 * P->nfield;
    *out = P;
    return SWD_OK;
}

extern "C" int swd_plan_destroy(void *plan)
{
    if (!plan) return SWD_OK;
    Plan *P = static_cast<Plan *>(plan);
    (void)hipSetDevice(P->dev);
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    for (hipStream_t st : P->side)
        if (st) (void)hipStreamSynchronize(st);
    for (auto &kv : P->graph_cache) (void)hipGraphExecDestroy(kv.second);
    P->graph_cache.clear();
    if (P->ws) (void)hipFree(P->ws);
    if (P->aux) (void)hipFree(P->aux);
    if (P->d_coef) (void)hipFree(P->d_coef);
    if (P->d_model) (void)hipFree(P->d_model);
    for (auto &e : P->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : P->tev)
        if (e) (void)hipEventDestroy(e);
    if (P->fork) (void)hipEventDestroy(P->fork);
    for (hipEvent_t e : P->join)
        if (e) (void)hipEventDestroy(e);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    for (hipStream_t st : P->side)
        if (st) (void)hipStreamDestroy(st);
    delete P;
    return SWD_OK;
}

extern "C" int swd_plan_set_coefs(void *plan, const swd_coefs *c)
{
    if (!plan || !c || !c->rc || !c->a_x || !c->a_x_half || !c->b_x || !c->b_x_half || !c->a_z || !c->a_z_half ||
        !c->b_z || !c->b_z_half)
        return swd_fail(SWD_E_ARG, "swd_plan_set_coefs: null argument");
    Plan *P = static_cast<Plan *>(plan);
    HIP_TRY(hipSetDevice(P->dev));
    std::vector<float> h(4 * (P->ntx + P->ntz));
    const float *xs[4] = {c->a_x, c->a_x_half, c->b_x, c->b_x_half};
    const float *zs[4] = {c->a_z, c->a_z_half, c->b_z, c->b_z_half};
    for (int i = 0; i < 4; i++) {
        std::copy(xs[i], xs[i] + P->ntx, h.begin() + i * P->ntx);
        std::copy(zs[i], zs[i] + P->ntz, h.begin() + 4 * P->ntx + i * P->ntz);
    }
    HIP_TRY(hipMemcpy(P->d_coef, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    for (int i = 0; i < P->g.Lc; i++) P->kg.rc[i] = c->rc[i];
    drop_graphs(*P);   // captured launches carry the grid constants
    // interior box where every damping coefficient is zero (the fused kernel's fast tiles)
    auto box = [&](const float *a, const float *ah, int n, int &lo, int &hi) {
        lo = P->pmlc;
        hi = n - P->pmlc - 1;
        while (lo <= hi && (a[lo] != 0.f || ah[lo] != 0.f)) lo++;
        while (hi >= lo && (a[hi] != 0.f || ah[hi] != 0.f)) hi--;
        for (int i = lo; i <= hi; i++)
            if (a[i] != 0.f || ah[i] != 0.f) {
                lo = 1;
                hi = 0;
                break;
            }
    };
    box(c->a_x, c->a_x_half, P->ntx, P->kg.fx0, P->kg.fx1);
    box(c->a_z, c->a_z_half, P->ntz, P->kg.fz0, P->kg.fz1);
    P->have_coefs = true;
    return SWD_OK;
}

extern "C" int swd_plan_set_model(void *plan, const float *vp, const float *rho)
{
    if (!plan || !vp || !rho) return swd_fail(SWD_E_ARG, "swd_plan_set_model: null argument");
    Plan *P = static_cast<Plan *>(plan);
    const int chk = swd_check_model(&P->g, P->have_coefs ? P->kg.rc : nullptr, vp, rho);
    if (chk != SWD_OK) return chk;
    HIP_TRY(hipSetDevice(P->dev));
    const int ntx = P->ntx, ntz = P->ntz, ldz = P->ldz;
    const float dt = P->g.dt;
    std::vector<float> h(4 * P->nfield, 0.f);
    float *Bx = h.data(), *Bz = Bx + P->nfield, *K = Bz + P->nfield, *K0 = K + P->nfield;
    for (int ix = 0; ix < ntx; ix++)
        for (int iz = 0; iz < ntz; iz++) {
            const size_t ip = static_cast<size_t>(ix) * ntz + iz, o = static_cast<size_t>(ix) * ldz + iz;
            // 1/(0.5*(rho+rho')) in double, stored as float, times dt (RK:62-64, RK:111-113)
            if (ix + 1 < ntx) Bx[o] = dt * static_cast<float>(1.0 / (0.5 * (static_cast<double>(rho[ip]) + rho[ip + ntz])));
            if (iz + 1 < ntz) Bz[o] = dt * static_cast<float>(1.0 / (0.5 * (static_cast<double>(rho[ip]) + rho[ip + 1])));
            K[o] = dt * rho[ip] * vp[ip] * vp[ip];   // RK:171, left to right in float
            K0[o] = rho[ip] * vp[ip] * vp[ip];       // RK:475
        }
    HIP_TRY(hipMemcpy(P->d_model, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    P->have_model = true;
    return SWD_OK;
}

extern "C" int swd_plan_set_path(void *plan, int path)
{
    if (!plan || (path != SWD_PATH_PLAIN && path != SWD_PATH_BATCHED && path != SWD_PATH_FUSED))
        return swd_fail(SWD_E_ARG, "swd_plan_set_path: bad path %d", path);
    static_cast<Plan *>(plan)->path = path;
    return SWD_OK;
}

extern "C" int swd_plan_set_groups(void *plan, int fwd_shots, int bwd_shots)
{
    if (!plan || fwd_shots < 0 || bwd_shots < 0) return swd_fail(SWD_E_ARG, "swd_plan_set_groups: bad arguments");
    Plan *P = static_cast<Plan *>(plan);
    P->group_fwd = fwd_shots;
    P->group_bwd = bwd_shots;
    return SWD_OK;
}

extern "C" int swd_plan_set_graphs(void *plan, int on)
{
    if (!plan) return swd_fail(SWD_E_ARG, "swd_plan_set_graphs: null plan");
    Plan *P = static_cast<Plan *>(plan);
    P->graphs = on != 0;
    if (!P->graphs) drop_graphs(*P);
    return SWD_OK;
}

extern "C" int swd_plan_set_streams(void *plan, int streams)
{
    if (!plan || streams < 1 || streams > 4)
        return swd_fail(SWD_E_ARG, "swd_plan_set_streams: %d streams, not 1 .. 4", streams);
    static_cast<Plan *>(plan)->streams = streams;
    return SWD_OK;
}

extern "C" int swd_plan_set_tile(void *plan, int tile_x)
{
    if (!plan || (tile_x != 0 && tile_x != 16 && tile_x != 32 && tile_x != 64))
        return swd_fail(SWD_E_ARG, "swd_plan_set_tile: tile width %d not in {0, 16, 32, 64}", tile_x);
    Plan *P = static_cast<Plan *>(plan);
    P->tile_req = tile_x;
    P->tuned_fwd.clear();
    P->tuned_bwd.clear();
    P->tuned_cig.clear();
    P->tuned_born.clear();
    return SWD_OK;
}

extern "C" int swd_plan_get_stats(void *plan, swd_stats *st)
{
    if (!plan || !st) return swd_fail(SWD_E_ARG, "swd_plan_get_stats: null argument");
    *st = static_cast<Plan *>(plan)->stats;
    return SWD_OK;
}

// ------------------------------------------------------------------ time loops
namespace {

// Upload of the batch's integer maps and instance table into the workspace.
int upload_batch_on(Plan &P, const Batch &b, size_t o_ints, size_t o_pairs, hipStream_t st)
{
    HIP_TRY(hipMemcpyAsync(P.ws + o_ints, b.ints.data(), sizeof(int) * b.ints.size(), hipMemcpyHostToDevice, st));
    if (!b.dup_pairs.empty())
        HIP_TRY(hipMemcpyAsync(P.ws + o_pairs, b.dup_pairs.data(), sizeof(int2) * b.dup_pairs.size(), hipMemcpyHostToDevice, st));
    return SWD_OK;
}

int upload_batch(Plan &P, const Batch &b, size_t o_ints, size_t o_pairs)
{
    HIP_TRY(hipMemcpyAsync(P.ws + o_ints, b.ints.data(), sizeof(int) * b.ints.size(), hipMemcpyHostToDevice, P.stream));
    if (!b.dup_pairs.empty())
        HIP_TRY(hipMemcpyAsync(P.ws + o_pairs, b.dup_pairs.data(), sizeof(int2) * b.dup_pairs.size(),
                               hipMemcpyHostToDevice, P.stream));
    return SWD_OK;
}

int upload_insts(Plan &P, const std::vector<KInst> &h, size_t o_inst)
{
    HIP_TRY(hipMemcpyAsync(P.ws + o_inst, h.data(), sizeof(KInst) * h.size(), hipMemcpyHostToDevice, P.stream));
    return SWD_OK;
}

// Arrays per wavefield: p, vx, vz, psx, psz (CPML memory of the velocity
// update), pvx, pvz (CPML memory of the pressure update).
constexpr int kFieldArrays = 7;

void zero_fields(Plan &P, float *base, size_t count) { (void)hipMemsetAsync(base, 0, sizeof(float) * count, P.stream); }

void bind_fields(KInst &I, float *f, size_t nfield)
{
    I.p = f; I.vx = f + nfield; I.vz = f + 2 * nfield;
    I.psx = f + 3 * nfield; I.psz = f + 4 * nfield; I.pvx = f + 5 * nfield; I.pvz = f + 6 * nfield;
}

// Arrays per wavefield of the fused path: p, vx, vz, psx, psz in two slots,
// then pvx, pvz.
constexpr int kStepArrays = 12;

void bind_step(KStep &S, float *f, size_t nf)
{
    for (int k = 0; k < 2; k++) {
        S.p[k] = f + k * nf;
        S.vx[k] = f + (2 + k) * nf;
        S.vz[k] = f + (4 + k) * nf;
        S.psx[k] = f + (6 + k) * nf;
        S.psz[k] = f + (8 + k) * nf;
    }
    S.pvx = f + 10 * nf;
    S.pvz = f + 11 * nf;
}

int upload_steps(Plan &P, const std::vector<KStep> &h, size_t o)
{
    HIP_TRY(hipMemcpyAsync(P.ws + o, h.data(), sizeof(KStep) * h.size(), hipMemcpyHostToDevice, P.stream));
    return SWD_OK;
}

// Runs the fused launches its[0], its[1], ... (time steps in execution order)
// of n instances on stream st (the plan's by default).  The tile width is the
// plan's requested one, else the one tuned earlier for n instances, else it is
// tuned here: after two warm-up launches, six launches with each candidate
// width are timed and the fastest width runs the rest (the results do not
// depend on the width).
// Work launched after every fused step (inside the same captured graph):
// run(it, stream) queues one kernel launch; key identifies its arguments for
// the graph cache.  The launch counts count it beside the fused step.
struct StepHook {
    std::function<int(int, hipStream_t)> run;
    std::vector<long long> key;
};

int fused_steps(Plan &P, const KStep *d, int n, const std::vector<int> &its, int first_it, std::map<int, int> &tuned,
                long long *launches, bool pair = false, hipStream_t st = nullptr, const StepHook *hook = nullptr,
                bool cig = false, bool born = false)
{
    if (!st) st = P.stream;
    int tx = P.tile_req;
    if (tx == 0) {
        const auto f = tuned.find(n);
        if (f != tuned.end()) tx = f->second;
    }
    size_t k = 0;
    auto launch = [&](int w) -> int {
        const int it = its[k++];
        HIP_TRY(launch_fstep(P.kg, d, n, it, it == first_it, w, pair, cig, born, st));
        (*launches)++;
        if (hook) {
            SWD_TRY(hook->run(it, st));
            (*launches)++;
        }
        return SWD_OK;
    };
    if (tx == 0) {
        constexpr int kWarm = 2, kTry = 6;
        const int cand[3] = {16, 32, 64};
        if (its.size() >= static_cast<size_t>(kWarm + 3 * kTry)) {
            for (int j = 0; j < kWarm; j++) SWD_TRY(launch(32));
            HIP_TRY(hipEventRecord(P.tev[0], st));
            for (int c = 0; c < 3; c++) {
                for (int j = 0; j < kTry; j++) SWD_TRY(launch(cand[c]));
                HIP_TRY(hipEventRecord(P.tev[c + 1], st));
            }
            HIP_TRY(hipEventSynchronize(P.tev[3]));
            // A wider tile must be more than 10% faster than the narrower
            // one: narrow tiles keep more workgroups in flight, which the
            // groups overlapping on the plan's streams need (c2, 16 shots on
            // two streams, whole runs at widths 16/32/64: AFor 505/548/686 ms,
            // RTM 785/820/1043 ms), while a short timing of one launch
            // sequence can favour a wider one.
            float best = 0.f;
            for (int c = 0; c < 3; c++) {
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, P.tev[c], P.tev[c + 1]));
                if (c == 0 || ms < 0.9f * best) {
                    best = ms;
                    tx = cand[c];
                }
            }
            tuned[n] = tx;
        } else {
            tx = 32;
        }
    }
    if (k < its.size() && P.graphs) {
        std::vector<long long> key = {1, reinterpret_cast<long long>(d), n, its[k], static_cast<long long>(its.size() - k),
                                      first_it, tx, (pair ? 1 : 0) + (cig ? 2 : 0) + (born ? 4 : 0),
                                      reinterpret_cast<long long>(P.ws)};
        if (hook) key.insert(key.end(), hook->key.begin(), hook->key.end());
        const size_t k0 = k;
        SWD_TRY(graph_run(
            P, key,
            [&]() -> int {
                for (size_t j = k0; j < its.size(); j++) {
                    HIP_TRY(launch_fstep(P.kg, d, n, its[j], its[j] == first_it, tx, pair, cig, born, st));
                    if (hook) SWD_TRY(hook->run(its[j], st));
                }
                return SWD_OK;
            },
            st));
        *launches += static_cast<long long>(its.size() - k0) * (hook ? 2 : 1);
        return SWD_OK;
    }
    while (k < its.size()) SWD_TRY(launch(tx));
    return SWD_OK;
}

std::vector<int> steps_up(int nt)
{
    std::vector<int> v(nt);
    for (int it = 0; it < nt; it++) v[it] = it;
    return v;
}

// Shots per group of the fused RTM loops.  The per-shot working set of a
// step (double-buffered state, CPML memories in the frame, and on the
// backward the reconstruction, images and dp/dt) is kept within a budget of
// the 256 MB last-level cache, leaving room for the model, the halos and the
// streamed strips; groups are then evened out.  The forward, which also
// writes its strips every step, runs best with a smaller budget (measured at
// c2: 4-shot groups at 287-290 ms per 16-shot step, 6-shot at 298-313,
// 8-shot at 317).
constexpr double kCacheBudgetBwd = 192.0 * (1 << 20);
constexpr double kCacheBudgetFwd = 128.0 * (1 << 20);

int even_groups(int n, int g)
{
    g = std::max(1, std::min(n, g));
    const int ng = (n + g - 1) / g;
    return (n + ng - 1) / ng;
}

// Group boundaries of n instances in groups of at most g: ceil(n/g) groups,
// their count raised to a multiple of the streams where there are instances
// enough, so that every stream gets as many groups (with an odd count on two
// streams one of them runs alone at the end: c2 AFor, 32 wavefields, five
// groups of 6-7 took 555 ms a batch, six groups of 5-6 504 ms); sizes differ
// by at most one (the larger ones first).
std::vector<int> group_bounds(int n, int g, int streams)
{
    int ng = (n + g - 1) / g;
    const int up = (ng + streams - 1) / streams * streams;
    if (up <= n) ng = up;
    const int base = n / ng, extra = n % ng;
    std::vector<int> b(1, 0);
    for (int i = 0; i < ng; i++) b.push_back(b.back() + base + (i < extra ? 1 : 0));
    return b;
}

// Per-step working set of one forward wavefield: double-buffered state and
// the CPML memories of the frame.
double fwd_group_bytes(const Plan &P)
{
    const double N = static_cast<double>(P.ntx) * P.ntz, Ncpml = N - static_cast<double>(P.g.nx) * P.g.nz;
    return 4.0 * (6 * N + 6 * Ncpml);
}

void rtm_groups(const Plan &P, int nshots, int &gf, int &gb)
{
    const double N = static_cast<double>(P.ntx) * P.ntz, Nint = static_cast<double>(P.g.nx) * P.g.nz;
    const double Ncpml = N - Nint;
    const double fwd = fwd_group_bytes(P);
    const double bwd = 4.0 * (9 * Nint + 6 * N + 6 * Ncpml);
    // groups in flight on the plan's streams share the cache
    gf = P.group_fwd > 0 ? P.group_fwd : static_cast<int>(kCacheBudgetFwd / P.streams / fwd);
    gb = P.group_bwd > 0 ? P.group_bwd : static_cast<int>(kCacheBudgetBwd / P.streams / bwd);
    gf = even_groups(nshots, gf);
    gb = even_groups(nshots, gb);
}

// With several streams, consecutive shot groups take them in turn so that
// one group's launch gaps and tails overlap the others' work.
hipStream_t rtm_group_stream(const Plan &P, int g)
{
    const int ns = P.streams;
    return g % ns == 0 ? P.stream : P.side[g % ns - 1];
}

// The other streams wait for the work queued so far on the plan's stream.
int fork_streams(Plan &P)
{
    if (P.streams == 1) return SWD_OK;
    HIP_TRY(hipEventRecord(P.fork, P.stream));
    for (int i = 0; i < P.streams - 1; i++) HIP_TRY(hipStreamWaitEvent(P.side[i], P.fork, 0));
    return SWD_OK;
}

// The plan's stream waits for the work queued so far on the other streams.
int join_streams(Plan &P)
{
    for (int i = 0; i < P.streams - 1; i++) {
        HIP_TRY(hipEventRecord(P.join[i], P.side[i]));
        HIP_TRY(hipStreamWaitEvent(P.stream, P.join[i], 0));
    }
    return SWD_OK;
}

// Forward time loop of nshots shots with per_shot wavefields each.  Plain
// path: shot after shot, one kernel per sub-step.  Batched path: every
// wavefield of a group of shots in each launch, a velocity and a pressure
// kernel.
int forward_loop(Plan &P, const KInst *d_inst, int per_shot, int nshots)
{
    const KGrid &g = P.kg;
    long long launches = 0;
    stats_begin(P, 0);
    if (P.path == SWD_PATH_PLAIN) {
        for (int s = 0; s < nshots; s++) {
            const KInst *di = d_inst + s * per_shot;
            for (int it = 0; it < g.nt; it++) {
                HIP_TRY(launch_v(g, di, per_shot, false, 1, P.stream));
                HIP_TRY(launch_v(g, di, per_shot, false, 2, P.stream));
                HIP_TRY(launch_p(g, di, per_shot, false, it, P.stream));
                launches += 3;
            }
        }
    } else {
        // shots in groups whose per-step working set stays in the last-level
        // cache, the groups taking the plan's streams in turn; a shot's
        // wavefields stay in one group (Born's scattered field reads the
        // background's velocities of the same step).  c5 Born forward, 8 shots
        // on one stream: one group 931 ms per 2000 steps, groups of 4 814 ms,
        // of 2 710 ms, of 1 797 ms.
        const double per = 4.0 * kFieldArrays * static_cast<double>(P.nfield) * per_shot;
        int gs = P.group_fwd > 0 ? std::max(1, P.group_fwd / per_shot)
                                 : static_cast<int>(kCacheBudgetBwd / P.streams / per);
        gs = even_groups(nshots, gs);
        SWD_TRY(fork_streams(P));
        const std::vector<int> b = group_bounds(nshots, gs, P.streams);
        for (size_t i = 0; i + 1 < b.size(); i++)
            SWD_TRY(batched_forward_steps(g, d_inst + per_shot * b[i], per_shot * (b[i + 1] - b[i]), 0, g.nt,
                                          rtm_group_stream(P, static_cast<int>(i)), &launches));
        SWD_TRY(join_streams(P));
    }
    P.stats.fwd_ms = stats_end(P, 0);
    P.stats.fwd_launches = launches;
    P.stats.fwd_cells = static_cast<long long>(P.ntx) * P.ntz * g.nt * per_shot * nshots;
    return SWD_OK;
}

// Forward time loop of the fused path: n independent wavefield instances in
// groups whose per-step working set stays in the last-level cache, the groups
// taking the plan's streams in turn (as the RTM forward).
int fused_forward_loop(Plan &P, const KStep *d, int n)
{
    long long launches = 0;
    // with no strips to write, the backward's budget (c2 AFor, 32 wavefields
    // on two streams: groups of 5-6 504 ms a batch, of 4 550 ms)
    int gf = P.group_fwd > 0 ? P.group_fwd : static_cast<int>(kCacheBudgetBwd / P.streams / fwd_group_bytes(P));
    gf = even_groups(n, gf);
    stats_begin(P, 0);
    SWD_TRY(fork_streams(P));
    const std::vector<int> bounds = group_bounds(n, gf, P.streams);
    for (size_t g = 0; g + 1 < bounds.size(); g++)
        SWD_TRY(fused_steps(P, d + bounds[g], bounds[g + 1] - bounds[g], steps_up(P.g.nt), -1, P.tuned_fwd, &launches,
                            false, rtm_group_stream(P, static_cast<int>(g))));
    SWD_TRY(join_streams(P));
    P.stats.fwd_ms = stats_end(P, 0);
    P.stats.fwd_launches = launches;
    P.stats.fwd_cells = static_cast<long long>(P.ntx) * P.ntz * P.g.nt * n;
    return SWD_OK;
}

}  // namespace

// ------------------------------------------------------------------ operators
namespace {

// Forward modelling of a batch; gathers stay on the device at d_gather.
int forward_dev_fused(Plan &P, int nshots, const swd_shot *shots, const Batch &b, const float *wavelet, float *d_gather)
{
    const int nt = P.g.nt;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KStep>(nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_f = c.add<float>(kStepArrays * P.nfield * nshots);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    zero_fields(P, at<float>(P, o_f), kStepArrays * P.nfield * nshots);
    std::vector<KStep> h(nshots);
    for (int k = 0; k < nshots; k++) {
        KStep &S = h[k];
        std::memset(&S, 0, sizeof S);
        S.mode = KSTEP_FWD;
        bind_step(S, at<float>(P, o_f) + kStepArrays * P.nfield * k, P.nfield);
        S.wavelet = at<float>(P, o_wav);
        S.s_ix = shots[k].s_ix; S.s_iz = shots[k].s_iz;
        if (shots[k].nrec > 0) {
            S.colmap = at<int>(P, o_ints) + b.col_off[k];
            S.rowmap = at<int>(P, o_ints) + b.row_off[k];
            S.gather = d_gather + static_cast<size_t>(b.toff[k]) * nt;
        }
    }
    SWD_TRY(upload_steps(P, h, o_inst));
    SWD_TRY(fused_forward_loop(P, at<KStep>(P, o_inst), nshots));
    HIP_TRY(launch_copy_rows(d_gather, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    return SWD_OK;
}

int forward_dev(Plan &P, int nshots, const swd_shot *shots, const float *wavelet, float *d_gather)
{
    Batch b;
    SWD_TRY(build_batch(P, nshots, shots, b));
    if (P.path == SWD_PATH_FUSED) return forward_dev_fused(P, nshots, shots, b, wavelet, d_gather);
    const int nt = P.g.nt;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KInst>(nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_f = c.add<float>(kFieldArrays * P.nfield * nshots);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    zero_fields(P, at<float>(P, o_f), kFieldArrays * P.nfield * nshots);
    std::vector<KInst> h(nshots);
    for (int k = 0; k < nshots; k++) {
        KInst &I = h[k];
        std::memset(&I, 0, sizeof I);
        bind_fields(I, at<float>(P, o_f) + kFieldArrays * P.nfield * k, P.nfield);
        I.wavelet = at<float>(P, o_wav);
        I.s_ix = shots[k].s_ix; I.s_iz = shots[k].s_iz;
        if (shots[k].nrec > 0) {
            I.colmap = at<int>(P, o_ints) + b.col_off[k];
            I.rowmap = at<int>(P, o_ints) + b.row_off[k];
            I.gather = d_gather + static_cast<size_t>(b.toff[k]) * nt;
        }
    }
    SWD_TRY(upload_insts(P, h, o_inst));
    SWD_TRY(forward_loop(P, at<KInst>(P, o_inst), 1, nshots));
    HIP_TRY(launch_copy_rows(d_gather, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    return SWD_OK;
}

int total_traces(int nshots, const swd_shot *shots)
{
    int n = 0;
    for (int k = 0; k < nshots; k++) n += shots[k].nrec > 0 ? shots[k].nrec : 0;
    return n;
}

// Device scratch for host-buffer wrappers (separate from the plan workspace).
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t bytes)
    {
        if (bytes == 0) bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return swd_fail(SWD_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return SWD_OK;
    }
    template <class T> T *as() { return static_cast<T *>(p); }
};

}  // namespace

// ------------------------------------------------------------------ shot waves
namespace {

// Shots per wave of a call whose workspace needs fixed + n * per_shot bytes:
// as many as fit the plan's memory limit (by default 85% of the free device
// memory plus the plan's current workspace, which the call may reallocate).
int wave_shots(Plan &P, int nshots, double per_shot, double fixed, int &wave)
{
    double budget;
    if (P.mem_limit > 0) {
        budget = static_cast<double>(P.mem_limit);
    } else {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        budget = 0.85 * (static_cast<double>(fr) + static_cast<double>(P.ws_cap));
    }
    const double room = budget - fixed;
    if (room < per_shot)
        return swd_fail(SWD_E_NOMEM, "one shot needs %.0f MB of device memory, the budget is %.0f MB",
                        (per_shot + fixed) / 1e6, budget / 1e6);
    wave = static_cast<int>(std::min<double>(nshots, std::floor(room / per_shot)));
    P.stats.waves = (nshots + wave - 1) / wave;
    return SWD_OK;
}

int traces_before(const swd_shot *shots, int k)
{
    int n = 0;
    for (int i = 0; i < k; i++) n += shots[i].nrec > 0 ? shots[i].nrec : 0;
    return n;
}

}  // namespace

extern "C" int swd_forward(void *plan, int nshots, const swd_shot *shots, const float *wavelet, float *gather)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !gather || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_forward: null argument");
    const size_t n = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    DevBuf d;
    SWD_TRY(d.alloc(sizeof(float) * n));
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, 4.0 * (P->path == SWD_PATH_FUSED ? 12.0 : 7.0) * P->nfield, 4.0 * P->g.nt * 4, wave));
    for (int s0 = 0; s0 < nshots; s0 += wave)
        SWD_TRY(forward_dev(*P, std::min(wave, nshots - s0), shots + s0, wavelet,
                            d.as<float>() + static_cast<size_t>(traces_before(shots, s0)) * P->g.nt));
    HIP_TRY(hipMemcpy(gather, d.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return SWD_OK;
}

namespace {

// AFor of a batch with the three gathers in device memory ([sum nrec][nt]
// each); returns when they are complete.
int afor_dev(Plan *P, int nshots, const swd_shot *shots, const float *wavelet, const float *hwavelet,
             const float *kzcoef, int nfftz, float *d_gather, float *d_up, float *d_down)
{
    Batch b;
    SWD_TRY(build_batch(*P, nshots, shots, b));
    const int nt = P->g.nt;
    // receiver columns of each shot and their trace lists (gather row, iz)
    std::vector<int> colcap(static_cast<size_t>(P->ntx) * nshots, -1);
    std::vector<KCapCol> cols;
    std::vector<int2> traces;
    std::vector<int> cap0(nshots);
    for (int k = 0; k < nshots; k++) {
        cap0[k] = static_cast<int>(cols.size());
        std::vector<std::vector<int2>> per(P->ntx);
        for (int tr = 0; tr < shots[k].nrec; tr++)
            per[shots[k].rec_ix[tr]].push_back(make_int2(b.toff[k] + tr, shots[k].rec_iz[tr]));
        for (int ix = 0; ix < P->ntx; ix++)
            if (!per[ix].empty()) {
                colcap[static_cast<size_t>(k) * P->ntx + ix] = static_cast<int>(cols.size()) - cap0[k];
                cols.push_back(KCapCol{static_cast<int>(traces.size()), static_cast<int>(per[ix].size())});
                traces.insert(traces.end(), per[ix].begin(), per[ix].end());
            }
    }
    const int ncap = std::max<int>(1, static_cast<int>(cols.size()));
    // twiddles w_N^k = exp(-2 pi i k / N) for k < N/2, evaluated in double
    std::vector<float2> tw(nfftz / 2);
    for (int k = 0; k < nfftz / 2; k++) {
        const double ang = -1 * 2.0 * 3.14159265358979323846 * k / nfftz;
        tw[k] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
    }
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const bool fused = P->path == SWD_PATH_FUSED;
    const int per_wf = fused ? kStepArrays : kFieldArrays;
    const size_t o_inst = fused ? c.add<KStep>(2 * nshots) : c.add<KInst>(2 * nshots);
    const size_t o_wav = c.add<float>(2 * nt);
    const size_t o_colcap = c.add<int>(colcap.size());
    const size_t o_cols = c.add<KCapCol>(ncap);
    const size_t o_tr = c.add<int2>(std::max<size_t>(1, traces.size()));
    const size_t o_tw = c.add<float2>(nfftz / 2);
    const size_t o_coef = c.add<float>(nfftz);
    const size_t o_cap = c.add<float2>(static_cast<size_t>(nt) * ncap * nfftz);
    const size_t o_f = c.add<float>(per_wf * P->nfield * 2 * nshots);
    SWD_TRY(ensure_ws(*P, c.off));
    SWD_TRY(upload_batch(*P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P->ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->ws + o_wav + sizeof(float) * nt, hwavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->ws + o_colcap, colcap.data(), sizeof(int) * colcap.size(), hipMemcpyHostToDevice, P->stream));
    if (!cols.empty()) {
        HIP_TRY(hipMemcpyAsync(P->ws + o_cols, cols.data(), sizeof(KCapCol) * cols.size(), hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(P->ws + o_tr, traces.data(), sizeof(int2) * traces.size(), hipMemcpyHostToDevice, P->stream));
    }
    HIP_TRY(hipMemcpyAsync(P->ws + o_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(P->ws + o_coef, kzcoef, sizeof(float) * nfftz, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemsetAsync(P->ws + o_cap, 0, sizeof(float2) * static_cast<size_t>(nt) * ncap * nfftz, P->stream));
    zero_fields(*P, at<float>(*P, o_f), per_wf * P->nfield * 2 * nshots);
    if (fused) {
        std::vector<KStep> hs(2 * nshots);
        for (int k = 0; k < nshots; k++)
            for (int w = 0; w < 2; w++) {
                KStep &S = hs[2 * k + w];
                std::memset(&S, 0, sizeof S);
                S.mode = KSTEP_FWD;
                bind_step(S, at<float>(*P, o_f) + kStepArrays * P->nfield * (2 * k + w), P->nfield);
                S.wavelet = at<float>(*P, o_wav) + w * nt;   // p: rick, htp: hrick (FK:707, 747)
                S.s_ix = shots[k].s_ix; S.s_iz = shots[k].s_iz;
                if (w == 0 && shots[k].nrec > 0) {
                    S.colmap = at<int>(*P, o_ints) + b.col_off[k];
                    S.rowmap = at<int>(*P, o_ints) + b.row_off[k];
                    S.gather = d_gather + static_cast<size_t>(b.toff[k]) * nt;
                }
                if (shots[k].nrec > 0) {
                    S.capture = at<float2>(*P, o_cap) + cap0[k] * static_cast<size_t>(nfftz);
                    S.colcap = at<int>(*P, o_colcap) + static_cast<size_t>(k) * P->ntx;
                    S.capcomp = w;
                    S.ncap = ncap;
                    S.nfftz = nfftz;
                }
            }
        SWD_TRY(upload_steps(*P, hs, o_inst));
        SWD_TRY(fused_forward_loop(*P, at<KStep>(*P, o_inst), 2 * nshots));
    }
    std::vector<KInst> h(fused ? 0 : 2 * nshots);
    for (int k = 0; !fused && k < nshots; k++)
        for (int w = 0; w < 2; w++) {
            KInst &I = h[2 * k + w];
            std::memset(&I, 0, sizeof I);
            bind_fields(I, at<float>(*P, o_f) + kFieldArrays * P->nfield * (2 * k + w), P->nfield);
            I.wavelet = at<float>(*P, o_wav) + w * nt;   // p: rick, htp: hrick (FK:707, 747)
            I.s_ix = shots[k].s_ix; I.s_iz = shots[k].s_iz;
            if (w == 0 && shots[k].nrec > 0) {
                I.colmap = at<int>(*P, o_ints) + b.col_off[k];
                I.rowmap = at<int>(*P, o_ints) + b.row_off[k];
                I.gather = d_gather + static_cast<size_t>(b.toff[k]) * nt;
            }
            if (shots[k].nrec > 0) {
                I.capture = at<float2>(*P, o_cap) + cap0[k] * static_cast<size_t>(nfftz);
                I.colcap = at<int>(*P, o_colcap) + static_cast<size_t>(k) * P->ntx;
                I.capcomp = w;
                I.ncap = ncap;
                I.nfftz = nfftz;
            }
        }
    if (!fused) {
        SWD_TRY(upload_insts(*P, h, o_inst));
        SWD_TRY(forward_loop(*P, at<KInst>(*P, o_inst), 2, nshots));
    }
    HIP_TRY(launch_copy_rows(d_gather, at<int2>(*P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P->stream));
    if (!cols.empty()) {
        stats_begin(*P, 1);
        HIP_TRY(launch_kzsep(at<float2>(*P, o_cap), ncap, nfftz, at<float2>(*P, o_tw), at<float>(*P, o_coef),
                             at<KCapCol>(*P, o_cols), at<int2>(*P, o_tr), d_up, d_down, nt, P->stream));
        P->stats.fft_ms = stats_end(*P, 1);
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SWD_OK;
}

// AFor of a batch in waves of shots that fit the memory limit: per shot two
// wavefields and the column capture ([nt][columns][nfftz] complex).
int afor_waves(Plan *P, int nshots, const swd_shot *shots, const float *wavelet, const float *hwavelet,
               const float *kzcoef, int nfftz, float *d_gather, float *d_up, float *d_down)
{
    int maxcols = 1;
    for (int k = 0; k < nshots; k++) {
        std::vector<char> seen(P->ntx, 0);
        int n = 0;
        for (int tr = 0; tr < shots[k].nrec; tr++) {
            const int ix = shots[k].rec_ix[tr];
            if (ix >= 0 && ix < P->ntx && !seen[ix]) {
                seen[ix] = 1;
                n++;
            }
        }
        maxcols = std::max(maxcols, n);
    }
    const int per_wf = P->path == SWD_PATH_FUSED ? kStepArrays : kFieldArrays;
    const double per_shot = 8.0 * P->g.nt * maxcols * nfftz + 8.0 * per_wf * P->nfield + 4.0 * P->ntx;
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, per_shot, 8.0 * P->g.nt + 12.0 * nfftz, wave));
    const long long waves = P->stats.waves;
    for (int s0 = 0; s0 < nshots; s0 += wave) {
        const size_t off = static_cast<size_t>(traces_before(shots, s0)) * P->g.nt;
        SWD_TRY(afor_dev(P, std::min(wave, nshots - s0), shots + s0, wavelet, hwavelet, kzcoef, nfftz, d_gather + off,
                         d_up + off, d_down + off));
    }
    P->stats.waves = waves;
    return SWD_OK;
}

int afor_args(Plan *P, const char *fn, int nshots, const swd_shot *shots, const float *wavelet, const float *hwavelet,
              const float *kzcoef, int nfftz, const float *gather, const float *up, const float *down)
{
    if (!wavelet || !hwavelet || !kzcoef || !gather || !up || !down || !shots || nshots < 1)
        return swd_fail(SWD_E_ARG, "%s: null argument", fn);
    if (nfftz < P->ntz || (nfftz & (nfftz - 1)) || nfftz > 8192)
        return swd_fail(SWD_E_ARG, "nfftz=%d must be a power of two in [ntz=%d, 8192]", nfftz, P->ntz);
    return SWD_OK;
}

}  // namespace

extern "C" int swd_afor(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *hwavelet,
                        const float *kzcoef, int nfftz, float *gather, float *up, float *down)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    SWD_TRY(afor_args(P, "swd_afor", nshots, shots, wavelet, hwavelet, kzcoef, nfftz, gather, up, down));
    const size_t ng = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    DevBuf d;
    SWD_TRY(d.alloc(sizeof(float) * 3 * ng));
    float *d_gather = d.as<float>(), *d_up = d_gather + ng, *d_down = d_up + ng;
    SWD_TRY(afor_waves(P, nshots, shots, wavelet, hwavelet, kzcoef, nfftz, d_gather, d_up, d_down));
    HIP_TRY(hipMemcpy(gather, d_gather, sizeof(float) * ng, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(up, d_up, sizeof(float) * ng, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(down, d_down, sizeof(float) * ng, hipMemcpyDeviceToHost));
    return SWD_OK;
}

extern "C" int swd_afor_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                            const float *hwavelet, const float *kzcoef, int nfftz, float *gather, float *up,
                            float *down)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    SWD_TRY(afor_args(P, "swd_afor_dev", nshots, shots, wavelet, hwavelet, kzcoef, nfftz, gather, up, down));
    return afor_waves(P, nshots, shots, wavelet, hwavelet, kzcoef, nfftz, gather, up, down);
}

namespace {

// Angle gathers asked of an RTM call: thetan one-degree bins, delt-step
// averaging, output [nshots][ntx*ntz][thetan] in device memory.
struct CigSpec {
    int thetan, delt;
    float *out;
};

// Bytes per shot the angle gathers add to an RTM call's workspace.
double cig_bytes(const Plan &P, const CigSpec *cig)
{
    if (!cig) return 0.0;
    const double nint = static_cast<double>(P.g.nx) * P.g.nz;
    return 4.0 * nint * (cig->thetan + (cig->delt > 1 ? 2.0 * cig->delt : 0.0)) + sizeof(KCig);
}

// RTM of a batch on the fused path: one launch per forward step, one launch
// per backward step for the paired reconstruction and receiver wavefields of
// every shot of a group, imaging in the same launch (k_fstep); with angle
// gathers, k_cig after every backward launch (in the same captured graph).
int rtm_dev_fused(Plan &P, int nshots, const swd_shot *shots, const Batch &b, const float *wavelet, const float *adj,
                  float *image, float *illum_src, float *illum_rcv, float *gather_syn, const CigSpec *cig = nullptr)
{
    const int nt = P.g.nt;
    const size_t nf = P.nfield;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_fwd = c.add<KStep>(nshots);
    const size_t o_bwd = c.add<KStep>(2 * nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_strips = c.add<float>(static_cast<size_t>(nt) * P.nframe * nshots);
    const size_t o_img = c.add<float>(3 * nf * nshots);   // image, illum_s, illum_r
    const size_t o_f = c.add<float>(kStepArrays * nf * 2 * nshots);   // source side, then receiver side
    const size_t nint = static_cast<size_t>(P.g.nx) * P.g.nz;
    const size_t n_ring = cig && cig->delt > 1 ? 2 * static_cast<size_t>(cig->delt) * nint : 0;
    const size_t o_kcig = cig ? c.add<KCig>(nshots) : 0;
    const size_t o_cig = cig ? c.add<float>(nint * cig->thetan * nshots) : 0;
    const size_t o_ring = cig ? c.add<float>(std::max<size_t>(1, n_ring * nshots)) : 0;
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    zero_fields(P, at<float>(P, o_f), kStepArrays * nf * 2 * nshots);
    zero_fields(P, at<float>(P, o_img), 3 * nf * nshots);
    float *fields = at<float>(P, o_f), *imgs = at<float>(P, o_img);
    std::vector<KStep> hf(nshots), hb(2 * nshots);
    for (int k = 0; k < nshots; k++) {
        float *fs = fields + kStepArrays * nf * k, *fr = fields + kStepArrays * nf * (nshots + k);
        float *im = imgs + 3 * nf * k;
        KStep &F = hf[k], &S = hb[2 * k], &R = hb[2 * k + 1];
        std::memset(&F, 0, sizeof F);
        std::memset(&S, 0, sizeof S);
        std::memset(&R, 0, sizeof R);
        F.mode = KSTEP_FWD;
        S.mode = KSTEP_SRC;
        R.mode = KSTEP_RCV;
        bind_step(F, fs, nf);
        bind_step(S, fs, nf);
        bind_step(R, fr, nf);
        F.wavelet = S.wavelet = at<float>(P, o_wav);
        F.s_ix = S.s_ix = R.s_ix = shots[k].s_ix;
        F.s_iz = S.s_iz = R.s_iz = shots[k].s_iz;
        F.strips = S.strips = at<float>(P, o_strips) + static_cast<size_t>(nt) * P.nframe * k;
        if (shots[k].nrec > 0) {
            F.colmap = R.colmap = at<int>(P, o_ints) + b.col_off[k];
            F.rowmap = R.rowmap = at<int>(P, o_ints) + b.row_off[k];
            if (gather_syn) F.gather = gather_syn + static_cast<size_t>(b.toff[k]) * nt;
            R.adj = adj + static_cast<size_t>(b.toff[k]) * nt;
        }
        R.image = im;
        R.illum_s = im + nf;
        R.illum_r = im + 2 * nf;
    }
    // bin weights exp(-ii^2/18) of the angle gathers, rounded to float
    const float w1 = static_cast<float>(std::exp(-static_cast<double>(static_cast<float>(1.0 / 18.0))));
    const float w2 = static_cast<float>(std::exp(-static_cast<double>(static_cast<float>(4.0 / 18.0))));
    // With delt = 1 the paired backward launch of step it bins the gathers of
    // step it+1 itself (the fields of that level are at hand there); k_cig
    // then only adds step 0 after the loop.  Averaging over delt > 1 steps
    // runs k_cig after every launch.
    const bool fuse_cig = cig && cig->delt == 1;
    // A bin needs half-angle <= thetan + 1.5 degrees, that is cos(full angle)
    // >= cos((thetan + 1.5) pi / 90); the test uses thetan + 2 and a margin far
    // above the float error of its estimate.
    const float cthr = !cig || cig->thetan + 2 >= 90
                           ? -2.f
                           : static_cast<float>(std::cos((cig->thetan + 2) * 3.14159265358979323846 / 90.0) - 1e-3);
    if (fuse_cig)
        for (int k = 0; k < nshots; k++) {
            KStep &R = hb[2 * k + 1];
            R.cig = at<float>(P, o_cig) + nint * cig->thetan * k;
            R.thetan = cig->thetan;
            R.cw1 = w1;
            R.cw2 = w2;
            R.cthr = cthr;
        }
    SWD_TRY(upload_steps(P, hf, o_fwd));
    SWD_TRY(upload_steps(P, hb, o_bwd));
    if (cig) {
        zero_fields(P, at<float>(P, o_cig), nint * cig->thetan * nshots);
        if (n_ring) zero_fields(P, at<float>(P, o_ring), n_ring * nshots);
        std::vector<KCig> hk(nshots);
        for (int k = 0; k < nshots; k++) {
            const KStep &S = hb[2 * k], &R = hb[2 * k + 1];
            KCig &K = hk[k];
            for (int q = 0; q < 2; q++) {
                K.ps[q] = S.p[q];
                K.vxs[q] = S.vx[q];
                K.vzs[q] = S.vz[q];
                K.pr[q] = R.p[q];
                K.vxr[q] = R.vx[q];
                K.vzr[q] = R.vz[q];
            }
            K.ring = n_ring ? at<float>(P, o_ring) + n_ring * k : nullptr;
            K.cig = at<float>(P, o_cig) + nint * cig->thetan * k;
        }
        HIP_TRY(hipMemcpyAsync(P.ws + o_kcig, hk.data(), sizeof(KCig) * nshots, hipMemcpyHostToDevice, P.stream));
    }
    // Shots run in groups whose per-step working set stays in the last-level
    // cache; every shot keeps its strips and final state, so the groups of the
    // backward need not match those of the forward.
    int gf = 0, gb = 0;
    rtm_groups(P, nshots, gf, gb);
    auto group_stream = [&](int g) { return rtm_group_stream(P, g); };
    auto fork = [&]() { return fork_streams(P); };
    auto join = [&]() { return join_streams(P); };
    {
        long long launches = 0;
        stats_begin(P, 0);
        SWD_TRY(fork());
        const std::vector<int> gbf = group_bounds(nshots, gf, P.streams);
        for (size_t g = 0; g + 1 < gbf.size(); g++) {
            const int k0 = gbf[g], n = gbf[g + 1] - gbf[g];
            SWD_TRY(fused_steps(P, at<KStep>(P, o_fwd) + k0, n, steps_up(nt), -1, P.tuned_fwd, &launches, false,
                                group_stream(static_cast<int>(g))));
        }
        SWD_TRY(join());
        P.stats.fwd_ms = stats_end(P, 0);
        P.stats.fwd_launches = launches;
        P.stats.fwd_cells = static_cast<long long>(P.ntx) * P.ntz * nt * nshots;
    }
    if (gather_syn)
        HIP_TRY(launch_copy_rows(gather_syn, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    // outside the interior the reconstruction keeps the final forward state: both slots hold it
    const int fin = (nt - 1) & 1, oth = (nt - 2) & 1;
    for (int k = 0; k < nshots; k++) {
        const KStep &F = hf[k];
        HIP_TRY(hipMemcpyAsync(F.p[oth], F.p[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
        HIP_TRY(hipMemcpyAsync(F.vx[oth], F.vx[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
        HIP_TRY(hipMemcpyAsync(F.vz[oth], F.vz[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
    }
    // backward: it = nt-2 .. 0 (RK:1073), then the imaging of step 0
    std::vector<int> its;
    for (int it = nt - 2; it >= 0; it--) its.push_back(it);
    long long launches = 0;
    stats_begin(P, 1);
    SWD_TRY(fork());
    const std::vector<int> gbb = group_bounds(nshots, gb, P.streams);
    for (size_t g = 0; g + 1 < gbb.size(); g++) {
        const int k0 = gbb[g], n = gbb[g + 1] - gbb[g];
        const KStep *d = at<KStep>(P, o_bwd) + 2 * k0;
        const hipStream_t gs = group_stream(static_cast<int>(g));
        if (fuse_cig) {
            SWD_TRY(fused_steps(P, d, 2 * n, its, nt - 2, P.tuned_cig, &launches, true, gs, nullptr, true));
            HIP_TRY(launch_cig(P.kg, at<KCig>(P, o_kcig) + k0, n, cig->thetan, 1, w1, w2, cthr, 0, gs));   // step 0
            launches++;
            continue;
        }
        StepHook hook;
        if (cig) {
            const KCig *dk = at<KCig>(P, o_kcig) + k0;
            const int thetan = cig->thetan, delt = cig->delt;
            hook.run = [&P, dk, n, thetan, delt, w1, w2, cthr](int it, hipStream_t st) -> int {
                HIP_TRY(launch_cig(P.kg, dk, n, thetan, delt, w1, w2, cthr, it, st));
                return SWD_OK;
            };
            hook.key = {3, reinterpret_cast<long long>(dk), thetan, delt};
        }
        SWD_TRY(fused_steps(P, d, 2 * n, its, nt - 2, P.tuned_bwd, &launches, true, gs, cig ? &hook : nullptr));
    }
    SWD_TRY(join());
    P.stats.bwd_ms = stats_end(P, 1);
    P.stats.bwd_launches = launches;
    P.stats.bwd_cells = static_cast<long long>(nt - 1) * nshots *
                        (static_cast<long long>(P.ntx) * P.ntz + static_cast<long long>(P.g.nx) * P.g.nz);
    const size_t ntp = static_cast<size_t>(P.ntx) * P.ntz;
    for (int k = 0; k < nshots; k++) {
        float *outs[3] = {image + ntp * k, illum_src + ntp * k, illum_rcv + ntp * k};
        for (int q = 0; q < 3; q++)
            HIP_TRY(hipMemcpy2DAsync(outs[q], sizeof(float) * P.ntz, imgs + 3 * nf * k + q * nf, sizeof(float) * P.ldz,
                                     sizeof(float) * P.ntz, P.ntx, hipMemcpyDeviceToDevice, P.stream));
    }
    if (cig) {
        // interior [nx][nz][thetan] -> [ntx][ntz][thetan], zero outside
        const size_t th = static_cast<size_t>(cig->thetan), per = ntp * th;
        HIP_TRY(hipMemsetAsync(cig->out, 0, sizeof(float) * per * nshots, P.stream));
        for (int k = 0; k < nshots; k++)
            HIP_TRY(hipMemcpy2DAsync(cig->out + per * k + (static_cast<size_t>(P.pmlc) * P.ntz + P.pmlc) * th,
                                     sizeof(float) * P.ntz * th, at<float>(P, o_cig) + nint * th * k,
                                     sizeof(float) * P.g.nz * th, sizeof(float) * P.g.nz * th, P.g.nx,
                                     hipMemcpyDeviceToDevice, P.stream));
    }
    HIP_TRY(hipStreamSynchronize(P.stream));
    P.stats.ws_bytes = static_cast<long long>(P.ws_cap);
    return SWD_OK;
}

}  // namespace

namespace {
int rtm_dev_impl(Plan *P, int nshots, const swd_shot *shots, const float *wavelet, const float *adj, float *image,
                 float *illum_src, float *illum_rcv, float *gather_syn)
{
    Batch b;
    SWD_TRY(build_batch(*P, nshots, shots, b));
    if (P->path == SWD_PATH_FUSED)
        return rtm_dev_fused(*P, nshots, shots, b, wavelet, adj, image, illum_src, illum_rcv, gather_syn);
    const int nt = P->g.nt;
    const size_t nf = P->nfield;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KInst>(2 * nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_strips = c.add<float>(static_cast<size_t>(nt) * P->nframe * nshots);
    const size_t o_img = c.add<float>(4 * nf * nshots);        // image, illum_s, illum_r, dpdt
    const size_t o_f = c.add<float>(kFieldArrays * nf * 2 * nshots);   // source-side then receiver-side fields
    SWD_TRY(ensure_ws(*P, c.off));
    SWD_TRY(upload_batch(*P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P->ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P->stream));
    zero_fields(*P, at<float>(*P, o_f), kFieldArrays * nf * 2 * nshots);
    zero_fields(*P, at<float>(*P, o_img), 4 * nf * nshots);
    std::vector<KInst> h(2 * nshots);
    float *fields = at<float>(*P, o_f), *imgs = at<float>(*P, o_img);
    for (int k = 0; k < nshots; k++) {
        KInst &S = h[k];        // source side: forward, then reconstruction
        KInst &R = h[nshots + k];   // receiver side
        std::memset(&S, 0, sizeof S);
        std::memset(&R, 0, sizeof R);
        bind_fields(S, fields + kFieldArrays * nf * k, nf);
        bind_fields(R, fields + kFieldArrays * nf * (nshots + k), nf);
        S.wavelet = at<float>(*P, o_wav);
        S.s_ix = R.s_ix = shots[k].s_ix;
        S.s_iz = R.s_iz = shots[k].s_iz;
        S.strips = at<float>(*P, o_strips) + static_cast<size_t>(nt) * P->nframe * k;
        if (shots[k].nrec > 0) {
            S.colmap = R.colmap = at<int>(*P, o_ints) + b.col_off[k];
            S.rowmap = R.rowmap = at<int>(*P, o_ints) + b.row_off[k];
            if (gather_syn) S.gather = gather_syn + static_cast<size_t>(b.toff[k]) * nt;
            R.adj = adj + static_cast<size_t>(b.toff[k]) * nt;
        }
        float *im = imgs + 4 * nf * k;
        R.image = im; R.illum_s = im + nf; R.illum_r = im + 2 * nf;
        S.dpdt = R.dpdt = im + 3 * nf;
        R.wavelet = nullptr;
    }
    SWD_TRY(upload_insts(*P, h, o_inst));
    const KInst *dS = at<KInst>(*P, o_inst), *dR = dS + nshots;
    SWD_TRY(forward_loop(*P, dS, 1, nshots));
    if (gather_syn)
        HIP_TRY(launch_copy_rows(gather_syn, at<int2>(*P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P->stream));
    // backward: it = nt-2 .. 0 (RK:1073)
    const KGrid &g = P->kg;
    long long launches = 0;
    stats_begin(*P, 1);
    if (P->path == SWD_PATH_PLAIN) {
        for (int k = 0; k < nshots; k++)
            for (int it = nt - 2; it >= 0; it--) {
                HIP_TRY(launch_src_p(g, dS + k, 1, it, P->stream));
                HIP_TRY(launch_src_v(g, dS + k, 1, 1, P->stream));
                HIP_TRY(launch_src_v(g, dS + k, 1, 2, P->stream));
                HIP_TRY(launch_v(g, dR + k, 1, true, 1, P->stream));
                HIP_TRY(launch_v(g, dR + k, 1, true, 2, P->stream));
                HIP_TRY(launch_p(g, dR + k, 1, true, it, P->stream));
                launches += 6;
            }
    } else {
        SWD_TRY(batched_backward_loop(P->kg, dS, dR, nshots, P->stream, &launches));
    }
    P->stats.bwd_ms = stats_end(*P, 1);
    P->stats.bwd_launches = launches;
    P->stats.bwd_cells = static_cast<long long>(nt - 1) * nshots *
                         (static_cast<long long>(P->ntx) * P->ntz + static_cast<long long>(P->g.nx) * P->g.nz);
    // compact [ntx][ldz] -> [ntx][ntz]
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz;
    for (int k = 0; k < nshots; k++) {
        float *outs[3] = {image + ntp * k, illum_src + ntp * k, illum_rcv + ntp * k};
        for (int q = 0; q < 3; q++)
            HIP_TRY(hipMemcpy2DAsync(outs[q], sizeof(float) * P->ntz, imgs + 4 * nf * k + q * nf, sizeof(float) * P->ldz,
                                     sizeof(float) * P->ntz, P->ntx, hipMemcpyDeviceToDevice, P->stream));
    }
    HIP_TRY(hipStreamSynchronize(P->stream));
    return SWD_OK;
}

}  // namespace

extern "C" int swd_plan_set_memory_limit(void *plan, size_t bytes)
{
    if (!plan) return swd_fail(SWD_E_ARG, "swd_plan_set_memory_limit: null plan");
    static_cast<Plan *>(plan)->mem_limit = bytes;
    return SWD_OK;
}

extern "C" int swd_rtm_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                           float *image, float *illum_src, float *illum_rcv, float *gather_syn)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !adj || !image || !illum_src || !illum_rcv || !shots || nshots < 1)
        return swd_fail(SWD_E_ARG, "swd_rtm: null argument");
    const double nt = P->g.nt, nf = static_cast<double>(P->nfield);
    const double per = 4.0 * (nt * P->nframe + (P->path == SWD_PATH_FUSED ? 27.0 : 18.0) * nf);
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, per, 4.0 * nt * 4, wave));
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz;
    for (int s0 = 0; s0 < nshots; s0 += wave) {
        const int n = std::min(wave, nshots - s0);
        const size_t tr = static_cast<size_t>(traces_before(shots, s0)) * P->g.nt;
        SWD_TRY(rtm_dev_impl(P, n, shots + s0, wavelet, adj + tr, image + ntp * s0, illum_src + ntp * s0,
                             illum_rcv + ntp * s0, gather_syn ? gather_syn + tr : nullptr));
    }
    return SWD_OK;
}

extern "C" int swd_rtm(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                       float *image, float *illum_src, float *illum_rcv, float *gather_syn)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!adj || !image || !illum_src || !illum_rcv || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_rtm: null argument");
    const size_t ng = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz;
    DevBuf d;
    SWD_TRY(d.alloc(sizeof(float) * (2 * ng + 3 * ntp * nshots)));
    float *d_adj = d.as<float>(), *d_gs = d_adj + ng, *d_img = d_gs + ng;
    HIP_TRY(hipMemcpy(d_adj, adj, sizeof(float) * ng, hipMemcpyHostToDevice));
    SWD_TRY(swd_rtm_dev(plan, nshots, shots, wavelet, d_adj, d_img, d_img + ntp * nshots, d_img + 2 * ntp * nshots,
                        gather_syn ? d_gs : nullptr));
    HIP_TRY(hipMemcpy(image, d_img, sizeof(float) * ntp * nshots, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(illum_src, d_img + ntp * nshots, sizeof(float) * ntp * nshots, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(illum_rcv, d_img + 2 * ntp * nshots, sizeof(float) * ntp * nshots, hipMemcpyDeviceToHost));
    if (gather_syn) HIP_TRY(hipMemcpy(gather_syn, d_gs, sizeof(float) * ng, hipMemcpyDeviceToHost));
    return SWD_OK;
}

// ------------------------------------------------------------------ angle gathers
extern "C" int swd_rtm_cig_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                               int thetan, int delt, float *image, float *illum_src, float *illum_rcv, float *cig)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !adj || !image || !illum_src || !illum_rcv || !cig || !shots || nshots < 1)
        return swd_fail(SWD_E_ARG, "swd_rtm_cig: null argument");
    if (thetan < 1 || thetan > 1024 || delt < 1 || delt > 64)
        return swd_fail(SWD_E_ARG, "swd_rtm_cig: thetan=%d must be in [1, 1024] and delt=%d in [1, 64]", thetan, delt);
    const double nt = P->g.nt, nf = static_cast<double>(P->nfield);
    CigSpec spec{thetan, delt, nullptr};
    const double per = 4.0 * (nt * P->nframe + 27.0 * nf) + cig_bytes(*P, &spec);
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, per, 4.0 * nt * 4, wave));
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz;
    for (int s0 = 0; s0 < nshots; s0 += wave) {
        const int n = std::min(wave, nshots - s0);
        Batch b;
        SWD_TRY(build_batch(*P, n, shots + s0, b));
        const size_t tr = static_cast<size_t>(traces_before(shots, s0)) * P->g.nt;
        spec.out = cig + ntp * thetan * s0;
        SWD_TRY(rtm_dev_fused(*P, n, shots + s0, b, wavelet, adj + tr, image + ntp * s0, illum_src + ntp * s0,
                              illum_rcv + ntp * s0, nullptr, &spec));
    }
    return SWD_OK;
}

extern "C" int swd_rtm_cig(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                           int thetan, int delt, float *image, float *illum_src, float *illum_rcv, float *cig)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!adj || !image || !illum_src || !illum_rcv || !cig || !shots || nshots < 1)
        return swd_fail(SWD_E_ARG, "swd_rtm_cig: null argument");
    if (thetan < 1 || thetan > 1024) return swd_fail(SWD_E_ARG, "swd_rtm_cig: thetan=%d must be in [1, 1024]", thetan);
    const size_t ng = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz, nc = ntp * thetan * nshots;
    DevBuf d;
    SWD_TRY(d.alloc(sizeof(float) * (ng + 3 * ntp * nshots + nc)));
    float *d_adj = d.as<float>(), *d_img = d_adj + ng, *d_cig = d_img + 3 * ntp * nshots;
    HIP_TRY(hipMemcpy(d_adj, adj, sizeof(float) * ng, hipMemcpyHostToDevice));
    SWD_TRY(swd_rtm_cig_dev(plan, nshots, shots, wavelet, d_adj, thetan, delt, d_img, d_img + ntp * nshots,
                            d_img + 2 * ntp * nshots, d_cig));
    HIP_TRY(hipMemcpy(image, d_img, sizeof(float) * ntp * nshots, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(illum_src, d_img + ntp * nshots, sizeof(float) * ntp * nshots, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(illum_rcv, d_img + 2 * ntp * nshots, sizeof(float) * ntp * nshots, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cig, d_cig, sizeof(float) * nc, hipMemcpyDeviceToHost));
    return SWD_OK;
}

// ------------------------------------------------------------------ up/down imaging
namespace {

// twiddles w_N^k = exp(-2 pi i k / N) for k < N/2, evaluated in double
std::vector<float2> kz_twiddles(int nfftz)
{
    std::vector<float2> tw(nfftz / 2);
    for (int k = 0; k < nfftz / 2; k++) {
        const double ang = -1 * 2.0 * 3.14159265358979323846 * k / nfftz;
        tw[k] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
    }
    return tw;
}

// Wavefield arrays of one shot of the up/down RTM: four wavefields (12 arrays
// each), then du, ud, uu, dd, image, illum_s, illum_r and the two dp/dt.
constexpr int kUdImages = 9;
constexpr int kUdPerShot = 4 * kStepArrays + kUdImages;

// Up/down-separated RTM of a batch on the fused kernels.  The forward runs two
// wavefields per shot (wavelet and hwavelet) and saves both strips and final
// states; every backward launch pairs the reconstruction and receiver
// wavefields of both (four per shot; the first pair also applies the plain
// imaging condition when image outputs are asked for) and stores the two
// reconstructions' dp/dt, and k_updown then splits and cross-correlates the
// columns of that step.  Outputs are device pointers [nshots][ntx][ntz].
int rtm_updown_fused(Plan &P, int nshots, const swd_shot *shots, const Batch &b, const float *wavelet,
                     const float *hwavelet, const float *adj, const float *hadj, const float *kzcoef, int nfftz,
                     const swd_updown_images &out)
{
    const int nt = P.g.nt;
    const size_t nf = P.nfield;
    const bool plain = out.image != nullptr, both = out.uu != nullptr;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_fwd = c.add<KStep>(2 * nshots);
    const size_t o_bwd = c.add<KStep>(4 * nshots);
    const size_t o_ud = c.add<KUpDown>(nshots);
    const size_t o_wav = c.add<float>(2 * nt);
    const size_t o_tw = c.add<float2>(nfftz / 2);
    const size_t o_coef = c.add<float>(nfftz);
    const size_t o_strips = c.add<float>(2 * static_cast<size_t>(nt) * P.nframe * nshots);
    const size_t o_f = c.add<float>(kUdPerShot * nf * nshots);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav + sizeof(float) * nt, hwavelet, sizeof(float) * nt, hipMemcpyHostToDevice,
                           P.stream));
    const std::vector<float2> tw = kz_twiddles(nfftz);
    HIP_TRY(hipMemcpyAsync(P.ws + o_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice, P.stream));
    HIP_TRY(hipMemcpyAsync(P.ws + o_coef, kzcoef, sizeof(float) * nfftz, hipMemcpyHostToDevice, P.stream));
    zero_fields(P, at<float>(P, o_f), kUdPerShot * nf * nshots);
    std::vector<KStep> hf(2 * nshots), hb(4 * nshots);
    std::vector<KUpDown> hu(nshots);
    for (int k = 0; k < nshots; k++) {
        float *base = at<float>(P, o_f) + kUdPerShot * nf * k;
        float *im = base + 4 * kStepArrays * nf;   // du, ud, uu, dd, image, illum_s, illum_r, dp/dt, dp_h/dt
        for (int h = 0; h < 2; h++) {
            KStep &F = hf[2 * k + h], &S = hb[4 * k + 2 * h], &R = hb[4 * k + 2 * h + 1];
            std::memset(&F, 0, sizeof F);
            std::memset(&S, 0, sizeof S);
            std::memset(&R, 0, sizeof R);
            F.mode = KSTEP_FWD;
            S.mode = KSTEP_SRC;
            R.mode = KSTEP_RCV;
            bind_step(F, base + kStepArrays * nf * h, nf);
            bind_step(S, base + kStepArrays * nf * h, nf);
            bind_step(R, base + kStepArrays * nf * (2 + h), nf);
            F.wavelet = S.wavelet = at<float>(P, o_wav) + static_cast<size_t>(nt) * h;
            F.s_ix = S.s_ix = R.s_ix = shots[k].s_ix;
            F.s_iz = S.s_iz = R.s_iz = shots[k].s_iz;
            F.strips = S.strips = at<float>(P, o_strips) + static_cast<size_t>(nt) * P.nframe * (2 * k + h);
            if (shots[k].nrec > 0) {
                F.colmap = R.colmap = at<int>(P, o_ints) + b.col_off[k];
                F.rowmap = R.rowmap = at<int>(P, o_ints) + b.row_off[k];
                R.adj = (h ? hadj : adj) + static_cast<size_t>(b.toff[k]) * nt;
            }
            S.dpdt = im + (7 + h) * nf;
            if (h == 0 && plain) {
                R.image = im + 4 * nf;
                R.illum_s = im + 5 * nf;
                R.illum_r = im + 6 * nf;
            }
        }
        KUpDown &U = hu[k];
        U.ds = im + 7 * nf;
        U.dh = im + 8 * nf;
        for (int q = 0; q < 2; q++) {
            U.pr[q] = hb[4 * k + 1].p[q];
            U.ph[q] = hb[4 * k + 3].p[q];
        }
        U.du = im;
        U.ud = im + nf;
        U.uu = both ? im + 2 * nf : nullptr;
        U.dd = both ? im + 3 * nf : nullptr;
        U.tw = at<float2>(P, o_tw);
        U.coef = at<float>(P, o_coef);
        U.ld = P.ldz;
    }
    SWD_TRY(upload_steps(P, hf, o_fwd));
    SWD_TRY(upload_steps(P, hb, o_bwd));
    HIP_TRY(hipMemcpyAsync(P.ws + o_ud, hu.data(), sizeof(KUpDown) * nshots, hipMemcpyHostToDevice, P.stream));
    // two wavefields per shot in both phases: the backward takes half the
    // plain RTM's shots per group; the forward keeps the plain RTM's count
    // (c2, 16 shots, nt 1000, three streams: one-shot forward groups 157 ms,
    // two-shot groups 131 ms -- nine groups, even over the streams; backward
    // one-shot groups 311 ms, two-shot 362 ms)
    int gf = 0, gb = 0;
    rtm_groups(P, nshots, gf, gb);
    if (P.group_bwd <= 0) gb = even_groups(nshots, std::max(1, gb / 2));
    {
        long long launches = 0;
        stats_begin(P, 0);
        SWD_TRY(fork_streams(P));
        const std::vector<int> gbf = group_bounds(nshots, gf, P.streams);
        for (size_t g = 0; g + 1 < gbf.size(); g++) {
            const int k0 = gbf[g], n = gbf[g + 1] - gbf[g];
            SWD_TRY(fused_steps(P, at<KStep>(P, o_fwd) + 2 * k0, 2 * n, steps_up(nt), -1, P.tuned_fwd, &launches, false,
                                rtm_group_stream(P, static_cast<int>(g))));
        }
        SWD_TRY(join_streams(P));
        P.stats.fwd_ms = stats_end(P, 0);
        P.stats.fwd_launches = launches;
        P.stats.fwd_cells = static_cast<long long>(P.ntx) * P.ntz * nt * 2 * nshots;
    }
    // outside the interior the reconstruction keeps the final forward state: both slots hold it
    const int fin = (nt - 1) & 1, oth = (nt - 2) & 1;
    for (const KStep &F : hf) {
        HIP_TRY(hipMemcpyAsync(F.p[oth], F.p[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
        HIP_TRY(hipMemcpyAsync(F.vx[oth], F.vx[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
        HIP_TRY(hipMemcpyAsync(F.vz[oth], F.vz[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
    }
    std::vector<int> its;
    for (int it = nt - 2; it >= 0; it--) its.push_back(it);
    long long launches = 0;
    stats_begin(P, 1);
    SWD_TRY(fork_streams(P));
    const std::vector<int> gbb = group_bounds(nshots, gb, P.streams);
    for (size_t g = 0; g + 1 < gbb.size(); g++) {
        const int k0 = gbb[g], n = gbb[g + 1] - gbb[g];
        const KUpDown *du = at<KUpDown>(P, o_ud) + k0;
        StepHook hook;
        hook.run = [&P, du, n, nfftz](int it, hipStream_t st) -> int {
            HIP_TRY(launch_updown(P.kg, du, n, nfftz, it, st));
            return SWD_OK;
        };
        hook.key = {2, reinterpret_cast<long long>(du), nfftz};
        SWD_TRY(fused_steps(P, at<KStep>(P, o_bwd) + 4 * k0, 4 * n, its, nt - 2, P.tuned_bwd, &launches, true,
                            rtm_group_stream(P, static_cast<int>(g)), &hook));
    }
    SWD_TRY(join_streams(P));
    P.stats.bwd_ms = stats_end(P, 1);
    P.stats.bwd_launches = launches;
    P.stats.bwd_cells = static_cast<long long>(nt - 1) * 2 * nshots *
                        (static_cast<long long>(P.ntx) * P.ntz + static_cast<long long>(P.g.nx) * P.g.nz);
    const size_t ntp = static_cast<size_t>(P.ntx) * P.ntz;
    for (int k = 0; k < nshots; k++) {
        const float *im = at<float>(P, o_f) + kUdPerShot * nf * k + 4 * kStepArrays * nf;
        float *dst[7] = {out.du, out.ud, out.uu, out.dd, out.image, out.illum_src, out.illum_rcv};
        for (int q = 0; q < 7; q++)
            if (dst[q])
                HIP_TRY(hipMemcpy2DAsync(dst[q] + ntp * k, sizeof(float) * P.ntz, im + q * nf, sizeof(float) * P.ldz,
                                         sizeof(float) * P.ntz, P.ntx, hipMemcpyDeviceToDevice, P.stream));
    }
    HIP_TRY(hipStreamSynchronize(P.stream));
    P.stats.ws_bytes = static_cast<long long>(P.ws_cap);
    return SWD_OK;
}

int check_updown_args(const Plan &P, const char *fn, int nshots, const swd_shot *shots, const float *wavelet,
                      const float *hwavelet, const float *adj, const float *hadj, const float *kzcoef, int nfftz,
                      const swd_updown_images *out)
{
    if (!shots || nshots < 1 || !wavelet || !hwavelet || !adj || !hadj || !kzcoef || !out)
        return swd_fail(SWD_E_ARG, "%s: null argument", fn);
    if (!out->du || !out->ud) return swd_fail(SWD_E_ARG, "%s: du and ud outputs are required", fn);
    if ((out->uu == nullptr) != (out->dd == nullptr))
        return swd_fail(SWD_E_ARG, "%s: uu and dd are both given or both null", fn);
    if ((out->image == nullptr) != (out->illum_src == nullptr) || (out->image == nullptr) != (out->illum_rcv == nullptr))
        return swd_fail(SWD_E_ARG, "%s: image, illum_src and illum_rcv are all given or all null", fn);
    if (nfftz < P.g.nz || nfftz < 128 || (nfftz & (nfftz - 1)) || nfftz > 4096)
        return swd_fail(SWD_E_ARG, "%s: nfftz=%d must be a power of two in [max(nz=%d, 128), 4096]", fn, nfftz,
                        P.g.nz);
    return SWD_OK;
}

}  // namespace

extern "C" int swd_rtm_updown_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                                  const float *hwavelet, const float *adj, const float *hadj, const float *kzcoef,
                                  int nfftz, const swd_updown_images *out)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    SWD_TRY(check_updown_args(*P, "swd_rtm_updown", nshots, shots, wavelet, hwavelet, adj, hadj, kzcoef, nfftz, out));
    const double nt = P->g.nt, nf = static_cast<double>(P->nfield);
    const double per = 4.0 * (2 * nt * P->nframe + kUdPerShot * nf);
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, per, 4.0 * (2 * nt + 1.5 * nfftz), wave));
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz;
    for (int s0 = 0; s0 < nshots; s0 += wave) {
        const int n = std::min(wave, nshots - s0);
        Batch b;
        SWD_TRY(build_batch(*P, n, shots + s0, b));
        const size_t tr = static_cast<size_t>(traces_before(shots, s0)) * P->g.nt, o = ntp * s0;
        swd_updown_images w = *out;
        for (float **q : {&w.du, &w.ud, &w.uu, &w.dd, &w.image, &w.illum_src, &w.illum_rcv})
            if (*q) *q += o;
        SWD_TRY(rtm_updown_fused(*P, n, shots + s0, b, wavelet, hwavelet, adj + tr, hadj + tr, kzcoef, nfftz, w));
    }
    return SWD_OK;
}

extern "C" int swd_updown_image_dev(void *plan, int nshots, const float *src, const float *hsrc, const float *rcv,
                                    const float *hrcv, const float *kzcoef, int nfftz, const swd_updown_images *out)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!src || !hsrc || !rcv || !hrcv || !kzcoef || !out || nshots < 1)
        return swd_fail(SWD_E_ARG, "swd_updown_image_dev: null argument");
    if (!out->du || !out->ud) return swd_fail(SWD_E_ARG, "swd_updown_image_dev: du and ud outputs are required");
    if ((out->uu == nullptr) != (out->dd == nullptr))
        return swd_fail(SWD_E_ARG, "swd_updown_image_dev: uu and dd are both given or both null");
    if (nfftz < P->g.nz || nfftz < 128 || (nfftz & (nfftz - 1)) || nfftz > 4096)
        return swd_fail(SWD_E_ARG, "swd_updown_image_dev: nfftz=%d must be a power of two in [max(nz=%d, 128), 4096]",
                        nfftz, P->g.nz);
    Carver c;
    const size_t o_ud = c.add<KUpDown>(nshots);
    const size_t o_tw = c.add<float2>(nfftz / 2);
    const size_t o_coef = c.add<float>(nfftz);
    SWD_TRY(ensure_ws(*P, c.off, true));
    const size_t ntp = static_cast<size_t>(P->ntx) * P->ntz;
    std::vector<KUpDown> hu(nshots);
    for (int k = 0; k < nshots; k++) {
        KUpDown &U = hu[k];
        const size_t o = ntp * k;
        U.ds = src + o;
        U.dh = hsrc + o;
        U.pr[0] = U.pr[1] = rcv + o;
        U.ph[0] = U.ph[1] = hrcv + o;
        U.du = out->du + o;
        U.ud = out->ud + o;
        U.uu = out->uu ? out->uu + o : nullptr;
        U.dd = out->dd ? out->dd + o : nullptr;
        U.tw = at<float2>(*P, o_tw);
        U.coef = at<float>(*P, o_coef);
        U.ld = P->ntz;
    }
    // descriptors, twiddles and taper go up only when they differ from the
    // last call's, so repeated calls queue nothing but the kernel
    std::vector<unsigned char> key(sizeof(KUpDown) * nshots + sizeof(float) * nfftz + sizeof(char *));
    std::memcpy(key.data(), hu.data(), sizeof(KUpDown) * nshots);
    std::memcpy(key.data() + sizeof(KUpDown) * nshots, kzcoef, sizeof(float) * nfftz);
    std::memcpy(key.data() + key.size() - sizeof(char *), &P->ws, sizeof(char *));
    if (key != P->updown_last) {
        const std::vector<float2> tw = kz_twiddles(nfftz);
        HIP_TRY(hipMemcpyAsync(P->ws + o_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(P->ws + o_coef, kzcoef, sizeof(float) * nfftz, hipMemcpyHostToDevice, P->stream));
        HIP_TRY(hipMemcpyAsync(P->ws + o_ud, hu.data(), sizeof(KUpDown) * nshots, hipMemcpyHostToDevice, P->stream));
        P->updown_last.swap(key);
    }
    HIP_TRY(launch_updown(P->kg, at<KUpDown>(*P, o_ud), nshots, nfftz, 0, P->stream));
    return SWD_OK;
}

extern "C" int swd_plan_get_stream(void *plan, void **stream)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!stream) return swd_fail(SWD_E_ARG, "swd_plan_get_stream: null argument");
    *stream = static_cast<void *>(P->stream);
    return SWD_OK;
}

extern "C" int swd_rtm_updown(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                              const float *hwavelet, const float *adj, const float *hadj, const float *kzcoef,
                              int nfftz, const swd_updown_images *out)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    SWD_TRY(check_updown_args(*P, "swd_rtm_updown", nshots, shots, wavelet, hwavelet, adj, hadj, kzcoef, nfftz, out));
    const size_t ng = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    const size_t ni = static_cast<size_t>(P->ntx) * P->ntz * nshots;
    float *host[7] = {out->du, out->ud, out->uu, out->dd, out->image, out->illum_src, out->illum_rcv};
    int nout = 0;
    for (float *h : host) nout += h ? 1 : 0;
    DevBuf d;
    SWD_TRY(d.alloc(sizeof(float) * (2 * ng + nout * ni)));
    float *d_adj = d.as<float>(), *d_hadj = d_adj + ng, *next = d_hadj + ng;
    HIP_TRY(hipMemcpy(d_adj, adj, sizeof(float) * ng, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_hadj, hadj, sizeof(float) * ng, hipMemcpyHostToDevice));
    float *dev[7] = {};
    for (int q = 0; q < 7; q++)
        if (host[q]) {
            dev[q] = next;
            next += ni;
        }
    const swd_updown_images w = {dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dev[6]};
    SWD_TRY(swd_rtm_updown_dev(plan, nshots, shots, wavelet, hwavelet, d_adj, d_hadj, kzcoef, nfftz, &w));
    for (int q = 0; q < 7; q++)
        if (host[q]) HIP_TRY(hipMemcpy(host[q], dev[q], sizeof(float) * ni, hipMemcpyDeviceToHost));
    return SWD_OK;
}

// ------------------------------------------------------------------ Born
namespace {

// Born forward of a batch: m on the device [nshots][nx][nz], d on the device
// [sum nrec][nt].  Each shot runs the background wavefield and the scattered
// wavefield as two instances of the same launches.
// Born forward on the fused path: the background (2k) and scattered (2k+1)
// fields of shot k in one workgroup per tile, the Born source from the
// background's divergence sums in registers; shots in cache-sized groups on
// the plan's streams.
int born_dev_fused(Plan &P, int nshots, const swd_shot *shots, const Batch &b, const float *wavelet, const float *m,
                   float *d)
{
    const int nt = P.g.nt;
    const size_t nf = P.nfield, nint = static_cast<size_t>(P.g.nx) * P.g.nz;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KStep>(2 * nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_f = c.add<float>(kStepArrays * nf * 2 * nshots);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    zero_fields(P, at<float>(P, o_f), kStepArrays * nf * 2 * nshots);
    std::vector<KStep> h(2 * nshots);
    for (int k = 0; k < nshots; k++) {
        KStep &B = h[2 * k], &U = h[2 * k + 1];
        std::memset(&B, 0, sizeof B);
        std::memset(&U, 0, sizeof U);
        B.mode = U.mode = KSTEP_FWD;
        bind_step(B, at<float>(P, o_f) + kStepArrays * nf * (2 * k), nf);
        bind_step(U, at<float>(P, o_f) + kStepArrays * nf * (2 * k + 1), nf);
        B.wavelet = at<float>(P, o_wav);
        B.s_ix = U.s_ix = shots[k].s_ix;
        B.s_iz = U.s_iz = shots[k].s_iz;
        U.born_m = m + nint * k;
        if (shots[k].nrec > 0) {
            U.colmap = at<int>(P, o_ints) + b.col_off[k];
            U.rowmap = at<int>(P, o_ints) + b.row_off[k];
            U.gather = d + static_cast<size_t>(b.toff[k]) * nt;
        }
    }
    SWD_TRY(upload_steps(P, h, o_inst));
    long long launches = 0;
    int gs = P.group_fwd > 0 ? std::max(1, P.group_fwd / 2)
                             : static_cast<int>(kCacheBudgetBwd / P.streams / (2 * fwd_group_bytes(P)));
    gs = even_groups(nshots, gs);
    stats_begin(P, 0);
    SWD_TRY(fork_streams(P));
    const std::vector<int> bounds = group_bounds(nshots, gs, P.streams);
    for (size_t i = 0; i + 1 < bounds.size(); i++)
        SWD_TRY(fused_steps(P, at<KStep>(P, o_inst) + 2 * bounds[i], 2 * (bounds[i + 1] - bounds[i]), steps_up(nt), -1,
                            P.tuned_born, &launches, true, rtm_group_stream(P, static_cast<int>(i)), nullptr, false,
                            true));
    SWD_TRY(join_streams(P));
    P.stats.fwd_ms = stats_end(P, 0);
    P.stats.fwd_launches = launches;
    P.stats.fwd_cells = static_cast<long long>(P.ntx) * P.ntz * nt * 2 * nshots;
    HIP_TRY(launch_copy_rows(d, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    P.stats.ws_bytes = static_cast<long long>(P.ws_cap);
    return SWD_OK;
}

int born_dev(Plan &P, int nshots, const swd_shot *shots, const float *wavelet, const float *m, float *d)
{
    Batch b;
    SWD_TRY(build_batch(P, nshots, shots, b));
    if (P.path == SWD_PATH_FUSED) return born_dev_fused(P, nshots, shots, b, wavelet, m, d);
    const int nt = P.g.nt;
    const size_t nf = P.nfield, nint = static_cast<size_t>(P.g.nx) * P.g.nz;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KInst>(2 * nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_f = c.add<float>(kFieldArrays * nf * 2 * nshots);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    zero_fields(P, at<float>(P, o_f), kFieldArrays * nf * 2 * nshots);
    std::vector<KInst> h(2 * nshots);
    for (int k = 0; k < nshots; k++) {
        KInst &B = h[2 * k], &U = h[2 * k + 1];
        std::memset(&B, 0, sizeof B);
        std::memset(&U, 0, sizeof U);
        bind_fields(B, at<float>(P, o_f) + kFieldArrays * nf * 2 * k, nf);
        bind_fields(U, at<float>(P, o_f) + kFieldArrays * nf * (2 * k + 1), nf);
        B.wavelet = at<float>(P, o_wav);
        B.s_ix = U.s_ix = shots[k].s_ix;
        B.s_iz = U.s_iz = shots[k].s_iz;
        U.bvx = B.vx; U.bvz = B.vz;
        U.born_m = m + nint * k;
        if (shots[k].nrec > 0) {
            U.colmap = at<int>(P, o_ints) + b.col_off[k];
            U.rowmap = at<int>(P, o_ints) + b.row_off[k];
            U.gather = d + static_cast<size_t>(b.toff[k]) * nt;
        }
    }
    SWD_TRY(upload_insts(P, h, o_inst));
    SWD_TRY(forward_loop(P, at<KInst>(P, o_inst), 2, nshots));
    HIP_TRY(launch_copy_rows(d, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    P.stats.ws_bytes = static_cast<long long>(P.ws_cap);
    return SWD_OK;
}

// Exact discrete adjoint of born_dev: d on the device [sum nrec][nt], m_grad
// on the device [nx][nz], summed over the batch in shot order.  The
// background is checkpointed every seg steps (seg = ceil(nt/nsnap), or
// ceil(sqrt(nt)) when nsnap <= 0) and recomputed segment by segment in
// reverse, which stores the background sums of one segment at a time instead
// of all nt steps.
int born_adjoint_dev(Plan &P, int nshots, const swd_shot *shots, const float *wavelet, const float *d, float *m_grad,
                     int nsnap, bool accumulate)
{
    Batch b;
    SWD_TRY(build_batch(P, nshots, shots, b));
    const int nt = P.g.nt;
    const size_t nf = P.nfield, nint = static_cast<size_t>(P.g.nx) * P.g.nz;
    const size_t ng = static_cast<size_t>(b.ntr) * nt;
    const int seg = nsnap > 0 ? (nt + nsnap - 1) / nsnap : static_cast<int>(std::ceil(std::sqrt(static_cast<double>(nt))));
    const int nseg = (nt + seg - 1) / seg;
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KInst>(2 * nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_d = c.add<float>(std::max<size_t>(ng, 1));
    const size_t o_grad = c.add<float>(nint * nshots);
    const size_t o_s0 = c.add<float>(nint * seg * nshots);
    const size_t o_snap = c.add<float>(kFieldArrays * nf * static_cast<size_t>(nseg) * nshots);
    // adjoint fields + ex, ez scratch (the tiled step of the fused path: the
    // second slots of pvx, pvz, psx, psz)
    const bool tiled = P.path == SWD_PATH_FUSED;
    const size_t per_a = kFieldArrays + (tiled ? 4 : 2);
    const size_t o_f = c.add<float>((kFieldArrays + per_a) * nf * nshots);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    // traces sharing a node are summed into the node's last trace (the adjoint of the shared recording)
    float *dsum = at<float>(P, o_d);
    HIP_TRY(hipMemcpyAsync(dsum, d, sizeof(float) * ng, hipMemcpyDeviceToDevice, P.stream));
    HIP_TRY(launch_sum_rows(dsum, d, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    zero_fields(P, at<float>(P, o_f), (kFieldArrays + per_a) * nf * nshots);
    zero_fields(P, at<float>(P, o_grad), nint * nshots);
    std::vector<KInst> h(2 * nshots);
    float *fb = at<float>(P, o_f), *fa = fb + kFieldArrays * nf * nshots;
    for (int k = 0; k < nshots; k++) {
        KInst &B = h[k], &A = h[nshots + k];
        std::memset(&B, 0, sizeof B);
        std::memset(&A, 0, sizeof A);
        bind_fields(B, fb + kFieldArrays * nf * k, nf);
        bind_fields(A, fa + per_a * nf * k, nf);
        A.ex = A.p + kFieldArrays * nf;
        A.ez = A.ex + nf;
        if (tiled) {
            A.e2 = A.ez + nf;
            A.e3 = A.e2 + nf;
        }
        B.wavelet = at<float>(P, o_wav);
        B.s_ix = A.s_ix = shots[k].s_ix;
        B.s_iz = A.s_iz = shots[k].s_iz;
        if (shots[k].nrec > 0) {
            A.colmap = at<int>(P, o_ints) + b.col_off[k];
            A.rowmap = at<int>(P, o_ints) + b.row_off[k];
            A.dsum = dsum + static_cast<size_t>(b.toff[k]) * nt;
        }
        A.grad = at<float>(P, o_grad) + nint * k;
        A.s0 = at<float>(P, o_s0) + nint * seg * k;
    }
    const KInst *dB = at<KInst>(P, o_inst), *dA = dB + nshots;
    const KGrid &g = P.kg;
    long long fl = 0, bl = 0;
    // a snapshot holds the background state before the first step of a segment
    auto snap = [&](int k, int s) -> float * {
        return at<float>(P, o_snap) + kFieldArrays * nf * (static_cast<size_t>(nseg) * k + s);
    };
    auto steps = [&](int it0, int it1) -> int { return batched_forward_steps(g, dB, nshots, it0, it1, P.stream, &fl); };
    // checkpoint pass
    SWD_TRY(upload_insts(P, h, o_inst));
    stats_begin(P, 0);
    for (int s = 0; s < nseg; s++) {
        const int t0 = s * seg, t1 = std::min(nt, t0 + seg);
        for (int k = 0; k < nshots; k++)
            HIP_TRY(hipMemcpyAsync(snap(k, s), h[k].p, sizeof(float) * kFieldArrays * nf, hipMemcpyDeviceToDevice, P.stream));
        SWD_TRY(steps(t0, t1));
    }
    // reverse segments: recompute the background sums, then the adjoint steps
    for (int s = nseg - 1; s >= 0; s--) {
        const int t0 = s * seg, t1 = std::min(nt, t0 + seg);
        for (int k = 0; k < nshots; k++) {
            HIP_TRY(hipMemcpyAsync(h[k].p, snap(k, s), sizeof(float) * kFieldArrays * nf, hipMemcpyDeviceToDevice, P.stream));
            h[k].s0_out = at<float>(P, o_s0) + nint * seg * k;
            h[k].s0_t0 = t0;
            h[nshots + k].s0_t0 = t0;
        }
        SWD_TRY(upload_insts(P, h, o_inst));
        SWD_TRY(steps(t0, t1));
        for (int it = t1 - 1; it >= t0; it--) {
            if (tiled)
                HIP_TRY(launch_adjoint_step_tiled(g, dA, nshots, it, it == nt - 1, P.stream));
            else
                HIP_TRY(launch_adjoint_step(g, dA, nshots, it, it == nt - 1, P.stream));
            bl += it == nt - 1 ? 1 : (tiled ? 2 : 4);
        }
    }
    P.stats.fwd_ms = stats_end(P, 0);
    P.stats.fwd_launches = fl;
    P.stats.bwd_launches = bl;
    HIP_TRY(launch_sum_shots(m_grad, at<float>(P, o_grad), nshots, nint, accumulate, P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    P.stats.ws_bytes = static_cast<long long>(P.ws_cap);
    P.stats.ckpt_bytes = static_cast<long long>(sizeof(float)) * nshots *
                         (static_cast<long long>(kFieldArrays) * nf * nseg + static_cast<long long>(nint) * seg);
    P.stats.full_bytes = static_cast<long long>(sizeof(float)) * nshots * static_cast<long long>(nint) * nt;
    return SWD_OK;
}

}  // namespace

extern "C" int swd_born_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *m,
                            float *d)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !m || !d || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_born_dev: null argument");
    const size_t nint = static_cast<size_t>(P->g.nx) * P->g.nz;
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, 4.0 * (P->path == SWD_PATH_FUSED ? 2 * kStepArrays : 14.0) * P->nfield,
                       4.0 * P->g.nt * 4, wave));
    for (int s0 = 0; s0 < nshots; s0 += wave) {
        const int n = std::min(wave, nshots - s0);
        SWD_TRY(born_dev(*P, n, shots + s0, wavelet, m + nint * s0, d + static_cast<size_t>(traces_before(shots, s0)) * P->g.nt));
    }
    return SWD_OK;
}

extern "C" int swd_born_adjoint_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet,
                                    const float *d, float *m_grad, int nsnap)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !m_grad || !d || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_born_adjoint_dev: null argument");
    const int nt = P->g.nt;
    const double nf = static_cast<double>(P->nfield), nint = static_cast<double>(P->g.nx) * P->g.nz;
    const int seg = nsnap > 0 ? (nt + nsnap - 1) / nsnap : static_cast<int>(std::ceil(std::sqrt(static_cast<double>(nt))));
    const int nseg = (nt + seg - 1) / seg;
    int ntr_max = 0;
    for (int k = 0; k < nshots; k++) ntr_max = std::max(ntr_max, shots[k].nrec);
    const double per = 4.0 * (7.0 * nf * nseg + nint * seg + 18.0 * nf + nint + static_cast<double>(ntr_max) * nt);
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, per, 4.0 * nt * 4, wave));
    // Waves also hold their per-step working set (both wavefields, the
    // background sums, the gradient) to about 768 MB, in waves of even size
    // (c5, 8 shots, nt 2000: one wave 1476 ms, waves of 4 1300 ms, of 2
    // 1390 ms, of 1 1669 ms).
    const double step_bytes = 4.0 * (18.0 * nf + 3.0 * nint);
    wave = std::min(wave, std::max(1, static_cast<int>(768.0 * (1 << 20) / step_bytes)));
    const int nwave = (nshots + wave - 1) / wave;
    wave = (nshots + nwave - 1) / nwave;
    P->stats.waves = nwave;
    for (int s0 = 0; s0 < nshots; s0 += wave) {
        const int n = std::min(wave, nshots - s0);
        SWD_TRY(born_adjoint_dev(*P, n, shots + s0, wavelet, d + static_cast<size_t>(traces_before(shots, s0)) * nt, m_grad,
                                 nsnap, s0 > 0));
    }
    return SWD_OK;
}

extern "C" int swd_born(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *m, float *d)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !m || !d || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_born: null argument");
    const size_t nint = static_cast<size_t>(P->g.nx) * P->g.nz;
    const size_t ng = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    DevBuf buf;
    SWD_TRY(buf.alloc(sizeof(float) * (nint * nshots + ng)));
    float *d_m = buf.as<float>(), *d_d = d_m + nint * nshots;
    HIP_TRY(hipMemcpy(d_m, m, sizeof(float) * nint * nshots, hipMemcpyHostToDevice));
    SWD_TRY(swd_born_dev(plan, nshots, shots, wavelet, d_m, d_d));
    HIP_TRY(hipMemcpy(d, d_d, sizeof(float) * ng, hipMemcpyDeviceToHost));
    return SWD_OK;
}

extern "C" int swd_born_adjoint(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *d,
                                float *m_grad, int nsnap)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !m_grad || !d || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_born_adjoint: null argument");
    const size_t nint = static_cast<size_t>(P->g.nx) * P->g.nz;
    const size_t ng = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    DevBuf buf;
    SWD_TRY(buf.alloc(sizeof(float) * (nint + std::max<size_t>(ng, 1))));
    float *d_g = buf.as<float>(), *d_d = d_g + nint;
    HIP_TRY(hipMemcpy(d_d, d, sizeof(float) * ng, hipMemcpyHostToDevice));
    SWD_TRY(swd_born_adjoint_dev(plan, nshots, shots, wavelet, d_d, d_g, nsnap));
    HIP_TRY(hipMemcpy(m_grad, d_g, sizeof(float) * nint, hipMemcpyDeviceToHost));
    return SWD_OK;
}

// ------------------------------------------------------------------ multi-grid forward
extern "C" int swd_multi_forward(int nplans, void *const *plans, const int *nshots, const swd_shot *const *shots,
                                 const float *const *wavelets, float *const *gathers)
{
    if (nplans < 1 || !plans || !nshots || !shots || !wavelets || !gathers)
        return swd_fail(SWD_E_ARG, "swd_multi_forward: null argument");
    std::vector<Plan *> Ps(nplans);
    for (int i = 0; i < nplans; i++) {
        for (int j = 0; j < i; j++)
            if (plans[j] == plans[i]) return swd_fail(SWD_E_ARG, "swd_multi_forward: plan %d is listed twice", i);
        SWD_TRY(check_plan(plans[i], Ps[i]));
        if (nshots[i] < 1 || !shots[i] || !wavelets[i] || !gathers[i])
            return swd_fail(SWD_E_ARG, "swd_multi_forward: plan %d has no shots or buffers", i);
        if (Ps[i]->dev != Ps[0]->dev || Ps[i]->g.Lc != Ps[0]->g.Lc || Ps[i]->g.nt != Ps[0]->g.nt)
            return swd_fail(SWD_E_ARG, "swd_multi_forward: plan %d differs from plan 0 in device, Lc or nt", i);
    }
    Plan &P0 = *Ps[0];
    HIP_TRY(hipSetDevice(P0.dev));
    const int nt = P0.g.nt;
    // per plan: its batch, fields and gathers in its own workspace
    std::vector<Batch> bs(nplans);
    std::vector<size_t> o_ints(nplans), o_pairs(nplans), o_wav(nplans), o_f(nplans), o_g(nplans);
    for (int i = 0; i < nplans; i++) {
        Plan &P = *Ps[i];
        SWD_TRY(build_batch(P, nshots[i], shots[i], bs[i]));
        Carver c;
        o_ints[i] = c.add<int>(bs[i].ints.size());
        o_pairs[i] = c.add<int2>(std::max<size_t>(1, bs[i].dup_pairs.size()));
        o_wav[i] = c.add<float>(nt);
        o_g[i] = c.add<float>(std::max<size_t>(1, static_cast<size_t>(bs[i].ntr) * nt));
        o_f[i] = c.add<float>(kStepArrays * P.nfield * nshots[i]);
        SWD_TRY(ensure_ws(P, c.off));
    }
    // instance, grid and block tables (tiles of 64 rows x 16 columns)
    std::vector<KGrid> grids(nplans);
    std::vector<KStep> h;
    std::vector<int2> btab;
    std::vector<int> inst0(nplans);
    for (int i = 0; i < nplans; i++) {
        grids[i] = Ps[i]->kg;
        inst0[i] = static_cast<int>(h.size());
        for (int k = 0; k < nshots[i]; k++) h.emplace_back();
    }
    Carver ca;
    const size_t oa_grid = ca.add<KGrid>(nplans);
    const size_t oa_inst = ca.add<KStep>(h.size());
    // Groups of whole instances (an instance's tiles exchange halos between
    // steps): as many as the cache budget asks, rounded up to a multiple of the
    // streams (at most eight), each closed once it holds its share of the
    // per-step working set, so the streams carry even loads (c4 on two streams:
    // four groups of 72-82 MB).  Each group runs its own time loop on the
    // plan's streams in turn.  Within a group the shots of one tile sit next to
    // each other (shared model in L2).
    struct Group {
        size_t off;
        int n;
    };
    std::vector<Group> groups;
    {
        double total = 0.0;
        for (int i = 0; i < nplans; i++) total += fwd_group_bytes(*Ps[i]) * nshots[i];
        const int ns = P0.streams;
        int ng = std::max(1, static_cast<int>(std::ceil(total / (kCacheBudgetBwd / ns))));
        ng = std::min(8, (ng + ns - 1) / ns * ns);
        const double target = total / ng;
        double used = 0.0;
        size_t off = 0;
        for (int i = 0; i < nplans; i++) {
            const int nzb = (Ps[i]->ntz + 63) / 64, nxb = (Ps[i]->ntx + 15) / 16;
            const double per = fwd_group_bytes(*Ps[i]);
            int k0 = 0;
            for (int k = 0; k < nshots[i]; k++) {
                used += per;
                const bool close = used >= target && static_cast<int>(groups.size()) + 1 < ng;
                if (close || k == nshots[i] - 1) {
                    for (int t = 0; t < nzb * nxb; t++)
                        for (int q = k0; q <= k; q++) btab.push_back(make_int2(inst0[i] + q, t));
                    k0 = k + 1;
                }
                if (close) {
                    groups.push_back(Group{off, static_cast<int>(btab.size() - off)});
                    off = btab.size();
                    used = 0.0;
                }
            }
        }
        if (btab.size() > off) groups.push_back(Group{off, static_cast<int>(btab.size() - off)});
    }
    const size_t oa_btab = ca.add<int2>(btab.size());
    SWD_TRY(ensure_aux(P0, ca.off));
    const KGrid *d_grids = reinterpret_cast<const KGrid *>(P0.aux + oa_grid);
    for (int i = 0; i < nplans; i++) {
        Plan &P = *Ps[i];
        SWD_TRY(upload_batch_on(P, bs[i], o_ints[i], o_pairs[i], P0.stream));
        HIP_TRY(hipMemcpyAsync(P.ws + o_wav[i], wavelets[i], sizeof(float) * nt, hipMemcpyHostToDevice, P0.stream));
        HIP_TRY(hipMemsetAsync(P.ws + o_f[i], 0, sizeof(float) * kStepArrays * P.nfield * nshots[i], P0.stream));
        for (int k = 0; k < nshots[i]; k++) {
            KStep &S = h[inst0[i] + k];
            std::memset(&S, 0, sizeof S);
            S.mode = KSTEP_FWD;
            S.grid = d_grids + i;
            bind_step(S, at<float>(P, o_f[i]) + kStepArrays * P.nfield * k, P.nfield);
            S.wavelet = at<float>(P, o_wav[i]);
            S.s_ix = shots[i][k].s_ix;
            S.s_iz = shots[i][k].s_iz;
            if (shots[i][k].nrec > 0) {
                S.colmap = at<int>(P, o_ints[i]) + bs[i].col_off[k];
                S.rowmap = at<int>(P, o_ints[i]) + bs[i].row_off[k];
                S.gather = at<float>(P, o_g[i]) + static_cast<size_t>(bs[i].toff[k]) * nt;
            }
        }
    }
    HIP_TRY(hipMemcpyAsync(P0.aux + oa_grid, grids.data(), sizeof(KGrid) * nplans, hipMemcpyHostToDevice, P0.stream));
    HIP_TRY(hipMemcpyAsync(P0.aux + oa_inst, h.data(), sizeof(KStep) * h.size(), hipMemcpyHostToDevice, P0.stream));
    HIP_TRY(hipMemcpyAsync(P0.aux + oa_btab, btab.data(), sizeof(int2) * btab.size(), hipMemcpyHostToDevice, P0.stream));
    const KStep *d_inst = reinterpret_cast<const KStep *>(P0.aux + oa_inst);
    const int2 *d_btab = reinterpret_cast<const int2 *>(P0.aux + oa_btab);
    stats_begin(P0, 0);
    SWD_TRY(fork_streams(P0));
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const Group G = groups[gi];
        const hipStream_t st = rtm_group_stream(P0, static_cast<int>(gi));
        auto record = [&]() -> int {
            for (int it = 0; it < nt; it++) HIP_TRY(launch_fstep_multi(P0.g.Lc, d_inst, d_btab + G.off, G.n, it, st));
            return SWD_OK;
        };
        if (P0.graphs) {
            std::vector<long long> key = {2, reinterpret_cast<long long>(P0.aux), static_cast<long long>(h.size()),
                                          static_cast<long long>(G.off), G.n, nt};
            for (int i = 0; i < nplans; i++) key.push_back(reinterpret_cast<long long>(Ps[i]->ws));
            SWD_TRY(graph_run(P0, key, record, st));
        } else {
            SWD_TRY(record());
        }
    }
    SWD_TRY(join_streams(P0));
    P0.stats.fwd_ms = stats_end(P0, 0);
    P0.stats.fwd_launches = static_cast<long long>(nt) * static_cast<long long>(groups.size());
    long long cells = 0;
    for (int i = 0; i < nplans; i++) cells += static_cast<long long>(Ps[i]->ntx) * Ps[i]->ntz * nt * nshots[i];
    P0.stats.fwd_cells = cells;
    for (int i = 0; i < nplans; i++) {
        Plan &P = *Ps[i];
        HIP_TRY(launch_copy_rows(at<float>(P, o_g[i]), at<int2>(P, o_pairs[i]), static_cast<int>(bs[i].dup_pairs.size()), nt,
                                 P0.stream));
        HIP_TRY(hipMemcpyAsync(gathers[i], at<float>(P, o_g[i]), sizeof(float) * bs[i].ntr * nt, hipMemcpyDeviceToHost,
                               P0.stream));
    }
    HIP_TRY(hipStreamSynchronize(P0.stream));
    return SWD_OK;
}
