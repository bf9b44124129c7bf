// swd_plan.h -- the plan runtime shared by libswd's operator translation units
// (swd_plan.hip: plans, workspace, graph replay, fused time loops, shot groups
// and waves; swd_forward.hip, swd_rtm.hip, swd_rtm_updown.hip, swd_born.hip,
// swd_multi.hip: the operators of the C-ABI).
//
// A plan owns one HIP stream (default priority) on one device, the model in
// the kernels' form (dt*buoyancy at the two half points, dt*rho*vp^2 and
// rho*vp^2, each rounded exactly as the reference rounds them per cell), the
// CPML vectors, and a growing workspace from which every call carves its
// wavefields, strips, gathers and scratch.  Operators run a batch of shots;
// on SWD_PATH_BATCHED every wavefield of the batch shares each launch, on
// SWD_PATH_PLAIN shots run one after another with one kernel per sub-step,
// which is the single-shot baseline.
#pragma once
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

namespace swd_rt {

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
    int tile_cfg = 0;           // fused tile configuration asked for (fstep_cfg; 0 = tune)
    int steps_per_launch = 0;   // fused loops: 1 or 2 time steps per launch where allowed (0 = tune)
    std::string arch;           // device architecture and compute units (the tuner cache's key)
    int cus = 0;
    int group_fwd = 0, group_bwd = 0;     // RTM shots per fused forward / backward loop (0 = sized to the cache)
    std::map<int, int> tuned_fwd, tuned_bwd;   // tuned tile configuration by instances per launch
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

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

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

template <class T>
T *at(Plan &P, size_t off) { return reinterpret_cast<T *>(P.ws + off); }

void drop_graphs(Plan &P);

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

// Receiver lookup of one shot: colmap[ntx] -> row map offset, rowmap -> trace
// of the node (the last trace on a node wins, as the reference's receiver
// loops do), dup = traces that share a node with a later trace.
struct RecvMap {
    std::vector<int> colmap, rowmap;
    std::vector<int> prim;   // primary trace of each trace
};

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

// Arrays per wavefield: p, vx, vz, psx, psz (CPML memory of the velocity
// update), pvx, pvz (CPML memory of the pressure update).
constexpr int kFieldArrays = 7;
// Arrays per wavefield of the fused path: p, vx, vz, psx, psz, pvx, pvz in
// two slots.
constexpr int kStepArrays = 14;
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

// Work launched after every fused step (inside the same captured graph):
// run(it, stream) queues one kernel launch; key identifies its arguments for
// the graph cache.  The launch counts count it beside the fused step.
struct StepHook {
    std::function<int(int, hipStream_t)> run;
    // after a two-step launch of steps it and it-1 (the paired backward), ph its place in
    // its pair: queues the work of both steps (launches2 launches); unset: no two-step launches
    std::function<int(int, int, hipStream_t)> run2;
    int launches2 = 1;
    // the two-step launches are the paired kernel's middle-level variant (default configuration)
    bool mid = false;
    std::vector<long long> key;
};

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

// Every entry point lays its own tables out in the workspace: what
// swd_updown_image_dev uploaded is stale after any other call (keep = false).
int ensure_ws(Plan &P, size_t bytes, bool keep_updown = false);

int ensure_aux(Plan &P, size_t bytes);

int check_plan(void *plan, Plan *&P);

int build_recv(const Plan &P, const swd_shot &s, RecvMap &m);

int build_batch(const Plan &P, int nshots, const swd_shot *shots, Batch &b);

void stats_begin(Plan &P, int which);

double stats_end(Plan &P, int which);

// Upload of the batch's integer maps and instance table into the workspace.
int upload_batch_on(Plan &P, const Batch &b, size_t o_ints, size_t o_pairs, hipStream_t st);

int upload_batch(Plan &P, const Batch &b, size_t o_ints, size_t o_pairs);

int upload_insts(Plan &P, const std::vector<KInst> &h, size_t o_inst);

void zero_fields(Plan &P, float *base, size_t count);

void bind_fields(KInst &I, float *f, size_t nfield);

void bind_step(KStep &S, float *f, size_t nf);

int upload_steps(Plan &P, const std::vector<KStep> &h, size_t o);

// Runs the fused time steps its[0], its[1], ... (in execution order) of n
// instances on stream st (the plan's by default).  The tile configuration is
// the plan's requested one, else the one tuned earlier for n instances, else
// it is tuned here: each candidate runs one warm-up unit, then kTry timed
// units of the loop's own steps, and the fastest per step runs the rest (the
// results do not depend on the configuration).  two: the steps may run two
// per launch (swd_fstep2.h; the caller's instances have no dp/dt output,
// there is no hook, and consecutive steps follow each other
// in its): the two-step configurations join the candidates (or alone run the
// loop, if the plan asks for two steps per launch).  A two-step schedule runs
// the first reconstruction step (first_it) and the remainder of the step
// count modulo 4 one step per launch first, then pairs of two-step launches.
int fused_steps(Plan &P, const KStep *d, int n, const std::vector<int> &its, int first_it, std::map<int, int> &tuned,
                long long *launches, bool pair = false, hipStream_t st = nullptr, const StepHook *hook = nullptr,
                bool cig = false, bool born = false, bool two = false);

std::vector<int> steps_up(int nt);

// The tile tuner's cache: key of a loop (device, grid, Lc, instances per
// launch n, kernel variant mode, streams), the configuration stored for it on
// disk (0: none) and a new choice to store.
std::string tune_key(const Plan &P, const char *mode, int n);
int tune_lookup(const std::string &key);
void tune_store(const std::string &key, int cfg);
// The tuner's pick from candidate timings ms[c] (candidate 0 the default).
int choose_cfg(const std::vector<float> &ms, const std::vector<int> &cand);

int even_groups(int n, int g);

// Group boundaries of n instances in groups of at most g: ceil(n/g) groups,
// their count raised to a multiple of the streams where there are instances
// enough, so that every stream gets as many groups (with an odd count on two
// streams one of them runs alone at the end: c2 AFor, 32 wavefields, five
// groups of 6-7 took 555 ms a batch, six groups of 5-6 504 ms); sizes differ
// by at most one (the larger ones first).
std::vector<int> group_bounds(int n, int g, int streams);

// Per-step working set of one forward wavefield: double-buffered state and
// the CPML memories of the frame.
double fwd_group_bytes(const Plan &P);

void rtm_groups(const Plan &P, int nshots, int &gf, int &gb);

// With several streams, consecutive shot groups take them in turn so that
// one group's launch gaps and tails overlap the others' work.
hipStream_t rtm_group_stream(const Plan &P, int g);

// The other streams wait for the work queued so far on the plan's stream.
int fork_streams(Plan &P);

// The plan's stream waits for the work queued so far on the other streams.
int join_streams(Plan &P);

// Forward time loop of nshots shots with per_shot wavefields each.  Plain
// path: shot after shot, one kernel per sub-step.  Batched path: every
// wavefield of a group of shots in each launch, a velocity and a pressure
// kernel.
int forward_loop(Plan &P, const KInst *d_inst, int per_shot, int nshots);

// Forward time loop of the fused path: n independent wavefield instances in
// groups whose per-step working set stays in the last-level cache, the groups
// taking the plan's streams in turn (as the RTM forward).  two: the instances
// may run two steps per launch (see fused_steps).
int fused_forward_loop(Plan &P, const KStep *d, int n, bool two);

int total_traces(int nshots, const swd_shot *shots);

// Shots per wave of a call whose workspace needs fixed + n * per_shot bytes:
// as many as fit the plan's memory limit (by default 85% of the free device
// memory plus the plan's current workspace, which the call may reallocate).
int wave_shots(Plan &P, int nshots, double per_shot, double fixed, int &wave);

int traces_before(const swd_shot *shots, int k);

int forward_dev(Plan &P, int nshots, const swd_shot *shots, const float *wavelet, float *d_gather);

// Twiddles w_N^k = exp(-2 pi i k / N) of the up/down split for k < N/2,
// evaluated in double and rounded to float, into tw[N/2].
void kz_twiddles_into(int nfftz, float2 *tw);

// The checks of swd_updown_image_dev's arguments (fn names the entry point).
int check_updown_image_args(const Plan &P, const char *fn, int nshots, const float *src, const float *hsrc,
                            const float *rcv, const float *hrcv, const float *kzcoef, int nfftz,
                            const swd_updown_images *out);

// Descriptors of one swd_updown_image_dev call: shot k reads and writes the
// [ntx][ntz] slices k of the caller's arrays; tw and coef are the device
// tables of the call.
void updown_image_descriptors(const Plan &P, int nshots, const float *src, const float *hsrc, const float *rcv,
                              const float *hrcv, const swd_updown_images *out, const float2 *tw, const float *coef,
                              KUpDown *hu);

}  // namespace swd_rt
