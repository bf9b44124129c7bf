// swd_plan.hip -- plans (creation, coefficients, model, settings), the
// workspace, graph replay, the fused and batched time loops, shot groups on
// the plan's streams and memory-sized shot waves.
#include <sys/stat.h>

#include <cstdlib>
#include <mutex>

#include "swd_plan.h"

// ------------------------------------------------------------------ errors
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

namespace swd_rt {

void drop_graphs(Plan &P)
{
    for (auto &kv : P.graph_cache) (void)hipGraphExecDestroy(kv.second);
    P.graph_cache.clear();
}

int ensure_ws(Plan &P, size_t bytes, bool keep_updown)
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

int check_plan(void *plan, Plan *&P)
{
    if (!plan) return swd_fail(SWD_E_ARG, "null plan");
    P = static_cast<Plan *>(plan);
    if (!P->have_coefs || !P->have_model)
        return swd_fail(SWD_E_STATE, "plan needs swd_plan_set_coefs and swd_plan_set_model before running");
    HIP_TRY(hipSetDevice(P->dev));
    return SWD_OK;
}

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

void zero_fields(Plan &P, float *base, size_t count) { (void)hipMemsetAsync(base, 0, sizeof(float) * count, P.stream); }

void bind_fields(KInst &I, float *f, size_t nfield)
{
    I.p = f; I.vx = f + nfield; I.vz = f + 2 * nfield;
    I.psx = f + 3 * nfield; I.psz = f + 4 * nfield; I.pvx = f + 5 * nfield; I.pvz = f + 6 * nfield;
}

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
    S.pvxb = f + 12 * nf;
    S.pvzb = f + 13 * nf;
}

int upload_steps(Plan &P, const std::vector<KStep> &h, size_t o)
{
    HIP_TRY(hipMemcpyAsync(P.ws + o, h.data(), sizeof(KStep) * h.size(), hipMemcpyHostToDevice, P.stream));
    return SWD_OK;
}

// ---- tile tuner and its cache on disk
//
// Chosen tile configurations are kept per plan by instances per launch and
// in a text file shared by processes: one line per choice, "<key>\t<cfg>",
// the key naming the device, the padded grid, Lc, the instances per launch,
// the kernel variant and the stream count.  The file is SWD_TUNE_CACHE if
// set (empty: no file), else $XDG_CACHE_HOME/swd_tiles.txt or
// $HOME/.cache/swd_tiles.txt.  Results never depend on the configuration.
namespace {
std::mutex g_tune_lock;
std::map<std::string, int> g_tune_disk;
bool g_tune_loaded = false;

std::string tune_path()
{
    if (const char *e = std::getenv("SWD_TUNE_CACHE")) return e;
    const char *x = std::getenv("XDG_CACHE_HOME"), *h = std::getenv("HOME");
    if (x && *x) return std::string(x) + "/swd_tiles.txt";
    if (h && *h) return std::string(h) + "/.cache/swd_tiles.txt";
    return "";
}

void tune_load_locked()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const std::string path = tune_path();
    if (path.empty()) return;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char *tab = std::strrchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        g_tune_disk[line] = std::atoi(tab + 1);
    }
    std::fclose(f);
}
}  // namespace

std::string tune_key(const Plan &P, const char *mode, int n)
{
    char k[256];
    std::snprintf(k, sizeof k, "%s cu%d grid %dx%d Lc %d n %d %s streams %d", P.arch.c_str(), P.cus, P.ntx, P.ntz,
                  P.g.Lc, n, mode, P.streams);
    return k;
}

int tune_lookup(const std::string &key)
{
    std::lock_guard<std::mutex> hold(g_tune_lock);
    tune_load_locked();
    const auto f = g_tune_disk.find(key);
    return f == g_tune_disk.end() ? 0 : f->second;
}

void tune_store(const std::string &key, int cfg)
{
    std::lock_guard<std::mutex> hold(g_tune_lock);
    tune_load_locked();
    g_tune_disk[key] = cfg;
    const std::string path = tune_path();
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos && slash > 0) ::mkdir(path.substr(0, slash).c_str(), 0755);   // one level, as ~/.cache
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f) return;
    std::fprintf(f, "%s\t%d\n", key.c_str(), cfg);
    std::fclose(f);
}

int choose_cfg(const std::vector<float> &ms, const std::vector<int> &cand)
{
    // The default (first) configuration stays unless another is more than
    // 10% faster: the timing is a short sample of one launch sequence, while
    // narrow 64-row tiles keep the most workgroups in flight, which the groups
    // overlapping on the plan's streams need (c2, 16 shots on two streams,
    // whole runs at widths 16/32/64: AFor 505/548/686 ms, RTM 785/820/1043 ms).
    int best = 0;
    for (size_t c = 1; c < cand.size(); c++)
        if (ms[c] < 0.9f * ms[0] && ms[c] < ms[best]) best = static_cast<int>(c);
    return cand[best];
}

int fused_steps(Plan &P, const KStep *d, int n, const std::vector<int> &its, int first_it, std::map<int, int> &tuned,
                long long *launches, bool pair, hipStream_t st, const StepHook *hook,
                bool cig, bool born, bool two)
{
    if (!st) st = P.stream;
    // two steps per launch: the forward goes up in time, the paired backward down, one step at a time
    // (the two-step kernel forms tile-relative offsets with a 24-bit multiply: column strides below 2^23)
    two = two && (!hook || hook->run2) && !cig && P.g.Lc == 4 && P.steps_per_launch != 1 && P.ldz < (1 << 23) &&
          (!born || pair);
    const int dir = pair && !born ? -1 : 1;   // the Born pair is two forward fields
    for (size_t j = 1; two && j < its.size(); j++)
        if (its[j] != its[j - 1] + dir) two = false;
    // steps = 2 asked: an asked tile runs two steps per launch in its shape (or the nearest)
    const bool only2 = two && P.steps_per_launch == 2;
    // the tuner's kind: two-step shapes wherever two steps per launch apply
    const int variant = two ? 2 : 0;
    const std::string mode = std::string(born ? "born" : cig ? "cig" : pair ? "pair" : "fwd") + (two ? " two" : "");
    int cfg = P.tile_cfg;
    if (cfg != 0) {
        const int tx = fstep_cfg_tx(cfg), tq = fstep_cfg_tq(cfg);
        if (only2)
            cfg = fstep2_cfg_supported(cfg | kFstepTwo) ? cfg | kFstepTwo
                  : fstep2_cfg_of_shape(tx, tq)         ? fstep2_cfg_of_shape(tx, tq)
                                                        : kFstep2Default;
        else if (fstep_cfg_two(cfg))
            cfg = fstep_cfg(tx, 16, 0);   // a split two-step configuration asked of a one-step loop
        else if (!fstep_cfg_supported(P.g.Lc, pair, cig, born, false, cfg))
            cfg = fstep_cfg(tx, 16, 0);   // the asked width with the variant's 64-row tiles
    }
    // the middle-level and Born variants have one shape (an asked one-step tile keeps one step)
    if (two && ((hook && hook->mid) || born) && (cfg == 0 || fstep_cfg_two(cfg))) cfg = kFstep2Default;
    const int tkey = 4 * n + variant;
    std::string key;
    if (cfg == 0) {
        const auto f = tuned.find(tkey);
        if (f != tuned.end()) cfg = f->second;
    }
    // a pick from the plan or the disk cache: of the kind the tuner would time
    auto usable = [&](int c) {
        return two ? fstep_cfg_two(c) && fstep2_cfg_supported(c)
                   : !fstep_cfg_two(c) && fstep_cfg_supported(P.g.Lc, pair, cig, born, false, c);
    };
    if (cfg == 0) {
        key = tune_key(P, mode.c_str(), n);
        cfg = tune_lookup(key);
        if (cfg != 0 && !usable(cfg)) cfg = 0;
        if (cfg != 0) tuned[tkey] = cfg;
    }
    size_t k = 0;
    long long nl = 0;
    // one step (it == first_it: the reconstruction keeps the forward's velocities)
    auto one = [&](int c) -> int {
        const int it = its[k++];
        HIP_TRY(launch_fstep(P.kg, d, n, it, it == first_it, fstep_cfg_two(c) ? kFstepDefault : c, pair, cig, born, st));
        nl++;
        if (hook) {
            SWD_TRY(hook->run(it, st));
            nl++;
        }
        return SWD_OK;
    };
    // four steps as a pair of two-step launches (slot parity as before afterwards)
    auto quad = [&](int c) -> int {
        for (int ph = 0; ph < 2; ph++) {
            const int it = its[k + 2 * ph];
            HIP_TRY(launch_fstep2(P.kg, d, n, it, ph, c, pair, st, hook && hook->mid, born));
            nl++;
            if (hook) {
                SWD_TRY(hook->run2(it, ph, st));
                nl += hook->launches2;
            }
        }
        k += 4;
        return SWD_OK;
    };
    auto unit = [&](int c) -> int { return fstep_cfg_two(c) ? quad(c) : one(c); };
    auto unit_steps = [](int c) -> size_t { return fstep_cfg_two(c) ? 4 : 1; };
    if (two && k < its.size() && its[k] == first_it) SWD_TRY(one(kFstepDefault));
    if (cfg == 0) {
        // time kTry units of every candidate after a warm-up unit of each, on
        // the loop's own steps, and keep the choice for later loops
        constexpr int kTry = 5;
        // Where two steps per launch apply (faster in every loop measured) the
        // candidates are the two-step shapes, else the one-step ones; the first
        // is the reference another must beat by 10%.  (With both kinds timed, a
        // one-step shape could win the short overlapped timing by 10% and lose
        // the run: AFor at c2, 470 against 424 ms.)
        const std::vector<int> cand = two ? fstep2_candidates() : fstep_candidates(P.g.Lc, pair, cig, born, false);
        size_t need = 1;
        for (int c : cand) need += unit_steps(c) * (kTry + 1);
        if (its.size() - k >= need) {
            std::vector<hipEvent_t> ev(cand.size() + 1);
            for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
            int rc = SWD_OK;
            std::vector<float> ms(cand.size(), 0.f);
            for (size_t c = 0; c < cand.size() && rc == SWD_OK; c++) {
                rc = unit(cand[c]);   // first use of the configuration, untimed
                if (rc == SWD_OK) rc = hipEventRecord(ev[c], st) == hipSuccess ? SWD_OK : SWD_E_DEVICE;
                for (int j = 0; j < kTry && rc == SWD_OK; j++) rc = unit(cand[c]);
                if (rc == SWD_OK) rc = hipEventRecord(ev[cand.size()], st) == hipSuccess ? SWD_OK : SWD_E_DEVICE;
                if (rc == SWD_OK) rc = hipEventSynchronize(ev[cand.size()]) == hipSuccess ? SWD_OK : SWD_E_DEVICE;
                if (rc == SWD_OK) rc = hipEventElapsedTime(&ms[c], ev[c], ev[cand.size()]) == hipSuccess ? SWD_OK
                                                                                                       : SWD_E_DEVICE;
                ms[c] /= static_cast<float>(kTry * unit_steps(cand[c]));   // per step
            }
            for (auto &e : ev) (void)hipEventDestroy(e);
            if (rc != SWD_OK) return swd_fail(rc, "tile tuner: a timed launch failed");
            cfg = choose_cfg(ms, cand);
            tuned[tkey] = cfg;
            tune_store(key, cfg);
        } else {
            cfg = two ? kFstep2Default : kFstepDefault;   // too few steps to time the candidates
        }
    }
    P.stats.tile_x = fstep_cfg_tx(cfg);
    P.stats.tile_rows = 4 * fstep_cfg_tq(cfg);
    P.stats.tile_waves = fstep_cfg_two(cfg) ? 0 : fstep_cfg_wb(cfg);
    P.stats.steps_per_launch = fstep_cfg_two(cfg) ? 2 : 1;
    P.stats.tile_group_waves = !fstep_cfg_two(cfg)      ? fstep_threads_of(cfg) / 64
                               : (cfg & kFstepSplit) ? -fstep_cfg_wb(cfg)
                                                     : fstep_cfg_wb(cfg);
    if (!pair) {
        P.stats.fwd_tile_x = P.stats.tile_x;
        P.stats.fwd_tile_rows = P.stats.tile_rows;
        P.stats.fwd_tile_group_waves = P.stats.tile_group_waves;
        P.stats.fwd_steps_per_launch = P.stats.steps_per_launch;
    }
    // the rest: with two steps per launch, the remainder modulo 4 one step at a time, then pairs
    auto rest = [&]() -> int {
        if (fstep_cfg_two(cfg)) {
            const size_t r = (its.size() - k) % 4;
            for (size_t j = 0; j < r; j++) SWD_TRY(one(kFstepDefault));
            while (k < its.size()) SWD_TRY(quad(cfg));
        } else {
            while (k < its.size()) SWD_TRY(one(cfg));
        }
        return SWD_OK;
    };
    if (k < its.size() && P.graphs) {
        std::vector<long long> gkey = {1, reinterpret_cast<long long>(d), n, its[k], static_cast<long long>(its.size() - k),
                                       first_it, cfg, (pair ? 1 : 0) + (cig ? 2 : 0) + (born ? 4 : 0),
                                       reinterpret_cast<long long>(P.ws)};
        if (hook) gkey.insert(gkey.end(), hook->key.begin(), hook->key.end());
        // launches of the rest, counted here since a replay does not record them
        const size_t left = its.size() - k;
        const long long h = hook ? 1 : 0;
        const long long nrest = fstep_cfg_two(cfg) ? static_cast<long long>(left % 4) * (1 + h) +
                                                         static_cast<long long>(left / 4) * 2 * (1 + (hook ? hook->launches2 : 0))
                                                   : static_cast<long long>(left) * (1 + h);
        const long long before = nl;   // the first step's and the tuner's launches
        SWD_TRY(graph_run(P, gkey, rest, st));
        *launches += before + nrest;
        return SWD_OK;
    }
    SWD_TRY(rest());
    *launches += nl;
    return SWD_OK;
}

std::vector<int> steps_up(int nt)
{
    std::vector<int> v(nt);
    for (int it = 0; it < nt; it++) v[it] = it;
    return v;
}

int even_groups(int n, int g)
{
    g = std::max(1, std::min(n, g));
    const int ng = (n + g - 1) / g;
    return (n + ng - 1) / ng;
}

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

hipStream_t rtm_group_stream(const Plan &P, int g)
{
    const int ns = P.streams;
    return g % ns == 0 ? P.stream : P.side[g % ns - 1];
}

int fork_streams(Plan &P)
{
    if (P.streams == 1) return SWD_OK;
    HIP_TRY(hipEventRecord(P.fork, P.stream));
    for (int i = 0; i < P.streams - 1; i++) HIP_TRY(hipStreamWaitEvent(P.side[i], P.fork, 0));
    return SWD_OK;
}

int join_streams(Plan &P)
{
    for (int i = 0; i < P.streams - 1; i++) {
        HIP_TRY(hipEventRecord(P.join[i], P.side[i]));
        HIP_TRY(hipStreamWaitEvent(P.stream, P.join[i], 0));
    }
    return SWD_OK;
}

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

int fused_forward_loop(Plan &P, const KStep *d, int n, bool two)
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
                            false, rtm_group_stream(P, static_cast<int>(g)), nullptr, false, false, two));
    SWD_TRY(join_streams(P));
    P.stats.fwd_ms = stats_end(P, 0);
    P.stats.fwd_launches = launches;
    P.stats.fwd_cells = static_cast<long long>(P.ntx) * P.ntz * P.g.nt * n;
    return SWD_OK;
}

int total_traces(int nshots, const swd_shot *shots)
{
    int n = 0;
    for (int k = 0; k < nshots; k++) n += shots[k].nrec > 0 ? shots[k].nrec : 0;
    return n;
}

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

}  // namespace swd_rt

using namespace swd_rt;

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
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, device);
        if (e == hipSuccess) {
            P->arch = prop.gcnArchName;
            P->cus = prop.multiProcessorCount;
        }
    }
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

extern "C" int swd_plan_set_tile_config(void *plan, int tile_x, int tile_rows, int waves)
{
    if (!plan) return swd_fail(SWD_E_ARG, "swd_plan_set_tile_config: null plan");
    Plan *P = static_cast<Plan *>(plan);
    int cfg = 0;
    if (tile_x != 0 || tile_rows != 0 || waves != 0) {
        // two-step shapes: waves is the wavefronts per workgroup, negative for split launches
        const int c2 = fstep_cfg2(tile_x, tile_rows / 4, waves < 0 ? -waves : waves) | (waves < 0 ? kFstepSplit : 0);
        const bool two = P->g.Lc == 4 && tile_x > 0 && tile_rows > 0 && fstep2_cfg_supported(c2);
        cfg = two && waves < 0 ? c2 : fstep_cfg(tile_x, tile_rows / 4, waves);
        if (tile_rows % 4 != 0 || (!two && !fstep_cfg_supported(P->g.Lc, false, false, false, false, cfg)))
            return swd_fail(SWD_E_ARG,
                            "swd_plan_set_tile_config: %d x %d rows, %d wavefronts is not a configuration of "
                            "the fused step at Lc %d",
                            tile_x, tile_rows, waves, P->g.Lc);
    }
    P->tile_cfg = cfg;
    P->tuned_fwd.clear();
    P->tuned_bwd.clear();
    P->tuned_cig.clear();
    P->tuned_born.clear();
    return SWD_OK;
}

extern "C" int swd_plan_set_steps_per_launch(void *plan, int steps)
{
    if (!plan || steps < 0 || steps > 2)
        return swd_fail(SWD_E_ARG, "swd_plan_set_steps_per_launch: %d steps per launch, not 0, 1 or 2", steps);
    Plan *P = static_cast<Plan *>(plan);
    P->steps_per_launch = steps;
    P->tuned_fwd.clear();
    P->tuned_bwd.clear();
    return SWD_OK;
}

extern "C" int swd_plan_set_tile(void *plan, int tile_x)
{
    if (!plan || (tile_x != 0 && tile_x != 16 && tile_x != 32 && tile_x != 64))
        return swd_fail(SWD_E_ARG, "swd_plan_set_tile: tile width %d not in {0, 16, 32, 64}", tile_x);
    return swd_plan_set_tile_config(plan, tile_x, tile_x ? 64 : 0, 0);
}

extern "C" int swd_plan_get_stats(void *plan, swd_stats *st)
{
    if (!plan || !st) return swd_fail(SWD_E_ARG, "swd_plan_get_stats: null argument");
    Plan *P = static_cast<Plan *>(plan);
    P->stats.ws_bytes = static_cast<long long>(P->ws_cap);
    *st = P->stats;
    return SWD_OK;
}

extern "C" int swd_plan_set_memory_limit(void *plan, size_t bytes)
{
    if (!plan) return swd_fail(SWD_E_ARG, "swd_plan_set_memory_limit: null plan");
    static_cast<Plan *>(plan)->mem_limit = bytes;
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

