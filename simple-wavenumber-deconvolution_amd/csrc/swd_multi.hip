// swd_multi.hip -- forward modelling of several plans of uneven grid sizes in one
// fused launch per step (swd_multi_forward).
#include "swd_plan.h"

using namespace swd_rt;

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
    // instance and grid tables
    std::vector<KGrid> grids(nplans);
    std::vector<KStep> h;
    std::vector<int> inst0(nplans);
    for (int i = 0; i < nplans; i++) {
        grids[i] = Ps[i]->kg;
        inst0[i] = static_cast<int>(h.size());
        for (int k = 0; k < nshots[i]; k++) h.emplace_back();
    }
    // Groups of whole instances (an instance's tiles exchange halos between
    // steps): as many as the cache budget asks, rounded up to a multiple of the
    // streams (at most eight), each closed once it holds its share of the
    // per-step working set, so the streams carry even loads (c4 on two streams:
    // four groups of 72-82 MB).  Each group runs its own time loop on the
    // plan's streams in turn.  A group is a list of (plan, first shot, end).
    struct Seg {
        int plan, k0, k1;
    };
    std::vector<std::vector<Seg>> groups(1);
    {
        double total = 0.0;
        for (int i = 0; i < nplans; i++) total += fwd_group_bytes(*Ps[i]) * nshots[i];
        const int ns = P0.streams;
        int ng = std::max(1, static_cast<int>(std::ceil(total / (kCacheBudgetBwd / ns))));
        ng = std::min(8, (ng + ns - 1) / ns * ns);
        const double target = total / ng;
        double used = 0.0;
        for (int i = 0; i < nplans; i++) {
            const double per = fwd_group_bytes(*Ps[i]);
            int k0 = 0;
            for (int k = 0; k < nshots[i]; k++) {
                used += per;
                const bool close = used >= target && static_cast<int>(groups.size()) < ng;
                if (close || k == nshots[i] - 1) {
                    groups.back().push_back(Seg{i, k0, k + 1});
                    k0 = k + 1;
                }
                if (close) {
                    groups.emplace_back();
                    used = 0.0;
                }
            }
        }
        if (groups.back().empty()) groups.pop_back();
    }
    // Block table of tile configuration cfg (16 columns, 4 tq rows): per group,
    // per segment, its grid's tiles numbered z-fastest, the shots of one tile
    // next to each other (shared model in L2); off[g], cnt[g] locate group g.
    struct BTab {
        int cfg;
        std::vector<int2> t;
        std::vector<size_t> off;
        std::vector<int> cnt;
    };
    auto make_btab = [&](int cfg) {
        BTab B;
        B.cfg = cfg;
        const int rows = 4 * fstep_cfg_tq(cfg);
        for (const auto &G : groups) {
            B.off.push_back(B.t.size());
            for (const Seg &sg : G) {
                const int nzb = (Ps[sg.plan]->ntz + rows - 1) / rows, nxb = (Ps[sg.plan]->ntx + 15) / 16;
                for (int t = 0; t < nzb * nxb; t++)
                    for (int q = sg.k0; q < sg.k1; q++) B.t.push_back(make_int2(inst0[sg.plan] + q, t));
            }
            B.cnt.push_back(static_cast<int>(B.t.size() - B.off.back()));
        }
        return B;
    };
    // the tile: the plan's, else remembered, else tuned on group 0's first steps
    const int ninst = static_cast<int>(h.size());
    long long cells_all = 0;
    for (int i = 0; i < nplans; i++) cells_all += static_cast<long long>(Ps[i]->ntx) * Ps[i]->ntz * nshots[i];
    const std::string tkey = tune_key(P0, "multi", ninst) + " plans " + std::to_string(nplans) + " cells " +
                             std::to_string(cells_all);
    int cfg = P0.tile_cfg;
    if (cfg != 0 && !fstep_cfg_supported(P0.g.Lc, false, false, false, true, cfg)) cfg = kFstepDefault;
    if (cfg == 0) {
        cfg = tune_lookup(tkey);
        if (cfg != 0 && !fstep_cfg_supported(P0.g.Lc, false, false, false, true, cfg)) cfg = 0;
    }
    constexpr int kTry = 5;
    std::vector<int> cand = fstep_candidates(P0.g.Lc, false, false, false, true);
    const bool tune = cfg == 0 && nt >= static_cast<int>(cand.size()) * (kTry + 1) + 1;
    if (cfg == 0 && !tune) cfg = kFstepDefault;
    if (!tune) cand.assign(1, cfg);
    std::vector<BTab> tabs;
    for (int c : cand) tabs.push_back(make_btab(c));
    // Two steps per launch (Lc 4, as fused_steps decides: not with one step or an asked one-step
    // tile): each group's leftover steps modulo 4 on the one-step kernel, the rest in pairs of
    // k_fstep2_multi launches over 32 x 32-row tiles; its table per group, per segment, the
    // grid's tiles z-fastest with the shots of a tile next to each other and the fast flag.
    bool two = P0.g.Lc == 4 && P0.steps_per_launch != 1 && (P0.tile_cfg == 0 || P0.steps_per_launch == 2);
    for (int i = 0; i < nplans; i++) two = two && Ps[i]->ldz < (1 << 23);
    std::vector<int4> t2;
    std::vector<size_t> off2;
    std::vector<int> cnt2;
    if (two)
        for (const auto &G : groups) {
            off2.push_back(t2.size());
            for (const Seg &sg : G) {
                const Plan &P = *Ps[sg.plan];
                const int nzb = (P.ntz + 31) / 32, nxb = (P.ntx + 31) / 32;
                const int4 fr = fstep2_fast_rect(P.kg, 32, 8, nxb, nzb);
                for (int t = 0; t < nzb * nxb; t++) {
                    const int xb = t / nzb, zb = t % nzb;
                    const int fast = xb >= fr.x && xb < fr.x + fr.y && zb >= fr.z && zb < fr.z + fr.w;
                    for (int q = sg.k0; q < sg.k1; q++) t2.push_back(make_int4(inst0[sg.plan] + q, xb, zb, fast));
                }
            }
            cnt2.push_back(static_cast<int>(t2.size() - off2.back()));
        }
    Carver ca;
    const size_t oa_grid = ca.add<KGrid>(nplans);
    const size_t oa_inst = ca.add<KStep>(h.size());
    std::vector<size_t> oa_btab;
    for (const BTab &B : tabs) oa_btab.push_back(ca.add<int2>(B.t.size()));
    const size_t oa_btab2 = ca.add<int4>(std::max<size_t>(1, t2.size()));
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
    for (size_t c = 0; c < tabs.size(); c++)
        HIP_TRY(hipMemcpyAsync(P0.aux + oa_btab[c], tabs[c].t.data(), sizeof(int2) * tabs[c].t.size(),
                               hipMemcpyHostToDevice, P0.stream));
    if (two)
        HIP_TRY(hipMemcpyAsync(P0.aux + oa_btab2, t2.data(), sizeof(int4) * t2.size(), hipMemcpyHostToDevice, P0.stream));
    const KStep *d_inst = reinterpret_cast<const KStep *>(P0.aux + oa_inst);
    stats_begin(P0, 0);
    SWD_TRY(fork_streams(P0));
    int it0 = 0;   // first step of group 0 after the tuning launches
    if (tune) {
        // group 0 runs kTry + 1 of its steps with every candidate (the first untimed)
        const hipStream_t st = rtm_group_stream(P0, 0);
        std::vector<hipEvent_t> ev(cand.size() + 1);
        for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
        std::vector<float> ms(cand.size(), 0.f);
        int rc = SWD_OK;
        for (size_t c = 0; c < cand.size() && rc == SWD_OK; c++) {
            const int2 *bt = reinterpret_cast<const int2 *>(P0.aux + oa_btab[c]) + tabs[c].off[0];
            for (int j = 0; j <= kTry && rc == SWD_OK; j++) {
                if (j == 1 && hipEventRecord(ev[c], st) != hipSuccess) rc = SWD_E_DEVICE;
                if (rc == SWD_OK && launch_fstep_multi(P0.g.Lc, cand[c], d_inst, bt, tabs[c].cnt[0], it0++, st) != hipSuccess)
                    rc = SWD_E_DEVICE;
            }
            if (rc == SWD_OK && hipEventRecord(ev[cand.size()], st) != hipSuccess) rc = SWD_E_DEVICE;
            if (rc == SWD_OK && hipEventSynchronize(ev[cand.size()]) != hipSuccess) rc = SWD_E_DEVICE;
            if (rc == SWD_OK && hipEventElapsedTime(&ms[c], ev[c], ev[cand.size()]) != hipSuccess) rc = SWD_E_DEVICE;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        if (rc != SWD_OK) return swd_fail(rc, "swd_multi_forward: a tuning launch failed");
        cfg = choose_cfg(ms, cand);
        tune_store(tkey, cfg);
    }
    size_t ci = 0;
    while (cand[ci] != cfg) ci++;
    const BTab &B = tabs[ci];
    const int2 *d_btab = reinterpret_cast<const int2 *>(P0.aux + oa_btab[ci]);
    const int4 *d_btab2 = reinterpret_cast<const int4 *>(P0.aux + oa_btab2);
    P0.stats.tile_x = two ? 32 : fstep_cfg_tx(cfg);
    P0.stats.tile_rows = two ? 32 : 4 * fstep_cfg_tq(cfg);
    P0.stats.tile_waves = two ? 0 : fstep_cfg_wb(cfg);
    P0.stats.steps_per_launch = two ? 2 : 1;
    P0.stats.tile_group_waves = two ? 4 : fstep_threads_of(cfg) / 64;
    long long nlaunch = 0;
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const hipStream_t st = rtm_group_stream(P0, static_cast<int>(gi));
        const int first = gi == 0 ? it0 : 0;
        const int r = two ? (nt - first) % 4 : nt - first;   // one-step launches
        nlaunch += r + (nt - first - r) / 2;
        auto record = [&]() -> int {
            int it = first;
            for (; it < first + r; it++)
                HIP_TRY(launch_fstep_multi(P0.g.Lc, cfg, d_inst, d_btab + B.off[gi], B.cnt[gi], it, st));
            for (; it < nt; it += 4)
                for (int ph = 0; ph < 2; ph++)
                    HIP_TRY(launch_fstep2_multi(d_inst, d_btab2 + off2[gi], cnt2[gi], it + 2 * ph, ph, st));
            return SWD_OK;
        };
        if (P0.graphs) {
            std::vector<long long> key = {2, reinterpret_cast<long long>(P0.aux), static_cast<long long>(h.size()),
                                          static_cast<long long>(B.off[gi]), B.cnt[gi], nt, first, cfg, two ? 1 : 0};
            for (int i = 0; i < nplans; i++) key.push_back(reinterpret_cast<long long>(Ps[i]->ws));
            SWD_TRY(graph_run(P0, key, record, st));
        } else {
            SWD_TRY(record());
        }
    }
    SWD_TRY(join_streams(P0));
    P0.stats.fwd_ms = stats_end(P0, 0);
    P0.stats.fwd_launches = nlaunch + it0;
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

