// swd_rtm.hip -- RTM and PSF imaging (swd_rtm, swd_rtm_dev) and RTM with angle-domain
// common-image gathers (swd_rtm_cig, swd_rtm_cig_dev).
#include "swd_plan.h"

using namespace swd_rt;

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
    // Angle gathers with two steps per launch: the paired kernel's middle-level variant
    // stores both fields' first level, and k_cig bins the launch's two steps from that and
    // from the slots (three descriptor tables a shot: the slots as after a pair of launches,
    // the slots swapped as after the first launch of a pair, the middle level).
    const bool two_cig = cig && P.g.Lc == 4 && P.steps_per_launch != 1 && P.ldz < (1 << 23) &&
                         (P.tile_cfg == 0 || P.steps_per_launch == 2) && nt >= 3;
    const size_t o_kcig = cig ? c.add<KCig>((two_cig ? 3 : 1) * nshots) : 0;
    const size_t o_mid = c.add<float>(two_cig ? 6 * nf * nshots : 1);
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
    if (two_cig)
        for (int k = 0; k < nshots; k++) {
            KStep &S = hb[2 * k], &R = hb[2 * k + 1];
            float *mid = at<float>(P, o_mid) + 6 * nf * k;
            S.pmid = mid;
            S.vxmid = mid + nf;
            S.vzmid = mid + 2 * nf;
            R.pmid = mid + 3 * nf;
            R.vxmid = mid + 4 * nf;
            R.vzmid = mid + 5 * nf;
        }
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
        std::vector<KCig> hk((two_cig ? 3 : 1) * nshots);
        for (int k = 0; k < nshots; k++) {
            const KStep &S = hb[2 * k], &R = hb[2 * k + 1];
            for (int v = 0; v < (two_cig ? 3 : 1); v++) {
                KCig &K = hk[v * nshots + k];
                for (int q = 0; q < 2; q++) {
                    const int u = v == 1 ? 1 - q : q;
                    K.ps[q] = v == 2 ? S.pmid : S.p[u];
                    K.vxs[q] = v == 2 ? S.vxmid : S.vx[u];
                    K.vzs[q] = v == 2 ? S.vzmid : S.vz[u];
                    K.pr[q] = v == 2 ? R.pmid : R.p[u];
                    K.vxr[q] = v == 2 ? R.vxmid : R.vx[u];
                    K.vzr[q] = v == 2 ? R.vzmid : R.vz[u];
                }
                K.ring = n_ring ? at<float>(P, o_ring) + n_ring * k : nullptr;
                K.cig = at<float>(P, o_cig) + nint * cig->thetan * k;
            }
        }
        HIP_TRY(hipMemcpyAsync(P.ws + o_kcig, hk.data(), sizeof(KCig) * hk.size(), hipMemcpyHostToDevice, P.stream));
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
                                group_stream(static_cast<int>(g)), nullptr, false, false, true));
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
        // the middle level's reconstructed pressure outside the interior too (its stencil reads there)
        if (two_cig)
            HIP_TRY(hipMemcpyAsync(hb[2 * k].pmid, F.p[fin], sizeof(float) * nf, hipMemcpyDeviceToDevice, P.stream));
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
        if (two_cig) {
            // k_cig bins every step after its launch: step it from the slots after a one-step
            // launch, steps it (middle level) and it-1 after a two-step launch (it, it-1), with
            // delt 1 both in one k_cig2 launch.  With delt 1 the first launch bins step nt-1
            // itself (it keeps the forward's velocities; with delt > 1 that step is not binned).
            const KCig *k1 = at<KCig>(P, o_kcig) + k0, *ksw = k1 + nshots, *kmid = k1 + 2 * nshots;
            const int thetan = cig->thetan, delt = cig->delt;
            std::vector<int> its2 = its;
            if (delt == 1) {
                HIP_TRY(launch_fstep(P.kg, d, 2 * n, nt - 2, true, kFstepDefault, true, true, false, gs));
                HIP_TRY(launch_cig(P.kg, k1, n, thetan, 1, w1, w2, cthr, nt - 2, gs));
                launches += 2;
                its2.erase(its2.begin());
            }
            StepHook hook;
            hook.mid = true;
            hook.launches2 = delt == 1 ? 1 : 2;
            hook.run = [&P, k1, n, thetan, delt, w1, w2, cthr](int it, hipStream_t st) -> int {
                HIP_TRY(launch_cig(P.kg, k1, n, thetan, delt, w1, w2, cthr, it, st));
                return SWD_OK;
            };
            hook.run2 = [&P, k1, ksw, kmid, n, thetan, delt, w1, w2, cthr](int it, int ph, hipStream_t st) -> int {
                if (delt == 1) {
                    HIP_TRY(launch_cig2(P.kg, kmid, ph ? k1 : ksw, n, thetan, w1, w2, cthr, it, it - 1, st));
                } else {
                    HIP_TRY(launch_cig(P.kg, kmid, n, thetan, delt, w1, w2, cthr, it, st));
                    HIP_TRY(launch_cig(P.kg, ph ? k1 : ksw, n, thetan, delt, w1, w2, cthr, it - 1, st));
                }
                return SWD_OK;
            };
            hook.key = {4, reinterpret_cast<long long>(k1), thetan, delt};
            if (!its2.empty())
                SWD_TRY(fused_steps(P, d, 2 * n, its2, nt - 2, P.tuned_bwd, &launches, true, gs, &hook, false, false,
                                    true));
            continue;
        }
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
        SWD_TRY(fused_steps(P, d, 2 * n, its, nt - 2, P.tuned_bwd, &launches, true, gs, cig ? &hook : nullptr, false,
                            false, !cig));
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


extern "C" int swd_rtm_dev(void *plan, int nshots, const swd_shot *shots, const float *wavelet, const float *adj,
                           float *image, float *illum_src, float *illum_rcv, float *gather_syn)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !adj || !image || !illum_src || !illum_rcv || !shots || nshots < 1)
        return swd_fail(SWD_E_ARG, "swd_rtm: null argument");
    const double nt = P->g.nt, nf = static_cast<double>(P->nfield);
    const double per = 4.0 * (nt * P->nframe + (P->path == SWD_PATH_FUSED ? 2.0 * kStepArrays + 3.0 : 18.0) * nf);
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
    const double per = 4.0 * (nt * P->nframe + (2.0 * kStepArrays + 3.0) * nf) + cig_bytes(*P, &spec);
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
