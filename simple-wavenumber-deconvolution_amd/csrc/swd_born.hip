// swd_born.hip -- Born modelling and its exact adjoint with sqrt(nt) checkpoints
// (swd_born, swd_born_adjoint and their device-pointer forms).
#include "swd_plan.h"

using namespace swd_rt;

// ------------------------------------------------------------------ Born
namespace {

// Snapshots of the fused path's background state: array j (p, vx, vz, psx,
// psz in the slot that holds the state, pvx, pvz) of shot k, slots[per k + j],
// and its snapshot at snap + kstride k + nf j, in float4s (nf a multiple of
// 4); save copies the state to the snapshot, else the snapshot back.
__global__ __launch_bounds__(256) void k_snap(float4 *__restrict__ snap, float4 *const *__restrict__ slots, int per,
                                              size_t nf4, size_t kstride4, int save)
{
    const int a = blockIdx.y, k = a / per, j = a - k * per;
    float4 *st = slots[a], *sn = snap + kstride4 * k + nf4 * j;
    const float4 *__restrict__ from = save ? st : sn;
    float4 *__restrict__ to = save ? sn : st;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < nf4; i += static_cast<size_t>(gridDim.x) * 256)
        to[i] = from[i];
}

// Born forward of a batch: m on the device [nshots][nx][nz], d on the device
// [sum nrec][nt].  Each shot runs the background wavefield and the scattered
// wavefield as two instances of the same launches.
// Born forward on the fused path: the background (2k) and scattered (2k+1)
// fields of shot k in one workgroup per tile, the Born source from the
// background's divergence sums in registers (two steps per launch where the
// plan allows); shots in cache-sized groups on the plan's streams.
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
                            true, true));
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
    // Fused path: the background runs as KStep instances (both slots), on
    // the two-step forward where the plan allows; the batched path keeps one
    // KInst state.  The adjoint fields carry ex, ez scratch, on the fused path
    // the second slots of pvx, pvz, psx, psz (the tiled adjoint step).
    const bool fused = P.path == SWD_PATH_FUSED;
    const bool two = fused && P.g.Lc == 4 && P.steps_per_launch != 1 && P.ldz < (1 << 23);
    const size_t per_b = fused ? kStepArrays : kFieldArrays, per_a = kFieldArrays + (fused ? 4 : 2);
    Carver c;
    const size_t o_ints = c.add<int>(b.ints.size());
    const size_t o_pairs = c.add<int2>(std::max<size_t>(1, b.dup_pairs.size()));
    const size_t o_inst = c.add<KInst>(2 * nshots);
    const size_t o_wav = c.add<float>(nt);
    const size_t o_d = c.add<float>(std::max<size_t>(ng, 1));
    const size_t o_grad = c.add<float>(nint * nshots);
    const size_t o_s0 = c.add<float>(nint * seg * nshots);
    const size_t o_snap = c.add<float>(kFieldArrays * nf * static_cast<size_t>(nseg) * nshots);
    const size_t o_f = c.add<float>((per_b + per_a) * nf * nshots);
    const size_t o_step = c.add<KStep>(fused ? nshots : 1);
    const size_t o_slots = c.add<float *>(fused ? 2 * kFieldArrays * nshots : 1);
    SWD_TRY(ensure_ws(P, c.off));
    SWD_TRY(upload_batch(P, b, o_ints, o_pairs));
    HIP_TRY(hipMemcpyAsync(P.ws + o_wav, wavelet, sizeof(float) * nt, hipMemcpyHostToDevice, P.stream));
    // traces sharing a node are summed into the node's last trace (the adjoint of the shared recording)
    float *dsum = at<float>(P, o_d);
    HIP_TRY(hipMemcpyAsync(dsum, d, sizeof(float) * ng, hipMemcpyDeviceToDevice, P.stream));
    HIP_TRY(launch_sum_rows(dsum, d, at<int2>(P, o_pairs), static_cast<int>(b.dup_pairs.size()), nt, P.stream));
    zero_fields(P, at<float>(P, o_f), (per_b + per_a) * nf * nshots);
    zero_fields(P, at<float>(P, o_grad), nint * nshots);
    std::vector<KInst> h(2 * nshots);
    std::vector<KStep> hs(fused ? nshots : 0);
    float *fb = at<float>(P, o_f), *fa = fb + per_b * nf * nshots;
    // the batched kernels' view of the fused background: the arrays of slot c
    auto view = [&](int k, int cs) {
        const KStep &S = hs[k];
        KInst &B = h[k];
        B.p = S.p[cs]; B.vx = S.vx[cs]; B.vz = S.vz[cs]; B.psx = S.psx[cs]; B.psz = S.psz[cs];
        B.pvx = S.pvx; B.pvz = S.pvz;
    };
    for (int k = 0; k < nshots; k++) {
        KInst &B = h[k], &A = h[nshots + k];
        std::memset(&B, 0, sizeof B);
        std::memset(&A, 0, sizeof A);
        if (fused) {
            KStep &S = hs[k];
            std::memset(&S, 0, sizeof S);
            S.mode = KSTEP_FWD;
            bind_step(S, fb + kStepArrays * nf * k, nf);
            S.wavelet = at<float>(P, o_wav);
            S.s_ix = shots[k].s_ix;
            S.s_iz = shots[k].s_iz;
            view(k, 0);
        } else {
            bind_fields(B, fb + kFieldArrays * nf * k, nf);
        }
        bind_fields(A, fa + per_a * nf * k, nf);
        A.ex = A.p + kFieldArrays * nf;
        A.ez = A.ex + nf;
        if (fused) {
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
    const KStep *dS = at<KStep>(P, o_step);
    const KGrid &g = P.kg;
    long long fl = 0, bl = 0;
    // a snapshot holds the background state before the first step of a segment
    auto snap = [&](int k, int s) -> float * {
        return at<float>(P, o_snap) + kFieldArrays * nf * (static_cast<size_t>(nseg) * k + s);
    };
    auto steps = [&](int it0, int it1) -> int { return batched_forward_steps(g, dB, nshots, it0, it1, P.stream, &fl); };
    // fused path: the state before step t0 is in slot (t0 + 1) & 1 after whole
    // loops; k_snap moves it to or from the KInst-layout snapshot of segment s
    float4 *const *dslots = at<float4 *>(P, o_slots);
    auto snap_copy = [&](int s, int t0, bool save) -> int {
        const size_t nf4 = nf / 4;
        const dim3 sg(static_cast<unsigned>(std::min<size_t>((nf4 + 255) / 256, 1024)), kFieldArrays * nshots);
        hipLaunchKernelGGL(k_snap, sg, dim3(256), 0, P.stream, reinterpret_cast<float4 *>(snap(0, s)),
                           dslots + kFieldArrays * nshots * ((t0 + 1) & 1), kFieldArrays, nf4,
                           kFieldArrays * nf4 * static_cast<size_t>(nseg), save ? 1 : 0);
        HIP_TRY(hipGetLastError());
        return SWD_OK;
    };
    if (fused) {
        std::vector<float *> slots(2 * kFieldArrays * nshots);
        for (int k = 0; k < nshots; k++)
            for (int cs = 0; cs < 2; cs++) {
                const KStep &S = hs[k];
                float **q = slots.data() + kFieldArrays * (nshots * cs + k);
                q[0] = S.p[cs]; q[1] = S.vx[cs]; q[2] = S.vz[cs]; q[3] = S.psx[cs]; q[4] = S.psz[cs];
                q[5] = S.pvx; q[6] = S.pvz;
            }
        HIP_TRY(hipMemcpyAsync(P.ws + o_slots, slots.data(), sizeof(float *) * slots.size(), hipMemcpyHostToDevice,
                               P.stream));
        SWD_TRY(upload_steps(P, hs, o_step));
    }
    // the two-step configuration of the recompute pass: the forward loops' pick for this many
    // instances, else the default
    int cfg2 = kFstep2Default;
    // each segment is its own loop: no graph is worth capturing for one replay
    const bool graphs = P.graphs;
    P.graphs = false;
    int rc = SWD_OK;
    // checkpoint pass
    SWD_TRY(upload_insts(P, h, o_inst));
    stats_begin(P, 0);
    for (int s = 0; s < nseg && rc == SWD_OK; s++) {
        const int t0 = s * seg, t1 = std::min(nt, t0 + seg);
        if (fused) {
            rc = snap_copy(s, t0, true);
            std::vector<int> its;
            for (int it = t0; it < t1; it++) its.push_back(it);
            if (rc == SWD_OK)
                rc = fused_steps(P, dS, nshots, its, -1, P.tuned_fwd, &fl, false, P.stream, nullptr, false, false, true);
        } else {
            for (int k = 0; k < nshots && rc == SWD_OK; k++)
                if (hipMemcpyAsync(snap(k, s), h[k].p, sizeof(float) * kFieldArrays * nf, hipMemcpyDeviceToDevice,
                                   P.stream) != hipSuccess)
                    rc = swd_fail(SWD_E_DEVICE, "born adjoint: snapshot copy failed");
            if (rc == SWD_OK) rc = steps(t0, t1);
        }
    }
    if (fused && rc == SWD_OK) {
        const auto f = P.tuned_fwd.find(4 * nshots + 2);   // fused_steps' key of a two-step loop
        if (f != P.tuned_fwd.end() && fstep_cfg_two(f->second)) cfg2 = f->second;
    }
    // reverse segments: recompute the background sums, then the adjoint steps.  On
    // the fused path whole pairs of two-step launches come first (the state keeps
    // its slot after four steps), then the rest on the batched kernels over the
    // slot that holds the state.
    for (int s = nseg - 1; s >= 0 && rc == SWD_OK; s--) {
        const int t0 = s * seg, t1 = std::min(nt, t0 + seg);
        const int nq = two ? (t1 - t0) / 4 : 0;
        if (fused) {
            rc = snap_copy(s, t0, false);
            for (int k = 0; k < nshots; k++) {
                hs[k].s0_out = at<float>(P, o_s0) + nint * seg * k;
                hs[k].s0_t0 = t0;
                view(k, (t0 + 1) & 1);
            }
            if (rc == SWD_OK && nq > 0) rc = upload_steps(P, hs, o_step);
            for (int q = 0; q < nq && rc == SWD_OK; q++)
                for (int ph = 0; ph < 2 && rc == SWD_OK; ph++) {
                    if (launch_fstep2(g, dS, nshots, t0 + 4 * q + 2 * ph, ph, cfg2, false, P.stream) != hipSuccess)
                        rc = swd_fail(SWD_E_DEVICE, "born adjoint: two-step recompute launch failed");
                    fl++;
                }
        } else {
            for (int k = 0; k < nshots && rc == SWD_OK; k++)
                if (hipMemcpyAsync(h[k].p, snap(k, s), sizeof(float) * kFieldArrays * nf, hipMemcpyDeviceToDevice,
                                   P.stream) != hipSuccess)
                    rc = swd_fail(SWD_E_DEVICE, "born adjoint: snapshot restore failed");
        }
        for (int k = 0; k < nshots; k++) {
            h[k].s0_out = at<float>(P, o_s0) + nint * seg * k;
            h[k].s0_t0 = t0;
            h[nshots + k].s0_t0 = t0;
        }
        if (rc == SWD_OK) rc = upload_insts(P, h, o_inst);
        if (rc == SWD_OK) rc = steps(t0 + 4 * nq, t1);
        for (int it = t1 - 1; it >= t0 && rc == SWD_OK; it--) {
            const hipError_t e = fused ? launch_adjoint_step_tiled(g, dA, nshots, it, it == nt - 1, P.stream)
                                       : launch_adjoint_step(g, dA, nshots, it, it == nt - 1, P.stream);
            if (e != hipSuccess) rc = swd_fail(SWD_E_DEVICE, "born adjoint: adjoint step launch failed");
            bl += it == nt - 1 ? 1 : (fused ? 2 : 4);
        }
    }
    P.graphs = graphs;
    SWD_TRY(rc);
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
    // per shot: snapshots, one segment of background sums, the background and
    // adjoint fields (kStepArrays + 11 on the fused path, 7 + 9 on the batched),
    // the gradient and the data
    const double nfa = P->path == SWD_PATH_FUSED ? kStepArrays + kFieldArrays + 4.0 : 2.0 * kFieldArrays + 2.0;
    const double per = 4.0 * (7.0 * nf * nseg + nint * seg + nfa * nf + nint + static_cast<double>(ntr_max) * nt);
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
