// swd_forward.hip -- forward modelling (swd_forward) and AFor with the kz
// up/down separation (swd_afor, swd_afor_dev).
#include "swd_plan.h"

namespace swd_rt {

// Forward modelling of a batch; gathers stay on the device at d_gather.
static int forward_dev_fused(Plan &P, int nshots, const swd_shot *shots, const Batch &b, const float *wavelet, float *d_gather)
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
    SWD_TRY(fused_forward_loop(P, at<KStep>(P, o_inst), nshots, true));
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

}  // namespace swd_rt

using namespace swd_rt;

extern "C" int swd_forward(void *plan, int nshots, const swd_shot *shots, const float *wavelet, float *gather)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    if (!wavelet || !gather || !shots || nshots < 1) return swd_fail(SWD_E_ARG, "swd_forward: null argument");
    const size_t n = static_cast<size_t>(total_traces(nshots, shots)) * P->g.nt;
    DevBuf d;
    SWD_TRY(d.alloc(sizeof(float) * n));
    int wave = nshots;
    SWD_TRY(wave_shots(*P, nshots, 4.0 * (P->path == SWD_PATH_FUSED ? double(kStepArrays) : 7.0) * P->nfield, 4.0 * P->g.nt * 4, wave));
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
        SWD_TRY(fused_forward_loop(*P, at<KStep>(*P, o_inst), 2 * nshots, true));
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
