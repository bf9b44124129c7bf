// swd_rtm_updown.hip -- up/down-separated RTM (swd_rtm_updown, swd_rtm_updown_dev) and its
// per-step imaging condition on caller fields (swd_updown_image_dev).
#include "swd_plan.h"

using namespace swd_rt;

void swd_rt::kz_twiddles_into(int nfftz, float2 *tw)
{
    for (int k = 0; k < nfftz / 2; k++) {
        const double ang = -1 * 2.0 * 3.14159265358979323846 * k / nfftz;
        tw[k] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
    }
}

// ------------------------------------------------------------------ up/down imaging
namespace {

std::vector<float2> kz_twiddles(int nfftz)
{
    std::vector<float2> tw(nfftz / 2);
    kz_twiddles_into(nfftz, tw.data());
    return tw;
}

// Wavefield arrays of one shot of the up/down RTM: four wavefields (12 arrays
// each), then du, ud, uu, dd, image, illum_s, illum_r and the two dp/dt.
// du, ud, uu, dd, image, illum_s, illum_r, the two reconstructions' dp/dt, their second
// levels and the two receiver fields' first-level pressures of two-step launches
constexpr int kUdImages = 13;
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
    // up/down descriptors: one step per launch; two steps, first level; two steps, second level
    // after the first and after the second launch of a pair
    const size_t o_ud = c.add<KUpDown>(4 * nshots);
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
    std::vector<KUpDown> hu(4 * nshots);
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
            S.dpdt_b = im + (9 + h) * nf;
            R.pmid = im + (11 + h) * nf;
            if (h == 0 && plain) {
                R.image = im + 4 * nf;
                R.illum_s = im + 5 * nf;
                R.illum_r = im + 6 * nf;
            }
        }
        for (int v = 0; v < 4; v++) {
            // k_updown of step it reads p of slot it & 1: one step per launch (v 0) and the second level of
            // the second launch of a pair (v 3) find it there; the first launch of a pair (v 2) wrote its
            // second level into the other slot; the first level of a two-step launch (v 1) is in pmid
            KUpDown &U = hu[v * nshots + k];
            const bool second = v >= 2;
            U.ds = im + (second ? 9 : 7) * nf;
            U.dh = im + (second ? 10 : 8) * nf;
            for (int q = 0; q < 2; q++) {
                const int sq = v == 2 ? 1 - q : q;
                U.pr[q] = v == 1 ? im + 11 * nf : hb[4 * k + 1].p[sq];
                U.ph[q] = v == 1 ? im + 12 * nf : hb[4 * k + 3].p[sq];
            }
            U.du = im;
            U.ud = im + nf;
            U.uu = both ? im + 2 * nf : nullptr;
            U.dd = both ? im + 3 * nf : nullptr;
            U.tw = at<float2>(P, o_tw);
            U.coef = at<float>(P, o_coef);
            U.ld = P.ldz;
        }
    }
    SWD_TRY(upload_steps(P, hf, o_fwd));
    SWD_TRY(upload_steps(P, hb, o_bwd));
    HIP_TRY(hipMemcpyAsync(P.ws + o_ud, hu.data(), sizeof(KUpDown) * hu.size(), hipMemcpyHostToDevice, P.stream));
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
                                rtm_group_stream(P, static_cast<int>(g)), nullptr, false, false, true));
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
        // a two-step launch of steps it and it-1: both levels in one launch, the first level's products first
        hook.run2 = [&P, du, n, nshots, nfftz](int it, int ph, hipStream_t st) -> int {
            HIP_TRY(launch_updown2(P.kg, du + nshots, du + (ph == 0 ? 2 : 3) * nshots, n, nfftz, it, it - 1, st));
            return SWD_OK;
        };
        hook.launches2 = updown2_launches(nfftz);
        hook.key = {2, reinterpret_cast<long long>(du), nfftz};
        SWD_TRY(fused_steps(P, at<KStep>(P, o_bwd) + 4 * k0, 4 * n, its, nt - 2, P.tuned_bwd, &launches, true,
                            rtm_group_stream(P, static_cast<int>(g)), &hook, false, false, true));
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

int swd_rt::check_updown_image_args(const Plan &P, const char *fn, int nshots, const float *src, const float *hsrc,
                                    const float *rcv, const float *hrcv, const float *kzcoef, int nfftz,
                                    const swd_updown_images *out)
{
    if (!src || !hsrc || !rcv || !hrcv || !kzcoef || !out || nshots < 1)
        return swd_fail(SWD_E_ARG, "%s: null argument", fn);
    if (!out->du || !out->ud) return swd_fail(SWD_E_ARG, "%s: du and ud outputs are required", fn);
    if ((out->uu == nullptr) != (out->dd == nullptr))
        return swd_fail(SWD_E_ARG, "%s: uu and dd are both given or both null", fn);
    if (nfftz < P.g.nz || nfftz < 128 || (nfftz & (nfftz - 1)) || nfftz > 4096)
        return swd_fail(SWD_E_ARG, "%s: nfftz=%d must be a power of two in [max(nz=%d, 128), 4096]", fn, nfftz,
                        P.g.nz);
    return SWD_OK;
}

void swd_rt::updown_image_descriptors(const Plan &P, int nshots, const float *src, const float *hsrc,
                                      const float *rcv, const float *hrcv, const swd_updown_images *out,
                                      const float2 *tw, const float *coef, KUpDown *hu)
{
    const size_t ntp = static_cast<size_t>(P.ntx) * P.ntz;
    for (int k = 0; k < nshots; k++) {
        KUpDown &U = hu[k];
        std::memset(&U, 0, sizeof U);
        const size_t o = ntp * k;
        U.ds = src + o;
        U.dh = hsrc + o;
        U.pr[0] = U.pr[1] = rcv + o;
        U.ph[0] = U.ph[1] = hrcv + o;
        U.du = out->du + o;
        U.ud = out->ud + o;
        U.uu = out->uu ? out->uu + o : nullptr;
        U.dd = out->dd ? out->dd + o : nullptr;
        U.tw = tw;
        U.coef = coef;
        U.ld = P.ntz;
    }
}

extern "C" int swd_updown_image_dev(void *plan, int nshots, const float *src, const float *hsrc, const float *rcv,
                                    const float *hrcv, const float *kzcoef, int nfftz, const swd_updown_images *out)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));
    SWD_TRY(check_updown_image_args(*P, "swd_updown_image_dev", nshots, src, hsrc, rcv, hrcv, kzcoef, nfftz, out));
    Carver c;
    const size_t o_ud = c.add<KUpDown>(nshots);
    const size_t o_tw = c.add<float2>(nfftz / 2);
    const size_t o_coef = c.add<float>(nfftz);
    SWD_TRY(ensure_ws(*P, c.off, true));
    std::vector<KUpDown> hu(nshots);
    updown_image_descriptors(*P, nshots, src, hsrc, rcv, hrcv, out, at<float2>(*P, o_tw), at<float>(*P, o_coef),
                             hu.data());
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
