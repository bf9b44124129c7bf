// swd_noalloc.hip -- swd_updown_image_dev_noalloc: the per-step up/down
// imaging condition of swd_updown_image_dev for callers that cannot allocate
// device memory.  Its tables (the shots' descriptors, the twiddles and the kz
// taper) go to a fixed array of this library on each GPU instead of the plan's
// workspace, so a call queues copies and one kernel and allocates nothing.
#include <mutex>

#include "swd_plan.h"

using namespace swd_rt;

namespace {

// Layout of the scratch array: twiddles [kMaxNfft/2] float2, taper [kMaxNfft]
// floats, descriptors [SWD_NOALLOC_MAX_SHOTS] KUpDown.
constexpr int kMaxNfft = 4096;
constexpr size_t kOffTw = 0;
constexpr size_t kOffCoef = kOffTw + sizeof(float2) * (kMaxNfft / 2);
constexpr size_t kOffDesc = kOffCoef + sizeof(float) * kMaxNfft;
constexpr size_t kScratchBytes = 64 * 1024;
static_assert(kOffDesc + sizeof(KUpDown) * SWD_NOALLOC_MAX_SHOTS <= kScratchBytes, "scratch too small");

constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char *g_ptr[kMaxDevices] = {};

}  // namespace

// Each GPU holds its own copy of the array for the life of the process; host
// code reaches it only through the address scratch_on() looks up per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
int scratch_on(int dev, unsigned char *&ptr);
}

extern "C" int swd_updown_image_dev_noalloc(void *plan, int nshots, const float *src, const float *hsrc,
                                            const float *rcv, const float *hrcv, const float *kzcoef, int nfftz,
                                            const swd_updown_images *out)
{
    Plan *P;
    SWD_TRY(check_plan(plan, P));   // makes the plan's device current
    SWD_TRY(check_updown_image_args(*P, "swd_updown_image_dev_noalloc", nshots, src, hsrc, rcv, hrcv, kzcoef, nfftz,
                                    out));
    if (nshots > SWD_NOALLOC_MAX_SHOTS)
        return swd_fail(SWD_E_RANGE, "swd_updown_image_dev_noalloc: %d shots, at most %d per call", nshots,
                        SWD_NOALLOC_MAX_SHOTS);
    unsigned char *base = nullptr;
    SWD_TRY(scratch_on(P->dev, base));
    float2 *d_tw = reinterpret_cast<float2 *>(base + kOffTw);
    float *d_coef = reinterpret_cast<float *>(base + kOffCoef);
    KUpDown *d_desc = reinterpret_cast<KUpDown *>(base + kOffDesc);
    // host staging on the stack: a copy from pageable memory has read its
    // source when hipMemcpyAsync returns
    float2 tw[kMaxNfft / 2];
    KUpDown hu[SWD_NOALLOC_MAX_SHOTS];
    kz_twiddles_into(nfftz, tw);
    updown_image_descriptors(*P, nshots, src, hsrc, rcv, hrcv, out, d_tw, d_coef, hu);
    HIP_TRY(hipMemcpyAsync(d_tw, tw, sizeof(float2) * (nfftz / 2), hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(d_coef, kzcoef, sizeof(float) * nfftz, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(hipMemcpyAsync(d_desc, hu, sizeof(KUpDown) * nshots, hipMemcpyHostToDevice, P->stream));
    HIP_TRY(launch_updown(P->kg, d_desc, nshots, nfftz, 0, P->stream));
    return SWD_OK;
}

namespace {

// The device address of scratch on the current device dev, looked up once.
int scratch_on(int dev, unsigned char *&ptr)
{
    if (dev < 0 || dev >= kMaxDevices) return swd_fail(SWD_E_DEVICE, "device %d beyond %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> hold(g_lock);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return swd_fail(SWD_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_ptr[dev] = static_cast<unsigned char *>(p);
    }
    ptr = g_ptr[dev];
    return SWD_OK;
}

}  // namespace
