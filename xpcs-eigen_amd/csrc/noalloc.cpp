// The allocation-free entry point: Corr::normalizeG2s of device-resident
// reference-unit arrays into device-resident results, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops).
//
// Its intermediates (FP64 sums, chunk partials, group values) live in one
// statically placed device array of kNoAllocScratchBytes.  Each device the
// library runs on holds its own copy of the array until the process ends; its
// address on a device is resolved once, with that device current, when the
// first plan is created there (hipGetSymbolAddress), and only that address is
// ever handed to kernels.  Every call on a device shares the device's array:
// calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace xpcs {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

// Resolves the device's scratch address; the caller has made `device` current.
xpcs_status noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(XPCS_ERR_UNSUPPORTED, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return XPCS_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(XPCS_ERR_HIP, "scratch address on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = ptr;
    return XPCS_OK;
}

static void* scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

// Bytes of scratch one call needs: FP64 sums [P][T] x 3, chunk partials
// [chunks][T], group values [groups][T], each rounded up to 256 bytes.
static size_t scratch_need(const HostTables& h, size_t* off_part, size_t* off_g2s)
{
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t PT = (size_t)h.P * (size_t)h.T;
    const size_t sums = up(3 * PT * sizeof(int64_t));
    const size_t part = up(h.chunks.size() * (size_t)h.T * sizeof(NormPartial));
    const size_t g2s = up((size_t)h.ngroups * (size_t)h.T * sizeof(double));
    *off_part = sums;
    *off_g2s = sums + part;
    return sums + part + g2s;
}

}  // namespace xpcs

using namespace xpcs;

extern "C" xpcs_status xpcs_mt_normalize_ref_f32_noalloc(xpcs_mt_plan* p, const float* G2, const float* IP,
                                                         const float* IF, int64_t pixels, int32_t taus, double* g2,
                                                         double* stderr_, void* stream)
{
    if (!p || !G2 || !IP || !IF || !g2 || !stderr_)
        return fail(XPCS_ERR_INVALID_ARG, "normalize_ref_f32_noalloc: NULL argument");
    const HostTables& h = p->h;
    const int64_t npd = (int64_t)h.W * h.H;
    if (pixels != npd || taus != h.T)
        return fail(XPCS_ERR_INVALID_ARG,
                    "normalize_ref_f32_noalloc: arrays are %d x %lld, the plan needs %d taus x %lld pixels", taus,
                    (long long)pixels, h.T, (long long)npd);
    size_t off_part = 0, off_g2s = 0;
    const size_t need = scratch_need(h, &off_part, &off_g2s);
    if (need > kNoAllocScratchBytes)
        return fail(XPCS_ERR_UNSUPPORTED,
                    "normalize_ref_f32_noalloc: %d pixels x %d taus need %zu scratch bytes, the maximum is %zu "
                    "(3 x 8 x pixels x taus plus partials); use xpcs_mt_normalize_ref_f32",
                    h.P, h.T, need, kNoAllocScratchBytes);
    unsigned char* base = static_cast<unsigned char*>(scratch_of(p->device));
    if (!base) return fail(XPCS_ERR_STATE, "normalize_ref_f32_noalloc: no scratch on device %d", p->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->stream;
    const size_t PT = (size_t)h.P * (size_t)h.T;
    int64_t* sums = reinterpret_cast<int64_t*>(base);
    launch_from_ref_f32(p, s, G2, IP, IF, sums, sums + PT, sums + 2 * PT);
    const NormOut out{reinterpret_cast<NormPartial*>(base + off_part), reinterpret_cast<double*>(base + off_g2s), g2,
                      stderr_};
    // no timer events here (they would be created on demand)
    const bool was = p->capturing;
    p->capturing = true;
    launch_normalize_sums(p, s, 0, h.T, h.T, p->d_taus, sums, sums + PT, sums + 2 * PT, true, false, nullptr, out);
    p->capturing = was;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(XPCS_ERR_HIP, "normalize_ref_f32_noalloc: %s", hipGetErrorString(e));
    return XPCS_OK;
}
