// Wavefront helpers for CDNA (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xpcs {

// Orders LDS traffic of the lanes of one wavefront: every LDS write issued
// before the call is visible to every lane's LDS read after it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

// Wavefront index within the workgroup, as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int mask_rank(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    return x;
}

__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t x)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int64_t t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// IP and IF per (pixel, tau), idx = p T + tau: int64 (or FP64 bit patterns),
// or (w32, integer sums of the batch path) the low 32 bits in 32-bit words
// [0, P T) and, for pixels whose total passes 2^32 - 1, the high 32 bits at
// hi + idx (hi = P T; 0: the pixel needs none).  The normaliser then reads 16
// instead of 24 bytes per pixel and tau, and a pixel's total says whether its
// high words exist: IP and IF never exceed it.
__device__ __forceinline__ void put_sum(int64_t* base, int64_t idx, int64_t v, bool w32, int64_t hi = 0)
{
    if (w32) {
        uint32_t* b = reinterpret_cast<uint32_t*>(base);
        b[idx] = (uint32_t)v;
        if (hi) b[hi + idx] = (uint32_t)((uint64_t)v >> 32);
    } else {
        base[idx] = v;
    }
}
__device__ __forceinline__ int64_t get_sum(const int64_t* base, int64_t idx, bool w32, int64_t hi = 0)
{
    if (!w32) return base[idx];
    const uint32_t* b = reinterpret_cast<const uint32_t*>(base);
    const uint64_t lo = b[idx];
    return (int64_t)(hi ? lo | ((uint64_t)b[hi + idx] << 32) : lo);
}
// the offset of a pixel's high words: hw = P T when the layout is w32
__device__ __forceinline__ int64_t sum_hi(uint64_t total, int64_t hw) { return total > 0xFFFFFFFFull ? hw : 0; }

__device__ __forceinline__ int next_pow2(int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

// First index in key[0, n) whose value is >= b.
__device__ __forceinline__ int lds_lower_bound(const uint32_t* key, int n, uint32_t b)
{
    int lo = 0, len = n;
    while (len > 0) {
        int half = len >> 1;
        if (key[lo + half] < b) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// 32-bit load from a 2-byte aligned address (IMM index arrays follow
// payloads of 6*dlen bytes, so they are only 2-byte aligned in general).
__device__ __forceinline__ uint32_t load_u32_a2(const uint8_t* p)
{
    const uint16_t* q = reinterpret_cast<const uint16_t*>(p);
    return (uint32_t)q[0] | ((uint32_t)q[1] << 16);
}

}  // namespace xpcs

namespace xpcs {

// ---- DPP (gfx9 data-parallel primitives; no LDS traffic) --------------------------
enum : int {
    kDppRowShl = 0x100,    // + n: lane i reads lane i+n of its row of 16
    kDppRowShr = 0x110,    // + n: lane i reads lane i-n of its row of 16
    kDppWaveShl1 = 0x130,  // lane i reads lane i+1 (lane 63 keeps `old`)
    kDppWaveShr1 = 0x138,  // lane i reads lane i-1 (lane 0 keeps `old`)
    kDppBcast15 = 0x142,   // rows 1..3 read lane 15 of the previous row
    kDppBcast31 = 0x143,   // rows 2..3 read lane 31
    kDppQuadXor1 = 0xB1,   // quad_perm [1,0,3,2]
    kDppQuadXor2 = 0x4E,   // quad_perm [2,3,0,1]
};

template <int CTRL, int ROW_MASK = 0xF, bool BOUND_ZERO = false>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, ROW_MASK, 0xF, BOUND_ZERO);
}

// Inclusive prefix sum over the 64 lanes (u32).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x)
{
    x += dpp_u32<kDppRowShr + 1, 0xF, true>(0u, x);
    x += dpp_u32<kDppRowShr + 2, 0xF, true>(0u, x);
    x += dpp_u32<kDppRowShr + 4, 0xF, true>(0u, x);
    x += dpp_u32<kDppRowShr + 8, 0xF, true>(0u, x);
    x += dpp_u32<kDppBcast15, 0xA>(0u, x);
    x += dpp_u32<kDppBcast31, 0xC>(0u, x);
    return x;
}

// Inclusive running maximum over the 64 lanes (u32).
__device__ __forceinline__ uint32_t wave_incl_max_u32(uint32_t x)
{
    x = max(x, dpp_u32<kDppRowShr + 1, 0xF, true>(0u, x));
    x = max(x, dpp_u32<kDppRowShr + 2, 0xF, true>(0u, x));
    x = max(x, dpp_u32<kDppRowShr + 4, 0xF, true>(0u, x));
    x = max(x, dpp_u32<kDppRowShr + 8, 0xF, true>(0u, x));
    x = max(x, dpp_u32<kDppBcast15, 0xA>(0u, x));
    x = max(x, dpp_u32<kDppBcast31, 0xC>(0u, x));
    return x;
}

// Value of lane i+1 (lane 63 gets `old`).
__device__ __forceinline__ uint32_t next_lane_u32(uint32_t x, uint32_t old)
{
    return dpp_u32<kDppWaveShl1>(old, x);
}

// Value of lane i-1 (lane 0 gets `old`).
__device__ __forceinline__ uint32_t prev_lane_u32(uint32_t x, uint32_t old)
{
    return dpp_u32<kDppWaveShr1>(old, x);
}

// Value of lane (i ^ J) for J in {1, 2, 4, 8} without LDS, else ds_bpermute.
template <int J>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t x)
{
    if constexpr (J == 1) {
        return dpp_u32<kDppQuadXor1>(0u, x);
    } else if constexpr (J == 2) {
        return dpp_u32<kDppQuadXor2>(0u, x);
    } else if constexpr (J == 4 || J == 8) {
        const uint32_t up = dpp_u32<kDppRowShl + J>(0u, x);    // from lane i+J
        const uint32_t dn = dpp_u32<kDppRowShr + J>(0u, x);    // from lane i-J
        return (lane_id() & J) ? dn : up;
    } else {
        return (uint32_t)__shfl_xor((int)x, J, 64);
    }
}

// Lane l receives `x` from the lane whose dst == l (lanes that nobody targets
// keep garbage).
__device__ __forceinline__ uint32_t push_to_lane(int dst, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_ds_permute(dst << 2, (int)x);
}

__device__ __forceinline__ uint32_t read_lane_u32(uint32_t x, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)x, l);
}

// DPP move of an FP64 value (two 32-bit halves); lanes the pattern does not
// write (row mask) or that read outside their row (BOUND_ZERO) get 0.0.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ double dpp_f64(double x)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK, BOUND_ZERO>(0u, (uint32_t)u);
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK, BOUND_ZERO>(0u, (uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Segmented inclusive FP64 sum over the 64 lanes: lane i gets the sum of x over
// lanes start..i of its segment, `start` being the segment's first lane (as
// from a running maximum of segment heads). Row-shift steps 1, 2, 4, 8 within
// each 16-lane row, then the row-15 and row-31 broadcasts; a lane takes a
// partner's partial sum only while that partner lies in its segment.
__device__ __forceinline__ double wave_seg_scan_f64(double x, uint32_t start)
{
    const uint32_t lane = (uint32_t)lane_id();
    double u;
    u = dpp_f64<kDppRowShr + 1, 0xF, true>(x);
    x += lane >= start + 1 ? u : 0.0;
    u = dpp_f64<kDppRowShr + 2, 0xF, true>(x);
    x += lane >= start + 2 ? u : 0.0;
    u = dpp_f64<kDppRowShr + 4, 0xF, true>(x);
    x += lane >= start + 4 ? u : 0.0;
    u = dpp_f64<kDppRowShr + 8, 0xF, true>(x);
    x += lane >= start + 8 ? u : 0.0;
    u = dpp_f64<kDppBcast15, 0xA, false>(x);   // rows 1, 3 from lanes 15, 47
    x += start < (lane & 0x30u) ? u : 0.0;
    u = dpp_f64<kDppBcast31, 0xC, false>(x);   // rows 2, 3 from lane 31
    x += start < 32u ? u : 0.0;
    return x;
}

// Segmented inclusive u32 sum over the 64 lanes, as wave_seg_scan_f64 (the
// caller guarantees that a segment's sum fits 32 bits).
__device__ __forceinline__ uint32_t wave_seg_scan_u32(uint32_t x, uint32_t start)
{
    const uint32_t lane = (uint32_t)lane_id();
    uint32_t u;
    u = dpp_u32<kDppRowShr + 1, 0xF, true>(0u, x);
    x += lane >= start + 1 ? u : 0u;
    u = dpp_u32<kDppRowShr + 2, 0xF, true>(0u, x);
    x += lane >= start + 2 ? u : 0u;
    u = dpp_u32<kDppRowShr + 4, 0xF, true>(0u, x);
    x += lane >= start + 4 ? u : 0u;
    u = dpp_u32<kDppRowShr + 8, 0xF, true>(0u, x);
    x += lane >= start + 8 ? u : 0u;
    u = dpp_u32<kDppBcast15, 0xA>(0u, x);
    x += start < (lane & 0x30u) ? u : 0u;
    u = dpp_u32<kDppBcast31, 0xC>(0u, x);
    x += start < 32u ? u : 0u;
    return x;
}

// Runs of equal `key` in neighbouring lanes: returns the first lane of this
// lane's run; `tail` marks the run's last lane. All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_run_start(uint32_t key, bool& tail)
{
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t kp = prev_lane_u32(key, ~key);
    const uint32_t kn = next_lane_u32(key, ~key);
    tail = kn != key;
    return wave_incl_max_u32(kp != key ? lane : 0u);
}

// Sum of x over each run of equal `key` in neighbouring lanes; `tail` marks the
// run's last lane, which holds the run's total. All 64 lanes must be active.
__device__ __forceinline__ double wave_run_sum_f64(double x, uint32_t key, bool& tail)
{
    return wave_seg_scan_f64(x, wave_run_start(key, tail));
}

// Lane l receives x of lane src (ds_bpermute).
__device__ __forceinline__ uint32_t pull_from_lane(int src, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)x);
}

}  // namespace xpcs
