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
