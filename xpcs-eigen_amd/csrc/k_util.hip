// Fill kernels: clear counters and accumulators inside a job's launch sequence
// (graph-capturable, unlike memset nodes).
#include "internal.h"

namespace xpcs {

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T* __restrict__ d, size_t n, T v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = v;
}

static unsigned fill_grid(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

void launch_fill_u32(uint32_t* d, size_t n, hipStream_t s)
{
    if (n) k_fill<uint32_t><<<fill_grid(n), 256, 0, s>>>(d, n, 0u);
}

void launch_fill_u64(uint64_t* d, size_t n, hipStream_t s)
{
    if (n) k_fill<uint64_t><<<fill_grid(n), 256, 0, s>>>(d, n, 0ull);
}

}  // namespace xpcs
