// Online multi-tau for streams of sparse frames (no reference counterpart: the
// reference correlates a finished file, Corr::multiTau2, src/xpcs/corr.cpp:315-431).
//
// Per stream and masked pixel the state lives in HBM, level-major with the
// pixel index fastest (every access coalesced):
//   win [L][D2][P] u32  the last D2 = 2*dpl values of level L (j = 0 newest)
//   head[L][D2][P] u32  the first D2 values of level L
//   acc [L][D2][P] u64  lag sums  sum_t x_L[t-d] x_L[t]  for d = 1..D2
//   pend[L][P]     u32  the last level-L value while its pair is incomplete
//   total[P]       u64  sum of all level-0 values
// After n frames, level L holds lf_L = floor(n / 2^L) values, and for a lag
// tau' of level L
//   G2 = acc[L][tau'],  IP = T_L - sum_{j<tau'} win[L][j],
//   IF = T_L - sum_{k<tau'} head[L][k],  T_L = sum_k x_L[k]
//      = total - sum_{l<L, lf_l odd} pend[l],
// which are exactly the batch sums for F = n: a lag product is added when its
// later value arrives, and the schedule for n frames is a prefix of the
// schedule for the stream's maximum length.
#include "device_util.h"
#include "internal.h"

namespace xpcs {

// Frames of a push: events into a dense [nb][P] value tile (summed per pixel).
__global__ void __launch_bounds__(256)
k_stream_scatter(const uint8_t* __restrict__ imm, const StreamFrames fr, int32_t nb,
                 const int32_t* __restrict__ pix2c, uint32_t npix_det, int32_t P, uint32_t* __restrict__ tile,
                 uint32_t* __restrict__ counters)
{
    for (int f = blockIdx.x; f < nb; f += gridDim.x) {
        const uint8_t* base = imm + fr.off[f];
        const uint32_t n = fr.dlen[f];
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
            const uint32_t idx = load_u32_a2(base + 4 * (size_t)j);
            const uint32_t v = reinterpret_cast<const uint16_t*>(base + 4 * (size_t)n)[j];
            if (idx >= npix_det) {
                atomicOr(&counters[1], 1u);
                continue;
            }
            const int32_t c = pix2c[idx];
            if (c >= 0 && v) atomicAdd(&tile[(int64_t)f * P + c], v);
        }
    }
}

// One chunk of nb frames through every level: thread per pixel, the level
// lists in LDS ([value][thread]), the window and lag sums of one level at a
// time in registers.
template <int D2>
__global__ void __launch_bounds__(256)
k_stream_update(uint32_t* __restrict__ tile, int32_t nb, int32_t P, int32_t n0, int32_t nlev,
                uint32_t* __restrict__ win, uint32_t* __restrict__ head, uint64_t* __restrict__ acc,
                uint32_t* __restrict__ pend, uint64_t* __restrict__ total)
{
    __shared__ uint32_t s_list[kStreamChunk][256];
    const int tid = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + tid;
    const bool active = c < P;
    uint64_t add = 0;
    for (int i = 0; i < nb; i++) {
        const uint32_t v = active ? tile[(int64_t)i * P + c] : 0u;
        // read entries go back to zero, ready for the next push's scatter
        if (active && v) tile[(int64_t)i * P + c] = 0u;
        s_list[i][tid] = v;
        add += v;
    }
    if (active) total[c] += add;
    uint32_t cnt_old = (uint32_t)n0;   // values of the level before this chunk
    int m = nb;                        // values entering the level in this chunk
    for (int L = 0; L < nlev && m > 0; L++) {
        const int64_t lb = (int64_t)L * D2 * P;
        uint32_t w[D2];
        uint64_t a[D2];
        uint32_t p = 0;
#pragma unroll
        for (int j = 0; j < D2; j++) {
            w[j] = active ? win[lb + (int64_t)j * P + c] : 0u;
            a[j] = active ? acc[lb + (int64_t)j * P + c] : 0ull;
        }
        if (active) p = pend[(int64_t)L * P + c];
        int m_next = 0;
        for (int i = 0; i < m; i++) {
            const uint32_t v = s_list[i][tid];
            const uint32_t k = cnt_old + (uint32_t)i;
#pragma unroll
            for (int d = 0; d < D2; d++) a[d] += (uint64_t)v * (uint64_t)w[d];
#pragma unroll
            for (int j = D2 - 1; j > 0; j--) w[j] = w[j - 1];
            w[0] = v;
            if (k < (uint32_t)D2 && active) head[lb + (int64_t)k * P + c] = v;
            if (k & 1u) {
                s_list[m_next][tid] = p + v;   // m_next <= i: this slot was read already
                m_next++;
            } else {
                p = v;
            }
        }
        if (active) {
#pragma unroll
            for (int j = 0; j < D2; j++) {
                win[lb + (int64_t)j * P + c] = w[j];
                acc[lb + (int64_t)j * P + c] = a[j];
            }
            pend[(int64_t)L * P + c] = p;
        }
        cnt_old >>= 1;
        m = m_next;
    }
}

// Sums for the schedule of n frames, taus [t0, t1) of it, into rows of stride
// S ([P][S]), and those taus' table entries (lf, N for n frames) that the
// normalisation reads.  Only the levels holding those taus are read.
template <int D2>
__global__ void __launch_bounds__(256)
k_stream_snapshot(const uint32_t* __restrict__ win, const uint32_t* __restrict__ head,
                  const uint64_t* __restrict__ acc, const uint32_t* __restrict__ pend,
                  const uint64_t* __restrict__ total, int32_t P, int32_t n, const LevelParam* __restrict__ levels,
                  int32_t nlev, const TauParam* __restrict__ taus_max, int32_t t0, int32_t t1, int32_t S,
                  TauParam* __restrict__ taus, int64_t* __restrict__ G2, int64_t* __restrict__ IP,
                  int64_t* __restrict__ IF)
{
    if (blockIdx.x == 0) {
        for (int i = t0 + (int)threadIdx.x; i < t1; i += blockDim.x) {
            TauParam t = taus_max[i];
            t.lf = n >> t.level;
            t.N = t.lf - t.tp;
            taus[i] = t;
        }
    }
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    uint64_t TL = total[c];
    for (int L = 0; L < nlev; L++) {
        if (L > 0 && ((n >> (L - 1)) & 1)) TL -= pend[(int64_t)(L - 1) * P + c];
        const LevelParam lp = levels[L];
        if (lp.first >= t1) break;
        const int a = max(lp.first, t0), b = min(lp.first + lp.count, t1);   // this level's taus in range
        if (a >= b) continue;
        const int64_t lb = (int64_t)L * D2 * P;
        uint64_t sw = 0, sh = 0;   // sums of the newest / oldest d values
#pragma unroll
        for (int d = 1; d <= D2; d++) {
            sw += win[lb + (int64_t)(d - 1) * P + c];
            sh += head[lb + (int64_t)(d - 1) * P + c];
            const int i = lp.first + d - lp.lo;
            if (i >= a && i < b) {
                const int64_t o = c * S + i;
                G2[o] = (int64_t)acc[lb + (int64_t)(d - 1) * P + c];
                IP[o] = (int64_t)(TL - sw);
                IF[o] = (int64_t)(TL - sh);
            }
        }
    }
}

void launch_stream_update(const StreamLaunch& a, hipStream_t s)
{
    const unsigned gp = (unsigned)((a.P + 255) / 256);
    // the tile is zero here: cleared when the stream opens, and every entry a
    // scatter sets is cleared again by the update that reads it
    k_stream_scatter<<<a.nb, 256, 0, s>>>(a.imm, a.fr, a.nb, a.pix2c, a.npix_det, a.P, a.tile, a.counters);
    if (a.d2 == 8)
        k_stream_update<8><<<gp, 256, 0, s>>>(a.tile, a.nb, a.P, a.n0, a.nlev, a.win, a.head, a.acc, a.pend,
                                              a.total);
    else
        k_stream_update<16><<<gp, 256, 0, s>>>(a.tile, a.nb, a.P, a.n0, a.nlev, a.win, a.head, a.acc, a.pend,
                                               a.total);
}

void launch_stream_snapshot(const StreamLaunch& a, int32_t n, const LevelParam* levels, int32_t nlev,
                            const TauParam* taus_max, int32_t t0, int32_t t1, int32_t S, TauParam* taus,
                            int64_t* G2, int64_t* IP, int64_t* IF, hipStream_t s)
{
    if (t1 <= t0) return;
    const unsigned gp = (unsigned)((a.P + 255) / 256) + 1;   // block 0 also writes the tau table
    if (a.d2 == 8)
        k_stream_snapshot<8><<<gp, 256, 0, s>>>(a.win, a.head, a.acc, a.pend, a.total, a.P, n, levels, nlev,
                                                taus_max, t0, t1, S, taus, G2, IP, IF);
    else
        k_stream_snapshot<16><<<gp, 256, 0, s>>>(a.win, a.head, a.acc, a.pend, a.total, a.P, n, levels, nlev,
                                                 taus_max, t0, t1, S, taus, G2, IP, IF);
}

}  // namespace xpcs
