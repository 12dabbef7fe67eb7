// (a) Frame -> pixel transpose kernels.
//
// Sparse IMM (SparseFilter::Apply, src/xpcs/filter/sparse_filter.cpp:115-193):
// a counting sort of the frame-major events by compact pixel id.
//   k_hist     one workgroup per selected raw frame; masked, non-zero events
//              bump their pixel's counter (the mask lookup drops the rest).
//   k_scan_*   exclusive scan of the counters -> 64-bit row offsets.
//   k_scatter  same traversal; each event claims a slot of its row with an
//              atomic decrement of the counter (which leaves the counters at 0
//              for the next job) and stores (frame << 32 | count).
// Rows come out in arbitrary slot order; the correlator sorts each row on chip.
//
// Dense IMM (DenseFilter::Apply, src/xpcs/filter/dense_filter.cpp:121-210):
// an LDS-tiled transpose of whole frames into pixel-major rows [P][F] of
// n-scaled integer values (dark subtraction and threshold fused).
#include "device_util.h"
#include "internal.h"

namespace xpcs {

__global__ void __launch_bounds__(256)
k_hist(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off,
       const uint32_t* __restrict__ sel_dlen, int64_t nsel, const int32_t* __restrict__ pix2c,
       uint32_t npix_det, uint32_t* __restrict__ cnt, uint32_t* __restrict__ counters)
{
    for (int64_t k = blockIdx.x; k < nsel; k += gridDim.x) {
        const uint8_t* base = imm + sel_off[k];
        const uint32_t dlen = sel_dlen[k];
        const uint16_t* vals = reinterpret_cast<const uint16_t*>(base + 4 * (size_t)dlen);
        const bool al4 = (reinterpret_cast<uintptr_t>(base) & 3u) == 0;
        for (uint32_t j = threadIdx.x; j < dlen; j += blockDim.x) {
            uint32_t idx = al4 ? reinterpret_cast<const uint32_t*>(base)[j] : load_u32_a2(base + 4 * (size_t)j);
            uint16_t v = vals[j];
            if (idx >= npix_det) {
                atomicOr(&counters[1], 1u);
                continue;
            }
            int32_t c = pix2c[idx];
            if (c < 0 || v == 0) continue;
            atomicAdd(&cnt[c], 1u);
        }
    }
}

__global__ void __launch_bounds__(256)
k_scatter(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off,
          const uint32_t* __restrict__ sel_dlen, const int32_t* __restrict__ sel_frame, int64_t nsel,
          const int32_t* __restrict__ pix2c, uint32_t npix_det, uint32_t* __restrict__ cnt,
          const int64_t* __restrict__ off, uint64_t* __restrict__ ev)
{
    for (int64_t k = blockIdx.x; k < nsel; k += gridDim.x) {
        const uint8_t* base = imm + sel_off[k];
        const uint32_t dlen = sel_dlen[k];
        const uint64_t f = (uint64_t)(uint32_t)sel_frame[k] << 32;
        const uint16_t* vals = reinterpret_cast<const uint16_t*>(base + 4 * (size_t)dlen);
        const bool al4 = (reinterpret_cast<uintptr_t>(base) & 3u) == 0;
        for (uint32_t j = threadIdx.x; j < dlen; j += blockDim.x) {
            uint32_t idx = al4 ? reinterpret_cast<const uint32_t*>(base)[j] : load_u32_a2(base + 4 * (size_t)j);
            uint16_t v = vals[j];
            if (idx >= npix_det) continue;
            int32_t c = pix2c[idx];
            if (c < 0 || v == 0) continue;
            uint32_t r = atomicSub(&cnt[c], 1u);
            ev[off[c] + (int64_t)r - 1] = f | (uint64_t)v;
        }
    }
}

// ---- exclusive scan of the row counters ------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ int64_t block_excl_scan(int64_t x, int64_t* s_warp, int64_t& total)
{
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    int64_t incl = wave_incl_scan_i64(x);
    if (lane == 63) s_warp[w] = incl;
    __syncthreads();
    if (w == 0) {
        int64_t v = lane < nw ? s_warp[lane] : 0;
        int64_t vi = wave_incl_scan_i64(v);
        if (lane < nw) s_warp[lane] = vi - v;
        if (lane == nw - 1) s_warp[32] = vi;
    }
    __syncthreads();
    total = s_warp[32];
    int64_t r = s_warp[w] + incl - x;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_reduce(const uint32_t* __restrict__ cnt, int64_t n, int64_t* __restrict__ bsum)
{
    __shared__ int64_t s[33];
    int64_t b0 = (int64_t)blockIdx.x * kScanTile;
    int64_t acc = 0;
    for (int i = 0; i < kScanItems; i++) {
        int64_t k = b0 + (int64_t)i * kScanBlock + threadIdx.x;
        if (k < n) acc += cnt[k];
    }
    int64_t tot;
    block_excl_scan(acc, s, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024)
k_scan_bsum(int64_t* __restrict__ bsum, int64_t nb)
{
    __shared__ int64_t s[33];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += blockDim.x) {
        int64_t k = b0 + threadIdx.x;
        int64_t v = k < nb ? bsum[k] : 0;
        int64_t tot;
        int64_t ex = block_excl_scan(v, s, tot);
        if (k < nb) bsum[k] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_tile(const uint32_t* __restrict__ cnt, int64_t n, const int64_t* __restrict__ bsum,
            int64_t* __restrict__ off)
{
    __shared__ int64_t s[33];
    int64_t b0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    int64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        int64_t k = b0 + i;
        v[i] = k < n ? cnt[k] : 0u;
        acc += v[i];
    }
    int64_t tot;
    int64_t ex = block_excl_scan(acc, s, tot) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        int64_t k = b0 + i;
        if (k < n) off[k] = ex;
        ex += v[i];
    }
}

__global__ void k_scan_total(const uint32_t* __restrict__ cnt, int64_t n, int64_t* __restrict__ off)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) off[n] = (n > 0 ? off[n - 1] + cnt[n - 1] : 0);
}

// ---- dense frames -> pixel-major rows ---------------------------------------------
constexpr int kDenseTF = 32;    // output frames per tile
constexpr int kDenseTP = 256;   // detector pixels per tile (one per thread)

__global__ void __launch_bounds__(kDenseTP)
k_dense_transpose(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off, int avg,
                  int32_t F, const int32_t* __restrict__ pix2c, uint32_t npix_det,
                  const int64_t* __restrict__ dark_sum, const int64_t* __restrict__ thr,
                  int64_t nscale, uint32_t* __restrict__ rows)
{
    __shared__ uint32_t tile[kDenseTF][kDenseTP + 1];
    __shared__ int32_t s_c[kDenseTP];
    const int64_t ntp = ((int64_t)npix_det + kDenseTP - 1) / kDenseTP;
    const int64_t ntf = ((int64_t)F + kDenseTF - 1) / kDenseTF;
    for (int64_t t = blockIdx.x; t < ntp * ntf; t += gridDim.x) {
        const int64_t tp = t % ntp;
        const int64_t tf = t / ntp;
        const uint32_t pix = (uint32_t)(tp * kDenseTP + threadIdx.x);
        const int f0 = (int)(tf * kDenseTF);
        int32_t c = pix < npix_det ? pix2c[pix] : -1;
        int64_t ds = 0, th = 0;
        if (c >= 0) { ds = dark_sum[c]; th = thr[c]; }
        s_c[threadIdx.x] = c;
        for (int i = 0; i < kDenseTF; i++) {
            int f = f0 + i;
            uint32_t acc = 0;
            if (c >= 0 && f < F) {
                for (int j = 0; j < avg; j++) {
                    const uint16_t* fr = reinterpret_cast<const uint16_t*>(imm + sel_off[(int64_t)f * avg + j]);
                    int64_t y = nscale * (int64_t)fr[pix] - ds;
                    if (y > th && y > 0) acc += (uint32_t)y;
                }
            }
            tile[i][threadIdx.x] = acc;
        }
        __syncthreads();
        // write: 8 threads per pixel, 4 frames each
        for (int k = threadIdx.x; k < kDenseTP * (kDenseTF / 4); k += kDenseTP) {
            int pl = k / (kDenseTF / 4);
            int q = k % (kDenseTF / 4);
            int32_t cc = s_c[pl];
            int f = f0 + q * 4;
            if (cc < 0) continue;
            uint32_t* dst = rows + (int64_t)cc * F + f;
            if (f + 4 <= F && ((reinterpret_cast<uintptr_t>(dst) & 15u) == 0)) {
                uint4 v = make_uint4(tile[q * 4 + 0][pl], tile[q * 4 + 1][pl], tile[q * 4 + 2][pl],
                                     tile[q * 4 + 3][pl]);
                *reinterpret_cast<uint4*>(dst) = v;
            } else {
                for (int u = 0; u < 4 && f + u < F; u++) dst[u] = tile[q * 4 + u][pl];
            }
        }
        __syncthreads();
    }
}

// ---- launchers ------------------------------------------------------------------
static int grid_for(int64_t n, int64_t cap)
{
    if (n <= 0) return 1;
    return (int)(n < cap ? n : cap);
}

void launch_hist(xpcs_mt_plan* p, hipStream_t s)
{
    timer_begin(p, s, "k_hist");
    k_hist<<<grid_for(p->nsel, 1 << 20), 256, 0, s>>>(p->d_imm, p->d_sel_off, p->d_sel_dlen, p->nsel,
                                                    p->d_pix2c, (uint32_t)((int64_t)p->h.W * p->h.H),
                                                    p->d_cnt, p->d_counters);
    timer_end(p, s);
}

void launch_scan(xpcs_mt_plan* p, hipStream_t s)
{
    const int64_t n = p->h.P;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    timer_begin(p, s, "k_scan");
    k_scan_reduce<<<(int)nb, kScanBlock, 0, s>>>(p->d_cnt, n, p->d_bsum);
    k_scan_bsum<<<1, 1024, 0, s>>>(p->d_bsum, nb);
    k_scan_tile<<<(int)nb, kScanBlock, 0, s>>>(p->d_cnt, n, p->d_bsum, p->d_off);
    k_scan_total<<<1, 64, 0, s>>>(p->d_cnt, n, p->d_off);
    timer_end(p, s);
}

void launch_scatter(xpcs_mt_plan* p, hipStream_t s)
{
    timer_begin(p, s, "k_scatter");
    k_scatter<<<grid_for(p->nsel, 1 << 20), 256, 0, s>>>(p->d_imm, p->d_sel_off, p->d_sel_dlen,
                                                       p->d_sel_frame, p->nsel, p->d_pix2c,
                                                       (uint32_t)((int64_t)p->h.W * p->h.H), p->d_cnt,
                                                       p->d_off, p->d_ev);
    timer_end(p, s);
}

void launch_dense_transpose(xpcs_mt_plan* p, hipStream_t s)
{
    const int64_t npd = (int64_t)p->h.W * p->h.H;
    const int64_t ntiles = ((npd + kDenseTP - 1) / kDenseTP) * ((p->h.F + kDenseTF - 1) / kDenseTF);
    timer_begin(p, s, "k_dense_transpose");
    k_dense_transpose<<<grid_for(ntiles, 1 << 16), kDenseTP, 0, s>>>(
        p->d_imm, p->d_sel_off, p->h.avg, p->h.F, p->d_pix2c, (uint32_t)npd, p->d_dark_sum, p->d_thr,
        p->nscale, p->d_rows);
    timer_end(p, s);
}

}  // namespace xpcs
