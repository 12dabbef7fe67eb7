// Two-time correlation (Corr::twotime, src/xpcs/corr.cpp:562-924): per listed
// q-bin, C2T[f0][f1] = sum_p x[p,f0] x[p,f1] / P_q for f0 <= f1, the Gram
// matrix of the q-bin's pixel x frame values -- the one dense contraction of
// the reference, computed here with FP64 matrix-core instructions
// (v_mfma_f64_16x16x4f64): integer counts give exact sums (every partial sum
// an integer below 2^53), smoothed values are FP64 throughout.
//
// Two ways to form the sums, chosen per q-bin by their cost:
//   event pairs  the reference's own loop (every pixel's event pairs, corr.cpp
//                :816-835), Sum_p n_p^2 / 2 products scattered into C with
//                FP64 atomics -- far less work for sparse (photon-counting) data;
//   Gram         P_q x F^2 dense FP64 products on the matrix cores -- for
//                dense data, where nearly every (pixel, frame) holds a value.
//
//   k_tt_pairs     one wave per pixel, all event pairs -> C (unscaled)
//   k_tt_sg_sparse per (bin, frame) sums of the values from the events
//   k_tt_values    dense X[i][f] (P_q x F, FP64, pixel-major) from the rows
//   k_tt_colsum    per (bin, frame) sums of X over a bin's pixels (smoothing)
//   k_tt_divide    X[i][f] /= sg (non-zero entries only, as the reference
//                  divides the events it holds)
//   k_tt_gram      upper-triangle 64x64 tiles of X^T X / P_q; each wave owns a
//                  32x32 quarter (2x2 MFMA blocks), pixels staged through LDS
//   k_tt_diag_*    g2full from segmented diagonal sums (scaling the event-pair
//                  sums by 1 / P_q in the same pass), k_tt_partials g2partials
#include "device_util.h"
#include "internal.h"

namespace xpcs {

typedef double v4d __attribute__((ext_vector_type(4)));

// One row of X per thread block: row i is compact pixel c0 + i.
__global__ void __launch_bounds__(256)
k_tt_values(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off, const int32_t* __restrict__ rowmap,
            const uint32_t* __restrict__ dense_rows, int32_t c0, int32_t n, int32_t F,
            const double* __restrict__ weight, double scale, const double* __restrict__ fw, double* __restrict__ X)
{
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int64_t c = c0 + i;
        const double w = (weight ? weight[c] : 1.0) * scale;
        double* xr = X + i * (int64_t)F;
        if (dense_rows) {
            const uint32_t* r = dense_rows + c * (int64_t)F;
            for (int f = threadIdx.x; f < F; f += blockDim.x) xr[f] = (double)r[f] * w * (fw ? fw[f] : 1.0);
        } else {
            const int64_t row = rowmap ? (int64_t)rowmap[c] : c;
            const int64_t b = off[row], e = off[row + 1];
            for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
                const uint64_t v = ev[k];
                const uint32_t f = (uint32_t)(v >> 32);
                // averaged raw frames share an output frame: several events may add here
                if (f < (uint32_t)F) atomicAdd(&xr[f], (double)(uint32_t)v * w * (fw ? fw[f] : 1.0));
            }
        }
    }
}

// sgsum[b][f] += X over the bin's pixels that lie in the chunk [lo, hi) (rows
// of X are chunk-relative); blockIdx.z takes a segment of kColSeg rows, whose
// partial sum is added atomically.
constexpr int kColSeg = 64;
__global__ void __launch_bounds__(256)
k_tt_colsum(const double* __restrict__ X, int32_t F, int32_t lo, int32_t hi, const int32_t* __restrict__ bin_begin,
            const int32_t* __restrict__ bin_end, int32_t nbins, double* __restrict__ sgsum)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int64_t r0 = lo + (int64_t)blockIdx.z * kColSeg; r0 < hi; r0 += (int64_t)gridDim.z * kColSeg) {
        const int r1 = (int)min<int64_t>(r0 + kColSeg, hi);
        for (int b = blockIdx.y; b < nbins; b += gridDim.y) {
            const int i0 = max(bin_begin[b], (int)r0), i1 = min(bin_end[b], r1);
            if (i0 >= i1) continue;
            double s = 0.0;
            for (int i = i0; i < i1; i++) s += X[(int64_t)(i - lo) * F + f];
            unsafeAtomicAdd(&sgsum[(int64_t)b * F + f], s);
        }
    }
}

// sg[b][f] = sgsum[b][f] / pixels of bin b (in place)
__global__ void __launch_bounds__(256)
k_tt_sgfinal(double* __restrict__ sg, int32_t F, const int32_t* __restrict__ bin_begin,
             const int32_t* __restrict__ bin_end, int32_t nbins)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int b = blockIdx.y; b < nbins; b += gridDim.y) sg[(int64_t)b * F + f] /= (double)(bin_end[b] - bin_begin[b]);
}

// Mean over frames of each bin's sg row (smoothing filter "Average").
__global__ void __launch_bounds__(256)
k_tt_sgavg(const double* __restrict__ sg, int32_t F, int32_t nbins, double* __restrict__ sga)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    double s = 0.0;
    for (int f = 0; f < F; f++) s += sg[(int64_t)b * F + f];
    sga[b] = s / (double)F;
}

// X[i][f] /= sg of the pixel's bin (chunk row i is q-bin pixel lo + i); only
// non-zero values are divided, as the reference divides the events it holds.
__global__ void __launch_bounds__(256)
k_tt_divide(double* __restrict__ X, int32_t F, int32_t lo, int32_t n, const int32_t* __restrict__ bin_of_row,
            const double* __restrict__ sg, int32_t per_frame)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int64_t i = blockIdx.y; i < n; i += gridDim.y) {
        const int b = bin_of_row[lo + i];
        const double d = per_frame ? sg[(int64_t)b * F + f] : sg[b];
        double& x = X[i * F + f];
        if (x != 0.0) x = x / d;
    }
}

// ---- event pairs ------------------------------------------------------------------
__device__ __forceinline__ double read_lane_f64(double x, int l)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Event k of the row [b, b + n): output frame and value (count x flat field x
// scale x frame weight when `fw` is set, divided by the smoothing image when
// `dv` is set).  Lanes past the row get value 0.
__device__ __forceinline__ void tt_event(const uint64_t* __restrict__ ev, int64_t b, int64_t n, int64_t k,
                                         int32_t F, double w, const double* __restrict__ fw,
                                         const double* __restrict__ dv, int32_t per_frame, uint32_t& f, double& v)
{
    f = 0u;
    v = 0.0;
    if (k < n) {
        const uint64_t e = ev[b + k];
        const uint32_t fe = (uint32_t)(e >> 32);
        if (fe < (uint32_t)F) {
            f = fe;
            v = (double)(uint32_t)e * w;
            if (fw) v *= fw[fe];
            if (dv) v = v / (per_frame ? dv[fe] : dv[0]);
        }
    }
}

// C[f_a][f_b] += v_a v_b over the unordered event pairs {a, b} of every pixel
// (f_a <= f_b; a == b gives v_a^2 on the diagonal).  Two raw events of one
// output frame (averaged frames) are one merged value in the reference, whose
// square holds their cross term twice: such a pair adds 2 v_a v_b.  One wave
// per pixel; events in blocks of 64, the first of each pair broadcast from a
// lane, the second held by every lane.
__global__ void __launch_bounds__(256)
k_tt_pairs(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off, const int32_t* __restrict__ rowmap,
           int32_t c0, int32_t Pq, int32_t F, const double* __restrict__ weight, double scale,
           const int32_t* __restrict__ bin_of_row, const double* __restrict__ div, int32_t per_frame,
           const double* __restrict__ fw, double* __restrict__ C)
{
    const int lane = lane_id();
    const int wpb = (int)(blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * wpb + wave_id(); i < Pq; i += (int64_t)gridDim.x * wpb) {
        const int64_t c = c0 + i;
        const int64_t row = rowmap ? (int64_t)rowmap[c] : c;
        const int64_t b = off[row];
        const int64_t n = off[row + 1] - b;
        if (n <= 0) continue;
        const double w = (weight ? weight[c] : 1.0) * scale;
        const double* dv = nullptr;
        if (div) dv = per_frame ? div + (int64_t)bin_of_row[i] * F : div + bin_of_row[i];
        for (int64_t I = 0; I < n; I += 64) {
            uint32_t fa;
            double va;
            tt_event(ev, b, n, I + lane, F, w, fw, dv, per_frame, fa, va);
            const int na = (int)min<int64_t>(64, n - I);
            for (int64_t J = I; J < n; J += 64) {
                uint32_t fb;
                double vb;
                tt_event(ev, b, n, J + lane, F, w, fw, dv, per_frame, fb, vb);
                const bool live = J + lane < n;
                for (int t = 0; t < na; t++) {
                    const uint32_t f1 = read_lane_u32(fa, t);
                    const double v1 = read_lane_f64(va, t);
                    if (!live || (J == I && lane < t)) continue;
                    double prod = v1 * vb;
                    uint32_t r = f1, col = fb;
                    if (J != I || lane != t) {
                        if (r == col) prod += prod;
                        else if (r > col) {
                            r = fb;
                            col = f1;
                        }
                    }
                    unsafeAtomicAdd(&C[(int64_t)r * F + col], prod);
                }
            }
        }
    }
}

// Event pairs Sum n (n + 1) / 2 of the compact pixels [c_begin[q], c_end[q])
// added into out[q] (cleared by the caller): blockIdx.x is the q-bin,
// blockIdx.y a stride of its pixels (the cost estimate that picks the method).
__global__ void __launch_bounds__(256)
k_tt_pair_count(const int64_t* __restrict__ off, const int32_t* __restrict__ rowmap,
                const int32_t* __restrict__ c_begin, const int32_t* __restrict__ c_end, double* __restrict__ out)
{
    __shared__ double s_red[256];
    const int q = blockIdx.x;
    double s = 0.0;
    const int stride = (int)(blockDim.x * gridDim.y);
    for (int c = c_begin[q] + (int)(blockIdx.y * blockDim.x + threadIdx.x); c < c_end[q]; c += stride) {
        const int64_t row = rowmap ? (int64_t)rowmap[c] : c;
        const double n = (double)(off[row + 1] - off[row]);
        s += n * (n + 1.0) * 0.5;
    }
    s_red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) unsafeAtomicAdd(&out[q], s_red[0]);
}

// sgsum[bin][f] += value of every event of the q-bin's pixels
__global__ void __launch_bounds__(256)
k_tt_sg_sparse(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off, const int32_t* __restrict__ rowmap,
               int32_t c0, int32_t Pq, int32_t F, const double* __restrict__ weight, double scale,
               const int32_t* __restrict__ bin_of_row, const double* __restrict__ fw, double* __restrict__ sgsum)
{
    const int lane = lane_id();
    const int wpb = (int)(blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * wpb + wave_id(); i < Pq; i += (int64_t)gridDim.x * wpb) {
        const int64_t c = c0 + i;
        const int64_t row = rowmap ? (int64_t)rowmap[c] : c;
        const int64_t b = off[row];
        const int64_t n = off[row + 1] - b;
        const double w = (weight ? weight[c] : 1.0) * scale;
        double* dst = sgsum + (int64_t)bin_of_row[i] * F;
        for (int64_t k = lane; k < n; k += 64) {
            uint32_t f;
            double v;
            tt_event(ev, b, n, k, F, w, fw, nullptr, 0, f, v);
            if (v != 0.0) unsafeAtomicAdd(&dst[f], v);
        }
    }
}

// ---- Gram ---------------------------------------------------------------------------
constexpr int kGramTile = 128;  // output tile (frames x frames)
constexpr int kGramK = 16;      // pixels per LDS stage
constexpr int kGramLoads = kGramK * kGramTile / 256;   // doubles per thread and operand per stage

// Upper-triangle tiles (ti <= tj) of C += X^T X * inv_pq over the pixels
// [p_begin, p_end) of split blockIdx.y; with several splits the tile sums are
// added atomically.  256 threads: wave w owns rows 64*(w/2)..+63 and columns
// 64*(w%2)..+63 of the tile as 4x4 MFMA blocks (on a diagonal tile the wave
// below the diagonal only loads).  The next stage's pixels are loaded into
// registers while the current one is multiplied from LDS.
// MFMA operand layout (16x16x4 FP64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; its result register r holds D[l/16 + 4r][l%16], r = 0..3.
// Two waves per SIMD: the accumulators move from AGPRs to VGPRs (210 in all,
// no spills), which hides the operand loads better than one wave with 332.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_tt_gram(const double* __restrict__ X, int32_t n, int32_t F, int32_t ntile, int32_t ksplit, double inv_pq,
          double* __restrict__ C)
{
    __shared__ double sA[kGramK][kGramTile];
    __shared__ double sB[kGramK][kGramTile];
    // blockIdx.x enumerates the upper-triangle tiles row by row
    int t = blockIdx.x, ti = 0;
    while (t >= ntile - ti) {
        t -= ntile - ti;
        ti++;
    }
    const int tj = ti + t;
    const bool diag = ti == tj;
    const int f0 = ti * kGramTile, f1 = tj * kGramTile;
    const int p_begin = (int)((int64_t)n * blockIdx.y / ksplit);
    const int p_end = (int)((int64_t)n * (blockIdx.y + 1) / ksplit);
    const int lane = lane_id();
    const int w = wave_id();
    const int rb0 = (w >> 1) * 64, cb0 = (w & 1) * 64;
    const bool idle = diag && rb0 > cb0;   // block entirely below the diagonal
    const double(*sBr)[kGramTile] = diag ? sA : sB;
    v4d acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    double ra[kGramLoads], rb[kGramLoads];
    auto load = [&](int p0) {
#pragma unroll
        for (int i = 0; i < kGramLoads; i++) {
            const int e = (int)threadIdx.x + i * 256;
            const int k = e / kGramTile, j = e % kGramTile;
            const int p = p0 + k;
            const bool pin = p < p_end;
            ra[i] = (pin && f0 + j < F) ? X[(int64_t)p * F + f0 + j] : 0.0;
            rb[i] = (!diag && pin && f1 + j < F) ? X[(int64_t)p * F + f1 + j] : 0.0;
        }
    };
    load(p_begin);
    for (int p0 = p_begin; p0 < p_end; p0 += kGramK) {
#pragma unroll
        for (int i = 0; i < kGramLoads; i++) {
            const int e = (int)threadIdx.x + i * 256;
            sA[e / kGramTile][e % kGramTile] = ra[i];
            if (!diag) sB[e / kGramTile][e % kGramTile] = rb[i];
        }
        __syncthreads();
        if (p0 + kGramK < p_end) load(p0 + kGramK);
        if (!idle) {
#pragma unroll
            for (int k4 = 0; k4 < kGramK; k4 += 4) {
                const int kk = k4 + (lane >> 4);
                double av[4], bv[4];
#pragma unroll
                for (int a = 0; a < 4; a++) av[a] = sA[kk][rb0 + a * 16 + (lane & 15)];
#pragma unroll
                for (int b = 0; b < 4; b++) bv[b] = sBr[kk][cb0 + b * 16 + (lane & 15)];
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (idle) return;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = f0 + rb0 + a * 16 + (lane >> 4) + 4 * r;
                const int col = f1 + cb0 + b * 16 + (lane & 15);
                if (row < F && col < F && row <= col) {
                    double* c = C + (int64_t)row * F + col;
                    const double v = acc[a][b][r] * inv_pq;
                    if (ksplit > 1)
                        unsafeAtomicAdd(c, v);
                    else
                        *c += v;
                }
            }
}

// Diagonal sums, pass 1: thread (ff, segment) sums C[fx][fx + ff] over the
// segment's fx (threads of a block read consecutive columns of one row).  With
// `scale` != 1 every upper-triangle element is scaled in the same pass (each is
// visited exactly once), which finishes the event-pair sums.
__global__ void __launch_bounds__(256)
k_tt_diag_seg(double* __restrict__ C, int32_t F, int32_t seg_len, double scale, int32_t rescale,
              double* __restrict__ partial)
{
    const int ff = blockIdx.x * blockDim.x + threadIdx.x;
    if (ff >= F) return;
    const int fx0 = blockIdx.y * seg_len;
    const int fx1 = min(fx0 + seg_len, F - ff);
    double s = 0.0;
    for (int fx = fx0; fx < fx1; fx++) {
        double* c = C + (int64_t)fx * F + fx + ff;
        double x = *c;
        if (rescale) {
            x *= scale;
            *c = x;
        }
        s += x;
    }
    partial[(int64_t)blockIdx.y * F + ff] = s;
}

// pass 2: g2full[ff] = sum of the segment sums / (F - ff)
__global__ void __launch_bounds__(256)
k_tt_diag_final(const double* __restrict__ partial, int32_t nseg, int32_t F, double* __restrict__ g2full)
{
    const int ff = blockIdx.x * blockDim.x + threadIdx.x;
    if (ff >= F) return;
    double s = 0.0;
    for (int k = 0; k < nseg; k++) s += partial[(int64_t)k * F + ff];
    g2full[ff] = s / (double)(F - ff);
}

// g2partials[ff][w] = sum of C[fx][fx + ff] over window w (fx / wsize == w) / wsize,
// ff < wsize; one block per window.
__global__ void __launch_bounds__(256)
k_tt_partials(const double* __restrict__ C, int32_t F, int32_t wsize, int32_t total, double* __restrict__ part)
{
    const int w = blockIdx.x;
    if (w >= total) return;
    for (int ff = threadIdx.x; ff < wsize; ff += blockDim.x) {
        double s = 0.0;
        for (int fx = w * wsize; fx < (w + 1) * wsize; fx++) s += C[(int64_t)fx * F + fx + ff];
        part[(int64_t)ff * total + w] = s / (double)wsize;
    }
}

// ---- host side --------------------------------------------------------------------
void launch_tt_values(xpcs_mt_plan* p, int32_t c0, int32_t n, double scale, double* X, hipStream_t s)
{
    const int32_t F = p->h.F;
    timer_begin(p, s, "k_tt_values");
    launch_fill_u64(reinterpret_cast<uint64_t*>(X), (size_t)n * F, s);
    const unsigned g = (unsigned)std::min<int64_t>(std::max<int32_t>(n, 1), 65535);
    k_tt_values<<<g, 256, 0, s>>>(p->mode == 2 ? nullptr : p->d_ev, p->d_off, p->d_rowmap,
                                  p->mode == 2 ? p->d_rows : nullptr, c0, n, F, p->d_weight, scale,
                                  p->fsnorm ? p->d_fw : nullptr, X);
    timer_end(p, s);
}

static unsigned wave_grid(int64_t items)
{
    return (unsigned)std::min<int64_t>(std::max<int64_t>((items + 3) / 4, 1), 65536);
}

void launch_tt_pairs(xpcs_mt_plan* p, int32_t c0, int32_t Pq, double scale, const int32_t* d_bin_of_row,
                     const double* d_div, bool per_frame, double* C, hipStream_t s)
{
    timer_begin(p, s, "k_tt_pairs");
    k_tt_pairs<<<wave_grid(Pq), 256, 0, s>>>(p->d_ev, p->d_off, p->d_rowmap, c0, Pq, p->h.F, p->d_weight, scale,
                                             d_bin_of_row, d_div, per_frame ? 1 : 0, p->fsnorm ? p->d_fw : nullptr,
                                             C);
    timer_end(p, s);
}

void launch_tt_pair_count(xpcs_mt_plan* p, const int32_t* d_c_begin, const int32_t* d_c_end, int32_t nq, double* d_out,
                          hipStream_t s)
{
    launch_fill_u64(reinterpret_cast<uint64_t*>(d_out), (size_t)nq, s);
    k_tt_pair_count<<<dim3(nq, 64), 256, 0, s>>>(p->d_off, p->d_rowmap, d_c_begin, d_c_end, d_out);
}

void launch_tt_sg_sparse(xpcs_mt_plan* p, int32_t c0, int32_t Pq, double scale, const int32_t* d_bin_of_row,
                         double* d_sgsum, hipStream_t s)
{
    timer_begin(p, s, "k_tt_sg_sparse");
    k_tt_sg_sparse<<<wave_grid(Pq), 256, 0, s>>>(p->d_ev, p->d_off, p->d_rowmap, c0, Pq, p->h.F, p->d_weight,
                                                 scale, d_bin_of_row, p->fsnorm ? p->d_fw : nullptr, d_sgsum);
    timer_end(p, s);
}

void launch_tt_colsum(xpcs_mt_plan* p, const double* X, int32_t F, int32_t lo, int32_t hi,
                      const int32_t* d_bin_begin, const int32_t* d_bin_end, int32_t nbins, double* d_sgsum,
                      hipStream_t s)
{
    dim3 g((F + 255) / 256, std::min(nbins, 65535),
           (unsigned)std::min<int64_t>(((int64_t)hi - lo + kColSeg - 1) / kColSeg, 65535));
    timer_begin(p, s, "k_tt_colsum");
    k_tt_colsum<<<g, 256, 0, s>>>(X, F, lo, hi, d_bin_begin, d_bin_end, nbins, d_sgsum);
    timer_end(p, s);
}

void launch_tt_sgfinal(double* d_sg, int32_t F, const int32_t* d_bin_begin, const int32_t* d_bin_end, int32_t nbins,
                       bool average, double* d_sga, hipStream_t s)
{
    dim3 g((F + 255) / 256, std::min(nbins, 65535));
    k_tt_sgfinal<<<g, 256, 0, s>>>(d_sg, F, d_bin_begin, d_bin_end, nbins);
    if (average) k_tt_sgavg<<<(nbins + 255) / 256, 256, 0, s>>>(d_sg, F, nbins, d_sga);
}

void launch_tt_divide(xpcs_mt_plan* p, double* X, int32_t F, int32_t lo, int32_t n, const int32_t* d_bin_of_row,
                      const double* d_div, bool per_frame, hipStream_t s)
{
    dim3 g((F + 255) / 256, std::min(std::max(n, 1), 65535));
    timer_begin(p, s, "k_tt_divide");
    k_tt_divide<<<g, 256, 0, s>>>(X, F, lo, n, d_bin_of_row, d_div, per_frame ? 1 : 0);
    timer_end(p, s);
}

void launch_tt_gram(xpcs_mt_plan* p, const double* X, int32_t n, int32_t F, double inv_pq, double* C, hipStream_t s)
{
    const int nt = (F + kGramTile - 1) / kGramTile;
    const int ntiles = nt * (nt + 1) / 2;
    // split the pixels when the tiles alone cannot fill the chip (>= 256 pixels a split)
    const int ksplit = std::max(1, std::min((1024 + ntiles - 1) / ntiles, n / 256));
    timer_begin(p, s, "k_tt_gram");
    k_tt_gram<<<dim3(ntiles, ksplit), 256, 0, s>>>(X, n, F, nt, ksplit, inv_pq, C);
    timer_end(p, s);
}

int tt_diag_segments(int32_t F)
{
    const int len = std::max(16, (F + 127) / 128);
    return (F + len - 1) / len;
}

void launch_tt_diag(xpcs_mt_plan* p, double* C, int32_t F, int32_t wsize, int32_t total, double rescale,
                    double* partial, double* g2full, double* g2part, hipStream_t s)
{
    const int nseg = tt_diag_segments(F);
    const int len = (F + nseg - 1) / nseg;
    timer_begin(p, s, "k_tt_diag");
    dim3 g((F + 255) / 256, nseg);
    k_tt_diag_seg<<<g, 256, 0, s>>>(C, F, len, rescale, rescale != 1.0 ? 1 : 0, partial);
    k_tt_diag_final<<<(F + 255) / 256, 256, 0, s>>>(partial, nseg, F, g2full);
    if (total > 0) k_tt_partials<<<total, 256, 0, s>>>(C, F, wsize, total, g2part);
    timer_end(p, s);
}

}  // namespace xpcs
