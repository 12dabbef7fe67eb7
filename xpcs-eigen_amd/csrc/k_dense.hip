// Dense path (DenseFilter, src/xpcs/filter/dense_filter.cpp:121-210, followed by
// Corr::multiTau2, src/xpcs/corr.cpp:315-431): thread-per-pixel streaming
// multi-tau over frame-major data.
//
// One thread owns one detector pixel and walks the frames in order.  A wave's
// loads of a frame are 64 consecutive pixels, so the raw IMM frames are read
// once, fully coalesced, with no transposed copy.  Each level keeps the last
// 2*dpl values in a register shift register; a new level-L value x at index k
// adds x * x_L[k-d] to G2 for every lag d of the level, then enters the
// register.  Level L+1 values are formed from pairs of level-L values; the
// pairing parity is the same for every lane (it only depends on the frame
// index), so the cascade never diverges.
//
// IP and IF come from level-0 prefix sums captured at the schedule's frame
// boundaries (IP = S0(2^L (lf - tau')), IF = S0(2^L lf) - S0(2^L tau')); the
// boundary list is sorted on the host and walked with a scalar pointer.
//
// Levels are processed K at a time (register budget).  The first pass reads the
// raw frames (dark subtraction, threshold and mask fused) and, when more
// levels remain, writes the level-K series [P][lf_K]; each further pass reads
// the previous series.
#include "device_util.h"
#include "internal.h"

namespace xpcs {

// Boundary events of the level-0 prefix: after frame `frame` (i.e. when the
// running sum covers frames < frame+1 ... see host), apply op to slot.
// op 0: IP[slot] = S;  op 1: IF[slot] = S;  op 2: IF[slot] = S - IF[slot].

template <int DPL, int K>
struct StreamState {
    uint32_t win[K][2 * DPL];    // win[l][j] = value j+1 steps back at level L0+l
    uint32_t pend[K];            // first value of an incomplete pair at level L0+l
    uint64_t acc[K][2 * DPL];    // G2 accumulators per (level, lag-1)
};

// Push one value into level l of the pass (static l), accumulate its lag
// products, and cascade the completed pair into level l+1.
template <int DPL, int K, bool FIRST, int l>
__device__ __forceinline__ void stream_push(StreamState<DPL, K>& s, uint32_t x, uint32_t (&cnt)[K + 1],
                                            const int (&nlag)[K], const int (&lo)[K], uint32_t* out_next,
                                            int lf_next)
{
    if constexpr (l < K) {
        // lag products with the last values of this level (level 0 of the
        // schedule has lags 1..2*dpl, every other level dpl+1..2*dpl)
        constexpr int dlo = (FIRST && l == 0) ? 1 : DPL + 1;
#pragma unroll
        for (int d = dlo; d <= 2 * DPL; d++) {
            const int rel = d - lo[l];
            if (rel >= 0 && rel < nlag[l]) s.acc[l][d - 1] += (uint64_t)x * (uint64_t)s.win[l][d - 1];
        }
#pragma unroll
        for (int j = 2 * DPL - 1; j > 0; j--) s.win[l][j] = s.win[l][j - 1];
        s.win[l][0] = x;
        const uint32_t k = cnt[l]++;
        if ((k & 1u) == 0u) {
            s.pend[l] = x;
        } else {
            const uint32_t y = s.pend[l] + x;
            if constexpr (l + 1 < K) {
                stream_push<DPL, K, FIRST, l + 1>(s, y, cnt, nlag, lo, out_next, lf_next);
            } else {
                const uint32_t kn = cnt[K]++;
                if (out_next && (int)kn < lf_next) out_next[kn] = y;
            }
        }
    }
}

template <int DPL, int K, bool RAW>
__global__ void __launch_bounds__(256)
k_mt_dense_pass(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off, int avg,
                const int32_t* __restrict__ pix2c, uint32_t npix_det, const int64_t* __restrict__ dark_sum,
                const int64_t* __restrict__ thr, int64_t nscale,
                const uint32_t* __restrict__ series_in, int32_t P, int32_t n_in,
                const LevelParam* __restrict__ levels, int32_t L0, int32_t nlev_pass,
                const int4* __restrict__ bnd, int32_t nbnd,
                int32_t T, int64_t* __restrict__ G2, int64_t* __restrict__ IP, int64_t* __restrict__ IF,
                uint64_t* __restrict__ pix_total, uint32_t* __restrict__ series_out, int32_t lf_out)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t c;
    if constexpr (RAW) {
        if (gid >= npix_det) return;
        c = pix2c[gid];
        if (c < 0) return;
    } else {
        if (gid >= P) return;
        c = gid;
    }
    StreamState<DPL, K> s;
#pragma unroll
    for (int l = 0; l < K; l++) {
        s.pend[l] = 0u;
#pragma unroll
        for (int j = 0; j < 2 * DPL; j++) {
            s.win[l][j] = 0u;
            s.acc[l][j] = 0ull;
        }
    }
    int nlag[K], lo[K];
#pragma unroll
    for (int l = 0; l < K; l++) {
        if (l < nlev_pass) {
            nlag[l] = levels[L0 + l].count;
            lo[l] = levels[L0 + l].lo;
        } else {
            nlag[l] = 0;
            lo[l] = 1;
        }
    }
    uint32_t cnt[K + 1];
#pragma unroll
    for (int l = 0; l <= K; l++) cnt[l] = 0u;
    uint32_t* out_next = (series_out && nlev_pass == K) ? series_out + c * (int64_t)lf_out : nullptr;

    uint64_t S = 0;
    int bi = 0;
    int64_t ds = 0, th = 0;
    if constexpr (RAW) {
        ds = dark_sum ? dark_sum[c] : 0;
        th = thr ? thr[c] : 0;
    }
    const uint32_t* in_row = RAW ? nullptr : series_in + c * (int64_t)n_in;
    // Frames are fetched kPrefetch at a time so that many independent loads
    // are in flight per thread; processing order stays frame order.
    constexpr int kPrefetch = 16;
    for (int t0 = 0; t0 < n_in; t0 += kPrefetch) {
        uint32_t xs[kPrefetch];
        if constexpr (RAW) {
            if (avg == 1) {
                uint16_t raw[kPrefetch];
#pragma unroll
                for (int u = 0; u < kPrefetch; u++) {
                    const int t = min(t0 + u, n_in - 1);
                    raw[u] = reinterpret_cast<const uint16_t*>(imm + sel_off[t])[gid];
                }
#pragma unroll
                for (int u = 0; u < kPrefetch; u++) {
                    const int64_t y = nscale * (int64_t)raw[u] - ds;
                    xs[u] = (y > th && y > 0) ? (uint32_t)y : 0u;
                }
            } else {
#pragma unroll
                for (int u = 0; u < kPrefetch; u++) {
                    const int t = min(t0 + u, n_in - 1);
                    uint32_t acc = 0;
                    for (int a = 0; a < avg; a++) {
                        const uint16_t* fr = reinterpret_cast<const uint16_t*>(imm + sel_off[(int64_t)t * avg + a]);
                        const int64_t y = nscale * (int64_t)fr[gid] - ds;
                        if (y > th && y > 0) acc += (uint32_t)y;
                    }
                    xs[u] = acc;
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < kPrefetch; u++) xs[u] = in_row[min(t0 + u, n_in - 1)];
        }
        const int nu = min(kPrefetch, n_in - t0);
        for (int u = 0; u < nu; u++) {
            const uint32_t x = xs[u];
            if constexpr (RAW) {
                const int t = t0 + u;
                S += x;
                // boundary captures after frame t (running sum covers frames <= t)
                while (bi < nbnd && bnd[bi].x == t + 1) {
                    const int4 b = bnd[bi];
                    if (b.y == 0) IP[c * T + b.z] = (int64_t)S;
                    else if (b.y == 1) IF[c * T + b.z] = (int64_t)S;
                    else IF[c * T + b.z] = (int64_t)S - IF[c * T + b.z];
                    bi++;
                }
            }
            stream_push<DPL, K, RAW, 0>(s, x, cnt, nlag, lo, out_next, lf_out);
        }
    }
    if constexpr (RAW) {
        pix_total[c] = S;
        // boundaries at frame 0 (b == 0) cover no frames
        for (int i = 0; i < nbnd; i++) {
            const int4 b = bnd[i];
            if (b.x != 0) break;
            if (b.y == 0) IP[c * T + b.z] = 0;
            else if (b.y == 1) IF[c * T + b.z] = 0;
        }
    }
#pragma unroll
    for (int l = 0; l < K; l++) {
        if (l < nlev_pass) {
            const LevelParam lp = levels[L0 + l];
            constexpr int dlo_any = 1;
#pragma unroll
            for (int d = dlo_any; d <= 2 * DPL; d++) {
                const int rel = d - lp.lo;
                if ((RAW && l == 0 ? d >= 1 : d >= DPL + 1) && rel >= 0 && rel < lp.count)
                    G2[c * T + lp.first + rel] = (int64_t)s.acc[l][d - 1];
            }
        }
    }
}

// Side statistics of the dense path straight from the raw frames: per
// (output frame, static bin) sums of the thresholded values.
__global__ void __launch_bounds__(256)
k_side_dense_raw(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off, int avg, int32_t F,
                 const int32_t* __restrict__ pix2c, uint32_t npix_det, const int64_t* __restrict__ dark_sum,
                 const int64_t* __restrict__ thr, int64_t nscale, const int32_t* __restrict__ sbin_of_c,
                 const double* __restrict__ weight, int32_t S, double* __restrict__ fs)
{
    __shared__ double s_acc[4096];
    for (int f = blockIdx.x; f < F; f += gridDim.x) {
        for (int b = threadIdx.x; b < S; b += blockDim.x) s_acc[b] = 0.0;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < npix_det; j += blockDim.x) {
            const int32_t c = pix2c[j];
            if (c < 0) continue;
            uint32_t acc = 0;
            for (int a = 0; a < avg; a++) {
                const uint16_t* fr = reinterpret_cast<const uint16_t*>(imm + sel_off[(int64_t)f * avg + a]);
                const int64_t y = nscale * (int64_t)fr[j] - (dark_sum ? dark_sum[c] : 0);
                if (y > (thr ? thr[c] : 0) && y > 0) acc += (uint32_t)y;
            }
            if (acc) atomicAdd(&s_acc[sbin_of_c[c] - 1], (double)acc * (weight ? weight[c] : 1.0));
        }
        __syncthreads();
        for (int b = threadIdx.x; b < S; b += blockDim.x)
            if (s_acc[b] != 0.0) atomicAdd(&fs[(int64_t)f * S + b], s_acc[b]);
        __syncthreads();
    }
}

// ---- host side --------------------------------------------------------------------
constexpr int kDenseK = 6;

bool dense_stream_supported(const HostTables& h)
{
    return h.dpl == 4 || h.dpl == 8;
}

template <int DPL>
static void run_passes(xpcs_mt_plan* p, hipStream_t s)
{
    const HostTables& h = p->h;
    const int nlev = (int)h.levels.size();
    const int npd = h.W * h.H;
    int L0 = 0;
    int n_in = h.F;
    int pass = 0;
    while (L0 < nlev) {
        const int k = std::min(kDenseK, nlev - L0);
        const bool more = L0 + k < nlev;
        const int lf_out = more ? h.levels[L0 + k].lf : 0;
        uint32_t* out = more ? (pass % 2 == 0 ? p->d_series_a : p->d_series_b) : nullptr;
        const uint32_t* in = pass == 0 ? nullptr : (pass % 2 == 1 ? p->d_series_a : p->d_series_b);
        if (pass == 0) {
            timer_begin(p, s, "k_mt_dense_raw");
            k_mt_dense_pass<DPL, kDenseK, true><<<(npd + 255) / 256, 256, 0, s>>>(
                p->d_imm, p->d_sel_off, h.avg, p->d_pix2c, (uint32_t)npd, p->d_dark_sum, p->d_thr, p->nscale,
                nullptr, h.P, n_in, p->d_levels, L0, k, p->d_bnd, p->nbnd, h.T, p->d_G2, p->d_IP, p->d_IF,
                p->d_pix_total, out, lf_out);
        } else {
            timer_begin(p, s, "k_mt_dense_series");
            k_mt_dense_pass<DPL, kDenseK, false><<<(h.P + 255) / 256, 256, 0, s>>>(
                nullptr, nullptr, 1, nullptr, 0, nullptr, nullptr, 1, in, h.P, n_in, p->d_levels, L0, k,
                nullptr, 0, h.T, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, out, lf_out);
        }
        timer_end(p, s);
        n_in = lf_out;
        L0 += k;
        pass++;
    }
}

void launch_mt_dense(xpcs_mt_plan* p, hipStream_t s)
{
    if (p->h.dpl == 4) run_passes<4>(p, s);
    else run_passes<8>(p, s);
}

void launch_side_dense_raw(xpcs_mt_plan* p, hipStream_t s)
{
    const HostTables& h = p->h;
    timer_begin(p, s, "k_side");
    k_side_dense_raw<<<h.F < 65536 ? h.F : 65536, 256, 0, s>>>(
        p->d_imm, p->d_sel_off, h.avg, h.F, p->d_pix2c, (uint32_t)(h.W * h.H), p->d_dark_sum, p->d_thr,
        p->nscale, p->d_sbin_of_c, p->d_weight, h.ns, p->d_fs);
    timer_end(p, s);
}

}  // namespace xpcs
