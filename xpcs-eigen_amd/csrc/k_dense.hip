// Dense path (DenseFilter, src/xpcs/filter/dense_filter.cpp:121-210, followed by
// Corr::multiTau2, src/xpcs/corr.cpp:315-431): thread-per-pixel streaming
// multi-tau over frame-major data.
//
// One thread owns one detector pixel and walks the frames in order.  A wave's
// loads of a frame are 64 consecutive pixels, so the raw IMM frames are read
// once, fully coalesced, with no transposed copy.  Frames are taken 16 at a
// time; within such a group the pairing of every level up to 2^4 is fixed at
// compile time, so levels 0..4 are fully unrolled (no shift-register moves,
// no branches) and only the levels above follow the group index (a
// wave-uniform branch).  Each level keeps its last 2*dpl values; a new level-L
// value x at index k adds x * x_L[k-d] to the lag-d sum for every lag d of the
// level.
//
// Lag sums are accumulated with FP64 fused multiply-adds, which is exact here:
// every value is a non-negative integer, and every product and partial sum of
// a pixel is at most S^2, S = the pixel's total count, so while S^2 < 2^53 no
// result is ever rounded.  A pixel with a larger total is listed by the first
// pass and recomputed with 64-bit integer sums by k_mt_dense_exact.
//
// IP and IF come from level-0 prefix sums captured at the schedule's frame
// boundaries (IP = S0(2^L (lf - tau')), IF = S0(2^L lf) - S0(2^L tau')); the
// boundary list is sorted on the host and walked with a wave-uniform index.
//
// Levels are processed K at a time (register budget).  The first pass reads the
// raw frames (dark subtraction, threshold and mask fused) and, when more
// levels remain, writes the level-K series frame-major ([lf_K][W*H], the
// column of a pixel is its raster index); each further pass reads the
// previous series with the same coalesced pattern.
#include <climits>
#include <type_traits>

#include "device_util.h"
#include "internal.h"

namespace xpcs {

// Largest pixel total S with S^2 < 2^53 (every lag sum exact in FP64).
constexpr uint32_t kExactF64Total = 94906265u;
constexpr int kGroup = 16;

// Levels per pass (register budget): dpl 4 at 4 levels keeps two waves per
// SIMD; dpl 8 at 3 levels runs one wave per SIMD and is as fast as 2 levels at
// two waves, with fewer passes over the coarsened series.
template <int DPL>
constexpr int dense_levels_per_pass()
{
    return DPL == 4 ? 4 : 3;
}

template <int DPL, bool FIRST, int l>
constexpr int lag_lo()
{
    // level 0 of the schedule has lags 1..2*dpl, every other level dpl+1..2*dpl
    return (FIRST && l == 0) ? 1 : DPL + 1;
}

template <int DPL, int K>
struct FState {
    double win[K][2 * DPL];   // win[l][j] = value j+1 steps back at level L0+l
    double acc[K][2 * DPL];   // lag sums, acc[l][d-1] for lag d
    double pend[K];           // first value of an incomplete pair (levels above the group)
};

// ST: uint32_t for counts, double for frame-weighted values (normalize_by_framesum)
template <typename ST>
struct SeriesOut {
    ST* col;                  // element k of the output series at col[k * stride]
    int64_t stride;
    uint32_t lf;              // values of the output level
};

// One value of level l (index k, wave-uniform): lag products, window shift,
// and pairing into level l+1 by the parity of k.
template <int DPL, int K, bool FIRST, int l, typename ST>
__device__ __forceinline__ void push_one(FState<DPL, K>& s, double v, uint32_t k, const SeriesOut<ST>& o)
{
    if constexpr (l < K) {
        constexpr int dlo = lag_lo<DPL, FIRST, l>();
#pragma unroll
        for (int d = dlo; d <= 2 * DPL; d++) s.acc[l][d - 1] = __builtin_fma(v, s.win[l][d - 1], s.acc[l][d - 1]);
#pragma unroll
        for (int j = 2 * DPL - 1; j > 0; j--) s.win[l][j] = s.win[l][j - 1];
        s.win[l][0] = v;
        if (k & 1u) push_one<DPL, K, FIRST, l + 1>(s, s.pend[l] + v, k >> 1, o);
        else s.pend[l] = v;
    } else {
        if (o.col && k < o.lf) o.col[(int64_t)k * o.stride] = (ST)v;
    }
}

// N consecutive values of level l, the first with index kb (N = 16 >> l).
// TAIL: the last, partial group; values with index >= lf of their level are
// zeroed (they do not exist in the series) and nothing reaches the levels
// above the group.
template <int DPL, int K, bool FIRST, bool TAIL, int l, int N, typename ST>
__device__ __forceinline__ void push_block(FState<DPL, K>& s, double (&v)[N], uint32_t kb, uint32_t n_in,
                                           const SeriesOut<ST>& o)
{
    if constexpr (TAIL) {
        const uint32_t lim = n_in >> l;
#pragma unroll
        for (int j = 0; j < N; j++)
            if (kb + j >= lim) v[j] = 0.0;
    }
    if constexpr (l < K) {
        constexpr int dlo = lag_lo<DPL, FIRST, l>();
#pragma unroll
        for (int j = 0; j < N; j++) {
#pragma unroll
            for (int d = dlo; d <= 2 * DPL; d++) {
                const double w = (j >= d) ? v[j >= d ? j - d : 0] : s.win[l][j >= d ? 0 : d - j - 1];
                s.acc[l][d - 1] = __builtin_fma(v[j], w, s.acc[l][d - 1]);
            }
        }
        double nw[2 * DPL];
#pragma unroll
        for (int i = 0; i < 2 * DPL; i++) nw[i] = (i < N) ? v[i < N ? N - 1 - i : 0] : s.win[l][i < N ? 0 : i - N];
#pragma unroll
        for (int i = 0; i < 2 * DPL; i++) s.win[l][i] = nw[i];
        if constexpr (N >= 2) {
            double u[N / 2];
#pragma unroll
            for (int j = 0; j < N / 2; j++) u[j] = v[2 * j] + v[2 * j + 1];
            push_block<DPL, K, FIRST, TAIL, l + 1, N / 2>(s, u, kb >> 1, n_in, o);
        } else if constexpr (!TAIL) {
            if (kb & 1u) push_one<DPL, K, FIRST, l + 1>(s, s.pend[l] + v[0], kb >> 1, o);
            else s.pend[l] = v[0];
        }
    } else {
#pragma unroll
        for (int j = 0; j < N; j++)
            if (o.col && kb + j < o.lf) o.col[(int64_t)(kb + j) * o.stride] = (ST)v[j];
    }
}

// Dark subtraction and threshold of one raw value (n-scaled integers: kept iff
// n*raw - sum(dark) > max(floor(n*threshold), 0)).
__device__ __forceinline__ uint32_t dense_value(uint32_t raw, int32_t nscale, int32_t nds, int32_t th)
{
    const int32_t y = __mul24((int32_t)raw, nscale) + nds;
    return y > th ? (uint32_t)y : 0u;
}

// Input of a pass: the level-K series of the previous pass, the raw frames
// one 2-byte load per lane and frame, or the raw frames staged through LDS
// (evenly spaced frames, avg 1): a wave fetches a 16-frame x 64-pixel tile
// with two 16-byte loads per lane, DEPTH tiles ahead of use.
enum DenseInput { kSeries = 0, kRawLanes = 1, kRawTiles = 2 };
constexpr int kTileDepth = 3;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// FP (normalize_by_framesum, main.cpp:313-336): level-0 values are the counts
// times fw[frame], the series between passes are FP64, and G2/IP/IF are FP64
// sums stored as bit patterns (no exactness bound, no integer redo);
// pix_total stays the raw count total.
template <int DPL, int K, int MODE, bool FP>
__global__ void __launch_bounds__(256)
k_mt_dense_f64(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off, int64_t tile_off0,
               int64_t tile_step, int avg, const int32_t* __restrict__ pix2c, uint32_t npix,
               const int64_t* __restrict__ dark_sum, const int64_t* __restrict__ thr, int32_t nscale,
               const std::conditional_t<FP, double, uint32_t>* __restrict__ series_in, uint32_t n_in,
               const LevelParam* __restrict__ levels, int32_t L0, int32_t nlev_pass, const int4* __restrict__ bnd,
               int32_t nbnd, int32_t T, int64_t* __restrict__ G2, int64_t* __restrict__ IP,
               int64_t* __restrict__ IF, uint64_t* __restrict__ pix_total,
               std::conditional_t<FP, double, uint32_t>* __restrict__ series_out, uint32_t lf_out,
               uint32_t* __restrict__ redo_count, int32_t* __restrict__ redo_list, const double* __restrict__ fw,
               bool w32)
{
    using ST = std::conditional_t<FP, double, uint32_t>;   // series element
    constexpr bool RAW = MODE != kSeries;
    // values of a group: integers from the raw frames, or the input series
    using XT = std::conditional_t<FP && !RAW, double, uint32_t>;
    constexpr bool TILES = MODE == kRawTiles;
    __shared__ u32x4 s_tile[TILES ? 4 : 1][2][64];

    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t c = gid < npix ? pix2c[gid] : -1;
    const bool active = c >= 0;
    if constexpr (TILES) {
        // every lane of a wave takes part in the tile loads; a wave without a
        // mapped pixel has nothing to do
        if (__ballot(active) == 0ull) return;
    } else {
        if (!active) return;
    }
    FState<DPL, K> s;
#pragma unroll
    for (int l = 0; l < K; l++) {
        s.pend[l] = 0.0;
#pragma unroll
        for (int j = 0; j < 2 * DPL; j++) {
            s.win[l][j] = 0.0;
            s.acc[l][j] = 0.0;
        }
    }
    const SeriesOut<ST> o{(active && series_out && nlev_pass == K) ? series_out + gid : nullptr, (int64_t)npix,
                          lf_out};
    int32_t nds = 0, th = 0;
    if constexpr (RAW) {
        if (active) {
            nds = dark_sum ? -(int32_t)dark_sum[c] : 0;
            const int64_t t = thr ? thr[c] : 0;
            th = t < 0 ? 0 : (t > INT32_MAX ? INT32_MAX : (int32_t)t);
        }
    }
    const int64_t rowT = (int64_t)(active ? c : 0) * T;
    uint32_t S = 0;
    double Sd = 0.0;   // FP: running sum of the weighted values
    int bi = 0;
    if constexpr (RAW) {
        while (bi < nbnd && bnd[bi].x == 0) bi++;
    }
    int nextb = (RAW && bi < nbnd) ? bnd[bi].x : INT_MAX;

    // boundary captures inside group t0 (wave-uniform test), then the running sums
    auto captures = [&](const XT (&x)[kGroup], const double (&v)[kGroup], uint32_t t0) {
        if constexpr (RAW) {
            while (nextb <= (int)(t0 + kGroup)) {
                const int4 b = bnd[bi];
                const int n = b.x - (int)t0;
                if constexpr (FP) {
                    double part = Sd;
#pragma unroll
                    for (int u = 0; u < kGroup; u++) part += (u < n) ? v[u] : 0.0;
                    if (active) {
                        if (b.y == 0) IP[rowT + b.z] = (int64_t)__double_as_longlong(part);
                        else if (b.y == 1) IF[rowT + b.z] = (int64_t)__double_as_longlong(part);
                        else
                            IF[rowT + b.z] =
                                (int64_t)__double_as_longlong(part - __longlong_as_double((long long)IF[rowT + b.z]));
                    }
                } else {
                    uint32_t part = S;
#pragma unroll
                    for (int u = 0; u < kGroup; u++) part += (u < n) ? x[u] : 0u;
                    if (active) {
                        if (b.y == 0) put_sum(IP, rowT + b.z, (int64_t)part, w32);
                        else if (b.y == 1) put_sum(IF, rowT + b.z, (int64_t)part, w32);
                        else put_sum(IF, rowT + b.z, (int64_t)part - get_sum(IF, rowT + b.z, w32), w32);
                    }
                }
                bi++;
                nextb = bi < nbnd ? bnd[bi].x : INT_MAX;
            }
#pragma unroll
            for (int u = 0; u < kGroup; u++) S += (uint32_t)x[u];
            if constexpr (FP) {
#pragma unroll
                for (int u = 0; u < kGroup; u++) Sd += v[u];
            }
        }
    };
    // the group's values (counts x fw[frame] for FP raw input)
    auto weigh = [&](const XT (&x)[kGroup], double (&v)[kGroup], uint32_t t0) {
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            v[u] = (double)x[u];
            if constexpr (FP && RAW) v[u] *= fw[min(t0 + (uint32_t)u, n_in - 1)];
        }
    };
    auto group = [&](XT (&x)[kGroup], uint32_t t0) {
        double v[kGroup];
        if constexpr (FP) {
            weigh(x, v, t0);
            captures(x, v, t0);
        } else {   // counts: captures read x only
            captures(x, v, t0);
            weigh(x, v, t0);
        }
        push_block<DPL, K, RAW, false, 0, kGroup>(s, v, t0, n_in, o);
    };
    auto tail_group = [&](XT (&x)[kGroup], uint32_t t0) {
#pragma unroll
        for (int u = 0; u < kGroup; u++)
            if (t0 + (uint32_t)u >= n_in) x[u] = XT(0);
        double v[kGroup];
        if constexpr (FP) {
            weigh(x, v, t0);
            captures(x, v, t0);
        } else {
            captures(x, v, t0);
            weigh(x, v, t0);
        }
        push_block<DPL, K, RAW, true, 0, kGroup>(s, v, t0, n_in, o);
    };
    const uint32_t nfull = n_in / kGroup * kGroup;

    if constexpr (TILES) {
        // lane = 8 * f + j loads frame t0 + 8 * i + f, pixels 8 j .. 8 j + 7 of
        // the wave, for i = 0, 1; in LDS the tile is [16 frames][64 pixels]
        const uint32_t lane = threadIdx.x & 63u;
        const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        const uint8_t* wsrc = imm + tile_off0 + 2 * (int64_t)(gid - lane) + 16 * (lane & 7u);
        auto issue = [&](u32x4 (&r)[2], uint32_t t0) {
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const uint32_t t = min(t0 + 8u * i + (lane >> 3), n_in - 1);
                r[i] = *reinterpret_cast<const u32x4*>(wsrc + (int64_t)t * tile_step);
            }
        };
        auto stage = [&](const u32x4 (&r)[2], uint32_t (&x)[kGroup]) {
            s_tile[w][0][lane] = r[0];
            s_tile[w][1][lane] = r[1];
            wave_sync();
            const uint16_t* t16 = reinterpret_cast<const uint16_t*>(&s_tile[w][0][0]);
#pragma unroll
            for (int u = 0; u < kGroup; u++) x[u] = dense_value(t16[u * 64 + lane], nscale, nds, th);
            wave_sync();
        };
        u32x4 ring[kTileDepth][2];
#pragma unroll
        for (int d = 0; d < kTileDepth; d++) issue(ring[d], (uint32_t)(d * kGroup));
        for (uint32_t g0 = 0; g0 < nfull; g0 += kTileDepth * kGroup) {
#pragma unroll
            for (int d = 0; d < kTileDepth; d++) {
                const uint32_t t0 = g0 + (uint32_t)(d * kGroup);
                if (t0 < nfull) {
                    uint32_t x[kGroup];
                    stage(ring[d], x);
                    issue(ring[d], t0 + kTileDepth * kGroup);
                    group(x, t0);
                }
            }
        }
        if (nfull < n_in) {
            u32x4 r[2];
            issue(r, nfull);
            uint32_t x[kGroup];
            stage(r, x);
            tail_group(x, nfull);
        }
    } else {
        const uint16_t* col16 = reinterpret_cast<const uint16_t*>(imm) + gid;
        const ST* col32 = RAW ? nullptr : series_in + gid;
        // loads of group t0 (frame indices clamped into range; the tail zeroes them)
        auto load = [&](XT (&r)[kGroup], uint32_t t0) {
#pragma unroll
            for (int u = 0; u < kGroup; u++) {
                const uint32_t t = min(t0 + (uint32_t)u, n_in - 1);
                if constexpr (RAW) r[u] = *reinterpret_cast<const uint16_t*>(
                                              reinterpret_cast<const uint8_t*>(col16) + sel_off[t]);
                else r[u] = col32[(int64_t)t * npix];
            }
        };
        auto values = [&](XT (&x)[kGroup], const XT (&r)[kGroup], uint32_t t0) {
            if constexpr (RAW) {
                if (avg == 1) {
#pragma unroll
                    for (int u = 0; u < kGroup; u++) x[u] = dense_value(r[u], nscale, nds, th);
                } else {
#pragma unroll
                    for (int u = 0; u < kGroup; u++) {
                        const uint32_t t = min(t0 + (uint32_t)u, n_in - 1);
                        uint32_t a_sum = 0;
                        for (int a = 0; a < avg; a++) {
                            const uint32_t raw = *reinterpret_cast<const uint16_t*>(
                                reinterpret_cast<const uint8_t*>(col16) + sel_off[(int64_t)t * avg + a]);
                            a_sum += dense_value(raw, nscale, nds, th);
                        }
                        x[u] = a_sum;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < kGroup; u++) x[u] = r[u];
            }
        };
        XT rcur[kGroup];
        if (RAW && avg != 1) {
#pragma unroll
            for (int u = 0; u < kGroup; u++) rcur[u] = XT(0);
        } else {
            load(rcur, 0);
        }
        for (uint32_t t0 = 0; t0 < nfull; t0 += kGroup) {
            // next group's loads are issued before this group is processed
            XT rnext[kGroup];
            if (RAW && avg != 1) {
#pragma unroll
                for (int u = 0; u < kGroup; u++) rnext[u] = XT(0);
            } else {
                load(rnext, t0 + kGroup);
            }
            XT x[kGroup];
            values(x, rcur, t0);
            group(x, t0);
#pragma unroll
            for (int u = 0; u < kGroup; u++) rcur[u] = rnext[u];
        }
        if (nfull < n_in) {
            XT x[kGroup];
            values(x, rcur, nfull);
            tail_group(x, nfull);
        }
    }
    if (!active) return;
    if constexpr (RAW) {
        pix_total[c] = S;
        // boundaries at frame 0 cover no frames
        for (int i = 0; i < nbnd; i++) {
            const int4 b = bnd[i];
            if (b.x != 0) break;
            if (b.y == 0) put_sum(IP, rowT + b.z, 0, !FP && w32);
            else if (b.y == 1) put_sum(IF, rowT + b.z, 0, !FP && w32);
        }
        if (!FP && S > kExactF64Total) redo_list[atomicAdd(redo_count, 1u)] = (int32_t)gid;
    }
#pragma unroll
    for (int l = 0; l < K; l++) {
        if (l < nlev_pass) {
            const LevelParam lp = levels[L0 + l];
            const int dlo = (RAW && l == 0) ? 1 : DPL + 1;
#pragma unroll
            for (int d = 1; d <= 2 * DPL; d++) {
                const int rel = d - lp.lo;
                if (d >= dlo && rel >= 0 && rel < lp.count)
                    G2[rowT + lp.first + rel] =
                        FP ? (int64_t)__double_as_longlong(s.acc[l][d - 1]) : (int64_t)s.acc[l][d - 1];
            }
        }
    }
}

// ---- exact 64-bit integer variant for the listed pixels ---------------------

template <int DPL, int K>
struct XState {
    uint32_t win[K][2 * DPL];
    uint32_t pend[K];
    uint64_t acc[K][2 * DPL];
};

template <int DPL, int K, bool FIRST, int l>
__device__ __forceinline__ void push_exact(XState<DPL, K>& s, uint32_t x, uint32_t (&cnt)[K])
{
    if constexpr (l < K) {
        constexpr int dlo = lag_lo<DPL, FIRST, l>();
#pragma unroll
        for (int d = dlo; d <= 2 * DPL; d++) s.acc[l][d - 1] += (uint64_t)x * (uint64_t)s.win[l][d - 1];
#pragma unroll
        for (int j = 2 * DPL - 1; j > 0; j--) s.win[l][j] = s.win[l][j - 1];
        s.win[l][0] = x;
        const uint32_t k = cnt[l]++;
        if (k & 1u) push_exact<DPL, K, FIRST, l + 1>(s, s.pend[l] + x, cnt);
        else s.pend[l] = x;
    }
}

// Lag sums of the pixels listed by k_mt_dense_f64 (totals too large for exact
// FP64), one thread per listed pixel, 64-bit integer accumulation.  Only G2
// is rewritten: IP, IF and the series are integer results of the FP64 pass.
template <int DPL, int K, bool RAW>
__global__ void __launch_bounds__(256)
k_mt_dense_exact(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off, int avg,
                 const int32_t* __restrict__ pix2c, uint32_t npix, const int64_t* __restrict__ dark_sum,
                 const int64_t* __restrict__ thr, int32_t nscale, const uint32_t* __restrict__ series_in,
                 uint32_t n_in, const LevelParam* __restrict__ levels, int32_t L0, int32_t nlev_pass, int32_t T,
                 int64_t* __restrict__ G2, const uint32_t* __restrict__ redo_count,
                 const int32_t* __restrict__ redo_list)
{
    const uint32_t n = *redo_count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t gid = (uint32_t)redo_list[i];
        const int32_t c = pix2c[gid];
        XState<DPL, K> s;
        uint32_t cnt[K];
#pragma unroll
        for (int l = 0; l < K; l++) {
            s.pend[l] = 0u;
            cnt[l] = 0u;
#pragma unroll
            for (int j = 0; j < 2 * DPL; j++) {
                s.win[l][j] = 0u;
                s.acc[l][j] = 0ull;
            }
        }
        int32_t nds = 0, th = 0;
        if constexpr (RAW) {
            nds = dark_sum ? -(int32_t)dark_sum[c] : 0;
            const int64_t t = thr ? thr[c] : 0;
            th = t < 0 ? 0 : (t > INT32_MAX ? INT32_MAX : (int32_t)t);
        }
        for (uint32_t t = 0; t < n_in; t++) {
            uint32_t x;
            if constexpr (RAW) {
                x = 0;
                for (int a = 0; a < avg; a++) {
                    const uint32_t raw = reinterpret_cast<const uint16_t*>(imm + sel_off[(int64_t)t * avg + a])[gid];
                    x += dense_value(raw, nscale, nds, th);
                }
            } else {
                x = series_in[(int64_t)t * npix + gid];
            }
            push_exact<DPL, K, RAW, 0>(s, x, cnt);
        }
        const int64_t rowT = (int64_t)c * T;
#pragma unroll
        for (int l = 0; l < K; l++) {
            if (l < nlev_pass) {
                const LevelParam lp = levels[L0 + l];
                const int dlo = (RAW && l == 0) ? 1 : DPL + 1;
#pragma unroll
                for (int d = 1; d <= 2 * DPL; d++) {
                    const int rel = d - lp.lo;
                    if (d >= dlo && rel >= 0 && rel < lp.count)
                        G2[rowT + lp.first + rel] = (int64_t)s.acc[l][d - 1];
                }
            }
        }
    }
}

// Side statistics of the dense path straight from the raw frames: per
// (output frame, static bin) sums of the thresholded, flat-field weighted
// values (the per-frame partition sums of dense_filter.cpp / sparse_filter.cpp
// :134-191). A lane owns 8 consecutive detector pixels (one 16-byte load per
// raw frame when VEC) and keeps their bin, dark, threshold and weight in
// registers for a chunk of FC frames. Pixels of equal bin are summed inside the
// lane (the runs before the lane's last bin change go to LDS directly), the
// last runs across lanes by a segmented scan whose segments (equal bins of
// neighbouring lanes) are fixed for the whole kernel; segment tails add into
// the block's LDS table [FC][S], which goes to fs with one global atomic per
// nonzero entry. Values in 32-bit arithmetic: the caller guarantees
// 65535 * nscale < 2^31. INT (no flat field, F >= 512): the sums stay u32 up to
// the LDS atomic -- one value is below 2^32 / F (F * 65535 * nscale * avg <
// 2^32), so a wave segment of at most 512 pixels sums below 2^32.
constexpr int kSidePix = 8;
constexpr int kSideLdsDoubles = 4096;   // LDS table of one block (32 KB)
constexpr int kSideRing = 3;            // raw frames loaded ahead (VEC, avg 1)

template <bool VEC, bool INT>
__global__ void __launch_bounds__(256)
k_side_dense_raw(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off, int avg, int32_t F,
                 const int32_t* __restrict__ pix2c, uint32_t npix_det, const int64_t* __restrict__ dark_sum,
                 const int64_t* __restrict__ thr, int32_t nscale, const int32_t* __restrict__ sbin_of_c,
                 const double* __restrict__ weight, int32_t S, int32_t FC, double* __restrict__ fs)
{
    extern __shared__ double s_acc[];   // [FC][S]
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t j0 = g * kSidePix;
    const bool live = j0 < npix_det;    // VEC: npix_det % 8 == 0
    int32_t bin[kSidePix], nds[kSidePix], th[kSidePix];
    double w[kSidePix];
#pragma unroll
    for (int k = 0; k < kSidePix; k++) {
        const uint32_t j = j0 + (uint32_t)k;
        const int32_t c = j < npix_det ? pix2c[j] : -1;
        if (c >= 0) {
            bin[k] = sbin_of_c[c] - 1;
            nds[k] = dark_sum ? -(int32_t)dark_sum[c] : 0;
            const int64_t t = thr ? thr[c] : 0;
            th[k] = t < 0 ? 0 : (t > INT32_MAX ? INT32_MAX : (int32_t)t);
            w[k] = weight ? weight[c] : 1.0;
        } else {
            bin[k] = -1;
            nds[k] = 0;
            th[k] = INT32_MAX;   // no value passes
            w[k] = 0.0;
        }
    }
    uint32_t brk = 0;   // bit k: pixel k starts a new run inside the lane
#pragma unroll
    for (int k = 1; k < kSidePix; k++)
        if (bin[k] != bin[k - 1]) brk |= 1u << k;
    // segments of equal last-run bins across the lanes
    const int32_t lb = bin[kSidePix - 1];
    const int32_t lb_prev = (int32_t)prev_lane_u32((uint32_t)lb, 0xFFFFFFFEu);
    const int32_t lb_next = (int32_t)next_lane_u32((uint32_t)lb, 0xFFFFFFFEu);
    const uint32_t start = wave_incl_max_u32(lane == 0 || lb != lb_prev ? lane : 0u);
    const bool tail = lane == 63 || lb != lb_next;

    const int32_t nchunks = (F + FC - 1) / FC;
    for (int32_t ch = blockIdx.y; ch < nchunks; ch += gridDim.y) {
        const int32_t f0 = ch * FC;
        const int32_t nf = min(FC, F - f0);
        for (int i = threadIdx.x; i < FC * S; i += 256) s_acc[i] = 0.0;
        __syncthreads();
        auto frame = [&](const uint32_t (&x)[kSidePix], int32_t fl) {
            double* row = s_acc + (int64_t)fl * S;
            if constexpr (INT) {
                uint32_t run = 0u;
#pragma unroll
                for (int k = 0; k < kSidePix; k++) {
                    if (k > 0 && ((brk >> k) & 1u)) {
                        if (bin[k - 1] >= 0 && run) atomicAdd(&row[bin[k - 1]], (double)run);
                        run = 0u;
                    }
                    run += x[k];
                }
                run = wave_seg_scan_u32(run, start);
                if (tail && lb >= 0 && run) atomicAdd(&row[lb], (double)run);
                return;
            }
            double run = 0.0;
#pragma unroll
            for (int k = 0; k < kSidePix; k++) {
                if (k > 0 && ((brk >> k) & 1u)) {
                    if (bin[k - 1] >= 0 && run != 0.0) atomicAdd(&row[bin[k - 1]], run);
                    run = 0.0;
                }
                run += (double)x[k] * w[k];
            }
            run = wave_seg_scan_f64(run, start);
            if (tail && lb >= 0 && run != 0.0) atomicAdd(&row[lb], run);
        };
        auto decode = [&](const u32x4& r, uint32_t (&x)[kSidePix]) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                x[2 * q] = dense_value(r[q] & 0xFFFFu, nscale, nds[2 * q], th[2 * q]);
                x[2 * q + 1] = dense_value(r[q] >> 16, nscale, nds[2 * q + 1], th[2 * q + 1]);
            }
        };
        if constexpr (VEC) {
            if (avg == 1) {
                auto issue = [&](u32x4& r, int32_t fl) {
                    const int32_t f = f0 + min(fl, nf - 1);
                    // each frame is streamed once: nontemporal (about 1 % faster)
                    r = live ? __builtin_nontemporal_load(
                                   reinterpret_cast<const u32x4*>(imm + sel_off[f] + 2 * (int64_t)j0))
                             : u32x4{0, 0, 0, 0};
                };
                u32x4 ring[kSideRing];
#pragma unroll
                for (int d = 0; d < kSideRing; d++) issue(ring[d], d);
                for (int32_t fb = 0; fb < nf; fb += kSideRing) {
#pragma unroll
                    for (int d = 0; d < kSideRing; d++) {
                        if (fb + d < nf) {
                            uint32_t x[kSidePix];
                            decode(ring[d], x);
                            issue(ring[d], fb + d + kSideRing);
                            frame(x, fb + d);
                        }
                    }
                }
            } else {
                for (int32_t fl = 0; fl < nf; fl++) {
                    uint32_t x[kSidePix] = {};
                    for (int a = 0; a < avg; a++) {
                        const u32x4 r = live ? *reinterpret_cast<const u32x4*>(
                                                   imm + sel_off[(int64_t)(f0 + fl) * avg + a] + 2 * (int64_t)j0)
                                             : u32x4{0, 0, 0, 0};
                        uint32_t y[kSidePix];
                        decode(r, y);
#pragma unroll
                        for (int k = 0; k < kSidePix; k++) x[k] += y[k];
                    }
                    frame(x, fl);
                }
            }
        } else {
            for (int32_t fl = 0; fl < nf; fl++) {
                uint32_t x[kSidePix] = {};
                for (int a = 0; a < avg; a++) {
                    const uint16_t* fr =
                        reinterpret_cast<const uint16_t*>(imm + sel_off[(int64_t)(f0 + fl) * avg + a]);
#pragma unroll
                    for (int k = 0; k < kSidePix; k++)
                        if (j0 + (uint32_t)k < npix_det) x[k] += dense_value(fr[j0 + k], nscale, nds[k], th[k]);
                }
                frame(x, fl);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nf * S; i += 256)
            if (s_acc[i] != 0.0) atomicAdd(&fs[(int64_t)f0 * S + i], s_acc[i]);
        __syncthreads();
    }
}

// ---- host side --------------------------------------------------------------------

bool dense_stream_supported(const HostTables& h)
{
    return h.dpl == 4 || h.dpl == 8;
}

bool dense_levels_allowed(int dpl, int k)
{
    return dpl == 4 ? (k >= 3 && k <= 5) : (dpl == 8 && (k == 2 || k == 3));
}

int dense_levels(const xpcs_mt_plan* p)
{
    const int k = p->tune.dense_levels;
    if (k > 0 && dense_levels_allowed(p->h.dpl, k)) return k;
    return p->h.dpl == 4 ? dense_levels_per_pass<4>() : dense_levels_per_pass<8>();
}

size_t dense_series_elems(const xpcs_mt_plan* p)
{
    const HostTables& h = p->h;
    const size_t K = (size_t)dense_levels(p);
    // 32-bit words: u32 level values, or FP64 frame-weighted values (2 words)
    const size_t words = p->fsnorm ? 2 : 1;
    return h.levels.size() > K ? words * (size_t)h.levels[K].lf * (size_t)h.W * (size_t)h.H : 1;
}

template <int DPL, int K, bool FP>
static void run_passes(xpcs_mt_plan* p, hipStream_t s)
{
    using ST = std::conditional_t<FP, double, uint32_t>;
    const HostTables& h = p->h;
    const int nlev = (int)h.levels.size();
    const uint32_t npd = (uint32_t)(h.W * h.H);
    const unsigned grid = (npd + 255) / 256;
    uint32_t* redo_count = p->d_counters + 3;
    ST* const sa = reinterpret_cast<ST*>(p->d_series_a);
    ST* const sb = reinterpret_cast<ST*>(p->d_series_b);
    const double* fw = FP ? p->d_fw : nullptr;
    int L0 = 0;
    uint32_t n_in = (uint32_t)h.F;
    int pass = 0;
    while (L0 < nlev) {
        const int k = std::min(K, nlev - L0);
        const bool more = L0 + k < nlev;
        const uint32_t lf_out = more ? (uint32_t)h.levels[L0 + k].lf : 0u;
        ST* out = more ? (pass % 2 == 0 ? sa : sb) : nullptr;
        const ST* in = pass == 0 ? nullptr : (pass % 2 == 1 ? sa : sb);
        if (pass == 0) {
            timer_begin(p, s, FP ? "k_mt_dense_raw_f64" : "k_mt_dense_raw");
            if (p->tile_off0 >= 0)
                k_mt_dense_f64<DPL, K, kRawTiles, FP><<<grid, 256, 0, s>>>(
                    p->d_imm, p->d_sel_off, p->tile_off0, p->tile_step, h.avg, p->d_pix2c, npd, p->d_dark_sum,
                    p->d_thr, (int32_t)p->nscale, nullptr, n_in, p->d_levels, L0, k, p->d_bnd, p->nbnd, h.T,
                    p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, out, lf_out, redo_count, p->d_long_list, fw, !FP && p->ipif32);
            else
                k_mt_dense_f64<DPL, K, kRawLanes, FP><<<grid, 256, 0, s>>>(
                    p->d_imm, p->d_sel_off, 0, 0, h.avg, p->d_pix2c, npd, p->d_dark_sum, p->d_thr,
                    (int32_t)p->nscale, nullptr, n_in, p->d_levels, L0, k, p->d_bnd, p->nbnd, h.T, p->d_G2,
                    p->d_IP, p->d_IF, p->d_pix_total, out, lf_out, redo_count, p->d_long_list, fw, !FP && p->ipif32);
            timer_end(p, s);
            if constexpr (!FP) {
                timer_begin(p, s, "k_mt_dense_exact");
                k_mt_dense_exact<DPL, K, true><<<256, 256, 0, s>>>(
                    p->d_imm, p->d_sel_off, h.avg, p->d_pix2c, npd, p->d_dark_sum, p->d_thr, (int32_t)p->nscale,
                    nullptr, n_in, p->d_levels, L0, k, h.T, p->d_G2, redo_count, p->d_long_list);
                timer_end(p, s);
            }
        } else {
            timer_begin(p, s, FP ? "k_mt_dense_series_f64" : "k_mt_dense_series");
            k_mt_dense_f64<DPL, K, kSeries, FP><<<grid, 256, 0, s>>>(
                nullptr, nullptr, 0, 0, 1, p->d_pix2c, npd, nullptr, nullptr, 1, in, n_in, p->d_levels, L0, k,
                nullptr, 0, h.T, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, out, lf_out, nullptr, nullptr, fw,
                !FP && p->ipif32);
            timer_end(p, s);
            if constexpr (!FP) {
                timer_begin(p, s, "k_mt_dense_exact");
                k_mt_dense_exact<DPL, K, false><<<256, 256, 0, s>>>(
                    nullptr, nullptr, 1, p->d_pix2c, npd, nullptr, nullptr, 1, in, n_in, p->d_levels, L0, k, h.T,
                    p->d_G2, redo_count, p->d_long_list);
                timer_end(p, s);
            }
        }
        n_in = lf_out;
        L0 += k;
        pass++;
    }
}

template <bool FP>
static void launch_dense_passes(xpcs_mt_plan* p, hipStream_t s)
{
    const int k = dense_levels(p);
    if (p->h.dpl == 4) {
        if (k == 3) run_passes<4, 3, FP>(p, s);
        else if (k == 5) run_passes<4, 5, FP>(p, s);
        else run_passes<4, 4, FP>(p, s);
    } else {
        if (k == 2) run_passes<8, 2, FP>(p, s);
        else run_passes<8, 3, FP>(p, s);
    }
}

void launch_mt_dense(xpcs_mt_plan* p, hipStream_t s, bool fsnorm)
{
    if (fsnorm) launch_dense_passes<true>(p, s);
    else launch_dense_passes<false>(p, s);
}

void launch_side_dense_raw(xpcs_mt_plan* p, hipStream_t s)
{
    const HostTables& h = p->h;
    const uint32_t npd = (uint32_t)(h.W * h.H);
    const uint32_t groups = (npd + kSidePix - 1) / kSidePix;
    const int32_t fc = std::max(1, std::min(64, kSideLdsDoubles / std::max(h.ns, 1)));
    const int32_t nchunks = (h.F + fc - 1) / fc;
    const dim3 grid((groups + 255) / 256, (unsigned)std::min(nchunks, 65535));
    const size_t lds = (size_t)fc * (size_t)h.ns * sizeof(double);
    timer_begin(p, s, "k_side");
    auto go = [&](auto vec, auto integer) {
        k_side_dense_raw<decltype(vec)::value, decltype(integer)::value><<<grid, 256, lds, s>>>(
            p->d_imm, p->d_sel_off, h.avg, h.F, p->d_pix2c, npd, p->d_dark_sum, p->d_thr, (int32_t)p->nscale,
            p->d_sbin_of_c, p->d_weight, h.ns, fc, p->d_fs);
    };
    const bool integer = !p->d_weight && h.F >= 512;
    if (p->side_vec) {
        if (integer) go(std::true_type{}, std::true_type{});
        else go(std::true_type{}, std::false_type{});
    } else {
        if (integer) go(std::false_type{}, std::true_type{});
        else go(std::false_type{}, std::false_type{});
    }
    timer_end(p, s);
}

}  // namespace xpcs
