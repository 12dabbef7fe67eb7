// (b)+(c) Multi-tau binning and lag sums (Corr::multiTau2,
// src/xpcs/corr.cpp:315-431), exact in integers.
//
// Level L frame k holds x_L[k] = x_{L-1}[2k] + x_{L-1}[2k+1], k < lf_L =
// floor(F/2^L).  For a schedule entry with lag tau' = tau >> L and
// N = lf_L - tau':
//   G2 = sum_{t<N} x_L[t] x_L[t+tau'],
//   IP = sum_{t<N} x_L[t]       = S0(2^L (lf_L - tau')),
//   IF = sum_{tau'<=t<lf_L} x_L[t] = S0(2^L lf_L) - S0(2^L tau'),
// where S0(b) is the sum of the level-0 counts of frames < b.
//
// k_mt_reg / k_mt_reg2: one wavefront per row of at most 64 / 128 events held
// in registers (one / two per lane).  The row is sorted (odd-even passes, a
// bitonic network when still unsorted), duplicates merged, IP/IF taken from
// the level-0 prefix sums at the plan's frame boundaries (per-frame boundary
// ranks, one running maximum per row), every level binned with ballot/permute
// compaction, and lag partners found by shifting the row along the lanes.
//
// k_mt_sparse: one wavefront per pixel row of at most CAP events.  The row is
// sorted on chip (register bitonic network up to 64 events, LDS network
// beyond), duplicates merged, IP/IF read off the level-0 prefix sums by binary
// search, and every level is binned in place in LDS (ballot/mbcnt
// compaction).  Pairs at lag tau' are found by looking ahead at most
// hi = max tau' entries; their products go to per-wave LDS accumulators.
//
// k_mt_long: one workgroup per row that is too long for LDS (and every row of
// the dense path).  The row becomes a dense series x_0[0..F) in HBM scratch;
// each level is a coalesced sweep with per-thread lag accumulators, then an
// in-place pairwise binning pass.
#include "device_util.h"
#include "internal.h"

namespace xpcs {

// In-place compaction of one level: keys are shifted by `shift`, entries with
// key >= lf are dropped, runs of equal keys are summed.  Returns the new count.
// Reads of a chunk complete before its writes, and writes only go to slots at
// or below the slot being read, so the pass is safe in place.  Each lane reads
// its entry and the next one together (runs of two, the rule after level 0,
// need no further read; longer runs, repeated frames at level 0, continue
// from there), and the neighbour keys move by DPP, not through LDS.
template <typename V>
__device__ int level_compact(uint32_t* key, V* val, int n, int shift, uint32_t lf)
{
    const int lane = lane_id();
    int carry = 0;
    uint32_t prevk = 0xFFFFFFFFu;
    for (int b0 = 0; b0 < n; b0 += kWave) {
        const int i = b0 + lane;
        const bool valid = i < n;
        const bool has_next = i + 1 < n;
        const uint32_t kr = valid ? key[i] : 0u;
        const V vi = valid ? val[i] : V(0);
        const uint32_t knr = has_next ? key[i + 1] : 0u;
        const V vn = has_next ? val[i + 1] : V(0);
        const uint32_t k = valid ? (kr >> shift) : 0xFFFFFFFFu;
        const uint32_t kp = prev_lane_u32(k, prevk);
        const bool head = valid && k < lf && k != kp;
        V sum = vi;
        if (head && has_next && (knr >> shift) == k) {
            sum += vn;
            for (int j = i + 2; j < n && (key[j] >> shift) == k; j++) sum += val[j];
        }
        prevk = read_lane_u32(k, kWave - 1);
        const uint64_t mask = __ballot(head);
        const int pos = carry + mask_rank(mask);
        const bool stop = __ballot(valid && k >= lf) != 0;
        wave_sync();
        if (head) {
            key[pos] = k;
            val[pos] = sum;
        }
        carry += __popcll(mask);
        wave_sync();
        if (stop) break;
    }
    return carry;
}

// ---- register path: rows of at most 64 events, one event per lane ------------------
// The row kernels run on counts (V = uint32_t, exact 64-bit lag sums) or, for
// normalize_by_framesum, on frame-weighted values (V = double, FP64 sums).
constexpr int kRegMaxTaus = 128;
constexpr int kMidSmall = 128;   // LDS rows up to this length run with 128-entry wave slices
constexpr int kLagStep = 4;      // lag-search candidates read per step in the LDS kernel
constexpr uint32_t kNoKey = 0xFFFFFFFFu;
// Lag accumulators are kept in kAccCopies interleaved copies per tau; lane l
// adds into copy l % kAccCopies, so the lanes of a wave that hit the same lag
// spread over banks instead of serialising on one address.
constexpr int kAccCopies = 4;
static_assert(kAccCopies >= 1 && (kAccCopies & (kAccCopies - 1)) == 0, "copies: a power of two");
// the accumulator slots of a wave first hold the prefix sums at the (at most
// 3 per tau) frame boundaries
static_assert(3 * kRegMaxTaus <= kRegMaxTaus * kAccCopies, "boundary slots");

// a sum as stored in the int64 [P][T] buffers: integers as such, FP64 as bits
__device__ __forceinline__ int64_t sum_bits(uint64_t v) { return (int64_t)v; }
__device__ __forceinline__ int64_t sum_bits(unsigned long long v) { return (int64_t)v; }
__device__ __forceinline__ int64_t sum_bits(uint32_t v) { return (int64_t)v; }
__device__ __forceinline__ int64_t sum_bits(double v) { return (int64_t)__double_as_longlong(v); }

// Cross-lane moves of a value: one 32-bit move for counts, two for FP64.
__device__ __forceinline__ uint32_t lo_half(double x) { return (uint32_t)(uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ uint32_t hi_half(double x) { return (uint32_t)((uint64_t)__double_as_longlong(x) >> 32); }
__device__ __forceinline__ double from_halves(uint32_t lo, uint32_t hi)
{
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint32_t next_lane_v(uint32_t x, uint32_t old) { return next_lane_u32(x, old); }
__device__ __forceinline__ double next_lane_v(double x, double old)
{
    return from_halves(next_lane_u32(lo_half(x), lo_half(old)), next_lane_u32(hi_half(x), hi_half(old)));
}
__device__ __forceinline__ uint32_t prev_lane_v(uint32_t x, uint32_t old) { return prev_lane_u32(x, old); }
__device__ __forceinline__ double prev_lane_v(double x, double old)
{
    return from_halves(prev_lane_u32(lo_half(x), lo_half(old)), prev_lane_u32(hi_half(x), hi_half(old)));
}
template <int J>
__device__ __forceinline__ uint32_t xor_lane_v(uint32_t x) { return xor_lane_u32<J>(x); }
template <int J>
__device__ __forceinline__ double xor_lane_v(double x)
{
    return from_halves(xor_lane_u32<J>(lo_half(x)), xor_lane_u32<J>(hi_half(x)));
}
__device__ __forceinline__ uint32_t push_lane_v(int dst, uint32_t x) { return push_to_lane(dst, x); }
__device__ __forceinline__ double push_lane_v(int dst, double x)
{
    return from_halves(push_to_lane(dst, lo_half(x)), push_to_lane(dst, hi_half(x)));
}
// Value of lane i+1, zero in lane 63 (DPP bound_ctrl: no old value to copy in).
__device__ __forceinline__ uint32_t next_lane_or_zero(uint32_t x) { return dpp_u32<kDppWaveShl1, 0xF, true>(0u, x); }
__device__ __forceinline__ double next_lane_or_zero(double x)
{
    return from_halves(next_lane_or_zero(lo_half(x)), next_lane_or_zero(hi_half(x)));
}
// Lanes (as a mask) whose x <= b, unsigned.
__device__ __forceinline__ uint64_t lanes_le(uint32_t x, uint32_t b) { return __builtin_amdgcn_uicmp(x, b, 37); }
// The first n lanes as a mask (n <= 64; 0 for n <= 0).
__device__ __forceinline__ uint64_t first_lanes(int n)
{
    return n >= kWave ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1ull));
}
__device__ __forceinline__ uint32_t read_lane_v(uint32_t x, int l) { return read_lane_u32(x, l); }
__device__ __forceinline__ double read_lane_v(double x, int l)
{
    return from_halves(read_lane_u32(lo_half(x), l), read_lane_u32(hi_half(x), l));
}
__device__ __forceinline__ uint32_t shfl_v(uint32_t x, int src) { return (uint32_t)__shfl((int)x, src, 64); }
__device__ __forceinline__ double shfl_v(double x, int src) { return __shfl(x, src, 64); }
__device__ __forceinline__ uint32_t wave_scan_v(uint32_t x) { return wave_incl_scan_u32(x); }
__device__ __forceinline__ double wave_scan_v(double x)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    return x;
}
// inclusive scans of the 64-bit accumulator type (prefix sums of a row)
__device__ __forceinline__ unsigned long long wave_scan_acc(unsigned long long x) { return wave_incl_scan_u64(x); }
__device__ __forceinline__ double wave_scan_acc(double x) { return wave_scan_v(x); }
// lag products: exact 64-bit integers for counts, FP64 for weighted values
__device__ __forceinline__ unsigned long long lag_prod(uint32_t a, uint32_t b)
{
    return (unsigned long long)a * (unsigned long long)b;
}
__device__ __forceinline__ double lag_prod(double a, double b) { return a * b; }
template <typename V>
struct LagAcc { using type = unsigned long long; };
template <>
struct LagAcc<double> { using type = double; };

template <int J, typename V>
__device__ __forceinline__ void bitonic_step(uint32_t& k, V& v, int kk)
{
    const uint32_t pk = xor_lane_u32<J>(k);
    const V pv = xor_lane_v<J>(v);
    const int lane = lane_id();
    const bool up = (lane & kk) == 0;
    const bool lower = (lane & J) == 0;
    const bool take = (lower == up) ? (pk < k) : (pk > k);
    if (take) {
        k = pk;
        v = pv;
    }
}

// Ascending sort of (k, v) over the first m lanes (m a power of two <= 64;
// lanes >= m hold kNoKey).
template <typename V>
__device__ __forceinline__ void bitonic_sort64(uint32_t& k, V& v, int m)
{
    if (m >= 2) { bitonic_step<1>(k, v, 2); }
    if (m >= 4) { bitonic_step<2>(k, v, 4); bitonic_step<1>(k, v, 4); }
    if (m >= 8) { bitonic_step<4>(k, v, 8); bitonic_step<2>(k, v, 8); bitonic_step<1>(k, v, 8); }
    if (m >= 16) {
        bitonic_step<8>(k, v, 16); bitonic_step<4>(k, v, 16);
        bitonic_step<2>(k, v, 16); bitonic_step<1>(k, v, 16);
    }
    if (m >= 32) {
        bitonic_step<16>(k, v, 32); bitonic_step<8>(k, v, 32); bitonic_step<4>(k, v, 32);
        bitonic_step<2>(k, v, 32); bitonic_step<1>(k, v, 32);
    }
    if (m >= 64) {
        bitonic_step<32>(k, v, 64); bitonic_step<16>(k, v, 64); bitonic_step<8>(k, v, 64);
        bitonic_step<4>(k, v, 64); bitonic_step<2>(k, v, 64); bitonic_step<1>(k, v, 64);
    }
}

// Rows leave the transpose in frame order up to local shuffles (events of one
// pixel placed concurrently).  A few odd-even transposition rounds sort such a
// row; the bitonic network handles whatever is still out of order.  Lanes
// past the row hold kNoKey, so the order test needs no bound.
__device__ __forceinline__ bool reg_unsorted(uint32_t k)
{
    return __ballot(next_lane_u32(k, kNoKey) < k) != 0ull;
}

template <typename V>
__device__ __forceinline__ void reg_sort(uint32_t& k, V& v, int n)
{
    if (!reg_unsorted(k)) return;
    const int lane = lane_id();
    for (int r = 0; r < 3; r++) {
        // even phase: (2j, 2j+1)
        {
            const uint32_t pk = xor_lane_u32<1>(k);
            const V pv = xor_lane_v<1>(v);
            const bool lower = (lane & 1) == 0;
            if (lower ? (pk < k) : (pk > k)) {
                k = pk;
                v = pv;
            }
        }
        // odd phase: (2j+1, 2j+2)
        {
            const uint32_t nk = next_lane_u32(k, kNoKey);
            const V nv = next_lane_v(v, V(0));
            const uint32_t bk = prev_lane_u32(k, 0u);
            const V bv = prev_lane_v(v, V(0));
            const bool odd = (lane & 1) != 0;
            if (odd && nk < k) {
                k = nk;
                v = nv;
            } else if (!odd && lane > 0 && bk > k) {
                k = bk;
                v = bv;
            }
        }
        if (!reg_unsorted(k)) return;
    }
    bitonic_sort64(k, v, next_pow2(n));
}

// One binning step in registers: keys >> shift, keys >= lf dropped, equal keys
// summed.  `pairs_only` when runs are at most two long (every level after the
// first, whose keys are distinct).  Returns the new count; lanes past it hold
// (kNoKey, 0).  Without repeated or dropped keys no data moves.
template <typename V>
__device__ __forceinline__ int reg_level(uint32_t& k, V& v, int nl, int shift, uint32_t lf, bool pairs_only)
{
    const int lane = lane_id();
    const bool valid = lane < nl;
    const uint32_t kk = valid ? (k >> shift) : kNoKey;
    const uint32_t kp = prev_lane_u32(kk, kNoKey);
    const bool dup = valid && kk == kp;
    const bool drop = valid && kk >= lf;
    const uint64_t dupmask = __ballot(dup);
    const uint64_t dropmask = __ballot(drop);
    if (dupmask == 0 && dropmask == 0) {
        k = kk;
        return nl;
    }
    const bool head = valid && !dup && !drop;
    V sum;
    if (pairs_only) {
        const uint32_t kn = next_lane_u32(kk, kNoKey);
        const V vn = next_lane_v(v, V(0));
        sum = v + ((lane + 1 < nl && kn == kk) ? vn : V(0));
    } else {
        const uint64_t starts = __ballot(valid && !dup);
        const uint64_t above = lane == 63 ? 0ull : (starts & ~((2ull << lane) - 1ull));
        const int j = above ? (int)__builtin_ctzll(above) : nl;      // end of this lane's run
        if constexpr (sizeof(V) == 4) {
            // integer run sums as differences of the (modular) prefix sums
            const uint32_t pre = wave_incl_scan_u32(valid ? v : 0u);
            const uint32_t pj = shfl_v(pre, j - 1 < 0 ? 0 : j - 1);
            const uint32_t pi = prev_lane_u32(pre, 0u);
            sum = pj - pi;
        } else {
            // FP64: a segmented scan, so each run is summed on its own
            const uint64_t upto = lane == 63 ? starts : (starts & ((2ull << lane) - 1ull));
            const int st = upto ? 63 - (int)__builtin_clzll(upto) : 0;   // start of this lane's run
            V x = valid ? v : V(0);
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const V t = __shfl_up(x, d, 64);
                if (lane - d >= st) x += t;
            }
            sum = shfl_v(x, j - 1 < 0 ? 0 : j - 1);
        }
    }
    const uint64_t hm = __ballot(head);
    const int newn = __popcll(hm);
    const int dst = head ? mask_rank(hm) : 63;
    uint32_t k2 = push_to_lane(dst, kk);
    V v2 = push_lane_v(dst, sum);
    if (lane >= newn) {
        k2 = kNoKey;
        v2 = V(0);
    }
    k = k2;
    v = v2;
    return newn;
}

// Level-0 prefix sums at the plan's distinct frame boundaries B_0 < B_1 < ...
// (nb of them): the event that is last among those of equal rank r (r = number
// of B_j <= its frame) leaves its inclusive prefix sum in M[r]; prefix sums
// never decrease, so a running maximum over M gives, in M[j], the sum of the
// values of the frames below B_j.  M holds nb slots (cleared here).
template <typename V>
__device__ __forceinline__ void boundary_slots_clear(V* M, int nb)
{
    for (int i = lane_id(); i < nb; i += kWave) M[i] = V(0);
    wave_sync();
}

__device__ __forceinline__ uint32_t running_max(uint32_t x) { return wave_incl_max_u32(x); }
__device__ __forceinline__ double running_max(double x)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double t = __shfl_up(x, d, 64);
        if (lane >= d) x = fmax(x, t);
    }
    return x;
}

// Running maximum over M[0, nb), then IP = S(B_ip), IF = S(B_end) - S(B_if)
// for every tau (indices per tau in tau_bnd).
template <typename V>
__device__ __forceinline__ void boundary_sums(V* M, int nb, const int4* __restrict__ tau_bnd, int T,
                                              int64_t* __restrict__ IP, int64_t* __restrict__ IF, int64_t row0,
                                              bool w32)
{
    const int lane = lane_id();
    wave_sync();
    V carry = V(0);
    for (int c0 = 0; c0 < nb; c0 += kWave) {
        const int i = c0 + lane;
        V x = i < nb ? M[i] : V(0);
        x = running_max(x);
        x = x > carry ? x : carry;
        if (i < nb) M[i] = x;
        carry = read_lane_v(x, 63);
    }
    wave_sync();
    for (int i = lane; i < T; i += kWave) {
        const int4 j = tau_bnd[i];
        put_sum(IP, row0 + i, sum_bits(M[j.x]), w32);
        put_sum(IF, row0 + i, sum_bits((V)(M[j.z] - M[j.y])), w32);
    }
    wave_sync();
}

// Lag products of one level, one event per lane (nl events): the partner of
// lane i is lane i+m, m = 1..hi; products go to the LDS accumulators.  Keys
// are complemented so that the zero a shift brings in past lane 63 reads as
// kNoKey: a valid lane's partner beyond the row then has d > hi, and only the
// lanes past nl need masking.
template <typename V>
__device__ __forceinline__ void reg_lags(uint32_t k, V v, int nl, const LevelParam& lp,
                                         typename LagAcc<V>::type* acc)
{
    const int lane = lane_id();
    const uint32_t lo = (uint32_t)lp.lo, hi = (uint32_t)lp.hi, span = hi - lo;
    const uint64_t rows = first_lanes(nl);
    const bool valid = lane < nl;
    // slot of lag d: (first + d - lo) copies + this lane's copy
    const int slot0 = (lp.first - (int)lo) * kAccCopies + (lane & (kAccCopies - 1));
    const uint32_t kc = ~k;
    uint32_t kj = kc;
    V vj = v;
    // two partners per pass: one exit test and one loop test for both (keys
    // are distinct, so d >= m and the loop ends by m = hi + 1)
    for (int m = 0; m < kWave; m += 2) {
        const uint32_t kj1 = next_lane_or_zero(kj);
        const V vj1 = next_lane_or_zero(vj);
        const uint32_t d1 = kc - kj1;   // = key(i+m+1) - key(i)
        if ((lanes_le(d1, hi) & rows) == 0ull) break;
        kj = next_lane_or_zero(kj1);
        vj = next_lane_or_zero(vj1);
        const uint32_t d2 = kc - kj;
        if (valid && d1 - lo <= span) atomicAdd(&acc[slot0 + (int)d1 * kAccCopies], lag_prod(v, vj1));
        if (valid && d2 - lo <= span) atomicAdd(&acc[slot0 + (int)d2 * kAccCopies], lag_prod(v, vj));
    }
}

// ---- two events per lane: rows of 65..128 events --------------------------------
// Element i of the row is lane i % 64 of register i / 64 (k0/v0, k1/v1); the
// row is sorted across both.  Levels run in this form until binning leaves at
// most 64 events, then in the one-register form.

// Compare-exchange of elements i and i^J inside one register (J < 64) for a
// bitonic network; `up` gives the direction of the element's sequence.
template <int J, typename V>
__device__ __forceinline__ void bitonic_step_dir(uint32_t& k, V& v, bool up)
{
    const uint32_t pk = xor_lane_u32<J>(k);
    const V pv = xor_lane_v<J>(v);
    const bool lower = (lane_id() & J) == 0;
    const bool take = (lower == up) ? (pk < k) : (pk > k);
    if (take) {
        k = pk;
        v = pv;
    }
}

template <int KK, typename V>
__device__ __forceinline__ void bitonic_merge_in_reg(uint32_t& k, V& v, bool up)
{
    if constexpr (KK >= 64) bitonic_step_dir<32>(k, v, up);
    if constexpr (KK >= 32) bitonic_step_dir<16>(k, v, up);
    if constexpr (KK >= 16) bitonic_step_dir<8>(k, v, up);
    if constexpr (KK >= 8) bitonic_step_dir<4>(k, v, up);
    if constexpr (KK >= 4) bitonic_step_dir<2>(k, v, up);
    if constexpr (KK >= 2) bitonic_step_dir<1>(k, v, up);
}

// Ascending bitonic sort of the 128 elements of (k0, v0), (k1, v1).
template <typename V>
__device__ void bitonic_sort128(uint32_t& k0, V& v0, uint32_t& k1, V& v1)
{
    const int lane = lane_id();
    // stages kk = 2..64 inside each register; element i = lane + 64 r sorts
    // ascending when (i & kk) == 0 (for kk = 64 that is register 0)
    bitonic_merge_in_reg<2>(k0, v0, (lane & 2) == 0);
    bitonic_merge_in_reg<2>(k1, v1, (lane & 2) == 0);
    bitonic_merge_in_reg<4>(k0, v0, (lane & 4) == 0);
    bitonic_merge_in_reg<4>(k1, v1, (lane & 4) == 0);
    bitonic_merge_in_reg<8>(k0, v0, (lane & 8) == 0);
    bitonic_merge_in_reg<8>(k1, v1, (lane & 8) == 0);
    bitonic_merge_in_reg<16>(k0, v0, (lane & 16) == 0);
    bitonic_merge_in_reg<16>(k1, v1, (lane & 16) == 0);
    bitonic_merge_in_reg<32>(k0, v0, (lane & 32) == 0);
    bitonic_merge_in_reg<32>(k1, v1, (lane & 32) == 0);
    bitonic_merge_in_reg<64>(k0, v0, true);
    bitonic_merge_in_reg<64>(k1, v1, false);
    // kk = 128: elements i and i + 64 (same lane, other register), then inside
    if (k1 < k0) {
        const uint32_t tk = k0;
        const V tv = v0;
        k0 = k1;
        v0 = v1;
        k1 = tk;
        v1 = tv;
    }
    bitonic_merge_in_reg<64>(k0, v0, true);
    bitonic_merge_in_reg<64>(k1, v1, true);
}

__device__ __forceinline__ bool reg2_unsorted(uint32_t k0, uint32_t k1)
{
    const uint32_t first1 = read_lane_u32(k1, 0);
    return __ballot(next_lane_u32(k0, first1) < k0 || next_lane_u32(k1, kNoKey) < k1) != 0ull;
}

// Nearly sorted rows: up to 3 odd-even transposition rounds over the 128
// elements, the bitonic network for whatever is still out of order.
template <typename V>
__device__ void reg2_sort(uint32_t& k0, V& v0, uint32_t& k1, V& v1)
{
    if (!reg2_unsorted(k0, k1)) return;
    const int lane = lane_id();
    const bool odd = (lane & 1) != 0;
    for (int r = 0; r < 3; r++) {
        // even phase: (2j, 2j+1), inside each register
        {
            const uint32_t pk0 = xor_lane_u32<1>(k0), pk1 = xor_lane_u32<1>(k1);
            const V pv0 = xor_lane_v<1>(v0), pv1 = xor_lane_v<1>(v1);
            if (!odd ? (pk0 < k0) : (pk0 > k0)) {
                k0 = pk0;
                v0 = pv0;
            }
            if (!odd ? (pk1 < k1) : (pk1 > k1)) {
                k1 = pk1;
                v1 = pv1;
            }
        }
        // odd phase: (2j+1, 2j+2); element 63 pairs with element 64 across the registers
        {
            const uint32_t f1k = read_lane_u32(k1, 0), l0k = read_lane_u32(k0, 63);
            const V f1v = read_lane_v(v1, 0), l0v = read_lane_v(v0, 63);
            const uint32_t nk0 = next_lane_u32(k0, f1k), bk0 = prev_lane_u32(k0, 0u);
            const uint32_t nk1 = next_lane_u32(k1, kNoKey), bk1 = prev_lane_u32(k1, l0k);
            const V nv0 = next_lane_v(v0, f1v), bv0 = prev_lane_v(v0, V(0));
            const V nv1 = next_lane_v(v1, V(0)), bv1 = prev_lane_v(v1, l0v);
            if (odd && nk0 < k0) {
                k0 = nk0;
                v0 = nv0;
            } else if (!odd && lane > 0 && bk0 > k0) {
                k0 = bk0;
                v0 = bv0;
            }
            if (odd && nk1 < k1) {
                k1 = nk1;
                v1 = nv1;
            } else if (!odd && bk1 > k1) {
                k1 = bk1;
                v1 = bv1;
            }
        }
        if (!reg2_unsorted(k0, k1)) return;
    }
    bitonic_sort128(k0, v0, k1, v1);
}

// Staging slots of one wave for reg2_level: keys and values of the shifted
// row, then of the compacted row (2 x 64 entries each).
template <typename V>
constexpr int reg2_stage_words() { return 2 * 2 * kWave + 2 * 2 * kWave * (int)(sizeof(V) / 4); }

// One binning step over both registers (keys >> shift, keys >= lf dropped,
// equal keys summed), compacted through the wave's staging slots when
// anything merges or drops.  Returns the new count.
template <typename V>
__device__ int reg2_level(uint32_t& k0, V& v0, uint32_t& k1, V& v1, int nl, int shift, uint32_t lf,
                          uint32_t* __restrict__ stage)
{
    const int lane = lane_id();
    const bool valid0 = lane < nl, valid1 = lane + kWave < nl;
    const uint32_t kk0 = valid0 ? (k0 >> shift) : kNoKey;
    const uint32_t kk1 = valid1 ? (k1 >> shift) : kNoKey;
    const uint32_t l0 = read_lane_u32(kk0, 63);
    const uint32_t kp0 = prev_lane_u32(kk0, kNoKey), kp1 = prev_lane_u32(kk1, l0);
    const bool dup0 = valid0 && kk0 == kp0, dup1 = valid1 && kk1 == kp1;
    const bool drop0 = valid0 && kk0 >= lf, drop1 = valid1 && kk1 >= lf;
    if (__ballot(dup0 || dup1 || drop0 || drop1) == 0ull) {
        k0 = kk0;
        k1 = kk1;
        return nl;
    }
    // stage the shifted row, then every run head sums its run
    constexpr int kVW = 2 * kWave * (int)(sizeof(V) / 4);   // words of 128 values
    uint32_t* sk = stage;
    V* sv = reinterpret_cast<V*>(stage + 2 * kWave);
    uint32_t* ok = stage + 2 * kWave + kVW;
    V* ov = reinterpret_cast<V*>(stage + 4 * kWave + kVW);
    sk[lane] = kk0;
    sv[lane] = v0;
    sk[lane + kWave] = kk1;
    sv[lane + kWave] = v1;
    wave_sync();
    const bool head0 = valid0 && !dup0 && !drop0, head1 = valid1 && !dup1 && !drop1;
    V sum0 = V(0), sum1 = V(0);
    if (head0) {
        sum0 = v0;
        for (int j = lane + 1; j < nl && sk[j] == kk0; j++) sum0 += sv[j];
    }
    if (head1) {
        sum1 = v1;
        for (int j = lane + kWave + 1; j < nl && sk[j] == kk1; j++) sum1 += sv[j];
    }
    const uint64_t hm0 = __ballot(head0), hm1 = __ballot(head1);
    const int n0 = __popcll(hm0);
    const int newn = n0 + __popcll(hm1);
    if (head0) {
        const int d = mask_rank(hm0);
        ok[d] = kk0;
        ov[d] = sum0;
    }
    if (head1) {
        const int d = n0 + mask_rank(hm1);
        ok[d] = kk1;
        ov[d] = sum1;
    }
    wave_sync();
    k0 = lane < newn ? ok[lane] : kNoKey;
    v0 = lane < newn ? ov[lane] : V(0);
    k1 = lane + kWave < newn ? ok[lane + kWave] : kNoKey;
    v1 = lane + kWave < newn ? ov[lane + kWave] : V(0);
    wave_sync();
    return newn;
}

// Lag products of one level over both registers (partner of element i is
// i+m), complemented keys as in reg_lags.
template <typename V>
__device__ __forceinline__ void reg2_lags(uint32_t k0, V v0, uint32_t k1, V v1, int nl, const LevelParam& lp,
                                          typename LagAcc<V>::type* acc)
{
    const int lane = lane_id();
    const uint32_t lo = (uint32_t)lp.lo, hi = (uint32_t)lp.hi, span = hi - lo;
    const uint64_t rows0 = first_lanes(nl), rows1 = first_lanes(nl - kWave);
    const bool valid0 = lane < nl, valid1 = lane + kWave < nl;
    const int slot0 = (lp.first - (int)lo) * kAccCopies + (lane & (kAccCopies - 1));
    const uint32_t kc0 = ~k0, kc1 = ~k1;
    uint32_t kj0 = kc0, kj1 = kc1;
    V vj0 = v0, vj1 = v1;
    // two partners per pass, as in reg_lags
    for (int m = 0; m < 2 * kWave; m += 2) {
        uint32_t f1k = read_lane_u32(kj1, 0);
        V f1v = read_lane_v(vj1, 0);
        const uint32_t ka0 = next_lane_u32(kj0, f1k), ka1 = next_lane_or_zero(kj1);
        const V va0 = next_lane_v(vj0, f1v), va1 = next_lane_or_zero(vj1);
        const uint32_t da0 = kc0 - ka0, da1 = kc1 - ka1;
        if (((lanes_le(da0, hi) & rows0) | (lanes_le(da1, hi) & rows1)) == 0ull) break;
        f1k = read_lane_u32(ka1, 0);
        f1v = read_lane_v(va1, 0);
        kj0 = next_lane_u32(ka0, f1k);
        kj1 = next_lane_or_zero(ka1);
        vj0 = next_lane_v(va0, f1v);
        vj1 = next_lane_or_zero(va1);
        const uint32_t db0 = kc0 - kj0, db1 = kc1 - kj1;
        if (valid0 && da0 - lo <= span) atomicAdd(&acc[slot0 + (int)da0 * kAccCopies], lag_prod(v0, va0));
        if (valid1 && da1 - lo <= span) atomicAdd(&acc[slot0 + (int)da1 * kAccCopies], lag_prod(v1, va1));
        if (valid0 && db0 - lo <= span) atomicAdd(&acc[slot0 + (int)db0 * kAccCopies], lag_prod(v0, vj0));
        if (valid1 && db1 - lo <= span) atomicAdd(&acc[slot0 + (int)db1 * kAccCopies], lag_prod(v1, vj1));
    }
}

// Boundary rank of frame f (the number of prefix-sum boundaries at or below
// it): from the per-64-frame words staged in LDS (bits of the boundary frames,
// rank before the word) when the launch gave them room, else gathered from the
// per-frame table (one dependent global load less per row with the words).
__device__ __forceinline__ uint32_t boundary_rank(const uint4* s_rw, bool staged,
                                                  const uint16_t* __restrict__ frame_rank, uint32_t f)
{
    if (staged) {
        const uint4 q = s_rw[f >> 6];
        const uint64_t bits = ((uint64_t)q.y << 32) | q.x;
        return q.z + (uint32_t)__popcll(bits & (~0ull >> (63u - (f & 63u))));
    }
    return (uint32_t)frame_rank[f];
}

// The level-0 value of an event: its count, or count x fw[frame].
template <typename V>
__device__ __forceinline__ V event_value(uint32_t c, uint32_t f, const double* __restrict__ fw, int32_t F)
{
    if constexpr (sizeof(V) == 4) {
        return c;
    } else {
        return f < (uint32_t)F ? (double)c * fw[f] : 0.0;
    }
}

// Hand-over of the deep levels to k_mt_deep: the series of the first level
// with at most kDeepMaxFrames binned frames, written by the register kernels
// as one dense row of `stride` u32 per pixel ([P][stride]).
constexpr int kDeepMaxFrames = 96;
constexpr int kDeepMaxTaus = 48;     // deep G2 staged through the series area: 48 x 8 B fit a pixel's 97 words
constexpr int kDeepMaxLevels = 6;
constexpr int kDeepLanes = 4;        // lanes per pixel in k_mt_deep (divides 4, the smallest dpl)

// One row's level series (keys < stride, distinct) as a dense row at dst,
// staged through the wave's LDS words ser[0, stride).
__device__ __forceinline__ void write_deep_row(uint32_t* ser, int stride, uint32_t k0, uint32_t v0, bool has0,
                                               uint32_t k1, uint32_t v1, bool has1, uint32_t* __restrict__ dst)
{
    const int lane = lane_id();
    for (int t = lane; t < stride; t += kWave) ser[t] = 0u;
    wave_sync();
    if (has0) ser[k0] = v0;
    if (has1) ser[k1] = v1;
    wave_sync();
    for (int t = lane; t < stride; t += kWave) dst[t] = ser[t];
    wave_sync();
}

// Rows r < *count of list[r] with 65..128 events (T <= kRegMaxTaus).  V is
// uint32_t (counts) or double (counts x fw[frame], normalize_by_framesum,
// main.cpp:313-336; pix_total stays the raw count total).
template <int WAVES, typename V>
__global__ void __launch_bounds__(kWave * WAVES) __attribute__((amdgpu_num_sgpr(88)))
k_mt_reg2(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off, const int32_t* __restrict__ list,
          const uint32_t* __restrict__ count, int32_t T, const int4* __restrict__ tau_bnd,
          const uint16_t* __restrict__ frame_rank, int32_t nb, const LevelParam* __restrict__ levels,
          int32_t nlev, int32_t F, int64_t* __restrict__ G2, int64_t* __restrict__ IP, int64_t* __restrict__ IF,
          uint64_t* __restrict__ pix_total, const int32_t* __restrict__ rowmap, const double* __restrict__ fw,
          int32_t ldeep, uint32_t* __restrict__ deep, int32_t dstride, bool w32, int32_t acc_words)
{
    using A = typename LagAcc<V>::type;
    // each wave's accumulators (acc_words entries: the taus before the
    // hand-over in kAccCopies copies, or the boundary slots if more) in
    // dynamic LDS, sized by the launch to what the schedule needs
    extern __shared__ __align__(16) unsigned char s_acc_dyn[];
    __shared__ uint32_t s_stage[WAVES][reg2_stage_words<V>()];
    const int tdeep = ldeep < nlev ? levels[ldeep].first : T;   // taus of the levels before the hand-over
    const int lane = lane_id();
    const int w = wave_id();
    A* acc = reinterpret_cast<A*>(s_acc_dyn) + (int64_t)w * acc_words;
    uint32_t* stage = s_stage[w];
    const int64_t nrows = (int64_t)*count;
    const int64_t nw = (int64_t)gridDim.x * WAVES;
    for (int64_t r = (int64_t)blockIdx.x * WAVES + w; r < nrows; r += nw) {
        const int64_t p = list[r];
        const int64_t row = rowmap ? (int64_t)rowmap[p] : p;
        const int64_t base = off[row];
        const int n = (int)min<int64_t>(off[row + 1] - base, 2 * kWave);   // 65..128 by construction
        int64_t* g2row = G2 + p * T;
        uint32_t k0 = kNoKey, k1 = kNoKey, c0 = 0u, c1 = 0u;
        if (lane < n) {
            const uint64_t e = ev[base + lane];
            k0 = (uint32_t)(e >> 32);
            c0 = (uint32_t)e;
        }
        if (lane + kWave < n) {
            const uint64_t e = ev[base + kWave + lane];
            k1 = (uint32_t)(e >> 32);
            c1 = (uint32_t)e;
        }
        V v0 = event_value<V>(c0, k0, fw, F), v1 = event_value<V>(c1, k1, fw, F);
        if constexpr (sizeof(V) != 4) {
            const uint32_t tot = read_lane_u32(wave_incl_scan_u32(c0 + c1), 63);
            if (lane == 0) pix_total[p] = tot;
        }
        reg2_sort(k0, v0, k1, v1);
        int nl = reg2_level(k0, v0, k1, v1, n, 0, (uint32_t)F, stage);
        // IP and IF from the level-0 prefix sums at the boundaries (the
        // accumulator slots serve as the boundary slots first)
        {
            V* M = reinterpret_cast<V*>(acc);
            boundary_slots_clear(M, nb);
            const V pre0 = wave_scan_v(v0);
            const V pre1 = wave_scan_v(v1) + read_lane_v(pre0, 63);
            if constexpr (sizeof(V) == 4) {
                if (lane == 0) pix_total[p] = read_lane_u32(pre1, 63);
            }
            const uint32_t r0 = lane < nl ? (uint32_t)frame_rank[k0] : kNoKey;
            const uint32_t r1 = lane + kWave < nl ? (uint32_t)frame_rank[k1] : kNoKey;
            const uint32_t rn0 = next_lane_u32(r0, read_lane_u32(r1, 0)), rn1 = next_lane_u32(r1, kNoKey);
            if (lane < nl && r0 != rn0 && r0 < (uint32_t)nb) M[r0] = pre0;
            if (lane + kWave < nl && r1 != rn1 && r1 < (uint32_t)nb) M[r1] = pre1;
            boundary_sums(M, nb, tau_bnd, T, IP, IF, p * T, w32);
        }
        for (int i = lane; i < tdeep * kAccCopies; i += kWave) acc[i] = A(0);   // the lags the row kernel sums
        wave_sync();
        bool two = nl > kWave;
        for (int L = 0; L < nlev; L++) {
            const LevelParam lp = levels[L];
            if (L > 0) {
                if (two) nl = reg2_level(k0, v0, k1, v1, nl, 1, (uint32_t)lp.lf, stage);
                else nl = reg_level(k0, v0, nl, 1, (uint32_t)lp.lf, true);
                two = nl > kWave;
            }
            if constexpr (sizeof(V) == 4) {
                if (L == ldeep) {   // the staging slots are free between levels
                    write_deep_row(stage, dstride, k0, v0, lane < nl, k1, v1, two && lane + kWave < nl,
                                   deep + p * dstride);
                    break;
                }
            }
            if (nl < 2) {
                if (ldeep >= nlev) break;
                continue;   // no more pairs; bin what is left down to the hand-over level
            }
            if (two) reg2_lags(k0, v0, k1, v1, nl, lp, acc);
            else reg_lags(k0, v0, nl, lp, acc);
        }
        wave_sync();
        for (int i = lane; i < tdeep; i += kWave) {
            A g = A(0);
#pragma unroll
            for (int c = 0; c < kAccCopies; c++) g += acc[i * kAccCopies + c];
            g2row[i] = sum_bits(g);
        }
        wave_sync();
    }
}

// Row classes of the sparse path, listed by k_mt_classify for the kernels
// after k_mt_reg (list: counter slot, front or back):
//   k_mt_reg2 / k_mt_sparse<kMidSmall>  rows of 65..kMidSmall events (every
//                                       non-empty row up to kMidSmall when
//                                       T > kRegMaxTaus): mid_list front, [2]
//   k_mt_sparse<kSparseCapUsed>         up to kSparseCapUsed: mid_list back, [4]
//   k_mt_sparse<kSparseCapMid>          up to kSparseCapMid: long_list front, [5]
//   k_mt_sparse<kSparseCapLong>,        longer rows: long_list back, [0]; the
//   k_mt_long                           LDS kernel takes those up to
//                                       kSparseCapLong, the series kernel the rest
constexpr int kSparseCapUsed = 512;
constexpr int kSparseCapMid = 1024;
constexpr int kSparseCapLong = 2048;
constexpr int kClassifyPer = 8;      // pixels per thread of the classification pass

// kClassifyPer x 256 consecutive compact pixels per workgroup: each row's
// class, appended to its list.  Positions come from LDS counters of the
// workgroup (one LDS atomic per wave and class) and one global atomic per
// class and workgroup, so the shared counters see few atomics.  The lists'
// order does not matter: every row writes only its own sums.
__global__ void __launch_bounds__(256)
k_mt_classify(const int64_t* __restrict__ off, const int32_t* __restrict__ rowmap, int32_t P, bool regs,
              int32_t* __restrict__ mid_list, int32_t* __restrict__ long_list, uint32_t* __restrict__ counters)
{
    __shared__ uint32_t s_cnt[4];
    __shared__ uint32_t s_base[4];
    if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int lane = lane_id();
    const int64_t p0 = (int64_t)blockIdx.x * (256 * kClassifyPer) + threadIdx.x;
    int cls[kClassifyPer];
    uint32_t pos[kClassifyPer];
#pragma unroll
    for (int i = 0; i < kClassifyPer; i++) {
        const int64_t p = p0 + (int64_t)i * 256;
        cls[i] = -1;
        pos[i] = 0u;
        if (p < P) {
            const int64_t row = rowmap ? (int64_t)rowmap[p] : p;
            const int64_t n = off[row + 1] - off[row];
            const bool reg = n == 0 || (n <= kWave && regs);
            if (!reg) cls[i] = n <= kMidSmall ? 0 : n <= kSparseCapUsed ? 1 : n <= kSparseCapMid ? 2 : 3;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint64_t m = __ballot(cls[i] == c);
            if (m == 0ull) continue;
            const int leader = (int)__builtin_ctzll(m);
            uint32_t b = 0u;
            if (lane == leader) b = atomicAdd(&s_cnt[c], (uint32_t)__popcll(m));
            b = read_lane_u32(b, leader);
            if (cls[i] == c) pos[i] = b + (uint32_t)mask_rank(m);
        }
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int c = (int)threadIdx.x;
        s_base[c] = s_cnt[c] ? atomicAdd(&counters[c == 0 ? 2 : c == 1 ? 4 : c == 2 ? 5 : 0], s_cnt[c]) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kClassifyPer; i++) {
        const int c = cls[i];
        if (c < 0) continue;
        const int64_t p = p0 + (int64_t)i * 256;
        const int64_t q = (int64_t)s_base[c] + pos[i];
        if (c == 0) mid_list[q] = (int32_t)p;
        else if (c == 1) mid_list[P - 1 - q] = (int32_t)p;
        else if (c == 2) long_list[q] = (int32_t)p;
        else long_list[P - 1 - q] = (int32_t)p;
    }
}

// Rows of at most 64 events, one wave each (T <= kRegMaxTaus), and the empty
// rows; k_mt_classify lists the others for the kernels after this one.  V as
// for k_mt_reg2.
// 88 scalar registers leave room for 8 waves per SIMD (106 would allow 7)
template <int WAVES, typename V>
__global__ void __launch_bounds__(kWave * WAVES) __attribute__((amdgpu_num_sgpr(88)))
k_mt_reg(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off, int32_t P, int32_t T,
         const int4* __restrict__ tau_bnd, const uint16_t* __restrict__ frame_rank, int32_t nb,
         const LevelParam* __restrict__ levels, int32_t nlev,
         int32_t F, int64_t* __restrict__ G2, int64_t* __restrict__ IP, int64_t* __restrict__ IF,
         uint64_t* __restrict__ pix_total, const int32_t* __restrict__ rowmap, const double* __restrict__ fw,
         int32_t ldeep, uint32_t* __restrict__ deep, int32_t dstride, bool w32, bool reg_ok,
         const uint4* __restrict__ rank_words, int32_t nrw)
{
    using A = typename LagAcc<V>::type;
    __shared__ A s_acc[WAVES][kRegMaxTaus * kAccCopies];
    __shared__ uint32_t s_ser[WAVES][sizeof(V) == 4 ? kDeepMaxFrames : 1];
    __shared__ int4 s_tb[kRegMaxTaus];   // each tau's boundary indices, read by every row
    extern __shared__ uint4 s_rw[];      // nrw rank words (dynamic LDS)
    const int lane = lane_id();
    const int w = wave_id();
    for (int i = (int)threadIdx.x; i < T && i < kRegMaxTaus; i += kWave * WAVES) s_tb[i] = tau_bnd[i];
    for (int i = (int)threadIdx.x; i < nrw; i += kWave * WAVES) s_rw[i] = rank_words[i];
    __syncthreads();
    A* acc = s_acc[w];
    const int tdeep = ldeep < nlev ? levels[ldeep].first : T;   // taus of the levels before the hand-over
    const int64_t nw = (int64_t)gridDim.x * WAVES;

    for (int64_t p = (int64_t)blockIdx.x * WAVES + w; p < P; p += nw) {
        const int64_t row = rowmap ? (int64_t)rowmap[p] : p;   // CSR row of pixel p
        const int64_t base = off[row];
        const int64_t n64 = off[row + 1] - base;
        if (n64 > kWave || (!reg_ok && n64 > 0)) continue;   // listed by k_mt_classify
        int64_t* g2row = G2 + p * T;
        const int n = (int)n64;
        if (n == 0) {
            for (int i = lane; i < T; i += kWave) {
                g2row[i] = 0;
                put_sum(IP, p * T + i, 0, w32);
                put_sum(IF, p * T + i, 0, w32);
            }
            if (lane == 0) pix_total[p] = 0;
            continue;
        }
        uint32_t k = kNoKey, c = 0u;
        if (lane < n) {
            const uint64_t e = ev[base + lane];
            k = (uint32_t)(e >> 32);
            c = (uint32_t)e;
        }
        V v = event_value<V>(c, k, fw, F);
        if constexpr (sizeof(V) != 4) {
            const uint32_t tot = read_lane_u32(wave_incl_scan_u32(c), 63);
            if (lane == 0) pix_total[p] = tot;
        }
        // each event's boundary rank, gathered before the sort: rows usually
        // arrive in order, and then it needs no second gather
        uint32_t r = lane < n ? boundary_rank(s_rw, nrw > 0, frame_rank, k) : kNoKey;
        const bool moved = reg_unsorted(k);
        reg_sort(k, v, n);
        int nl = reg_level(k, v, n, 0, (uint32_t)F, false);
        if (moved || nl != n) r = lane < nl ? boundary_rank(s_rw, nrw > 0, frame_rank, k) : kNoKey;
        // IP and IF of every tau from the level-0 prefix sums:
        // IP = S0(2^L (lf - tau')), IF = S0(2^L lf) - S0(2^L tau'), read at the
        // boundaries (the accumulator slots serve as the boundary slots first)
        {
            V* M = reinterpret_cast<V*>(acc);
            boundary_slots_clear(M, nb);
            const V pre = wave_scan_v(v);
            if constexpr (sizeof(V) == 4) {
                if (lane == 0) pix_total[p] = read_lane_u32(pre, 63);
            }
            const uint32_t rn = next_lane_u32(r, kNoKey);
            if (lane < nl && r != rn && r < (uint32_t)nb) M[r] = pre;
            boundary_sums(M, nb, s_tb, T, IP, IF, p * T, w32);
        }
        for (int i = lane; i < tdeep * kAccCopies; i += kWave) acc[i] = A(0);   // the lags the row kernel sums
        wave_sync();
        for (int L = 0; L < nlev; L++) {
            const LevelParam lp = levels[L];
            if (L > 0) nl = reg_level(k, v, nl, 1, (uint32_t)lp.lf, true);
            if constexpr (sizeof(V) == 4) {
                if (L == ldeep) {
                    write_deep_row(s_ser[w], dstride, k, v, lane < nl, 0u, 0u, false, deep + p * dstride);
                    break;
                }
            }
            if (nl < 2) {
                if (ldeep >= nlev) break;
                continue;   // no more pairs; bin what is left down to the hand-over level
            }
            reg_lags(k, v, nl, lp, acc);
        }
        wave_sync();
        for (int i = lane; i < tdeep; i += kWave) {
            A g = A(0);
#pragma unroll
            for (int c = 0; c < kAccCopies; c++) g += acc[i * kAccCopies + c];
            g2row[i] = sum_bits(g);
        }
        wave_sync();
    }
}

// Levels ldeep .. nlev-1 of one pixel's series x (in place), 64-bit lag sums,
// shared by LPP lanes: lane h of the pixel sums lags dpl + 1 + h NE + e
// (e < NE = dpl / LPP).  Every lane indexes one register window of the last
// 2 dpl values with the same constants (modulo 2 dpl over an unrolled block);
// lane h's window is rotated by h NE (each value stored h NE slots further
// on), so the same slots give it values h NE frames older.  Only lane 0
// writes the pair sums x[t-1] + x[t] at odd t to x[t/2], the next level's
// series; the lanes are at the same t, so every overwritten value has been
// read by all of them.
template <int DPL, int LPP>
__device__ __forceinline__ void deep_levels(uint32_t* x, const LevelParam* __restrict__ levels, int ldeep, int nlev,
                                            int h, unsigned long long (&res)[kDeepMaxLevels][DPL / LPP])
{
    constexpr int W = 2 * DPL, NE = DPL / LPP;
#pragma unroll
    for (int j = 0; j < kDeepMaxLevels; j++) {
        unsigned long long acc[NE];
#pragma unroll
        for (int e = 0; e < NE; e++) acc[e] = 0ull;
        if (ldeep + j < nlev) {
            const int lf = levels[ldeep + j].lf;
            uint32_t buf[W];
#pragma unroll
            for (int u = 0; u < W; u++) buf[u] = 0u;
            uint32_t last = 0u;   // x[t-1] for the pair sums
            for (int t0 = 0; t0 < lf; t0 += W) {
#pragma unroll
                for (int u = 0; u < W; u++) {
                    const int t = t0 + u;
                    if (t < lf) {
                        const uint32_t val = x[t];
#pragma unroll
                        for (int e = 0; e < NE; e++)
                            acc[e] += (unsigned long long)buf[(u + W - (DPL + 1 + e)) % W] * val;
                        if ((u & 1) && h == 0) x[t >> 1] = last + val;
                        last = val;
#pragma unroll
                        for (int g = 0; g < LPP; g++) {
                            const int ur = (u + g * NE) % W;   // constants once the block is unrolled
                            buf[ur] = h == g ? val : buf[ur];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < NE; e++) res[j][e] = acc[e];
    }
}

// The deep levels of the rows the register kernels handled (1..kMidSmall
// events) and of the LDS kernels' rows (up to kSparseCapLong events) whose
// total fits 32 bits, kDeepLanes lanes per pixel: levels ldeep .. nlev-1,
// whose series are short, cost a wave per row one lag search each but only a
// few instructions per value here.  A wave takes 64 / kDeepLanes consecutive
// pixels: their series rows (contiguous in HBM, all 16-byte loads in flight)
// go to LDS (odd pitch 97; with four lanes per pixel 6 KB per wave, so seven
// waves share a SIMD, against 1.5 with one lane per pixel: 0.194 -> 0.163 ms
// with two lanes, 0.150 ms with four on the 1 Mpx job); the lanes of a
// pixel stream its series once per level, each a share of the level's lags
// (see deep_levels), and the deep taus' G2 leave through the same LDS area,
// one contiguous run per pixel.  Exact: u32 values, 64-bit integer sums
// (32-bit sums for waves of small totals measured no faster).
template <int DPL>
__global__ void __launch_bounds__(kWave)
k_mt_deep(const uint32_t* __restrict__ series, int32_t stride, const int64_t* __restrict__ off,
          const int32_t* __restrict__ rowmap, int32_t P, int32_t T, const LevelParam* __restrict__ levels,
          int32_t ldeep, int32_t nlev, const uint64_t* __restrict__ pix_total, int64_t* __restrict__ G2)
{
    constexpr int SP = kDeepMaxFrames + 1;   // odd pitch: the pixels' reads of one index hit distinct banks
    constexpr int LPP = kDeepLanes;
    constexpr int NPW = kWave / LPP;         // pixels per wave
    constexpr int NE = DPL / LPP;            // lags per lane and level
    __shared__ uint32_t s_x[NPW * SP];
    const int lane = lane_id();
    const int q = lane / LPP;                // the lanes' pixel
    const int h = lane % LPP;
    const int64_t p0 = (int64_t)blockIdx.x * NPW;
    const int np = (int)min<int64_t>(NPW, (int64_t)P - p0);
    {
        // the wave's np rows are one contiguous run of np * stride u32 (stride
        // % 4 == 0): every 16-byte load in flight at once, then into LDS
        typedef uint32_t v4u __attribute__((ext_vector_type(4)));
        constexpr int kV4 = NPW * (kDeepMaxFrames / 4) / kWave;   // 16-byte loads per lane at most
        const int s4 = stride >> 2;
        const int n4 = np * s4;
        const v4u* src = reinterpret_cast<const v4u*>(series + p0 * (int64_t)stride);
        v4u v[kV4];
#pragma unroll
        for (int k = 0; k < kV4; k++) {
            const int i = k * kWave + lane;
            if (i < n4) v[k] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int k = 0; k < kV4; k++) {
            const int i = k * kWave + lane;
            if (i < n4) {
                const int r = i / s4;
                uint32_t* d = s_x + r * SP + (i - r * s4) * 4;
                d[0] = v[k].x;
                d[1] = v[k].y;
                d[2] = v[k].z;
                d[3] = v[k].w;
            }
        }
    }
    bool mine = false;
    if (q < np) {
        const int64_t p = p0 + q;
        const int64_t row = rowmap ? (int64_t)rowmap[p] : p;
        const int64_t n = off[row + 1] - off[row];
        // the rows handed over: those of the register kernels (totals below
        // 2^32) and the LDS rows whose total fits 32 bits
        mine = n >= 1 && n <= kSparseCapLong && pix_total[p] <= 0xFFFFFFFFull;
    }
    wave_sync();
    uint32_t* x = s_x + min(q, NPW - 1) * SP;
    unsigned long long res[kDeepMaxLevels][NE];
    deep_levels<DPL, LPP>(x, levels, ldeep, nlev, h, res);
    wave_sync();
    // results to LDS (two u32 per value, odd pitch), then one run per pixel
    const int tdeep = levels[ldeep].first;
    const int nd = T - tdeep;
    const int gp = 2 * nd + 1;
#pragma unroll
    for (int j = 0; j < kDeepMaxLevels; j++) {
        if (ldeep + j >= nlev) continue;
        const LevelParam lp = levels[ldeep + j];
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const int ee = e + h * NE;
            if (ee < lp.count && q < np) {
                const int i = lp.first + ee - tdeep;
                s_x[q * gp + 2 * i] = (uint32_t)res[j][e];
                s_x[q * gp + 2 * i + 1] = (uint32_t)(res[j][e] >> 32);
            }
        }
    }
    wave_sync();
    const uint64_t mine_mask = __ballot(mine && h == 0);
    for (int r = 0; r < np; r++) {
        if (((mine_mask >> (LPP * r)) & 1ull) == 0ull) continue;
        int64_t* dst = G2 + (p0 + r) * (int64_t)T + tdeep;
        for (int i = lane; i < nd; i += kWave)
            dst[i] = (int64_t)(((uint64_t)s_x[r * gp + 2 * i + 1] << 32) | s_x[r * gp + 2 * i]);
    }
}

// Dense levels of an LDS row (counts): once a level's events fill at least a
// quarter of its frames (and the frames fit the row's 2 CAP words), the row
// becomes the series x[0, lf) over those words.  A lag sum is then
// sum_t x[t] x[t+d] with lanes over t: independent, pipelined LDS reads and
// one wave reduction per lag, instead of each lane scanning its events'
// partners one dependent read at a time (which many partners per event make
// slow).  Binning a dense level is a pairwise sum in place.  Values are at
// most the row's total, which the caller keeps below 2^32.

// The nl events (keys < lf ascending, values at key + CAP) as x[0, lf).
template <int CAP>
__device__ void lds_to_dense(uint32_t* key, int nl, int lf)
{
    constexpr int R = CAP / kWave;
    const int lane = lane_id();
    uint32_t rk[R], rv[R];
#pragma unroll
    for (int c = 0; c < R; c++) {
        const int i = c * kWave + lane;
        rk[c] = i < nl ? key[i] : kNoKey;
        rv[c] = i < nl ? key[CAP + i] : 0u;
    }
    wave_sync();
    for (int t = lane; t < lf; t += kWave) key[t] = 0u;
    wave_sync();
#pragma unroll
    for (int c = 0; c < R; c++)
        if (rk[c] != kNoKey) key[rk[c]] = rv[c];
    wave_sync();
}

// x[0, lf) -> x[0, lf / 2), pair sums (lf / 2 <= CAP).
template <int CAP>
__device__ void lds_dense_bin(uint32_t* x, int lf)
{
    constexpr int R = CAP / kWave;
    const int lane = lane_id();
    const int h = lf >> 1;
    uint32_t sm[R];
#pragma unroll
    for (int c = 0; c < R; c++) {
        const int t = c * kWave + lane;
        sm[c] = t < h ? x[2 * t] + x[2 * t + 1] : 0u;
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < R; c++) {
        const int t = c * kWave + lane;
        if (t < h) x[t] = sm[c];
    }
    wave_sync();
}

// Sum over the wave of each of 8 per-lane values: a butterfly that halves the
// set at each of the first three exchanges (lanes ^ 32, 16, 8 keep 4, 2, 1
// of them), then sums the one left over lanes ^ 4, 2, 1.  Lane l ends with
// the total of value (l >> 3) & 7, in 10 exchanges instead of 48.
__device__ __forceinline__ unsigned long long wave_sum8_u64(unsigned long long (&v)[8])
{
    const int lane = lane_id();
    unsigned long long a[4], b[2], c;
    const bool h32 = (lane & 32) != 0, h16 = (lane & 16) != 0, h8 = (lane & 8) != 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        // keep v[j] (lanes < 32) or v[j + 4] (lanes >= 32), add the partner's copy
        const unsigned long long keep = h32 ? v[j + 4] : v[j], give = h32 ? v[j] : v[j + 4];
        a[j] = keep + __shfl_xor(give, 32, 64);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const unsigned long long keep = h16 ? a[j + 2] : a[j], give = h16 ? a[j] : a[j + 2];
        b[j] = keep + __shfl_xor(give, 16, 64);
    }
    {
        const unsigned long long keep = h8 ? b[1] : b[0], give = h8 ? b[0] : b[1];
        c = keep + __shfl_xor(give, 8, 64);
    }
    c += __shfl_xor(c, 4, 64);
    c += __shfl_xor(c, 2, 64);
    c += __shfl_xor(c, 1, 64);
    return c;
}

// The level's lags lo..hi over x[0, lf) into copy 0 of their accumulators,
// 8 lags at a time (t + d < lf checked only in the last hi + 1 frames).
__device__ void lds_dense_lags(const uint32_t* x, int lf, const LevelParam& lp, unsigned long long* acc, int nc)
{
    constexpr int kLagBlock = 8;
    const int lane = lane_id();
    for (int d0 = lp.lo; d0 <= lp.hi; d0 += kLagBlock) {
        unsigned long long sd[kLagBlock];
#pragma unroll
        for (int j = 0; j < kLagBlock; j++) sd[j] = 0ull;
        const int tmain = lf - (d0 + kLagBlock - 1);   // every lag of the block in range below it
        int t = lane;
        for (; t < tmain; t += kWave) {
            const unsigned long long xt = x[t];
#pragma unroll
            for (int j = 0; j < kLagBlock; j++) sd[j] += xt * x[t + d0 + j];
        }
        for (; t < lf; t += kWave) {
            const unsigned long long xt = x[t];
#pragma unroll
            for (int j = 0; j < kLagBlock; j++)
                if (t + d0 + j < lf) sd[j] += xt * x[t + d0 + j];
        }
        const unsigned long long g = wave_sum8_u64(sd);
        // lane 8 j holds lag d0 + j
        const int j = lane >> 3;
        if ((lane & 7) == 0 && d0 + j <= lp.hi) acc[(lp.first + d0 + j - lp.lo) * nc] += g;
    }
    wave_sync();
}

// LDS rows whose events are out of order only locally (the transpose places
// the events of one pixel within one slice of its bucket concurrently, so a
// long row is a run of short unordered stretches in frame order): 64-entry
// blocks sorted in registers (DPP bitonic network), alternately aligned at 0
// and at 32, so an entry moves up to 32 places per round.  True when the row
// ends up sorted within kBlockRounds rounds.
constexpr int kBlockRounds = 4;

// A row held in registers, element c * 64 + lane in kk[c] / vv[c] (rc
// registers in use, kNoKey past the row): up to `rounds` odd-even
// transposition rounds across the whole row, the pairs that straddle two
// registers exchanged through lane reads.  Rows leave the transpose sorted
// except among a pixel's events of one transpose chunk, which stand next to
// each other, so a few rounds usually finish the row.  Returns whether it is
// sorted.
template <int R, typename V>
__device__ __forceinline__ bool regs_oddeven(uint32_t (&kk)[R], V (&vv)[R], int rc, int rounds)
{
    const int lane = lane_id();
    auto unsorted = [&]() {
        bool bad = false;
#pragma unroll
        for (int c = 0; c < R; c++) {
            if (c < rc) {
                const uint32_t first = c + 1 < rc ? read_lane_u32(kk[c + 1], 0) : kNoKey;
                bad |= next_lane_u32(kk[c], first) < kk[c];
            }
        }
        return __ballot(bad) != 0ull;
    };
    if (!unsorted()) return true;
    for (int r = 0; r < rounds; r++) {
        // even phase: (2j, 2j+1), inside one register
#pragma unroll
        for (int c = 0; c < R; c++) {
            if (c < rc) {
                const uint32_t pk = xor_lane_u32<1>(kk[c]);
                const V pv = xor_lane_v<1>(vv[c]);
                if ((lane & 1) == 0 ? (pk < kk[c]) : (pk > kk[c])) {
                    kk[c] = pk;
                    vv[c] = pv;
                }
            }
        }
        // odd phase: (2j+1, 2j+2); lane 63 of register c pairs with lane 0 of
        // register c + 1 (neighbours taken before any of them moves)
        uint32_t lastk[R], firstk[R];
        V lastv[R], firstv[R];
#pragma unroll
        for (int c = 0; c < R; c++) {
            lastk[c] = read_lane_u32(kk[c], kWave - 1);
            lastv[c] = read_lane_v(vv[c], kWave - 1);
            firstk[c] = read_lane_u32(kk[c], 0);
            firstv[c] = read_lane_v(vv[c], 0);
        }
#pragma unroll
        for (int c = 0; c < R; c++) {
            if (c < rc) {
                const uint32_t nk = next_lane_u32(kk[c], c + 1 < rc ? firstk[c + 1] : kNoKey);
                const V nv = next_lane_v(vv[c], c + 1 < rc ? firstv[c + 1] : V(0));
                const uint32_t bk = prev_lane_u32(kk[c], c > 0 ? lastk[c - 1] : 0u);
                const V bv = prev_lane_v(vv[c], c > 0 ? lastv[c - 1] : V(0));
                if ((lane & 1) != 0) {
                    if (nk < kk[c]) {
                        kk[c] = nk;
                        vv[c] = nv;
                    }
                } else if ((c > 0 || lane > 0) && bk > kk[c]) {
                    kk[c] = bk;
                    vv[c] = bv;
                }
            }
        }
        if (!unsorted()) return true;
    }
    return false;
}
constexpr int kRowRounds = 4;   // odd-even rounds on a loaded row before the LDS block sorts
template <typename V>
__device__ bool lds_block_sort(uint32_t* key, V* val, int m)
{
    const int lane = lane_id();
    auto sorted = [&]() {
        bool bad = false;
        for (int i = lane; i + 1 < m; i += kWave) bad |= key[i] > key[i + 1];
        return __builtin_amdgcn_ballot_w64(bad) == 0ull;
    };
    if (sorted()) return true;
    for (int r = 0; r < kBlockRounds; r++) {
        for (int b = (r & 1) ? kWave / 2 : 0; b < m; b += kWave) {
            const int i = b + lane;
            uint32_t k = i < m ? key[i] : kNoKey;
            V v = i < m ? val[i] : V(0);
            if (reg_unsorted(k)) {
                bitonic_sort64(k, v, kWave);
                if (i < m) {
                    key[i] = k;
                    val[i] = v;
                }
            }
        }
        wave_sync();
        if (sorted()) return true;
    }
    return false;
}

// Rows r < *count of the list rows[r * step] (step -1 walks a list filled from
// the back).  AUX slots hold the level-0 prefix sums (n <= CAP), then the lag
// accumulators (T taus x copies).
// Counts (V = uint32_t): dense levels as above, and the levels from ldeep on
// handed to k_mt_deep like the register kernels' rows (rows whose total is
// below 2^32).
template <int CAP, int WAVES, typename V>
__global__ void __launch_bounds__(kWave * WAVES)
k_mt_sparse(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off,
            const int32_t* __restrict__ rows, int32_t step, const uint32_t* __restrict__ count,
            const TauParam* __restrict__ taus, int32_t T, const LevelParam* __restrict__ levels,
            int32_t nlev, int32_t F, int64_t* __restrict__ G2, int64_t* __restrict__ IP,
            int64_t* __restrict__ IF, uint64_t* __restrict__ pix_total, const int32_t* __restrict__ rowmap,
            const double* __restrict__ fw, int32_t ldeep, uint32_t* __restrict__ deep, int32_t dstride, bool w32,
            int64_t hw)
{
    using A = typename LagAcc<V>::type;
    constexpr bool U = sizeof(V) == 4;
    constexpr int AUX = CAP > 512 ? CAP : 512;
    // keys, then values: 2 CAP contiguous words per wave for counts
    __shared__ __align__(8) uint32_t s_kv[WAVES][CAP * (1 + sizeof(V) / 4)];
    __shared__ A s_aux[WAVES][AUX];   // prefix sums, then lag accumulators
    const int lane = lane_id();
    const int w = wave_id();
    uint32_t* key = s_kv[w];
    V* val = reinterpret_cast<V*>(s_kv[w] + CAP);
    A* aux = s_aux[w];
    const int tdeep = ldeep < nlev ? levels[ldeep].first : T;   // taus before the hand-over
    const int64_t nw = (int64_t)gridDim.x * WAVES;

    const int64_t nrows = (int64_t)*count;
    for (int64_t r = (int64_t)blockIdx.x * WAVES + w; r < nrows; r += nw) {
        const int64_t p = rows[r * step];
        const int64_t row = rowmap ? (int64_t)rowmap[p] : p;
        const int64_t base = off[row];
        const int64_t n64 = off[row + 1] - base;
        int64_t* g2row = G2 + p * T;
        if (n64 > CAP) continue;        // a list shared with the series kernel (its longer rows)
        const int n = (int)n64;         // 0 < n <= CAP
        const int m = next_pow2(n);
        bool m_sorted = false;
        // ---- load and sort by frame ----
        uint32_t counts = 0u;   // the row's raw count total (FP64 values: pixelSum precedes the normalisation)
        if (m <= kWave) {
            uint32_t k = 0xFFFFFFFFu, c = 0u;
            if (lane < n) {
                uint64_t e = ev[base + lane];
                k = (uint32_t)(e >> 32);
                c = (uint32_t)e;
            }
            counts = c;
            V v = event_value<V>(c, k, fw, F);
            reg_sort(k, v, n);
            key[lane] = k;
            val[lane] = v;
        } else {
            // every load of the row in flight at once, then the LDS stores
            constexpr int R = CAP / kWave;
            uint64_t e[R];
#pragma unroll
            for (int c = 0; c < R; c++) {
                const int i = c * kWave + lane;
                e[c] = i < n ? ev[base + i] : ~0ull;
            }
            uint32_t kk[R];
            V vv[R];
#pragma unroll
            for (int c = 0; c < R; c++) {
                const int i = c * kWave + lane;
                kk[c] = i < n ? (uint32_t)(e[c] >> 32) : kNoKey;
                const uint32_t cnt = i < n ? (uint32_t)e[c] : 0u;
                counts += cnt;
                vv[c] = i < n ? event_value<V>(cnt, kk[c], fw, F) : V(0);
            }
            // locally permuted rows: odd-even rounds in registers first, then
            // (what is still out of order) register block sorts in LDS and the
            // bitonic network over the whole row
            const bool reg_ok = regs_oddeven<R>(kk, vv, (m + kWave - 1) / kWave, kRowRounds);
#pragma unroll
            for (int c = 0; c < R; c++) {
                const int i = c * kWave + lane;
                if (i < m) {
                    key[i] = kk[c];
                    val[i] = vv[c];
                }
            }
            wave_sync();
            if (reg_ok || lds_block_sort(key, val, m)) m_sorted = true;
            if (!m_sorted)
            for (int kk = 2; kk <= m; kk <<= 1) {
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    for (int i = lane; i < m; i += kWave) {
                        const int l = i ^ j;
                        if (l > i) {
                            const uint32_t ki = key[i], kl = key[l];
                            const bool up = (i & kk) == 0;
                            if (ki != kl && ((ki > kl) == up)) {
                                key[i] = kl;
                                key[l] = ki;
                                const V vi = val[i];
                                val[i] = val[l];
                                val[l] = vi;
                            }
                        }
                    }
                    wave_sync();
                }
            }
        }
        wave_sync();
        // ---- level 0: merge repeated frames, prefix sums, IP/IF ----
        const int n0 = level_compact(key, val, n, 0, (uint32_t)F);
        A carry = A(0);
        for (int b0 = 0; b0 < n0; b0 += kWave) {
            const int i = b0 + lane;
            const A v = i < n0 ? (A)val[i] : A(0);
            const A incl = wave_scan_acc(v) + carry;
            if (i < n0) aux[i] = incl;
            carry = __shfl(incl, kWave - 1, 64);
        }
        wave_sync();
        // S0(lf_L 2^L) depends on the level only: lane L finds it once
        A s_lvl = A(0);
        if (lane < nlev) {
            const int j_end = lds_lower_bound(key, n0, (uint32_t)levels[lane].lf << lane);
            s_lvl = j_end > 0 ? aux[j_end - 1] : A(0);
        }
        const int64_t hi = U ? sum_hi((uint64_t)carry, hw) : 0;
        for (int i0 = 0; i0 < T; i0 += kWave) {
            const int i = i0 + lane;
            const TauParam tp = taus[i < T ? i : T - 1];
            const A s_end = __shfl(s_lvl, tp.level, 64);
            if (i < T) {
                const uint32_t b_ip = (uint32_t)(tp.lf - tp.tp) << tp.level;
                const uint32_t b_if = (uint32_t)tp.tp << tp.level;
                const int j_ip = lds_lower_bound(key, n0, b_ip);
                const int j_if = lds_lower_bound(key, n0, b_if);
                const A s_ip = j_ip > 0 ? aux[j_ip - 1] : A(0);
                const A s_if = j_if > 0 ? aux[j_if - 1] : A(0);
                put_sum(IP, p * T + i, sum_bits(s_ip), w32, hi);
                put_sum(IF, p * T + i, sum_bits((A)(s_end - s_if)), w32, hi);
            }
        }
        // counts: the row total bounds every level value (32 bits for the
        // dense levels and the hand-over)
        bool small = false;
        if constexpr (sizeof(V) == 4) {
            if (lane == 0) pix_total[p] = carry;
            small = carry <= 0xFFFFFFFFull;
        } else {
            const uint64_t tot = wave_sum_u64((uint64_t)counts);
            if (lane == 0) pix_total[p] = tot;
        }
        const bool hand = U && small && ldeep < nlev;
        wave_sync();
        // lag accumulators: nc copies per tau while T * nc fits the CAP slots
        const int nc = T * kAccCopies <= AUX ? kAccCopies : 1;
        for (int i = lane; i < T * nc; i += kWave) aux[i] = A(0);
        wave_sync();
        // ---- levels: bin in place, then lag products ----
        int nl = n0;
        bool dense = false;
        for (int L = 0; L < nlev; L++) {
            const LevelParam lp = levels[L];
            if constexpr (U) {
                if (dense) {
                    lds_dense_bin<CAP>(key, levels[L - 1].lf);
                } else {
                    if (L > 0) nl = level_compact(key, val, nl, 1, (uint32_t)lp.lf);
                    if (small && ((lp.lf <= 2 * CAP && 4 * nl >= lp.lf) || (hand && L == ldeep))) {
                        lds_to_dense<CAP>(key, nl, lp.lf);
                        dense = true;
                    }
                }
                if (hand && L == ldeep) {
                    // the series of this level, zero-padded to the row length
                    uint32_t* dst = deep + p * dstride;
                    for (int t = lane; t < dstride; t += kWave) dst[t] = t < lp.lf ? key[t] : 0u;
                    wave_sync();
                    break;
                }
                if (dense) {
                    lds_dense_lags(key, lp.lf, lp, reinterpret_cast<unsigned long long*>(aux), nc);
                    continue;
                }
            } else {
                if (L > 0) nl = level_compact(key, val, nl, 1, (uint32_t)lp.lf);
            }
            if (nl < 2) {
                if (!hand) break;
                continue;   // no more pairs; bin what is left down to the hand-over level
            }
            // each event's partners up to hi frames later, kLagStep candidates
            // at a time: their keys and values are read together (one LDS
            // round trip per step instead of one per candidate); keys are
            // sorted and distinct, so a step whose last key is past hi ends it
            const uint32_t lo = (uint32_t)lp.lo, hi = (uint32_t)lp.hi;
            for (int i = lane; i < nl; i += kWave) {
                const uint32_t ki = key[i];
                const V vi = val[i];
                for (int j0 = i + 1; j0 < nl; j0 += kLagStep) {
                    uint32_t kj[kLagStep];
                    V vj[kLagStep];
#pragma unroll
                    for (int u = 0; u < kLagStep; u++) {
                        const int j = min(j0 + u, nl - 1);
                        kj[u] = key[j];
                        vj[u] = val[j];
                    }
#pragma unroll
                    for (int u = 0; u < kLagStep; u++) {
                        const uint32_t d = kj[u] - ki;
                        if (j0 + u < nl && d >= lo && d <= hi)
                            atomicAdd(&aux[(lp.first + (int)(d - lo)) * nc + (lane & (nc - 1))], lag_prod(vi, vj[u]));
                    }
                    if (kj[kLagStep - 1] - ki > hi) break;
                }
            }
            wave_sync();
        }
        const int tg = hand ? tdeep : T;   // the deep kernel writes the others
        for (int i = lane; i < tg; i += kWave) {
            A g = A(0);
            for (int c = 0; c < nc; c++) g += aux[i * nc + c];
            g2row[i] = sum_bits(g);
        }
        wave_sync();
    }
}

// ---- rows too long for LDS, and the dense path ------------------------------------
constexpr int kLongBlock = 256;
constexpr int kMaxLagBlock = 16;

__device__ uint64_t block_sum_u64(uint64_t x, unsigned long long* s_red)
{
    x = wave_sum_u64(x);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (lane_id() == 0) s_red[w] = x;
    __syncthreads();
    uint64_t t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += s_red[i];
    __syncthreads();
    return t;
}

// FP64 block sum through the same 64-bit LDS slots (wave order fixed).
__device__ double block_sum_f64(double x, unsigned long long* s_red)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (lane_id() == 0) s_red[w] = (unsigned long long)__double_as_longlong(x);
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += __longlong_as_double((long long)s_red[i]);
    __syncthreads();
    return t;
}

__device__ __forceinline__ uint64_t block_sum(uint64_t x, unsigned long long* s_red) { return block_sum_u64(x, s_red); }
__device__ __forceinline__ double block_sum(double x, unsigned long long* s_red) { return block_sum_f64(x, s_red); }
__device__ __forceinline__ unsigned long long slot_bits(uint64_t v) { return v; }
__device__ __forceinline__ unsigned long long slot_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
template <typename AccT>
__device__ __forceinline__ AccT slot_value(unsigned long long v);
template <>
__device__ __forceinline__ uint64_t slot_value<uint64_t>(unsigned long long v) { return v; }
template <>
__device__ __forceinline__ double slot_value<double>(unsigned long long v) { return __longlong_as_double((long long)v); }

// Lag sums of one level over a dense series src[0, lf) for lags lo..hi; the
// series is read from `src` (level 0 of the dense path) or scratch.  AccT is
// uint64_t (exact integer sums) or double (frame-sum normalised values).
template <typename AccT, typename SrcT>
__device__ void long_level(const SrcT* __restrict__ src, int lf, const LevelParam& lp,
                           int64_t* __restrict__ g2row, int64_t* __restrict__ IP,
                           int64_t* __restrict__ IF, int64_t row0, bool w32, int64_t hi, unsigned long long* s_red,
                           unsigned long long* s_lag, uint64_t* total0)
{
    // sum over the level
    AccT s = 0;
    for (int t = threadIdx.x; t < lf; t += blockDim.x) s += (AccT)src[t];
    const AccT S = block_sum(s, s_red);
    if (total0 && threadIdx.x == 0) *total0 = (uint64_t)sum_bits(S);
    // products, kMaxLagBlock lags at a time
    for (int l0 = lp.lo; l0 <= lp.hi; l0 += kMaxLagBlock) {
        AccT acc[kMaxLagBlock];
#pragma unroll
        for (int d = 0; d < kMaxLagBlock; d++) acc[d] = 0;
        const int nlag = min(kMaxLagBlock, lp.hi - l0 + 1);
        for (int t = threadIdx.x; t < lf; t += blockDim.x) {
            const AccT xt = (AccT)src[t];
            if (xt == 0) continue;
#pragma unroll
            for (int d = 0; d < kMaxLagBlock; d++) {
                const int u = t + l0 + d;
                if (d < nlag && u < lf) acc[d] += xt * (AccT)src[u];
            }
        }
#pragma unroll
        for (int d = 0; d < kMaxLagBlock; d++) {
            const AccT v = block_sum(acc[d], s_red);
            if (threadIdx.x == 0 && d < nlag) s_lag[d] = slot_bits(v);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < nlag; d += blockDim.x)
            g2row[lp.first + (l0 - lp.lo) + d] = sum_bits(slot_value<AccT>(s_lag[d]));
        __syncthreads();
    }
    // IP = S - sum of the last tau' values, IF = S - sum of the first tau' values
    for (int d = threadIdx.x; d < lp.count; d += blockDim.x) {
        const int tp = lp.lo + d;
        AccT head = 0, tail = 0;
        for (int t = 0; t < tp && t < lf; t++) head += (AccT)src[t];
        for (int t = max(lf - tp, 0); t < lf; t++) tail += (AccT)src[t];
        put_sum(IP, row0 + lp.first + d, sum_bits((AccT)(S - tail)), w32, hi);
        put_sum(IF, row0 + lp.first + d, sum_bits((AccT)(S - head)), w32, hi);
    }
    __syncthreads();
}

// x[k] = src[2k] + src[2k+1] for k < lf_out, chunked so it may run in place.
template <typename AccT, typename SrcT>
__device__ void long_bin(const SrcT* src, AccT* x, int lf_out)
{
    for (int k0 = 0; k0 < lf_out; k0 += blockDim.x) {
        const int k = k0 + threadIdx.x;
        AccT v = 0;
        if (k < lf_out) v = (AccT)src[2 * k] + (AccT)src[2 * k + 1];
        __syncthreads();
        if (k < lf_out) x[k] = v;
        __syncthreads();
    }
}

// One workgroup per pixel row: the listed long rows of the sparse events, or
// every dense row.  With frame weights fw the
// level-0 value of frame f is count x fw[f] and the sums are FP64
// (AccT = double, normalize_by_framesum, main.cpp:313-336).
template <typename AccT>
__global__ void __launch_bounds__(kLongBlock)
k_mt_long(const uint64_t* __restrict__ ev, const int64_t* __restrict__ off,
          const uint32_t* __restrict__ rows, int32_t P, int32_t T,
          const LevelParam* __restrict__ levels, int32_t nlev, int32_t F,
          const int32_t* __restrict__ long_list, int32_t step, int64_t min_len, const uint32_t* __restrict__ counters,
          int dense,
          const double* __restrict__ fw, uint64_t* __restrict__ scratch, int64_t* __restrict__ G2,
          int64_t* __restrict__ IP, int64_t* __restrict__ IF, uint64_t* __restrict__ pix_total,
          const int32_t* __restrict__ rowmap, bool w32)
{
    __shared__ unsigned long long s_red[kLongBlock / 64];
    __shared__ unsigned long long s_lag[kMaxLagBlock];
    // w32: the long rows write every high word (their totals are found late)
    const int64_t hi = w32 ? (int64_t)P * T : 0;
    AccT* x = reinterpret_cast<AccT*>(scratch + (size_t)blockIdx.x * (size_t)F);
    const int64_t nrows = dense ? (int64_t)P : (int64_t)counters[0];
    for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int64_t p = dense ? r : (int64_t)long_list[r * step];
        if (!dense) {
            // the list is shared with the LDS kernel, which takes rows up to min_len
            const int64_t row = rowmap ? (int64_t)rowmap[p] : p;
            if (off[row + 1] - off[row] <= min_len) continue;
        }
        int64_t* g2row = G2 + p * T;
        for (int i = threadIdx.x; i < T; i += blockDim.x) {
            g2row[i] = sum_bits((AccT)0);
            put_sum(IP, p * T + i, sum_bits((AccT)0), w32, hi);
            put_sum(IF, p * T + i, sum_bits((AccT)0), w32, hi);
        }
        __syncthreads();
        if (dense && !fw) {
            const uint32_t* row = rows + p * (int64_t)F;
            long_level<AccT, uint32_t>(row, F, levels[0], g2row, IP, IF, p * T, w32, hi, s_red, s_lag, pix_total + p);
            if (nlev > 1) long_bin<AccT, uint32_t>(row, x, levels[1].lf);
        } else {
            uint64_t counts = 0;   // the pixel's raw count total (pixelSum precedes the normalisation)
            if (dense) {
                const uint32_t* row = rows + p * (int64_t)F;
                for (int t = threadIdx.x; t < F; t += blockDim.x) {
                    x[t] = (AccT)row[t] * (AccT)fw[t];
                    counts += row[t];
                }
            } else {
                for (int t = threadIdx.x; t < F; t += blockDim.x) x[t] = 0;
                __syncthreads();
                const int64_t row = rowmap ? (int64_t)rowmap[p] : p;
                const int64_t b = off[row], e = off[row + 1];
                for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
                    const uint64_t v = ev[i];
                    const uint32_t f = (uint32_t)(v >> 32);
                    if (fw)
                        atomicAdd(reinterpret_cast<double*>(&x[f]), (double)(uint32_t)v * fw[f]);
                    else
                        atomicAdd(reinterpret_cast<unsigned long long*>(&x[f]), (unsigned long long)(uint32_t)v);
                    counts += (uint32_t)v;
                }
            }
            __threadfence_block();
            __syncthreads();
            if (fw) {
                const uint64_t c = block_sum_u64(counts, s_red);
                if (threadIdx.x == 0) pix_total[p] = c;
            }
            long_level<AccT, AccT>(x, F, levels[0], g2row, IP, IF, p * T, w32, hi, s_red, s_lag, fw ? nullptr : pix_total + p);
            if (nlev > 1) long_bin<AccT, AccT>(x, x, levels[1].lf);
        }
        for (int L = 1; L < nlev; L++) {
            const LevelParam lp = levels[L];
            long_level<AccT, AccT>(x, lp.lf, lp, g2row, IP, IF, p * T, w32, hi, s_red, s_lag, nullptr);
            if (L + 1 < nlev) long_bin<AccT, AccT>(x, x, levels[L + 1].lf);
        }
        __syncthreads();
    }
}

// ---- launchers ------------------------------------------------------------------
// Waves per workgroup of the LDS row kernel (counts: 8 / 16 / 32 KB of LDS per
// wave for CAP 512 / 1024 / 2048).  These leave a quarter of the 160 KB of a
// CU unused, but filling it measured slower: 10 or 5 waves for CAP 512, 5 for
// 1024 and 2048 took the largest batched jobs' k_mt_lds from 12.0 and 11.7 ms
// to 12.9-16.4 and 17.8-17.9 ms.
constexpr int kSparseWaves = 8;
constexpr int kSparseMidWaves = 4;    // 4 x 16 KB of LDS rows per workgroup (counts): 2 waves per SIMD
constexpr int kSparseLongWaves = 2;   // 2 x 32 KB of LDS rows per workgroup (counts)
constexpr int kRegWaves = 8;
constexpr int kReg2Waves = 4;

// Dynamic LDS bytes for the rank words of k_mt_reg: all of them, when that
// leaves the kernel's workgroups per CU unchanged, else 0 (the kernel gathers
// the per-frame table).  1 Mpx job: k_mt_reg 1.134 -> 1.121 ms; k_mt_reg2,
// whose rows are fewer and longer, took 0.371 against 0.356 ms with them and
// gathers.
static int rank_lds_bytes(const void* kernel, int block, int words)
{
    if (words <= 0) return 0;
    const int bytes = words * (int)sizeof(uint4);
    int base = 0, with = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&base, kernel, block, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&with, kernel, block, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return base > 0 && with >= base ? bytes : 0;
}

void rank_lds_setup(xpcs_mt_plan* p)
{
    if (p->rank_lds[0] >= 0) return;   // the tables are fixed at plan creation
    const int w = p->nrank_words;
    p->rank_lds[0] = rank_lds_bytes(reinterpret_cast<const void*>(k_mt_reg<kRegWaves, uint32_t>), kWave * kRegWaves, w);
    p->rank_lds[1] = rank_lds_bytes(reinterpret_cast<const void*>(k_mt_reg<kRegWaves, double>), kWave * kRegWaves, w);
}

// The row kernels of the sparse path, counts (V = uint32_t, exact 64-bit
// sums) or frame-weighted values (V = double, normalize_by_framesum): the
// classification pass, one wave per register row, then the listed rows.
template <typename V>
static void launch_rows(xpcs_mt_plan* p, hipStream_t s, const char* t_reg, const char* t_lds, const char* t_long,
                        const double* fw)
{
    const HostTables& h = p->h;
    // register rows (32-bit level values): T <= kRegMaxTaus, and no row of at
    // most kMidSmall events totals 2^32 or more (pixel-major rows of u32 counts)
    const bool regs = h.T <= kRegMaxTaus && !(p->mode == 3 && p->wide_rows);
    timer_begin(p, s, "k_mt_classify");
    if (h.P > 0)
        k_mt_classify<<<(unsigned)((h.P + 256 * kClassifyPer - 1) / (256 * kClassifyPer)), 256, 0, s>>>(
            p->d_off, p->d_rowmap, h.P, regs, p->d_mid_list, p->d_long_list, p->d_counters);
    timer_end(p, s);
    timer_begin(p, s, t_reg);
    // the deep levels' hand-over (counts only)
    const int nlev = (int)h.levels.size();
    int dstride = 0;
    const int ldeep = sizeof(V) == 4 ? deep_level(p, &dstride) : nlev;
    const bool w32 = sizeof(V) == 4 && p->ipif32;
    const int64_t hw = (int64_t)h.P * h.T;   // high words of the w32 layout
    const int vi = sizeof(V) == 4 ? 0 : 1;
    // k_mt_reg2's accumulators per wave: the taus it sums (before the
    // hand-over level) in kAccCopies copies, at least the nb boundary slots,
    // even (16-byte steps between waves)
    using A2 = typename LagAcc<V>::type;
    const int tsum = ldeep < nlev ? h.levels[(size_t)ldeep].first : h.T;
    int reg2_acc_words = std::max(tsum * kAccCopies, (int)((p->nrank * (int)sizeof(V) + (int)sizeof(A2) - 1) / (int)sizeof(A2)));
    reg2_acc_words = (reg2_acc_words + 1) & ~1;
    const int rl_reg = std::max(p->rank_lds[vi], 0);
    k_mt_reg<kRegWaves, V><<<reg_blocks(p), kWave * kRegWaves, rl_reg, s>>>(
        p->d_ev, p->d_off, h.P, h.T, p->d_tau_bnd, p->d_frame_rank, p->nrank, p->d_levels, nlev, h.F, p->d_G2,
        p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap, fw, ldeep, p->d_deep, dstride, w32, regs, p->d_rank_words,
        rl_reg / (int)sizeof(uint4));
    timer_end(p, s);
    timer_begin(p, s, t_lds);
    if (regs)
        k_mt_reg2<kReg2Waves, V><<<lds_blocks(p), kWave * kReg2Waves, reg2_acc_words * (int)sizeof(A2) * kReg2Waves, s>>>(
            p->d_ev, p->d_off, p->d_mid_list, p->d_counters + 2, h.T, p->d_tau_bnd, p->d_frame_rank, p->nrank,
            p->d_levels, nlev, h.F, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap, fw, ldeep, p->d_deep,
            dstride, w32, reg2_acc_words);
    else
        k_mt_sparse<kMidSmall, kSparseWaves, V><<<lds_blocks(p), kWave * kSparseWaves, 0, s>>>(
            p->d_ev, p->d_off, p->d_mid_list, 1, p->d_counters + 2, p->d_taus, h.T, p->d_levels,
            (int32_t)h.levels.size(), h.F, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap, fw, ldeep,
            p->d_deep, dstride, w32, hw);
    k_mt_sparse<kSparseCapUsed, kSparseWaves, V><<<std::max(1, lds_blocks(p) / 4), kWave * kSparseWaves, 0, s>>>(
        p->d_ev, p->d_off, p->d_mid_list + (h.P - 1), -1, p->d_counters + 4, p->d_taus, h.T, p->d_levels,
        (int32_t)h.levels.size(), h.F, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap, fw, ldeep,
        p->d_deep, dstride, w32, hw);
    k_mt_sparse<kSparseCapMid, kSparseMidWaves, V><<<std::max(1, lds_blocks(p) / 4), kWave * kSparseMidWaves, 0, s>>>(
        p->d_ev, p->d_off, p->d_long_list, 1, p->d_counters + 5, p->d_taus, h.T, p->d_levels,
        (int32_t)h.levels.size(), h.F, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap, fw, ldeep,
        p->d_deep, dstride, w32, hw);
    k_mt_sparse<kSparseCapLong, kSparseLongWaves, V><<<std::max(1, lds_blocks(p) / 4), kWave * kSparseLongWaves, 0, s>>>(
        p->d_ev, p->d_off, p->d_long_list + (h.P - 1), -1, p->d_counters, p->d_taus, h.T, p->d_levels,
        (int32_t)h.levels.size(), h.F, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap, fw, ldeep,
        p->d_deep, dstride, w32, hw);
    timer_end(p, s);
    // rows longer than kSparseCapLong: the dense-series kernel
    timer_begin(p, s, t_long);
    if (sizeof(V) == 4)
        k_mt_long<uint64_t><<<p->long_blocks, kLongBlock, 0, s>>>(
            p->d_ev, p->d_off, p->d_rows, h.P, h.T, p->d_levels, (int32_t)h.levels.size(), h.F,
            p->d_long_list + (h.P - 1), -1, (int64_t)kSparseCapLong, p->d_counters, 0, nullptr, p->d_scratch,
            p->d_G2, p->d_IP, p->d_IF,
            p->d_pix_total, p->d_rowmap, w32);
    else
        k_mt_long<double><<<p->long_blocks, kWave, 0, s>>>(
            p->d_ev, p->d_off, p->d_rows, h.P, h.T, p->d_levels, (int32_t)h.levels.size(), h.F,
            p->d_long_list + (h.P - 1), -1, (int64_t)kSparseCapLong, p->d_counters, 0, fw, p->d_scratch,
            p->d_G2, p->d_IP, p->d_IF,
            p->d_pix_total, p->d_rowmap, false);
    timer_end(p, s);
    if (ldeep < nlev && h.P > 0) {
        timer_begin(p, s, "k_mt_deep");
        const int npw = kWave / kDeepLanes;   // pixels per wave
        const unsigned nb = (unsigned)((h.P + npw - 1) / npw);
        if (h.dpl == 4)
            k_mt_deep<4><<<nb, kWave, 0, s>>>(p->d_deep, dstride, p->d_off, p->d_rowmap, h.P, h.T, p->d_levels,
                                              ldeep, nlev, p->d_pix_total, p->d_G2);
        else
            k_mt_deep<8><<<nb, kWave, 0, s>>>(p->d_deep, dstride, p->d_off, p->d_rowmap, h.P, h.T, p->d_levels,
                                              ldeep, nlev, p->d_pix_total, p->d_G2);
        timer_end(p, s);
    }
}

// First level of the one-lane-per-pixel kernel: the first level L >= 1 with
// at most deep_frames binned frames (tuning; 0: kDeepMaxFrames, negative:
// none), for counts (not the FP64 frame-sum mode), dpl 4 or 8 (the kernel's
// instantiations), T <= kRegMaxTaus (the register kernels run) and at most
// kDeepMaxTaus taus / kDeepMaxLevels levels from L on.
int deep_level(const xpcs_mt_plan* p, int* stride)
{
    const HostTables& h = p->h;
    const int nlev = (int)h.levels.size();
    *stride = 0;
    const int lim = p->tune.deep_frames < 0 ? 0 : (p->tune.deep_frames > 0 ? p->tune.deep_frames : kDeepMaxFrames);
    if (lim == 0 || p->fsnorm || (h.dpl != 4 && h.dpl != 8) || h.T > kRegMaxTaus) return nlev;
    for (int L = 1; L < nlev; L++) {
        const LevelParam& lp = h.levels[(size_t)L];
        if (lp.lf > lim) continue;
        if (h.T - lp.first > kDeepMaxTaus || nlev - L > kDeepMaxLevels || lp.lo != h.dpl + 1) return nlev;
        *stride = (lp.lf + 3) & ~3;
        return L;
    }
    return nlev;
}

void launch_mt_sparse(xpcs_mt_plan* p, hipStream_t s)
{
    launch_rows<uint32_t>(p, s, "k_mt_reg", "k_mt_lds", "k_mt_long", nullptr);
}

// Every dense row through the dense-series kernel (the fallback of the dense
// path for delays per level outside the streaming kernel's instantiations).
void launch_mt_long(xpcs_mt_plan* p, hipStream_t s, bool dense_rows)
{
    const HostTables& h = p->h;
    if (!dense_rows) return;   // sparse rows: launch_mt_sparse runs the long rows too
    timer_begin(p, s, "k_mt_long_dense");
    k_mt_long<uint64_t><<<p->long_blocks, kLongBlock, 0, s>>>(
        p->d_ev, p->d_off, p->d_rows, h.P, h.T, p->d_levels, (int32_t)h.levels.size(), h.F, p->d_long_list, 1, 0,
        p->d_counters, 1, nullptr, p->d_scratch, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap,
        p->ipif32);
    timer_end(p, s);
}

// normalize_by_framesum: FP64 sums of count x fw[frame], through the same row
// kernels as the exact path (register rows, two per lane, the LDS kernels)
// with V = double; rows over the LDS capacity and every dense row run through
// the series kernel, one wavefront per workgroup so the per-lag reductions
// stay inside the wave.
void launch_mt_fsnorm(xpcs_mt_plan* p, hipStream_t s, bool dense_rows)
{
    const HostTables& h = p->h;
    if (!dense_rows) {
        launch_rows<double>(p, s, "k_mt_reg_f64", "k_mt_lds_f64", "k_mt_long_f64", p->d_fw);
        return;
    }
    timer_begin(p, s, "k_mt_long_f64");
    k_mt_long<double><<<p->long_blocks, kWave, 0, s>>>(
        p->d_ev, p->d_off, p->d_rows, h.P, h.T, p->d_levels, (int32_t)h.levels.size(), h.F, p->d_long_list, 1, 0,
        p->d_counters, 1, p->d_fw, p->d_scratch, p->d_G2, p->d_IP, p->d_IF, p->d_pix_total, p->d_rowmap,
        false);
    timer_end(p, s);
}

int sparse_cap_long() { return kSparseCapLong; }
int reg_cap() { return kMidSmall; }
int sparse_cap() { return kSparseCapUsed; }

// Workgroups of the row kernels: one row per wave up to 16384 workgroups by
// default (both kernels loop over their rows), or the tuned count.
int reg_blocks(const xpcs_mt_plan* p)
{
    if (p->tune.reg_blocks > 0) return p->tune.reg_blocks;
    const int64_t b = ((int64_t)p->h.P + kRegWaves - 1) / kRegWaves;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

int lds_blocks(const xpcs_mt_plan* p)
{
    return p->tune.lds_blocks > 0 ? p->tune.lds_blocks : 2048;
}

}  // namespace xpcs
