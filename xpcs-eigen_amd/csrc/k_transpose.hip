// (a) Frame -> pixel transpose kernels.
//
// Sparse IMM (SparseFilter::Apply, src/xpcs/filter/sparse_filter.cpp:115-193):
// a two-level counting sort of the frame-major events by compact pixel id
// (buckets of pixels first, then rows within a bucket), with masked pixels and
// zero counts dropped on the way.  Rows come out in arbitrary order within a
// row; the correlator sorts each row on chip.
//
// Dense IMM (DenseFilter::Apply, src/xpcs/filter/dense_filter.cpp:121-210):
// an LDS-tiled transpose of whole frames into pixel-major rows [P][F] of
// n-scaled integer values (dark subtraction and threshold fused).
#include <type_traits>

#include "device_util.h"
#include "internal.h"

namespace xpcs {

// ---- bucketed transpose (no per-event global atomics) -------------------------------
// Pixels are grouped into buckets of 2^bshift consecutive compact ids.
//   k_tr_count   per chunk of CF raw frames: LDS histogram of events per bucket,
//                written to hist[chunk][bucket].
//   (scan)       column scan of hist into bucket-major order: the start of
//                every (bucket, chunk) segment of the bucketed array B1.
//   k_tr_bucket  per chunk: each event takes a slot of its (bucket, chunk)
//                segment through an LDS cursor and is written to B1 as
//                (frame << 32 | local pixel << 16 | count).
//   k_tr_rows    per bucket: LDS counting sort of its events by local pixel;
//                the bucket's rows are contiguous, so the global row offsets
//                follow from the bucket start, and events land in their rows
//                through LDS cursors.
// Only LDS atomics are used; every global write goes to a small, hot region.
constexpr int kTrChunkBlock = 1024;   // workgroup of the per-chunk kernels
__device__ int64_t block_excl_scan(int64_t x, int64_t* s_warp, int64_t& total);

// Element j of a frame record: raster pixel index and count, by word format.
//   IMM      int32 index[dlen] then int16 value[dlen] (src/xpcs/io/imm.cpp:87-99)
//   UFXC     u32 words: value (w >> 15) & 3, column-major pixel w & 0x7fff
//            (src/xpcs/io/ufxc.cpp:140-148)
//   Rigaku   u64 words: value w & 0x7ff, column-major pixel (w >> 16) & 0xfffff
//            (src/xpcs/io/rigaku.cpp:201-208)
// Column-major pixels p become (p % H) * W + p / H.
__device__ __forceinline__ void tr_decode(const uint8_t* base, uint32_t dlen, bool al4, uint32_t j, EvFormat fm,
                                          uint32_t& idx, uint32_t& v)
{
    if (fm.kind == kFmtImm) {
        idx = al4 ? reinterpret_cast<const uint32_t*>(base)[j] : load_u32_a2(base + 4 * (size_t)j);
        v = reinterpret_cast<const uint16_t*>(base + 4 * (size_t)dlen)[j];
    } else if (fm.kind == kFmtUfxc) {
        const uint32_t w = reinterpret_cast<const uint32_t*>(base)[j];
        const uint32_t pix = w & 0x7fffu;
        idx = (pix % fm.H) * fm.W + pix / fm.H;
        v = (w >> 15) & 3u;
    } else {
        const uint64_t w = reinterpret_cast<const uint64_t*>(base)[j];
        const uint32_t pix = (uint32_t)(w >> 16) & 0xfffffu;
        idx = (pix % fm.H) * fm.W + pix / fm.H;
        v = (uint32_t)w & 0x7ffu;
    }
}

// The events of a chunk's raw frames form one flat index space, visited in
// groups of up to kTrGroupFrames frames: the prefix sums of the group's record
// lengths and the records' offsets sit in LDS, and each thread keeps
// kTrInFlight events in flight across frame boundaries (frames hold a few
// thousand events, so a thread's frame index advances by at most a step or two
// per event).  Measured on a megapixel job: the count pass 0.125 ms with 2
// (3: 0.130, 4: 0.132); the bucket pass 0.318 ms with 1 and 0.413 ms with 2,
// whose slot reservations and scattered writes do not gain from overlapped
// reads.
constexpr int kTrGroupFrames = 256;
constexpr int kTrInFlight = 2;
constexpr int kTrBucketInFlight = 1;

struct ChunkGroup {
    uint32_t pre[kTrGroupFrames + 1];   // events before each frame of the group
    int64_t rec[kTrGroupFrames];        // byte offset of each frame's record
    int32_t frame[kTrGroupFrames];      // output frame of each raw frame
    int64_t warp[33];
};

// Loads group [g0, g0 + ng) into LDS; returns its event count.  Every thread
// of the workgroup calls it (it synchronises).
__device__ uint32_t chunk_group_load(ChunkGroup& cg, const int64_t* __restrict__ sel_off,
                                     const uint32_t* __restrict__ sel_dlen, const int32_t* __restrict__ sel_frame,
                                     int64_t g0, int ng)
{
    const int t = threadIdx.x;
    const uint32_t dl = t < ng ? sel_dlen[g0 + t] : 0u;
    int64_t tot;
    const int64_t ex = block_excl_scan((int64_t)dl, cg.warp, tot);
    if (t < ng) {
        cg.pre[t] = (uint32_t)ex;
        cg.rec[t] = sel_off[g0 + t];
        if (sel_frame) cg.frame[t] = sel_frame[g0 + t];
    }
    if (t == 0) cg.pre[ng] = (uint32_t)tot;
    __syncthreads();
    return (uint32_t)tot;
}

// Events q0 + u * step (u < N) of the group, each with its own frame pointer
// kf[u]: raster pixel r, count v, and ok when the event is inside the group,
// its pixel on the detector and mapped, and its count nonzero; `bad` records
// a pixel index outside the detector.  The word format is fixed at compile
// time and the N events' loads are issued together, after every address is
// known and before any result is used (then the N mask words).  IMM indices
// are only 2-byte aligned in general: the aligned 8 bytes around one are read
// and joined (inside the record, whose values follow its indices; an aligned
// word never straddles a page).  Needs total >= 1.
template <int KIND, int N>
__device__ __forceinline__ void chunk_events(const ChunkGroup& cg, const uint8_t* __restrict__ imm, uint32_t q0,
                                             uint32_t step, uint32_t total, int (&kf)[N], const EvFormat& fm,
                                             const uint32_t* __restrict__ mapped, uint32_t npix_det, bool& bad,
                                             uint32_t (&r)[N], uint32_t (&v)[N], bool (&ok)[N])
{
    bool in[N];
    const uint8_t* wp[N];   // the event's word (IMM: the aligned 8 bytes around its index)
    const uint8_t* vp[N];   // IMM: its value
    uint32_t sh[N];         // IMM: 2 when the index starts in the middle of a word
#pragma unroll
    for (int u = 0; u < N; u++) {
        const uint32_t q = q0 + (uint32_t)u * step;
        in[u] = q < total;
        const uint32_t qc = in[u] ? q : total - 1u;
        while (cg.pre[kf[u] + 1] <= qc) kf[u]++;
        const uint32_t b = cg.pre[kf[u]];
        const uint32_t j = qc - b;
        const uint8_t* base = imm + cg.rec[kf[u]];
        if constexpr (KIND == kFmtImm) {
            const uint8_t* at = base + 4 * (size_t)j;
            sh[u] = (uint32_t)reinterpret_cast<uintptr_t>(at) & 2u;
            wp[u] = at - sh[u];
            vp[u] = base + 4 * (size_t)(cg.pre[kf[u] + 1] - b) + 2 * (size_t)j;
        } else if constexpr (KIND == kFmtUfxc) {
            wp[u] = base + 4 * (size_t)j;
        } else {
            wp[u] = base + 8 * (size_t)j;
        }
    }
#pragma unroll
    for (int u = 0; u < N; u++) {
        if constexpr (KIND == kFmtImm) {
            const uint2 w = *reinterpret_cast<const uint2*>(wp[u]);
            r[u] = sh[u] ? __builtin_amdgcn_alignbyte(w.y, w.x, 2) : w.x;
            v[u] = *reinterpret_cast<const uint16_t*>(vp[u]);
        } else if constexpr (KIND == kFmtUfxc) {
            const uint32_t w = *reinterpret_cast<const uint32_t*>(wp[u]);
            const uint32_t pix = w & 0x7fffu;
            r[u] = (pix % fm.H) * fm.W + pix / fm.H;
            v[u] = (w >> 15) & 3u;
        } else {
            const uint64_t w = *reinterpret_cast<const uint64_t*>(wp[u]);
            const uint32_t pix = (uint32_t)(w >> 16) & 0xfffffu;
            r[u] = (pix % fm.H) * fm.W + pix / fm.H;
            v[u] = (uint32_t)w & 0x7ffu;
        }
    }
    uint32_t m[N];
#pragma unroll
    for (int u = 0; u < N; u++) {
        const bool inb = r[u] < npix_det;
        bad |= in[u] && !inb;
        m[u] = mapped[(inb ? r[u] : 0u) >> 5];
    }
#pragma unroll
    for (int u = 0; u < N; u++) ok[u] = in[u] && r[u] < npix_det && v[u] != 0u && ((m[u] >> (r[u] & 31u)) & 1u);
}

// The bucket pass stages a group's IMM events in LDS a piece of kPieceEvents
// at a time: each frame's share of the piece is one range of index bytes and
// one of value bytes, read as whole aligned 16-byte blocks (every load
// independent of the others and of any event, where reading event by event
// chained each load to the frame-table lookup before it); events then decode
// from LDS.  A block that holds any byte of a range lies in that range's page,
// so no load leaves the caller's pages.  Megapixel job: k_tr_bucket 0.316 ->
// 0.296 ms; the count pass, whose per-event work is a single LDS atomic, took
// 0.158 ms staged against 0.126 ms reading in place, and reads in place.
constexpr int kPieceEvents = 4096;
// blocks of a piece: a frame of e events in the piece spans at most
// (14 + 4 e + 15) / 16 index and (14 + 2 e + 15) / 16 value blocks (ranges
// start 2-byte aligned), so nf <= kTrGroupFrames frames and kPieceEvents
// events need at most (58 nf + 6 kPieceEvents) / 16
constexpr int kPieceBlocks = (58 * kTrGroupFrames + 6 * kPieceEvents + 15) / 16;
constexpr int kPieceCopy = (kPieceBlocks + kTrChunkBlock - 1) / kTrChunkBlock;

struct NoStage {};

struct PieceStage {
    uint4 blk[kPieceBlocks];                  // the staged blocks (38.5 KB)
    uint32_t rstart[2 * kTrGroupFrames + 1];  // first block of each range (index, value per frame)
    int64_t rgaddr[2 * kTrGroupFrames];       // address of each range's first block
    int32_t ibase[kTrGroupFrames];            // staged byte of index j of piece frame t: ibase[t] + 4 j
    int32_t vbase[kTrGroupFrames];            // staged byte of value j: vbase[t] + 2 j
    int32_t fa, fb;                           // first and last frame of the piece (in the group)
};

// Stages events [q0, q1) of the group (q0 < q1 <= its event count); every
// thread of the workgroup calls it (it synchronises).
__device__ void stage_piece(ChunkGroup& cg, const uint8_t* __restrict__ imm, uint32_t q0, uint32_t q1, int ng,
                            PieceStage& ps)
{
    const int t = threadIdx.x;
    // the frames holding q0 and q1 - 1 (each is the one frame whose events
    // include it), found by the threads of the group's frames at once
    if (t < ng) {
        const uint32_t p0 = cg.pre[t], p1 = cg.pre[t + 1];
        if (p0 <= q0 && q0 < p1) ps.fa = t;
        if (p0 <= q1 - 1u && q1 - 1u < p1) ps.fb = t;
    }
    __syncthreads();
    const int fa = ps.fa, nf = ps.fb - fa + 1;
    const int64_t img = (int64_t)reinterpret_cast<uintptr_t>(imm);   // blocks are aligned in address space
    uint32_t nbi = 0u, nbv = 0u;
    int64_t a0 = 0, b0 = 0;
    if (t < nf) {
        const int f = fa + t;
        const uint32_t p0 = cg.pre[f], p1 = cg.pre[f + 1];
        const uint32_t ja = max(q0, p0) - p0, jb = min(q1, p1) - p0;
        if (jb > ja) {
            const int64_t rec = img + cg.rec[f];
            const int64_t i0 = rec + 4 * (int64_t)ja, i1 = rec + 4 * (int64_t)jb;
            a0 = i0 & ~(int64_t)15;
            nbi = (uint32_t)((((i1 + 15) & ~(int64_t)15) - a0) >> 4);
            const int64_t vr = rec + 4 * (int64_t)(p1 - p0);
            const int64_t v0 = vr + 2 * (int64_t)ja, v1 = vr + 2 * (int64_t)jb;
            b0 = v0 & ~(int64_t)15;
            nbv = (uint32_t)((((v1 + 15) & ~(int64_t)15) - b0) >> 4);
        }
    }
    int64_t tot;
    const uint32_t ex = (uint32_t)block_excl_scan((int64_t)(nbi + nbv), cg.warp, tot);
    if (t < nf) {
        const int f = fa + t;
        const int64_t rec = img + cg.rec[f];
        ps.rstart[2 * t] = ex;
        ps.rgaddr[2 * t] = a0;
        ps.rstart[2 * t + 1] = ex + nbi;
        ps.rgaddr[2 * t + 1] = b0;
        ps.ibase[t] = (int32_t)((int64_t)ex * 16 + rec - a0);
        ps.vbase[t] = (int32_t)((int64_t)(ex + nbi) * 16 + rec + 4 * (int64_t)(cg.pre[f + 1] - cg.pre[f]) - b0);
    }
    if (t == 0) ps.rstart[2 * nf] = (uint32_t)tot;
    __syncthreads();
    const uint32_t nblk = (uint32_t)tot;
    const int nr = 2 * nf;
    // the address of block b (of the range holding it: the last r with
    // rstart[r] <= b, empty ranges sharing a start), -1 past the piece
    auto block_addr = [&](uint32_t b) -> int64_t {
        if (b >= nblk) return -1;
        int lo = 0, hi = nr - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ps.rstart[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        return ps.rgaddr[lo] + ((int64_t)(b - ps.rstart[lo]) << 4);
    };
    static_assert(kPieceCopy == 3, "three blocks per thread");
    const int64_t g0 = block_addr((uint32_t)t), g1 = block_addr((uint32_t)(kTrChunkBlock + t)),
                  g2 = block_addr((uint32_t)(2 * kTrChunkBlock + t));
    uint4 w0 = {0u, 0u, 0u, 0u}, w1 = {0u, 0u, 0u, 0u}, w2 = {0u, 0u, 0u, 0u};
    if (g0 >= 0) w0 = *reinterpret_cast<const uint4*>(imm + (g0 - img));
    if (g1 >= 0) w1 = *reinterpret_cast<const uint4*>(imm + (g1 - img));
    if (g2 >= 0) w2 = *reinterpret_cast<const uint4*>(imm + (g2 - img));
    if (g0 >= 0) ps.blk[t] = w0;
    if (g1 >= 0) ps.blk[kTrChunkBlock + t] = w1;
    if (g2 >= 0) ps.blk[2 * kTrChunkBlock + t] = w2;
    __syncthreads();
}

// Events q0 + u * step (u < N) of a staged piece ending at q1, as chunk_events
// gives them, decoded from LDS (kf[u]: frame pointers in the group).
template <int N>
__device__ __forceinline__ void piece_events(const ChunkGroup& cg, const PieceStage& ps, uint32_t q0, uint32_t step,
                                             uint32_t q1, int (&kf)[N], const uint32_t* __restrict__ mapped,
                                             uint32_t npix_det, bool& bad, uint32_t (&r)[N], uint32_t (&v)[N],
                                             bool (&ok)[N])
{
    const uint16_t* h = reinterpret_cast<const uint16_t*>(ps.blk);
    bool in[N];
#pragma unroll
    for (int u = 0; u < N; u++) {
        const uint32_t q = q0 + (uint32_t)u * step;
        in[u] = q < q1;
        const uint32_t qc = in[u] ? q : q1 - 1u;
        while (cg.pre[kf[u] + 1] <= qc) kf[u]++;
        const int tf = kf[u] - ps.fa;
        const uint32_t j = qc - cg.pre[kf[u]];
        const uint32_t ib = (uint32_t)(ps.ibase[tf] + 4 * (int32_t)j) >> 1;
        r[u] = (uint32_t)h[ib] | ((uint32_t)h[ib + 1] << 16);
        v[u] = h[(uint32_t)(ps.vbase[tf] + 2 * (int32_t)j) >> 1];
    }
    uint32_t m[N];
#pragma unroll
    for (int u = 0; u < N; u++) {
        const bool inb = r[u] < npix_det;
        bad |= in[u] && !inb;
        m[u] = mapped[(inb ? r[u] : 0u) >> 5];
    }
#pragma unroll
    for (int u = 0; u < N; u++) ok[u] = in[u] && r[u] < npix_det && v[u] != 0u && ((m[u] >> (r[u] & 31u)) & 1u);
}

// SIDE: the side statistics' per-(output frame, static bin) sums of the
// chunk's events in the same pass (sparse_filter.cpp:168-191; no flat field,
// static bins < 256 from the 1-byte per-pixel table): an LDS table
// [output frames of the chunk][S], runs of equal (frame, bin) summed across
// the wave first, flushed into fs with one global atomic per nonzero entry.
template <bool SIDE, int KIND>
__global__ void __launch_bounds__(kTrChunkBlock)
k_tr_count(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off,
           const uint32_t* __restrict__ sel_dlen, int64_t nsel, int64_t cf, EvFormat fm,
           const uint32_t* __restrict__ mapped, uint32_t npix_det, int bshift, int32_t nb,
           int32_t nblk, uint32_t* __restrict__ hist, uint32_t* __restrict__ counters,
           const int32_t* __restrict__ sel_frame, const uint8_t* __restrict__ sbin_pix, int32_t S,
           double* __restrict__ fs)
{
    extern __shared__ uint32_t s_h[];   // [nb] bucket counts; SIDE: then [frames][S] doubles
    __shared__ ChunkGroup cg;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s_h[b] = 0u;
    const int64_t k0 = (int64_t)blockIdx.x * cf;
    const int64_t k1 = min(nsel, k0 + cf);
    const uint32_t bd = blockDim.x;
    double* s_fs = reinterpret_cast<double*>(s_h + ((nb + 1) & ~1));
    int32_t f_lo = 0, f_span = 0;
    if constexpr (SIDE) {
        f_lo = sel_frame[k0];
        f_span = sel_frame[k1 - 1] - f_lo + 1;
        for (int i = threadIdx.x; i < f_span * S; i += blockDim.x) s_fs[i] = 0.0;
    }
    bool bad = false;
    for (int64_t g0 = k0; g0 < k1; g0 += kTrGroupFrames) {
        const int ng = (int)min<int64_t>(kTrGroupFrames, k1 - g0);
        const uint32_t total = chunk_group_load(cg, sel_off, sel_dlen, SIDE ? sel_frame : nullptr, g0, ng);
        if constexpr (SIDE) {
            int kf[kTrInFlight] = {};
            // uniform trip count: every lane takes part in the wave sums
            for (uint32_t qb = 0; qb < total; qb += kTrInFlight * bd) {
                uint32_t c[kTrInFlight], v[kTrInFlight];
                bool ok[kTrInFlight];
                chunk_events<KIND, kTrInFlight>(cg, imm, qb + threadIdx.x, bd, total, kf, fm, mapped, npix_det, bad,
                                                c, v, ok);
#pragma unroll
                for (int u = 0; u < kTrInFlight; u++) {
                    if (ok[u]) atomicAdd(&s_h[c[u] >> bshift], 1u);
                    const uint32_t bin = ok[u] ? (uint32_t)sbin_pix[c[u]] : 0u;   // 1..S, 0 unmapped
                    const uint32_t key =
                        bin ? (uint32_t)(cg.frame[kf[u]] - f_lo) * 256u + (bin - 1u) : 0xFFFFFFFFu;
                    bool tail;
                    const double run = wave_run_sum_f64(bin ? (double)v[u] : 0.0, key, tail);
                    if (tail && bin) atomicAdd(&s_fs[(key >> 8) * (uint32_t)S + (key & 255u)], run);
                }
            }
            __syncthreads();   // the group tables are reloaded
            continue;
        }
        int kf[kTrInFlight] = {};
        // (one atomic per event: aggregating runs of equal buckets across the
        // wave, as k_tr_bucket does, measured no faster here)
        for (uint32_t q0 = threadIdx.x; q0 < total; q0 += kTrInFlight * bd) {
            uint32_t c[kTrInFlight], v[kTrInFlight];
            bool ok[kTrInFlight];
            chunk_events<KIND, kTrInFlight>(cg, imm, q0, bd, total, kf, fm, mapped, npix_det, bad, c, v, ok);
#pragma unroll
            for (int u = 0; u < kTrInFlight; u++)
                if (ok[u]) atomicAdd(&s_h[c[u] >> bshift], 1u);
        }
        __syncthreads();   // the group tables are reloaded
    }
    if (bad && counters) atomicOr(&counters[1], 1u);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[(int64_t)blockIdx.x * nb + b] = s_h[b];
    if constexpr (SIDE) {
        for (int i = threadIdx.x; i < f_span * S; i += blockDim.x)
            if (s_fs[i] != 0.0) atomicAdd(&fs[(int64_t)f_lo * S + i], s_fs[i]);
    }
}

template <int KIND>
__global__ void __launch_bounds__(kTrChunkBlock)
k_tr_bucket(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off,
            const uint32_t* __restrict__ sel_dlen, const int32_t* __restrict__ sel_frame, int64_t nsel,
            int64_t cf, EvFormat fm, const uint32_t* __restrict__ mapped, uint32_t npix_det, int bshift, int32_t nb,
            int32_t nblk, const int64_t* __restrict__ hoff, uint64_t* __restrict__ b1)
{
    extern __shared__ unsigned long long s_base[];   // [nb] segment starts of this chunk
    uint32_t* s_cur = reinterpret_cast<uint32_t*>(s_base + nb);
    __shared__ ChunkGroup cg;
    __shared__ std::conditional_t<KIND == kFmtImm, PieceStage, NoStage> ps;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        s_cur[b] = 0u;
        s_base[b] = (unsigned long long)hoff[(int64_t)blockIdx.x * nb + b];
    }
    const uint32_t lmask = (1u << bshift) - 1u;
    const int64_t k0 = (int64_t)blockIdx.x * cf;
    const int64_t k1 = min(nsel, k0 + cf);
    const uint32_t bd = blockDim.x;
    for (int64_t g0 = k0; g0 < k1; g0 += kTrGroupFrames) {
        const int ng = (int)min<int64_t>(kTrGroupFrames, k1 - g0);
        const uint32_t total = chunk_group_load(cg, sel_off, sel_dlen, sel_frame, g0, ng);
        int kf[kTrBucketInFlight] = {};
        const uint32_t piece = KIND == kFmtImm ? (uint32_t)kPieceEvents : total;
        for (uint32_t pb = 0; pb < total; pb += piece) {
        const uint32_t pe = min(total, pb + piece);
        if constexpr (KIND == kFmtImm) stage_piece(cg, imm, pb, pe, ng, ps);
        // uniform trip count: every lane takes part in the run detection
        for (uint32_t qb = pb; qb < pe; qb += kTrBucketInFlight * bd) {
            uint32_t c[kTrBucketInFlight], v[kTrBucketInFlight], slot[kTrBucketInFlight];
            bool ok[kTrBucketInFlight];
            bool bad = false;   // reported by the count pass
            if constexpr (KIND == kFmtImm)
                piece_events<kTrBucketInFlight>(cg, ps, qb + threadIdx.x, bd, pe, kf, mapped, npix_det, bad, c, v, ok);
            else
                chunk_events<KIND, kTrBucketInFlight>(cg, imm, qb + threadIdx.x, bd, pe, kf, fm, mapped, npix_det, bad,
                                                      c, v, ok);
            // one LDS cursor atomic per run of equal buckets: the run's last
            // lane takes the run's slots and hands the base to the run's first
            // lane, from which every lane of the run reads it (slots in lane order)
#pragma unroll
            for (int u = 0; u < kTrBucketInFlight; u++) {
                const uint32_t key = ok[u] ? c[u] >> bshift : 0xFFFFFFFFu;
                bool tail;
                const uint32_t lane = (uint32_t)lane_id();
                const uint32_t start = wave_run_start(key, tail);
                uint32_t base = 0u;
                if (tail && ok[u]) base = atomicAdd(&s_cur[key], lane - start + 1u);
                // the lanes of a run push in a rotation (lane -> lane + 1, last ->
                // first) so that no two lanes target one lane; lane 63 always
                // ends a run
                const uint32_t at_head = push_to_lane(tail ? (int)start : (int)lane + 1, base);
                slot[u] = pull_from_lane((int)start, at_head) + (lane - start);
            }
#pragma unroll
            for (int u = 0; u < kTrBucketInFlight; u++)
                if (ok[u])
                    b1[s_base[c[u] >> bshift] + slot[u]] = ((uint64_t)(uint32_t)cg.frame[kf[u]] << 32) |
                                                           ((uint64_t)(c[u] & lmask) << 16) | (uint64_t)v[u];
        }
        __syncthreads();   // the piece (or the group tables) are reloaded
        }
    }
}

// Row placement, one bucket at a time per workgroup.  The bucket's events are
// split into G contiguous slices, one per group of waves (G = 16 single waves
// for buckets of up to 256 pixels); every group counts its slice per pixel,
// and its cursors start after the earlier slices' events of the same pixel,
// so each row receives its events in bucket order (frame order, up to the
// order inside a transpose chunk) without barriers between the groups.  A wave
// keeps the first kTrRowsRegs * 64 events of its share in registers (read from
// HBM once, by the counting pass).  Rows are assembled in LDS a pixel range at
// a time: the range holds the pixels whose row starts fall in a window of
// kTrRowsCap - (longest row) events, so its rows fit the staging buffer, and
// it is written out as one contiguous, coalesced span.  A bucket whose longest
// row exceeds half the buffer, or that would need more than four ranges,
// places its events straight into HBM through the same cursors.
constexpr int kTrRowsBlock = 1024;
constexpr int kTrRowsWaves = kTrRowsBlock / 64;
constexpr int kTrRowsCap = 15872;     // staged events per range (124 KB of LDS)
constexpr int kTrRowsTable = 4096;    // per-group count and cursor entries
constexpr int kTrRowsRegs = 48;       // register-held events per lane (2 VGPRs each; 128 in all)
// batches of the large buckets (8 events per thread; the batch's per-pixel
// group starts, 4096 u32, sit in the staging area behind them)
constexpr int kTrRowsBatch = 8 * kTrRowsBlock;
static_assert(kTrRowsBatch + kTrRowsTable / 2 <= kTrRowsCap, "batch staging");
constexpr int kTrBatchSteps = (kTrRowsBatch + kTrRowsBlock - 1) / kTrRowsBlock;

// Index of the first of the n nondecreasing values s[] that is >= x.
__device__ __forceinline__ int lds_first_ge(const uint32_t* s, int n, uint32_t x)
{
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (s[lo + half] < x) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// One workgroup per CU: 124 KB of staging, 16 KB of cursors and up to 16 KB
// of row starts (4096-pixel buckets) stay within the 160 KB of LDS.
__global__ void __launch_bounds__(kTrRowsBlock)
k_tr_rows(const uint64_t* __restrict__ b1, const int64_t* __restrict__ hoff, int32_t nblk, int bshift,
          int32_t nb, int32_t P, int64_t* __restrict__ off, uint64_t* __restrict__ ev)
{
    __shared__ unsigned long long s_stage[kTrRowsCap];
    __shared__ uint32_t s_tab[kTrRowsTable];   // [G][2^bshift] counts, then cursors
    __shared__ int64_t s_warp[33];
    __shared__ uint32_t s_max;
    extern __shared__ uint32_t s_start[];      // [2^bshift] row starts (from the bucket start)
    const int nloc = 1 << bshift;
    const uint32_t lmask = (uint32_t)nloc - 1u;
    const int G = nloc >= kTrRowsTable ? 1 : min(kTrRowsWaves, kTrRowsTable / nloc);
    const int wpg = kTrRowsWaves / G;          // waves per group
    const int w = wave_id(), lane = lane_id();
    const int g = w / wpg, wj = w % wpg;
    uint32_t* cur = s_tab + g * nloc;
    const uint32_t step = (uint32_t)wpg * 64u;
    const int per = (nloc + kTrRowsBlock - 1) / kTrRowsBlock;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t s0 = hoff[b];   // chunk 0's segment starts the bucket
        const int64_t s1 = b + 1 < nb ? hoff[b + 1] : hoff[(int64_t)nblk * nb];
        const uint32_t n = (uint32_t)(s1 - s0);
        const uint64_t* seg = b1 + s0;
        const uint32_t L = (n + (uint32_t)G - 1u) / (uint32_t)G;
        const uint32_t ga = min(n, (uint32_t)g * L), gb = min(n, ga + L);   // this group's slice
        const uint32_t e_first = ga + (uint32_t)wj * 64u + (uint32_t)lane;
        const uint32_t e_rest = ga + (uint32_t)wj * 64u + (uint32_t)kTrRowsRegs * step;   // beyond the registers
        for (int i = threadIdx.x; i < G * nloc; i += kTrRowsBlock) s_tab[i] = 0u;
        if (threadIdx.x == 0) s_max = 0u;
        uint64_t xr[kTrRowsRegs];
#pragma unroll
        for (int t = 0; t < kTrRowsRegs; t++) {
            const uint32_t e = e_first + (uint32_t)t * step;
            // read once into registers: nontemporal (0.232 -> 0.222 ms on 1 Mpx)
            xr[t] = e < gb ? __builtin_nontemporal_load(seg + e) : ~0ull;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kTrRowsRegs; t++)
            if (xr[t] != ~0ull) atomicAdd(&cur[(uint32_t)(xr[t] >> 16) & lmask], 1u);
        for (uint32_t e0 = e_rest; e0 < gb; e0 += step) {
            const uint32_t e = e0 + (uint32_t)lane;
            if (e < gb) atomicAdd(&cur[(uint32_t)(seg[e] >> 16) & lmask], 1u);
        }
        __syncthreads();
        // row lengths, their exclusive scan (row offsets), the longest row
        const int i0 = threadIdx.x * per;
        uint32_t sum = 0u, mx = 0u;
        for (int u = 0; u < per; u++)
            if (i0 + u < nloc) {
                uint32_t c = 0u;
                for (int q = 0; q < G; q++) c += s_tab[q * nloc + i0 + u];
                sum += c;
                mx = max(mx, c);
            }
        mx = read_lane_u32(wave_incl_max_u32(mx), 63);
        if (lane == 0 && mx) atomicMax(&s_max, mx);
        int64_t tot;
        uint32_t ex = (uint32_t)block_excl_scan((int64_t)sum, s_warp, tot);
        for (int u = 0; u < per; u++) {
            const int i = i0 + u;
            if (i < nloc) {
                s_start[i] = ex;
                const int64_t c_id = ((int64_t)b << bshift) + i;
                if (c_id < P) off[c_id] = s0 + ex;
                for (int q = 0; q < G; q++) {   // cursors: after the earlier slices' events
                    const uint32_t c = s_tab[q * nloc + i];
                    s_tab[q * nloc + i] = ex;
                    ex += c;
                }
            }
        }
        if (b == nb - 1 && threadIdx.x == 0) off[P] = s1;
        __syncthreads();
        const uint32_t maxrow = s_max;
        if (2 * maxrow > (uint32_t)kTrRowsCap || n > 4 * ((uint32_t)kTrRowsCap - maxrow)) {
            // long rows or a large bucket: the bucket in batches of kTrRowsBatch
            // events in bucket order; a batch is grouped by pixel in LDS and
            // each pixel's group is written to its row as one run, where
            // scattering single events straight to their rows wrote 3.3 times
            // the bytes (largest batched job: 4.1 -> 2.7 ms).  Ranks are taken
            // one event per thread and step with a barrier between steps, so a
            // row's events keep bucket order except among those of one step,
            // which the row kernels' sorts put right (sorting each group here
            // cost more than it saved there)
            uint32_t* bc = s_tab;                                                // [nloc] batch counts
            uint32_t* bs = reinterpret_cast<uint32_t*>(s_stage + kTrRowsBatch);  // [nloc] batch group starts
            uint32_t* rowpos = s_start;                                          // [nloc] next slot of each row
            for (uint32_t b0 = 0; b0 < n; b0 += kTrRowsBatch) {
                const uint32_t nbt = min(n - b0, (uint32_t)kTrRowsBatch);
                for (int i = threadIdx.x; i < nloc; i += kTrRowsBlock) bc[i] = 0u;
                __syncthreads();
                // (the events are read again for the staging below, from cache:
                // holding them here as well would spill)
                uint32_t rk[kTrBatchSteps];
#pragma unroll
                for (int k = 0; k < kTrBatchSteps; k++) {
                    const uint32_t i = (uint32_t)k * kTrRowsBlock + threadIdx.x;
                    rk[k] = 0u;
                    if (i < nbt) rk[k] = atomicAdd(&bc[(uint32_t)(seg[b0 + i] >> 16) & lmask], 1u);
                    __syncthreads();
                }
                uint32_t bsum = 0u;
                for (int u = 0; u < per; u++)
                    if (i0 + u < nloc) bsum += bc[i0 + u];
                int64_t btot;
                uint32_t bex = (uint32_t)block_excl_scan((int64_t)bsum, s_warp, btot);
                for (int u = 0; u < per; u++)
                    if (i0 + u < nloc) {
                        bs[i0 + u] = bex;
                        bex += bc[i0 + u];
                    }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kTrBatchSteps; k++) {
                    const uint32_t i = (uint32_t)k * kTrRowsBlock + threadIdx.x;
                    if (i < nbt) {
                        const uint64_t x = seg[b0 + i];
                        s_stage[bs[(uint32_t)(x >> 16) & lmask] + rk[k]] = x;
                    }
                }
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < nbt; i += kTrRowsBlock) {
                    const uint64_t x = s_stage[i];
                    const uint32_t p = (uint32_t)(x >> 16) & lmask;
                    ev[s0 + rowpos[p] + (i - bs[p])] = (x & 0xFFFFFFFF00000000ull) | (x & 0xFFFFull);
                }
                __syncthreads();
                for (int i = threadIdx.x; i < nloc; i += kTrRowsBlock) rowpos[i] += bc[i];
                __syncthreads();
            }
            continue;   // (the batches end on a barrier)
        }
        const uint32_t span = (uint32_t)kTrRowsCap - maxrow;
        for (uint32_t lo = 0; lo < n; lo += span) {
            const uint32_t hi = lo + span;
            // pixels whose rows start in [lo, hi) and the window their rows fill
            const int pa = lds_first_ge(s_start, nloc, lo);
            const int pb = lds_first_ge(s_start, nloc, hi);
            const uint32_t wlo = pa < nloc ? s_start[pa] : n;
            const uint32_t whi = pb < nloc ? s_start[pb] : n;
            if (whi == wlo) continue;
            auto place = [&](uint64_t x) {
                const uint32_t p = (uint32_t)(x >> 16) & lmask;
                const uint64_t y = (x & 0xFFFFFFFF00000000ull) | (x & 0xFFFFull);
                const uint32_t st = s_start[p];
                if (st >= lo && st < hi) s_stage[atomicAdd(&cur[p], 1u) - wlo] = y;
            };
            // an empty register barrier per held event: without it the
            // compiler hoists each event's pixel and output word out of this
            // loop, 2-3 more VGPRs per event, and the held events spill
#pragma unroll
            for (int t = 0; t < kTrRowsRegs; t++) asm volatile("" : "+v"(xr[t]));
#pragma unroll
            for (int t = 0; t < kTrRowsRegs; t++)
                if (xr[t] != ~0ull) place(xr[t]);
            for (uint32_t e0 = e_rest; e0 < gb; e0 += step) {
                const uint32_t e = e0 + (uint32_t)lane;
                if (e < gb) place(seg[e]);
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < whi - wlo; i += kTrRowsBlock) ev[s0 + wlo + i] = s_stage[i];
            __syncthreads();
        }
        __syncthreads();   // s_max and the row tables are rewritten for the next bucket
    }
}

// ---- bucket offsets --------------------------------------------------------------
constexpr int kScanBlock = 256;

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

// The per-chunk histograms are chunk-major, hist[c][b], so each chunk writes
// and reads its nb counters contiguously.  Segment (bucket b, chunk c) of the
// bucketed array starts at hoff[c][b] = (events of buckets < b) + (events of
// bucket b in chunks < c); hoff[nblk][0] holds the total.  The column scan
// runs over segments of kColChunks chunks: per-segment sums, their scan down
// each bucket's column with the bucket totals, one workgroup's scan of the
// totals (bucket bases), then the offsets.
constexpr int kColChunks = 16;

__global__ void __launch_bounds__(kScanBlock)
k_tr_colsum(const uint32_t* __restrict__ hist, int32_t nb, int32_t nblk, int64_t* __restrict__ part)
{
    const int32_t b = blockIdx.x * kScanBlock + threadIdx.x;
    if (b >= nb) return;
    const int32_t c0 = (int32_t)blockIdx.y * kColChunks, c1 = min(nblk, c0 + kColChunks);
    int64_t sum = 0;
    for (int32_t c = c0; c < c1; c++) sum += hist[(int64_t)c * nb + b];
    part[(int64_t)blockIdx.y * nb + b] = sum;
}

__global__ void __launch_bounds__(kScanBlock)
k_tr_colscan(int64_t* __restrict__ part, int32_t nb, int32_t nseg, int64_t* __restrict__ total)
{
    const int32_t b = blockIdx.x * kScanBlock + threadIdx.x;
    if (b >= nb) return;
    int64_t run = 0;
    for (int32_t r = 0; r < nseg; r++) {
        const int64_t v = part[(int64_t)r * nb + b];
        part[(int64_t)r * nb + b] = run;
        run += v;
    }
    total[b] = run;
}

// In place: total[b] becomes the exclusive scan, total[nb] the sum.
__global__ void __launch_bounds__(1024)
k_tr_bucketscan(int64_t* __restrict__ total, int32_t nb)
{
    __shared__ int64_t s[33];
    int64_t carry = 0;
    for (int32_t b0 = 0; b0 < nb; b0 += blockDim.x) {
        const int32_t b = b0 + threadIdx.x;
        const int64_t v = b < nb ? total[b] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan(v, s, tot);
        if (b < nb) total[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) total[nb] = carry;
}

__global__ void __launch_bounds__(kScanBlock)
k_tr_coloff(const uint32_t* __restrict__ hist, const int64_t* __restrict__ part, const int64_t* __restrict__ base,
            int32_t nb, int32_t nblk, int64_t* __restrict__ hoff)
{
    const int32_t b = blockIdx.x * kScanBlock + threadIdx.x;
    if (b >= nb) return;
    const int32_t c0 = (int32_t)blockIdx.y * kColChunks, c1 = min(nblk, c0 + kColChunks);
    int64_t run = base[b] + part[(int64_t)blockIdx.y * nb + b];
    for (int32_t c = c0; c < c1; c++) {
        hoff[(int64_t)c * nb + b] = run;
        run += hist[(int64_t)c * nb + b];
    }
    if (b == 0 && blockIdx.y == 0) hoff[(int64_t)nblk * nb] = base[nb];
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

// ---- side statistics: per (output frame, static bin) sums ---------------------------
// fs[f][s-1] = sum over masked pixels with sqmap == s of count * flatfield
// (SparseFilter/DenseFilter accumulate partitions_mean_ / partial means and the
// frame sums from the same values, src/xpcs/filter/sparse_filter.cpp:168-191).
constexpr int kSideEv = 4;   // events per thread in flight (decode, pixel id, bin)

// One workgroup per selected raw frame; an LDS table of the S bins (dynamic,
// S doubles). Each thread keeps kSideEv events in flight: their decodes, then
// their bins (a 1- or 2-byte per-pixel table; with a flat field also the row
// id for the weight) are loaded before the sums. The events of a
// frame come in raster order, so neighbouring lanes mostly share a bin: runs of
// equal bins are summed across the wave first and only each run's last lane
// adds into LDS (one lane per address instead of up to 64).
template <bool WEIGHT, typename BT>
__global__ void __launch_bounds__(256)
k_side_sparse(const uint8_t* __restrict__ imm, const int64_t* __restrict__ sel_off,
              const uint32_t* __restrict__ sel_dlen, const int32_t* __restrict__ sel_frame, int64_t nsel,
              EvFormat fm, const int32_t* __restrict__ pix2c, uint32_t npix_det,
              const BT* __restrict__ sbin_pix,
              const double* __restrict__ weight, int32_t S, double* __restrict__ fs)
{
    extern __shared__ double s_acc[];   // [S]
    for (int64_t k = blockIdx.x; k < nsel; k += gridDim.x) {
        for (int b = threadIdx.x; b < S; b += blockDim.x) s_acc[b] = 0.0;
        __syncthreads();
        const uint8_t* base = imm + sel_off[k];
        const uint32_t dlen = sel_dlen[k];
        const bool al4 = (reinterpret_cast<uintptr_t>(base) & 3u) == 0;
        // the loop bound is uniform: every lane takes part in the wave sums
        for (uint32_t jb = 0; jb < dlen; jb += blockDim.x * kSideEv) {
            uint32_t idx[kSideEv], v[kSideEv];
            int32_t c[kSideEv], b[kSideEv];
#pragma unroll
            for (int e = 0; e < kSideEv; e++) {
                const uint32_t j = jb + (uint32_t)e * blockDim.x + threadIdx.x;
                idx[e] = npix_det;
                v[e] = 0;
                if (j < dlen) tr_decode(base, dlen, al4, j, fm, idx[e], v[e]);
            }
#pragma unroll
            for (int e = 0; e < kSideEv; e++) {
                const bool ok = v[e] != 0 && idx[e] < npix_det;
                b[e] = ok ? (int32_t)sbin_pix[idx[e]] - 1 : -1;
                if constexpr (WEIGHT) c[e] = ok ? pix2c[idx[e]] : 0;
            }
#pragma unroll
            for (int e = 0; e < kSideEv; e++) {
                double x = b[e] >= 0 ? (double)v[e] : 0.0;
                if constexpr (WEIGHT) x *= b[e] >= 0 ? weight[c[e]] : 0.0;
                bool tail;
                const double run = wave_run_sum_f64(x, (uint32_t)b[e], tail);
                if (tail && b[e] >= 0 && run != 0.0) atomicAdd(&s_acc[b[e]], run);
            }
        }
        __syncthreads();
        double* row = fs + (int64_t)sel_frame[k] * S;
        for (int b = threadIdx.x; b < S; b += blockDim.x)
            if (s_acc[b] != 0.0) atomicAdd(&row[b], s_acc[b]);
        __syncthreads();
    }
}

// Dense rows [P][F]: thread = frame, loop over a chunk of compact pixels (which
// are grouped by static bin), flushing a running sum whenever the bin changes.
__global__ void __launch_bounds__(256)
k_side_dense(const uint32_t* __restrict__ rows, int32_t P, int32_t F, const int32_t* __restrict__ sbin_of_c,
             const double* __restrict__ weight, int32_t S, int32_t chunk, double* __restrict__ fs)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = blockIdx.y * chunk;
    const int c1 = min(P, c0 + chunk);
    if (f >= F) return;
    double acc = 0.0;
    int cur = c0 < c1 ? sbin_of_c[c0] : 0;
    for (int c = c0; c < c1; c++) {
        const int sb = sbin_of_c[c];
        if (sb != cur) {
            if (acc != 0.0) atomicAdd(&fs[(int64_t)f * S + cur - 1], acc);
            acc = 0.0;
            cur = sb;
        }
        acc += (double)rows[(int64_t)c * F + f] * (weight ? weight[c] : 1.0);
    }
    if (acc != 0.0) atomicAdd(&fs[(int64_t)f * S + cur - 1], acc);
}

EvFormat ev_format(const xpcs_mt_plan* p);

void launch_side(xpcs_mt_plan* p, hipStream_t s)
{
    const HostTables& h = p->h;
    timer_begin(p, s, "k_side");
    if (p->mode == 1) {
        const unsigned grid = p->nsel < 65536 ? (unsigned)p->nsel : 65536u;
        const size_t lds = (size_t)h.ns * sizeof(double);
        const uint32_t npd = (uint32_t)((int64_t)h.W * h.H);
        auto go = [&](auto weighted, const auto* bins) {
            k_side_sparse<decltype(weighted)::value><<<grid, 256, lds, s>>>(
                p->d_imm, p->d_sel_off, p->d_sel_dlen, p->d_sel_frame, p->nsel, ev_format(p), p->d_pix2c, npd,
                bins, p->d_weight, h.ns, p->d_fs);
        };
        const bool one = p->sbin_pix_bytes == 1;
        const auto* b8 = static_cast<const uint8_t*>(p->d_sbin_pix);
        const auto* b16 = static_cast<const uint16_t*>(p->d_sbin_pix);
        if (p->d_weight) {
            if (one) go(std::true_type{}, b8);
            else go(std::true_type{}, b16);
        } else {
            if (one) go(std::false_type{}, b8);
            else go(std::false_type{}, b16);
        }
    } else if (p->mode == 2) {
        const int chunk = 256;
        dim3 grid((unsigned)((h.F + 255) / 256), (unsigned)((h.P + chunk - 1) / chunk));
        k_side_dense<<<grid, 256, 0, s>>>(p->d_rows, h.P, h.F, p->d_sbin_of_c, p->d_weight, h.ns, chunk,
                                          p->d_fs);
    }
    timer_end(p, s);
}

// normalize_by_framesum (main.cpp:313-336): fw[f] = mean over frames of the
// frame sums / frame sum of f (0 for a frame without counts).  The frame sums
// are the per-(frame, static bin) side sums; the reference's 1/(avg * pixels)
// factors cancel in the ratio.
__global__ void __launch_bounds__(256)
k_frame_sums(const double* __restrict__ fs, int32_t F, int32_t S, double* __restrict__ fw)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double t = 0.0;
    for (int b = 0; b < S; b++) t += fs[(int64_t)f * S + b];
    fw[f] = t;
}

__global__ void __launch_bounds__(1024)
k_frame_weights(double* __restrict__ fw, int32_t F)
{
    __shared__ double s_red[16];
    double t = 0.0;
    for (int f = threadIdx.x; f < F; f += blockDim.x) t += fw[f];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = t;
    __syncthreads();
    double total = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) total += s_red[i];
    const double mean = total / (double)F;
    __syncthreads();
    for (int f = threadIdx.x; f < F; f += blockDim.x) fw[f] = fw[f] > 0.0 ? mean / fw[f] : 0.0;
}

void launch_frame_weights(xpcs_mt_plan* p, hipStream_t s)
{
    const HostTables& h = p->h;
    k_frame_sums<<<(h.F + 255) / 256, 256, 0, s>>>(p->d_fs, h.F, h.ns, p->d_fw);
    k_frame_weights<<<1, 1024, 0, s>>>(p->d_fw, h.F);
}

// ---- launchers ------------------------------------------------------------------
EvFormat ev_format(const xpcs_mt_plan* p)
{
    EvFormat f;
    f.kind = p->fmt;
    f.W = (uint32_t)p->h.W;
    f.H = (uint32_t)p->h.H;
    return f;
}

static int grid_for(int64_t n, int64_t cap)
{
    if (n <= 0) return 1;
    return (int)(n < cap ? n : cap);
}

// Bucket-major segment offsets from the chunk-major histograms; scratch holds
// bucket_scan_scratch(nb, nblk) values.
static void bucket_scan(const uint32_t* hist, int32_t nb, int32_t nblk, int64_t* hoff, int64_t* scratch,
                        hipStream_t s)
{
    const int32_t nseg = (nblk + kColChunks - 1) / kColChunks;
    int64_t* part = scratch;
    int64_t* total = scratch + (int64_t)nseg * nb;
    const dim3 grid((nb + kScanBlock - 1) / kScanBlock, nseg);
    k_tr_colsum<<<grid, kScanBlock, 0, s>>>(hist, nb, nblk, part);
    k_tr_colscan<<<grid.x, kScanBlock, 0, s>>>(part, nb, nseg, total);
    k_tr_bucketscan<<<1, 1024, 0, s>>>(total, nb);
    k_tr_coloff<<<grid, kScanBlock, 0, s>>>(hist, part, total, nb, nblk, hoff);
}

int64_t bucket_scan_scratch(int32_t nb, int32_t nblk)
{
    return (int64_t)((nblk + kColChunks - 1) / kColChunks) * nb + nb + 1;
}

// Buckets of 2^bshift pixel rows (at least 256: k_tr_rows scans a bucket with
// 256 threads; at most 4096 buckets: the per-chunk LDS histogram; by default
// about 1024, which keeps each chunk's bucket segments long), cf raw frames
// per chunk workgroup (default: at most 1024 chunks).  bucket_shift and
// chunk_frames override the defaults when non-zero.
TransposeShape transpose_shape(int64_t P, int64_t nsel, int32_t bucket_shift, int32_t chunk_frames)
{
    TransposeShape t;
    t.bshift = 8;
    if (bucket_shift > 8) t.bshift = bucket_shift;
    else if (bucket_shift == 0)
        while (((P + (int64_t(1) << t.bshift) - 1) >> t.bshift) > 1024) t.bshift++;
    while (((P + (int64_t(1) << t.bshift) - 1) >> t.bshift) > 4096) t.bshift++;
    t.nb = (int32_t)((P + (int64_t(1) << t.bshift) - 1) >> t.bshift);
    t.cf = nsel > 1024 ? (nsel + 1023) / 1024 : 1;
    if (chunk_frames > 0) t.cf = chunk_frames;
    t.nblk = (int32_t)((nsel + t.cf - 1) / t.cf);
    if (t.nblk < 1) t.nblk = 1;
    return t;
}

bool side_fused(const xpcs_mt_plan* p)
{
    return p->mode == 1 && p->d_fs && !p->d_weight && p->sbin_pix_bytes == 1 &&
           p->tshape.cf * (int64_t)p->h.ns <= kSideFuseDoubles;
}

void launch_transpose(xpcs_mt_plan* p, hipStream_t s, bool with_side)
{
    const TransposeShape& t = p->tshape;
    const uint32_t npd = (uint32_t)((int64_t)p->h.W * p->h.H);
    timer_begin(p, s, "k_tr_count");
    const EvFormat fm = ev_format(p);
    // the kernels are instantiated per word format (no decode branches)
    auto by_format = [&](auto launch) {
        if (fm.kind == kFmtUfxc) launch(std::integral_constant<int, kFmtUfxc>{});
        else if (fm.kind == kFmtRigaku) launch(std::integral_constant<int, kFmtRigaku>{});
        else launch(std::integral_constant<int, kFmtImm>{});
    };
    if (with_side) {
        // the chunk's output frames never exceed its cf raw frames
        const size_t lds = (size_t)((t.nb + 1) & ~1) * sizeof(uint32_t) + (size_t)t.cf * p->h.ns * sizeof(double);
        by_format([&](auto kind) {
            k_tr_count<true, decltype(kind)::value><<<t.nblk, kTrChunkBlock, lds, s>>>(
                p->d_imm, p->d_sel_off, p->d_sel_dlen, p->nsel, t.cf, fm, p->d_mapped, npd, t.bshift, t.nb, t.nblk,
                p->d_hist, p->d_counters, p->d_sel_frame, static_cast<const uint8_t*>(p->d_sbin_pix), p->h.ns,
                p->d_fs);
        });
    } else {
        by_format([&](auto kind) {
            k_tr_count<false, decltype(kind)::value><<<t.nblk, kTrChunkBlock, (size_t)t.nb * sizeof(uint32_t), s>>>(
                p->d_imm, p->d_sel_off, p->d_sel_dlen, p->nsel, t.cf, fm, p->d_mapped, npd, t.bshift, t.nb, t.nblk,
                p->d_hist, p->d_counters, nullptr, nullptr, 0, nullptr);
        });
    }
    timer_end(p, s);
    timer_begin(p, s, "k_tr_scan");
    bucket_scan(p->d_hist, t.nb, t.nblk, p->d_hoff, p->d_bsum, s);
    timer_end(p, s);
    timer_begin(p, s, "k_tr_bucket");
    by_format([&](auto kind) {
        k_tr_bucket<decltype(kind)::value>
            <<<t.nblk, kTrChunkBlock, (size_t)t.nb * (sizeof(uint32_t) + sizeof(unsigned long long)), s>>>(
                p->d_imm, p->d_sel_off, p->d_sel_dlen, p->d_sel_frame, p->nsel, t.cf, fm, p->d_mapped, npd, t.bshift,
                t.nb, t.nblk, p->d_hoff, p->d_b1);
    });
    timer_end(p, s);
    timer_begin(p, s, "k_tr_rows");
    k_tr_rows<<<t.nb < 256 ? t.nb : 256, kTrRowsBlock, (size_t)sizeof(uint32_t) << t.bshift, s>>>(
        p->d_b1, p->d_hoff, t.nblk, t.bshift, t.nb, (int32_t)p->nrows, p->d_off, p->d_ev);
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
