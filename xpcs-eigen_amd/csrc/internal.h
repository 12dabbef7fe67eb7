// Internal declarations of the MI355X multi-tau library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "xpcs_mt.h"

namespace xpcs {

constexpr int kWave = 64;
// Longest per-pixel event list the LDS kernel keeps on chip; longer rows go to
// the dense-row kernel.
constexpr int kSparseCap = 1024;
// Largest schedule supported by the per-wave LDS accumulators.
constexpr int kMaxTaus = 256;

// One schedule entry as the kernels need it.
struct TauParam {
    int32_t level;   // L
    int32_t tp;      // tau' = tau >> L
    int32_t lf;      // lastframe_L = floor(F / 2^L)
    int32_t N;       // lastframe_L - tau'
};

// One level of the schedule: its taus are [first, first+count), with
// contiguous tau' values lo..hi.
struct LevelParam {
    int32_t first;
    int32_t count;
    int32_t lo;
    int32_t hi;
    int32_t lf;
    int32_t shift;   // 2^shift = level-0 frames per level frame (== level)
};

// A normalisation chunk: pixels [begin, end) of one (q, s) group.
struct NormChunk {
    int32_t group;
    int32_t begin;
    int32_t end;
    int32_t q;
};

struct NormPartial {   // per (chunk, tau)
    int64_t g2, ip, ifs;   // integer sums (bit patterns of doubles when weighted)
    double n, mean, m2;    // Welford of x_p = N G2/(IP IF)
};

// Host-side tables derived from the description.
struct HostTables {
    int W = 0, H = 0, F = 0, dpl = 0, stride = 1, avg = 1, frame_from = 0, darks = 0;
    float lld = 0.f, sigma = 0.f;
    uint32_t flags = 0;
    // schedule
    int T = 0, max_level = 0;
    std::vector<int32_t> level, tau;
    std::vector<TauParam> taus;
    std::vector<LevelParam> levels;
    // q map
    int P = 0, n_mapped = 0, ngroups = 0, nq = 0, ns = 0;
    std::vector<int32_t> order, pix2c, group_q, group_s, group_begin, group_end, q_first_group;
    std::vector<int32_t> pixels_per_bin, sbin_of_c;
    std::vector<double> ff_det;      // [W*H] or empty
    std::vector<double> weight;      // [P] flat-field weight per compact pixel or empty
    std::vector<NormChunk> chunks;
    std::vector<int32_t> group_chunk_first;   // [ngroups+1]
    std::vector<int32_t> q_chunk_first;       // [nq+1]
};

// Launch shape of the bucketed sparse transpose.
struct TransposeShape {
    int bshift = 8;       // 2^bshift pixels per bucket
    int32_t nb = 0;       // buckets
    int64_t cf = 1;       // raw frames per chunk
    int32_t nblk = 1;     // chunks
};

struct KernelTimer {
    std::string name;
    hipEvent_t a, b;
};

// Word formats of sparse frame records (the transpose decodes them on the device).
constexpr int32_t kFmtImm = 0;
constexpr int32_t kFmtUfxc = 1;
constexpr int32_t kFmtRigaku = 2;
struct EvFormat {
    int32_t kind;
    uint32_t W, H;   // detector width and height (column-major pixel ids of UFXC / Rigaku)
};

}  // namespace xpcs

struct xpcs_mt_plan {
    xpcs::HostTables h;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream_last = nullptr;   // stream of the last execute
    bool own_stream = false;


    // device tables
    int32_t* d_pix2c = nullptr;
    xpcs::TauParam* d_taus = nullptr;
    xpcs::LevelParam* d_levels = nullptr;
    xpcs::NormChunk* d_chunks = nullptr;
    int32_t* d_group_chunk_first = nullptr;
    int32_t* d_group_q = nullptr;
    int32_t* d_group_size = nullptr;
    int32_t* d_q_chunk_first = nullptr;
    int32_t* d_q_first_group = nullptr;
    double* d_weight = nullptr;
    int32_t* d_c2pix = nullptr;
    int32_t* d_sbin_of_c = nullptr;
    void* d_sbin_pix = nullptr;          // static bin per raster pixel (0: unmapped), side statistics only
    int sbin_pix_bytes = 0;              // its element size: 1 (S < 256) or 2

    // ingest state: 0 none, 1 sparse IMM, 2 dense IMM, 3 rows
    int mode = 0;
    int32_t fmt = 0;                     // word format of the sparse frame records (kFmt*)
    bool fsnorm = false;                 // normalize_by_framesum: FP64 sums in the G2/IP/IF buffers
    double* d_fw = nullptr;              // frame weights mean(frameSum) / frameSum[f]
    const uint8_t* d_imm = nullptr;      // raw IMM bytes (owned when own_imm)
    uint8_t* d_imm_owned = nullptr;
    size_t imm_cap = 0;
    // dense frames evenly spaced in d_imm (payload of frame t at tile_off0 +
    // t * tile_step, 16-byte aligned, readable up to the next 64-pixel
    // boundary): the streaming kernel loads them as tiles; -1 otherwise
    // launch shapes (0 = automatic) and a counter bumped whenever the data or
    // the shapes change (batches recapture their graph then)
    xpcs_mt_tuning tune{};
    uint64_t generation = 0;
    bool capturing = false;              // inside a graph capture: no timer events
    // pixel-major event rows (CSR): nrows rows; pixel c's row is d_rowmap[c]
    // (its raster index, rows of the IMM transpose) or c (d_rowmap == NULL,
    // rows handed in by xpcs_mt_ingest_rows)
    int64_t nrows = 0;
    const int32_t* d_rowmap = nullptr;
    uint32_t* d_mapped = nullptr;        // bit per detector pixel: dqmap, sqmap >= 1
    int64_t tile_off0 = -1;
    int64_t tile_step = 0;
    bool side_vec = false;   // every selected dense frame 16-byte aligned, pixels % 8 == 0
    int64_t nsel = 0;                    // selected raw frames
    int64_t* d_sel_off = nullptr;        // payload byte offset per selected raw frame
    uint32_t* d_sel_dlen = nullptr;
    int32_t* d_sel_frame = nullptr;      // output frame per selected raw frame
    int64_t sel_cap = 0;
    int64_t raw_events = 0;              // sum of dlen over selected frames (sparse)
    std::vector<double> clock, ticks;    // timestamps of the data frames (host)
    // dense ingest
    int64_t nscale = 1;                  // value scale (dark frames)
    int64_t* d_dark_sum = nullptr;       // [P] per compact pixel
    int64_t* d_thr = nullptr;            // [P]
    std::vector<double> dark_avg, dark_std;   // [W*H] (host, --darkout)

    // events (pixel-major CSR over compact pixels)
    int64_t ev_cap = 0;
    uint64_t* d_ev = nullptr;            // (frame << 32) | count
    int64_t* d_off = nullptr;            // [P+1]
    uint64_t* d_b1 = nullptr;            // bucketed events (transpose stage)
    uint32_t* d_hist = nullptr;          // [nblk][nb] events per (chunk, bucket)
    int64_t* d_hoff = nullptr;           // [nblk][nb] segment starts, then the total
    int64_t hist_cap = 0;
    int64_t* d_bsum = nullptr;           // bucket scan scratch
    int64_t bsum_cap = 0;
    xpcs::TransposeShape tshape;
    int64_t events = 0;                  // known after execute (host copy)
    // dense rows (fallback path)
    uint32_t* d_rows = nullptr;          // [P][F]
    // two-time results (per processed q-bin, device FP64)
    std::vector<int32_t> tt_q;
    std::vector<double*> tt_c2t, tt_full, tt_part, tt_sg;
    std::vector<int32_t> tt_sg_rows;
    int32_t tt_sg_cols = 0, tt_total = 0, tt_wsize = 0;
    std::vector<int32_t> tt_method;      // per processed q-bin: 1 Gram, 2 event pairs
    double* tt_x = nullptr;              // scratch: values of one pixel chunk [rows][F]
    size_t tt_x_cap = 0;
    double* tt_dpart = nullptr;          // scratch: diagonal segment sums
    size_t tt_dpart_cap = 0;
    // result buffers of the previous call, reused by size by the next one
    std::vector<std::pair<double*, size_t>> tt_pool;
    size_t rows_cap = 0;
    // dense streaming path: coarsened series between passes, boundary events
    uint32_t* d_series_a = nullptr;
    uint32_t* d_series_b = nullptr;
    size_t series_cap = 0;
    int4* d_bnd = nullptr;               // (frame boundary, op, tau slot, 0), sorted
    int32_t nbnd = 0;
    // distinct prefix-sum boundaries B_j of the sparse row kernels (sorted)
    uint16_t* d_frame_rank = nullptr;    // [F] number of B_j <= frame
    int4* d_tau_bnd = nullptr;           // [T] (j of IP, j of the IF head, j of the level end, 0)
    int32_t nrank = 0;                   // number of distinct B_j
    // the same ranks per 64 frames: (boundary bits lo, hi, rank before the
    // word, 0); k_mt_reg stages them in LDS when that costs no occupancy
    // (rank_lds: dynamic LDS bytes, 0 = gather d_frame_rank)
    uint4* d_rank_words = nullptr;
    int32_t nrank_words = 0;
    int32_t rank_lds[2] = {-1, -1};   // [counts / FP64]; -1: undecided

    // long pixels
    int32_t* d_long_list = nullptr;      // [P]
    int32_t* d_mid_list = nullptr;       // [P] rows of 65..CAP events
    uint32_t* d_counters = nullptr;      // [0] rows over 1024 events, [1] error flags, [2] rows of 65..128,
                                         // [3] dense exact-redo pixels, [4] rows of 129..512, [5] 513..1024
    uint32_t* d_deep = nullptr;          // [P][deep stride] level series handed to the one-lane-per-pixel kernel
    size_t deep_cap = 0;
    uint64_t* d_scratch = nullptr;
    size_t scratch_cap = 0;
    int long_blocks = 0;

    // outputs
    int64_t* d_G2 = nullptr;             // [P][T]
    int64_t* d_IP = nullptr;             // [P][T] int64, or with ipif32 the 32-bit layout
    int64_t* d_IF = nullptr;             //   (low words [P*T], high words of pixels with totals >= 2^32 after them)
    bool ipif32 = false;                 // integer sums of the batch path (set by enqueue_execute)
    bool wide_rows = false;              // ingest_rows: a row of <= 128 events totals 2^32 or more
    uint64_t* d_pix_total = nullptr;     // [P] sum of level-0 values
    xpcs::NormPartial* d_part = nullptr; // [nchunks][T]
    double* d_g2s = nullptr;             // [ngroups][T]
    double* d_g2 = nullptr;              // [T][Q]
    double* d_se = nullptr;              // [T][Q]
    // side statistics
    double* d_fs = nullptr;              // [F][S] side statistics (XPCS_MT_SIDE_STATS)
    bool executed = false;
    bool normalized = false;             // g2/stderr hold an xpcs_mt_normalize_ref_f32 result

    // timing
    std::vector<xpcs::KernelTimer> pending;
    std::map<std::string, std::pair<double, int64_t>> times;
    std::vector<hipEvent_t> event_pool;
    std::string timer_only;              // time only this timer (empty: all)
    bool timer_skipped = false;          // the open timer is filtered out
};

namespace xpcs {

// host_tables.cpp
xpcs_status build_tables(const xpcs_mt_desc& d, HostTables& h);
// formats.cpp: frame records of packed detector words
struct PackedRecord {
    int64_t word;     // first word of the run
    uint32_t count;   // words in the run
    int32_t frame;    // output frame
};
xpcs_status ufxc_records(const uint32_t* w, int64_t n, const HostTables& h, std::vector<PackedRecord>& out);
xpcs_status rigaku_records(const uint64_t* w, int64_t n, const HostTables& h, std::vector<PackedRecord>& out);
xpcs_status packed_records(const void* words, size_t nbytes, int32_t format, const HostTables& h,
                           std::vector<PackedRecord>& rec);
xpcs_status build_qmap(const int32_t* dq, const int32_t* sq, HostTables& h);
void set_error(const char* fmt, ...);
xpcs_status fail(xpcs_status st, const char* fmt, ...);

// kernel launchers (kernels_*.hip); all asynchronous on `s`
TransposeShape transpose_shape(int64_t P, int64_t nsel, int32_t bucket_shift, int32_t chunk_frames);
int64_t bucket_scan_scratch(int32_t nb, int32_t nblk);
// with_side: also accumulate the side statistics' frame sums into d_fs (which
// the caller has zeroed); only when side_fused(p)
void launch_transpose(xpcs_mt_plan* p, hipStream_t s, bool with_side = false);
constexpr int64_t kSideFuseDoubles = 4096;   // LDS frame-sum table of the fused count pass (32 KB)
bool side_fused(const xpcs_mt_plan* p);
EvFormat ev_format(const xpcs_mt_plan* p);
void launch_mt_sparse(xpcs_mt_plan* p, hipStream_t s);
// first level of the one-lane-per-pixel kernel of the sparse path (the level
// count: none) and its series row length in u32 (0: none)
int deep_level(const xpcs_mt_plan* p, int* stride);
void rank_lds_setup(xpcs_mt_plan* p);
void launch_mt_long(xpcs_mt_plan* p, hipStream_t s, bool dense_rows);
void launch_dense_transpose(xpcs_mt_plan* p, hipStream_t s);
void launch_side(xpcs_mt_plan* p, hipStream_t s);
bool dense_stream_supported(const HostTables& h);
size_t dense_series_elems(const xpcs_mt_plan* p);
int dense_levels(const xpcs_mt_plan* p);
bool dense_levels_allowed(int dpl, int k);
int reg_blocks(const xpcs_mt_plan* p);
int lds_blocks(const xpcs_mt_plan* p);
void launch_fill_u32(uint32_t* d, size_t n, hipStream_t s);
void launch_fill_u64(uint64_t* d, size_t n, hipStream_t s);
xpcs_status prepare_execute(xpcs_mt_plan* p);
void launch_normalize_with(xpcs_mt_plan* p, hipStream_t s, int32_t t0, int32_t t1, int32_t S,
                           const TauParam* taus);
// where a normalisation writes: chunk partials [chunks][S], group values
// [groups][S], g2 and stderr [tau][Q]
struct NormOut {
    NormPartial* part;
    double* g2s;
    double* g2;
    double* se;
};
NormOut plan_norm_out(xpcs_mt_plan* p);
void launch_normalize_sums(xpcs_mt_plan* p, hipStream_t s, int32_t t0, int32_t t1, int32_t S,
                           const TauParam* taus, const int64_t* G2, const int64_t* IP, const int64_t* IF, bool fp,
                           bool w32, const double* weight, const NormOut& out);
// the statically placed scratch of the allocation-free entry point
// (noalloc.cpp): its address is resolved once per device, at plan creation
xpcs_status noalloc_setup(int device);
// reference-unit [T][W*H] floats -> FP64 sums [P][T] (summed-bin units)
void launch_from_ref_f32(xpcs_mt_plan* p, hipStream_t s, const float* G2, const float* IP, const float* IF,
                         int64_t* oG2, int64_t* oIP, int64_t* oIF);
void launch_mt_fsnorm(xpcs_mt_plan* p, hipStream_t s, bool dense_rows);
void launch_frame_weights(xpcs_mt_plan* p, hipStream_t s);
// two-time (k_twotime.hip)
void launch_tt_values(xpcs_mt_plan* p, int32_t c0, int32_t n, double scale, double* X, hipStream_t s);
void launch_tt_colsum(xpcs_mt_plan* p, const double* X, int32_t F, int32_t lo, int32_t hi,
                      const int32_t* d_bin_begin, const int32_t* d_bin_end, int32_t nbins, double* d_sgsum,
                      hipStream_t s);
void launch_tt_sgfinal(double* d_sg, int32_t F, const int32_t* d_bin_begin, const int32_t* d_bin_end, int32_t nbins,
                       bool average, double* d_sga, hipStream_t s);
void launch_tt_divide(xpcs_mt_plan* p, double* X, int32_t F, int32_t lo, int32_t n, const int32_t* d_bin_of_row,
                      const double* d_div, bool per_frame, hipStream_t s);
void launch_tt_gram(xpcs_mt_plan* p, const double* X, int32_t n, int32_t F, double inv_pq, double* C, hipStream_t s);
void launch_tt_pairs(xpcs_mt_plan* p, int32_t c0, int32_t Pq, double scale, const int32_t* d_bin_of_row,
                     const double* d_div, bool per_frame, double* C, hipStream_t s);
void launch_tt_pair_count(xpcs_mt_plan* p, const int32_t* d_c_begin, const int32_t* d_c_end, int32_t nq, double* d_out,
                          hipStream_t s);
void launch_tt_sg_sparse(xpcs_mt_plan* p, int32_t c0, int32_t Pq, double scale, const int32_t* d_bin_of_row,
                         double* d_sgsum, hipStream_t s);
int tt_diag_segments(int32_t F);
void launch_tt_diag(xpcs_mt_plan* p, double* C, int32_t F, int32_t wsize, int32_t total, double rescale,
                    double* partial, double* g2full, double* g2part, hipStream_t s);
void twotime_release(xpcs_mt_plan* p);

// online multi-tau (k_stream.hip)
constexpr int kStreamChunk = 64;   // frames per update launch
struct StreamFrames {             // passed by value: no upload per update
    int64_t off[kStreamChunk];     // payload offsets
    uint32_t dlen[kStreamChunk];   // events per frame
};
struct StreamLaunch {
    const uint8_t* imm;            // device IMM bytes of the pushed frames
    StreamFrames fr;
    int32_t nb;                    // frames of this update (<= kStreamChunk)
    const int32_t* pix2c;
    uint32_t npix_det;
    int32_t P;
    uint32_t* tile;                // [kStreamChunk][P]
    uint32_t* counters;
    int32_t d2;                    // 2 * dpl (8 or 16)
    int32_t n0;                    // frames before this update
    int32_t nlev;                  // levels with state
    uint32_t* win;
    uint32_t* head;
    uint64_t* acc;
    uint32_t* pend;
    uint64_t* total;
};
void launch_stream_update(const StreamLaunch& a, hipStream_t s);
void launch_stream_snapshot(const StreamLaunch& a, int32_t n, const LevelParam* levels, int32_t nlev,
                            const TauParam* taus_max, int32_t t0, int32_t t1, int32_t S, TauParam* taus,
                            int64_t* G2, int64_t* IP, int64_t* IF, hipStream_t s);
void enqueue_execute(xpcs_mt_plan* p, hipStream_t s);
void launch_mt_dense(xpcs_mt_plan* p, hipStream_t s, bool fsnorm);
void launch_side_dense_raw(xpcs_mt_plan* p, hipStream_t s);
void launch_normalize(xpcs_mt_plan* p, hipStream_t s);
void launch_ref_f32(xpcs_mt_plan* p, hipStream_t s, float* G2, float* IP, float* IF);

// timing helpers (api.cpp)
void timer_begin(xpcs_mt_plan* p, hipStream_t s, const char* name);
void timer_end(xpcs_mt_plan* p, hipStream_t s);
void collect_timers(xpcs_mt_plan* p);   // after the stream has drained

}  // namespace xpcs
