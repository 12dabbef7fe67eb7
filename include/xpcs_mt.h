/*
 * xpcs_mt.h -- C ABI of the MI355X multi-tau correlator.
 *
 * Plain pointers and sizes only; no C++ types cross this boundary and no
 * exception escapes it.  Every entry point returns an xpcs_status (0 = OK);
 * xpcs_last_error() gives a thread-local message for the last failure.
 *
 * The reference (AdvancedPhotonSource/xpcs-eigen) has no FFI for this path; its
 * in-process interface is the static class xpcs::Corr (src/xpcs/corr.h:66-126)
 * fed by a Reader -> Filter -> SparseData chain and parameterised by the
 * Configuration singleton.  Each entry point below names the reference
 * interface it replaces.
 *
 * Units.  Per-pixel sums are exact integers in "summed-bin" units: at level L
 * the binned value of frame k is x_L[k] = sum of the 2^L level-0 counts it
 * covers (the reference stores x_L / 2^L), with N = lastframe_L - tau' and
 * tau' = tau / 2^L:
 *     G2 = sum_{t<N} x_L[t] x_L[t+tau'],  IP = sum_{t<N} x_L[t],
 *     IF = sum_{tau'<=t<lastframe_L} x_L[t].
 * The reference's float outputs are G2/4^L/N, IP/2^L/N, IF/2^L/N (times the
 * flat-field and 1/avg factors of each pixel).
 *
 * Ownership.  A plan owns all device memory it allocates.  Host buffers passed
 * in are read during the call and never retained.  Inputs are never mutated.
 *
 * Threading.  One plan per host thread; plans on different streams may run
 * concurrently.  The library keeps no mutable global state besides the
 * thread-local error message.
 */
#ifndef XPCS_MT_H
#define XPCS_MT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t xpcs_status;
#define XPCS_OK 0
#define XPCS_ERR_INVALID_ARG 1   /* bad sizes, NULL pointers, inconsistent desc */
#define XPCS_ERR_HIP 2           /* a HIP runtime call failed */
#define XPCS_ERR_FORMAT 3        /* malformed / truncated IMM data */
#define XPCS_ERR_RANGE 4         /* pixel index outside the detector */
#define XPCS_ERR_UNSUPPORTED 5   /* option combination outside the exact path */
#define XPCS_ERR_STATE 6         /* call out of order (e.g. read before execute) */
#define XPCS_ERR_NOMEM 7         /* device allocation failed */
#define XPCS_ERR_NO_DEVICE 8     /* no HIP device visible */

/* desc.flags */
#define XPCS_MT_TIMING 0x1u      /* record HIP events around every kernel */
#define XPCS_MT_SIDE_STATS 0x2u  /* also sum counts per (frame, static bin) for frameSum /
                                    partition-mean-partial (xpcs_mt_read_side) */
/* normalize_by_framesum (src/xpcs/main.cpp:313-336): every value of frame f is
 * scaled by mean_f(frameSum) / frameSum[f] before the correlation, so G2/IP/IF
 * are FP64 sums (xpcs_mt_read_sums_f64; xpcs_mt_read_sums_i64 refuses) formed
 * by the per-pixel dense-series kernel; two-time uses the same values.  Side
 * statistics stay those of the raw counts, as in the reference.  Not for rows
 * ingests (no frame sums) or streams. */
#define XPCS_MT_NORMALIZE_BY_FRAMESUM 0x4u

typedef struct xpcs_mt_plan xpcs_mt_plan;

/*
 * Analysis description.  Replaces the Configuration singleton fields the
 * multi-tau path reads (src/xpcs/configuration.cpp:80-242, :594-606).
 */
typedef struct xpcs_mt_desc {
    int32_t width;              /* x_dimension (columns)                      */
    int32_t height;             /* y_dimension (rows)                         */
    int32_t frames;             /* output frames = getFrameTodoCount()        */
    int32_t delays_per_level;   /* delays_per_level (>= 1)                    */
    int32_t stride;             /* stride_frames (>= 1)                       */
    int32_t average;            /* avg_frames (>= 1)                          */
    int32_t frame_from;         /* getFrameStartTodo(): first data frame, 0-based */
    int32_t dark_frames;        /* getDarkFrames(): leading dark frames (dense files) */
    float lld;                  /* getDarkThreshold()                         */
    float sigma;                /* getDarkSigma()                             */
    const int32_t* dqmap;       /* [height*width] dynamic q-map, row-major    */
    const int32_t* sqmap;       /* [height*width] static q-map, row-major     */
    const double* flatfield;    /* [height*width] or NULL (flatfield_enabled) */
    uint32_t flags;             /* XPCS_MT_* */
} xpcs_mt_desc;

typedef struct xpcs_mt_info {
    int32_t taus;               /* T = Corr::delaysPerLevel(...).size()       */
    int32_t max_level;          /* Corr::calculateLevelMax(frames, dpl)        */
    int32_t dynamic_bins;       /* Q (rows of norm-0-g2 per tau)               */
    int32_t static_bins;        /* S (largest static bin)                      */
    int32_t pixels;             /* P masked pixels (dqmap>=1 && sqmap>=1)      */
    int32_t mapped_pixels;      /* pixels that reach normalisation (m_mapping) */
    int32_t groups;             /* (q, s) groups of m_mapping                  */
    int32_t exact_pixels;       /* dense streaming path: pixels whose lag sums
                                   were recomputed in 64-bit integers (total
                                   count above 94906265, beyond exact FP64) */
    int64_t events;             /* level-0 events after the last ingest        */
    int64_t long_pixels;        /* dense ingests: every pixel (dense-row or
                                   streaming kernels); sparse: pixels whose rows
                                   hold more than 1024 events                 */
    int32_t deep_level;         /* sparse path: first level of the one-lane-per-pixel
                                   kernel (levels above it; = level count: none) */
    int32_t deep_stride;        /* its series row, u32 values per pixel (0: none) */
} xpcs_mt_info;

/* Library version string. */
const char* xpcs_version(void);

/* Thread-local message for the last non-OK status. */
const char* xpcs_last_error(void);

/*
 * Tau schedule.  Replace Corr::calculateLevelMax (src/xpcs/corr.cpp:1156-1160),
 * Corr::delaysPerLevel (:1133-1154) and Corr::calculateDelayCount (:1162-1164).
 * xpcs_delays_per_level writes up to cap entries and stores the full length in
 * *count.
 */
int32_t xpcs_level_max(int32_t frames, int32_t dpl);
int32_t xpcs_delay_count(int32_t dpl, int32_t level);
xpcs_status xpcs_delays_per_level(int32_t frames, int32_t dpl, int32_t max_level,
                                  int32_t* level, int32_t* tau, int32_t cap, int32_t* count);

/*
 * q-partition tables, no device needed (Configuration::BuildQMap,
 * src/xpcs/configuration.cpp:244-381).  order[] lists the flat index of every
 * masked pixel (dqmap >= 1 && sqmap >= 1): first the pixels of m_mapping in its
 * iteration order (q ascending, s ascending, raster order), then the pixels the
 * sbin->qbin dedupe drops (raster order).  Groups (q, s, [begin, end)) index
 * order[].  counts[0..4] = pixels, mapped pixels, groups, max q, max s.  Pass
 * NULL arrays to query counts first; order must hold `pixels` entries and the
 * group arrays group_cap entries.
 */
xpcs_status xpcs_build_qmap(int32_t width, int32_t height, const int32_t* dqmap,
                            const int32_t* sqmap, int32_t* order, int32_t* group_q,
                            int32_t* group_s, int32_t* group_begin, int32_t* group_end,
                            int32_t group_cap, int32_t* counts);

/*
 * Plan: validates the description, builds the q-partition tables
 * (Configuration::BuildQMap, src/xpcs/configuration.cpp:244-381) and the
 * schedule, and allocates the device tables on the current HIP device.
 */
xpcs_status xpcs_mt_plan_create(const xpcs_mt_desc* desc, xpcs_mt_plan** plan);
xpcs_status xpcs_mt_plan_destroy(xpcs_mt_plan* plan);
xpcs_status xpcs_mt_plan_info(const xpcs_mt_plan* plan, xpcs_mt_info* info);
/* Schedule (level, tau) of the plan, T entries each. */
xpcs_status xpcs_mt_schedule(const xpcs_mt_plan* plan, int32_t* level, int32_t* tau);
/* Flat detector index of each compact pixel, P entries (m_mapping order). */
xpcs_status xpcs_mt_pixel_order(const xpcs_mt_plan* plan, int32_t* flat_index);

/*
 * Ingest a whole IMM file held in host memory.  Replaces the load loop of
 * src/xpcs/main.cpp:222-268 with io::Imm (src/xpcs/io/imm.cpp:59-129): dark
 * frames first (uncompressed files), skip to frame_from, then
 * frames*stride*average raw frames.  Compressed files take the sparse path
 * (SparseFilter, src/xpcs/filter/sparse_filter.cpp:115-193), uncompressed files
 * the dense path (DenseFilter, src/xpcs/filter/dense_filter.cpp:121-210).
 * The needed bytes are copied to the device on the plan's stream.
 */
xpcs_status xpcs_mt_ingest_imm(xpcs_mt_plan* plan, const void* imm, size_t nbytes);

/*
 * Ingest an IMM image already resident in device memory.  payload_off[i] and
 * dlen[i] describe raw frame i (byte offset of its payload in dev_imm and its
 * entry count, as Imm::NextFrames reads them); compressed != 0 selects the
 * sparse layout.  The device buffer must stay valid until execute returns.
 */
xpcs_status xpcs_mt_ingest_imm_device(xpcs_mt_plan* plan, const void* dev_imm, size_t nbytes,
                                      const int64_t* payload_off, const uint32_t* dlen,
                                      int64_t raw_frames, int32_t compressed);

/*
 * Packed detector words (the readers of src/xpcs/io/ufxc.cpp:59-168 and
 * src/xpcs/io/rigaku.cpp:63-273, selected by the reference's --ufxc / --rigaku):
 *   XPCS_FMT_UFXC    u32 words: 11-bit wrapping frame counter (w >> 21),
 *                    2-bit value ((w >> 15) & 3), column-major pixel (w & 0x7fff)
 *   XPCS_FMT_RIGAKU  u64 words: frame (w >> 40), column-major pixel
 *                    ((w >> 16) & 0xfffff), 11-bit value (w & 0x7ff)
 * Column-major pixel p is detector pixel (p % height) * width + p / height.
 * The frame logic of each reader (UFXC: counter unwrap, frames from 0 since
 * its SkipFrames does nothing; Rigaku: the flush/average state machine that
 * makes no output frame for a file frame without words, stops at `frames`
 * output frames and keeps the last one only with more than one pixel) runs on
 * the host; the words are copied to the device unchanged and decoded there.
 * Both formats are sparse: dark_frames of the plan must be 0.
 */
#define XPCS_FMT_UFXC 1
#define XPCS_FMT_RIGAKU 2
xpcs_status xpcs_mt_ingest_packed(xpcs_mt_plan* plan, const void* words, size_t nbytes, int32_t format);

/*
 * Raw frames (the HDF5 reader, src/xpcs/io/hdf5.cpp:62-221: dataset
 * /entry/data/data, frames x height x width): nframes frames of width*height
 * values, 2 or 4 bytes each, unsigned, frame-major.  Non-zero values are the
 * events (the reader keeps them, then the sparse filter); frames are skipped to
 * frame_from, no dark frames (the plan's dark_frames must be 0).  Values must
 * be below 65536.
 */
xpcs_status xpcs_mt_ingest_frames(xpcs_mt_plan* plan, const void* frames, int64_t nframes, int32_t bytes_per_value);

/*
 * Host-only: the frame records xpcs_mt_ingest_packed would build for `desc`
 * (run of words first_word[i] .. + count[i] - 1 in output frame frame[i]).
 * Writes up to cap records, total in *n.
 */
xpcs_status xpcs_packed_records(const xpcs_mt_desc* desc, const void* words, size_t nbytes, int32_t format,
                                int64_t* first_word, uint32_t* count, int32_t* frame, int64_t cap, int64_t* n);

/*
 * Index the headers of an IMM file in host memory (io::Header fields of
 * src/xpcs/io/imm.h:63-144).  Writes up to cap entries, total in *frames.
 */
xpcs_status xpcs_imm_index(const void* imm, size_t nbytes, int64_t* payload_off, uint32_t* dlen,
                           double* elapsed, int32_t* corecotick, int64_t cap, int64_t* frames,
                           int32_t* compressed);

/*
 * Ingest frame-major events from a caller's own decoder (what io::Imm hands a
 * SparseFilter for a compressed file, src/xpcs/io/imm.cpp:87-99): raw frame k
 * holds the events frame_off[k] .. frame_off[k+1]-1 of pixel[] (raster index
 * y*W+x) and count[].  Frame selection as for a compressed IMM file
 * (frame_from, stride, average; no dark frames); pixels outside the detector
 * are reported by execute, as for IMM input.
 */
xpcs_status xpcs_mt_ingest_events(xpcs_mt_plan* plan, const int32_t* pixel, const uint16_t* count,
                                  const int64_t* frame_off, int64_t nframes);

/*
 * Ingest pixel-major rows (the SparseData layout of
 * src/xpcs/data_structure/sparse_data.h:57-79 as filled by a Filter):
 * row_off[W*H+1] into frame[] (ascending per row) and count[] (integer
 * values).  Rows of masked-out pixels are ignored.
 */
xpcs_status xpcs_mt_ingest_rows(xpcs_mt_plan* plan, const int64_t* row_off, const int32_t* frame,
                                const uint32_t* count);

/*
 * Run the path on `stream` (a hipStream_t; NULL = the plan's own stream):
 * (a) frame->pixel transpose, (b)+(c) multi-tau binning and lag sums
 * (Corr::multiTau2, src/xpcs/corr.cpp:315-431), (d) partition reduction and
 * normalisation (Corr::normalizeG2s, src/xpcs/corr.cpp:927-1091).
 * Asynchronous; the read functions synchronise.
 */
xpcs_status xpcs_mt_execute(xpcs_mt_plan* plan, void* stream);
xpcs_status xpcs_mt_synchronize(xpcs_mt_plan* plan);

/* Exact per-pixel sums, [P][T] int64 in compact pixel order (summed-bin units). */
xpcs_status xpcs_mt_read_sums_i64(xpcs_mt_plan* plan, int64_t* G2, int64_t* IP, int64_t* IF);
/* G2/IP/IF as FP64, [pixels][taus] (the normalize_by_framesum sums, or the
 * exact integer sums converted). */
xpcs_status xpcs_mt_read_sums_f64(xpcs_mt_plan* plan, double* G2, double* IP, double* IF);

/*
 * Per-pixel sums in the reference's float units and layout [T][W*H]
 * (the G2s/IPs/IFs arrays of src/xpcs/main.cpp:193-201 after multiTau2; the
 * --g2out datasets).  Pixels without events are 0.
 */
xpcs_status xpcs_mt_read_ref_f32(xpcs_mt_plan* plan, float* G2, float* IP, float* IF);

/* g2 and stderr, [T][Q] float64 (norm-0-g2 / norm-0-stderr before the f32 cast). */
xpcs_status xpcs_mt_read_g2(xpcs_mt_plan* plan, double* g2, double* stderr_);

/*
 * Normalise a caller's per-pixel arrays: Corr::normalizeG2s(G2, IP, IF)
 * (src/xpcs/corr.h:106-108, src/xpcs/corr.cpp:927-1091, called at
 * src/xpcs/main.cpp:441-445 on the accumulated G2s/IPs/IFs).  G2, IP and IF
 * are float arrays in the reference's units and layout, [taus][pixels] with
 * pixels = width*height (the column-major pixels x taus matrices the
 * reference maps over them; what multiTau2 accumulated, or
 * xpcs_mt_read_ref_f32 returned).  Uses only the plan's q-map and schedule
 * (no ingest needed): static-bin sums of each array, g2s = (sum G2 / c) /
 * ((sum IP / c) (sum IF / c)), the NaN-skipping mean per dynamic bin, and the
 * stderr of G2/(IP IF) over the pixels of each dynamic bin with NaN counted
 * as 0.  The values already carry any flat field, so it is not applied again.
 * FP64 on the device; g2 and stderr are [taus][Q] and also become the plan's
 * xpcs_mt_read_g2 results.  Synchronous; the caller's arrays are not kept.
 */
xpcs_status xpcs_mt_normalize_ref_f32(xpcs_mt_plan* plan, const float* G2, const float* IP, const float* IF,
                                      int64_t pixels, int32_t taus, double* g2, double* stderr_);

/*
 * xpcs_mt_normalize_ref_f32 for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  No counterpart in the
 * reference.  G2, IP, IF are DEVICE arrays [taus][pixels] and g2, stderr_
 * DEVICE arrays [taus][Q]; the work is enqueued on `stream` (a hipStream_t;
 * NULL = the plan's stream) and the call returns without waiting, allocating
 * or freeing anything.  Intermediates go to a fixed 64 MiB device scratch
 * area per GPU, held from library load until the process ends: jobs needing
 * more (about 24 x pixels x taus bytes of sums plus the partials) are refused
 * with XPCS_ERR_UNSUPPORTED.  Every call on one GPU shares that area, so calls
 * that overlap on one GPU (other threads, other streams) race; serialise them.
 * Results equal xpcs_mt_normalize_ref_f32's bit for bit; the plan's own
 * xpcs_mt_read_g2 results are not touched.
 */
xpcs_status xpcs_mt_normalize_ref_f32_noalloc(xpcs_mt_plan* plan, const float* G2, const float* IP,
                                              const float* IF, int64_t pixels, int32_t taus, double* g2,
                                              double* stderr_, void* stream);

/*
 * Dark image of a dense ingest (DarkImage, src/xpcs/data_structure/dark_image.cpp:81-106):
 * per-pixel Welford mean and standard deviation, [W*H] float64 each (zeros
 * without dark frames); the --darkout DarkAvg/DarkStd datasets.
 */
xpcs_status xpcs_mt_read_dark(xpcs_mt_plan* plan, double* dark_avg, double* dark_std);

/* Device pointers of the [T][Q] results (valid until the next execute/destroy). */
xpcs_status xpcs_mt_device_results(xpcs_mt_plan* plan, const double** g2, const double** stderr_);

/*
 * Side statistics of the load loop (SparseFilter/DenseFilter accessors,
 * src/xpcs/filter/filter.h:62-96, scaled as src/xpcs/main.cpp:339-411 does):
 * pixel_sum [W*H] (pixelSum), frame_sum [2][frames] (frameSum),
 * partition_mean [S] (partition-mean-total),
 * partition_partial [floor(frames/window)][S] (partition-mean-partial).
 * window = static_mean_window_size.  Any pointer may be NULL.
 */
xpcs_status xpcs_mt_read_side(xpcs_mt_plan* plan, int32_t window, float* pixel_sum,
                              float* frame_sum, float* partition_mean, float* partition_partial);

/*
 * Row classes of the last sparse execute (no reference counterpart: the
 * reference treats every pixel row alike, src/xpcs/corr.cpp:329-430).  For
 * class k the compact pixels and level-0 events whose rows hold: 0 no event,
 * 1 1..64 events (one event per lane), 2 65..128 (two per lane), 3 129..512
 * (LDS rows, 8 waves a workgroup), 4 513..2048 (LDS rows, 2 waves a
 * workgroup), 5 more (dense-series rows).  Six entries each; either array may
 * be NULL.  XPCS_ERR_STATE for dense ingests (no rows are formed).
 */
xpcs_status xpcs_mt_row_stats(xpcs_mt_plan* plan, int64_t* pixels, int64_t* events);

/*
 * Kernel timing (desc.flags & XPCS_MT_TIMING): per kernel name, the summed
 * event-measured milliseconds and launch count since the last reset.
 * Names are written as a '\n'-separated list into names (cap bytes).
 */
xpcs_status xpcs_mt_kernel_times(xpcs_mt_plan* plan, char* names, size_t cap, double* ms,
                                 int64_t* launches, int32_t max_kernels, int32_t* n_kernels);
xpcs_status xpcs_mt_kernel_times_reset(xpcs_mt_plan* plan);
/*
 * Restrict kernel timing to the timers called `name` (NULL or "": all).  No
 * reference counterpart: an instrumentation knob, so that a run can time one
 * kernel without the other timers' events (two per timer, each at a kernel
 * boundary: all of them add about 3 % to the sparse 1 Mpx job).
 */
xpcs_status xpcs_mt_set_timing_filter(xpcs_mt_plan* plan, const char* name);

/*
 * Launch shapes of a plan (no reference counterpart: the reference's only knob
 * is the OpenMP schedule of multiTau2, src/xpcs/corr.cpp:329).  0 = automatic.
 *   tr_bucket_shift  pixels per transpose bucket = 2^shift (8..12)
 *   tr_chunk_frames  raw frames per transpose workgroup (>= 1)
 *   reg_blocks       workgroups of the register-resident row kernel (1..65535)
 *   lds_blocks       workgroups of the LDS row kernel (1..65535)
 *   dense_levels     levels per pass of the dense streaming kernel, i.e. the
 *                    depth of its register-resident lag window
 *                    (dpl 4: 3, 4, 5; dpl 8: 2, 3)
 *   deep_frames      sparse rows of at most 128 events hand their levels over
 *                    to a one-lane-per-pixel kernel from the first level with
 *                    at most deep_frames binned frames (16..96; negative: no
 *                    hand-over; 0: 96).  Only for dpl 4 or 8, T <= 128 and at
 *                    most 48 taus from that level on (otherwise no hand-over).
 *   sums64           0: integer IP and IF sums are kept on the device as 32-bit
 *                    words plus the high words of the pixels whose total
 *                    passes 2^32 - 1 (the normalisation reads 16 instead of
 *                    24 bytes per pixel and tau); 1: as int64.  Results are
 *                    the same either way; read_sums returns int64.
 */
typedef struct xpcs_mt_tuning {
    int32_t tr_bucket_shift;
    int32_t tr_chunk_frames;
    int32_t reg_blocks;
    int32_t lds_blocks;
    int32_t dense_levels;
    int32_t deep_frames;
    int32_t sums64;
    int32_t reserved;
} xpcs_mt_tuning;

xpcs_status xpcs_mt_set_tuning(xpcs_mt_plan* plan, const xpcs_mt_tuning* tuning);
/* The shapes the next execute uses (automatic values resolved). */
xpcs_status xpcs_mt_get_tuning(xpcs_mt_plan* plan, xpcs_mt_tuning* tuning);

/*
 * Autotuner keyed on occupancy: the plan's key is (ingest mode, dpl,
 * log2 of events per pixel and frame in half octaves, log2 pixels, log2
 * frames).  A key seen before in this process takes the cached shapes;
 * otherwise candidate shapes are timed on the plan's own data (coordinate
 * descent, HIP events, best of 2 runs each) and the fastest is cached.  The
 * chosen shapes are set on the plan and written to *best (may be NULL).
 * Runs jobs on the plan's internal stream; results of the last run are valid.
 */
xpcs_status xpcs_mt_autotune(xpcs_mt_plan* plan, xpcs_mt_tuning* best);

/*
 * Batched execution (no reference counterpart; the reference runs one job per
 * process).  Plans already ingested are added to a batch; the first execute
 * (re)captures every job into one hipGraph -- the jobs spread over `branches`
 * parallel graph branches, longest first -- and every execute is one graph
 * launch on `stream`.  Re-ingesting or re-tuning a plan makes the next
 * execute recapture.  Results are read from each plan as after
 * xpcs_mt_execute.  flags: XPCS_BATCH_AUTOTUNE tunes every plan before the
 * capture.  Plans must outlive the batch and not be executed on their own
 * while a batch launch is in flight.
 */
typedef struct xpcs_mt_batch xpcs_mt_batch;
#define XPCS_BATCH_AUTOTUNE 0x1u

xpcs_status xpcs_mt_batch_create(int32_t branches, xpcs_mt_batch** out);
xpcs_status xpcs_mt_batch_add(xpcs_mt_batch* batch, xpcs_mt_plan* plan, int32_t* index);
xpcs_status xpcs_mt_batch_execute(xpcs_mt_batch* batch, void* stream, uint32_t flags);
xpcs_status xpcs_mt_batch_synchronize(xpcs_mt_batch* batch);
xpcs_status xpcs_mt_batch_destroy(xpcs_mt_batch* batch);

/*
 * Two-time analysis (Corr::twotime, src/xpcs/corr.cpp:562-924; configuration
 * keys analysis_type "Twotime", qphi_bin_to_process, smoothing_method,
 * smoothing_filter, twotime2onetime_window_size, configuration.cpp:151,216-232).
 * For every listed dynamic bin present in the q-map (ascending):
 *   C2T[f0][f1] = sum_p x[p,f0] x[p,f1] / P_q  for f0 <= f1, 0 below,
 *   g2full[f][i] = mean of diagonal f, g2partials[f][w][i] (f < window,
 *   w < (frames - window) / window) = sum of diagonal f over window w / window,
 * with x the pixel values (counts / avg_frames, dense values per dark scale,
 * times the flat field) after smoothing: smoothing 1 = "StaticMap" (divide by
 * the mean of the pixel's static bin, per frame or, with average, its mean
 * over frames), 2 = "symmetric" (the same per dynamic bin), 0 = none.  FP64 on
 * the device; integer counts give exact sums while every C2T sum stays below
 * 2^53.  The sums are formed per q-bin either from the event pairs of every
 * pixel (the reference's loop, Sum_p n_p^2 / 2 products, FP64 atomics) or as
 * a dense Gram product on the FP64 matrix cores (P_q x F^2 products);
 * method 0 picks the cheaper one from the row lengths, 1 forces the Gram
 * product, 2 the event pairs (sparse ingests only: dense frames always use
 * the Gram product).  The Gram product builds a q-bin's values in pixel chunks
 * of at most chunk_bytes device bytes (0: 4 GiB).  Results stay on the device
 * until xpcs_mt_twotime runs again or the plan is destroyed.
 */
typedef struct xpcs_tt_desc {
    const int32_t* qbins;       /* qphi_bin_to_process (1-based dynamic bins)  */
    int32_t nq;
    int32_t smoothing;          /* 0 none, 1 StaticMap, 2 symmetric             */
    int32_t average;            /* smoothing_filter == "Average"                */
    int32_t window;             /* twotime2onetime_window_size (>= 1)           */
    int32_t method;             /* 0 auto, 1 Gram product, 2 event pairs        */
    int64_t chunk_bytes;        /* Gram value chunk size; 0 = default (4 GiB)   */
} xpcs_tt_desc;

xpcs_status xpcs_mt_twotime(xpcs_mt_plan* plan, const xpcs_tt_desc* desc, void* stream);
/* Processed q-bins (ascending) and the method each used (1 Gram, 2 event
 * pairs; either array may be NULL), rows/columns of sg, windows of g2partials. */
xpcs_status xpcs_mt_twotime_info(xpcs_mt_plan* plan, int32_t* nq, int32_t* qbins, int32_t* methods, int32_t cap,
                                 int32_t* sg_rows, int32_t* sg_cols, int32_t* total_partials);
/* C2T of the i-th processed q-bin, [frames][frames]. */
xpcs_status xpcs_mt_read_twotime(xpcs_mt_plan* plan, int32_t i, double* c2t);
/* g2full [frames][nq], g2partials [window][total_partials][nq] (the
 * reference's layouts, corr.cpp:882-912), sg [sg_rows][sg_cols]; any may be NULL. */
xpcs_status xpcs_mt_read_twotime_1d(xpcs_mt_plan* plan, double* g2full, double* g2partials, double* sg);

/*
 * Streams (no reference counterpart: the reference correlates a finished IMM
 * file, src/xpcs/main.cpp:222-268 then Corr::multiTau2 / Corr::normalizeG2s).
 * Frames arrive in pieces; the per-pixel level state (last and first 2*dpl
 * values of every level, 64-bit lag sums, pending pair values, totals) stays
 * in HBM, and a snapshot after n frames gives g2/stderr and sums identical to
 * a plan of n frames.  Compressed (sparse) IMM frames only; dpl 4 or 8; no
 * stride/average/skip/dark frames/flags.  Each stream has its own HIP stream,
 * so work of different streams overlaps on the device.
 */
typedef struct xpcs_mt_stream xpcs_mt_stream;

/* Device bytes a stream of at most max_frames frames holds (admission). */
xpcs_status xpcs_mt_stream_bytes(const xpcs_mt_desc* desc, int32_t max_frames, uint64_t* bytes);
xpcs_status xpcs_mt_stream_open(const xpcs_mt_desc* desc, int32_t max_frames, xpcs_mt_stream** out);
/* Whole compressed IMM frames in host memory (free again on return). */
xpcs_status xpcs_mt_stream_push_imm(xpcs_mt_stream* stream, const void* imm, size_t nbytes);
/* Frames already in device memory: payload offsets / event counts per frame
 * (host arrays); the device bytes must stay valid until the next snapshot
 * has finished. */
xpcs_status xpcs_mt_stream_push_imm_device(xpcs_mt_stream* stream, const void* dev_imm, size_t nbytes,
                                           const int64_t* payload_off, const uint32_t* dlen, int32_t nframes);
/* Frames pushed so far and the taus of their schedule. */
xpcs_status xpcs_mt_stream_frames(xpcs_mt_stream* stream, int32_t* frames, int32_t* taus);
/* Enqueue g2 for the frames pushed so far (asynchronous). */
xpcs_status xpcs_mt_stream_snapshot(xpcs_mt_stream* stream);
/* Frames covered by the newest snapshot known to have finished (non-blocking). */
xpcs_status xpcs_mt_stream_poll(xpcs_mt_stream* stream, int32_t* frames_done);
/* Results of the newest snapshot (waits for it): g2/stderr [T_n][Q],
 * sums [P][T_n] in the plan's pixel order. */
xpcs_status xpcs_mt_stream_read_g2(xpcs_mt_stream* stream, double* g2, double* stderr_, int32_t* taus);
xpcs_status xpcs_mt_stream_read_sums_i64(xpcs_mt_stream* stream, int64_t* G2, int64_t* IP, int64_t* IF);
xpcs_status xpcs_mt_stream_close(xpcs_mt_stream* stream);

#ifdef __cplusplus
}
#endif
#endif
