/*
 * xpcs_oracle.h -- CPU restatement of the reference multi-tau path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links,
 * loads or calls it.
 *
 * Each function cites the reference file:line range it restates (paths are
 * relative to the AdvancedPhotonSource/xpcs-eigen source tree).
 *
 * Two arithmetic modes are provided for the correlator:
 *   - exact:  int64 photon-count sums, the closed form of multiTau2
 *             (G2 = sum x_L[t] x_L[t+tau'], IP, IF over binned frames x_L),
 *             and an FP64 normaliser.
 *   - compat: a float32 replica of the reference arithmetic, including its
 *             in-place binning and its whole-vector lower_bound search over the
 *             stale tail of each row (src/xpcs/corr.cpp:356-410), and a float32
 *             replica of Corr::normalizeG2s (src/xpcs/corr.cpp:927-1091).
 *
 * Parity status of each function is listed in the oracle section of the
 * design notes; every function here is pinned by at least one test under
 * tests/ (closed forms, brute force, hand-built known answers, invariants).
 */
#ifndef XPCS_ORACLE_H
#define XPCS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tau schedule: src/xpcs/corr.cpp:1133-1164 ------------------------- */
int xo_level_max(int frames, int dpl);
int xo_delay_count(int dpl, int level);
/* Writes up to cap (level, tau) entries; returns the schedule length. */
int xo_delays_per_level(int frames, int dpl, int max_level,
                        int32_t* level, int32_t* tau, int cap);

/* ---- q-partition map: src/xpcs/configuration.cpp:244-381 ---------------- */
/*
 * order[]       : flat pixel indices of all masked pixels, first the pixels
 *                 of m_mapping in its iteration order (q asc, s asc, pixel asc),
 *                 then the pixels the sbin->qbin dedupe drops (asc).
 * group_*[]     : one entry per (q, s) group of m_mapping, in iteration order;
 *                 [begin, end) indexes order[].
 * q_first_group : per q = 1..nq_out, first group index (nq_out+1 entries).
 * pixels_per_bin: per sbin 1..static_out (counts every masked pixel).
 * counts[0..5]  : npix, n_mapped, ngroups, nq_out (= max q), static_out
 *                 (= max sbin), n_qkeys (= m_mapping.size()).
 */
int xo_build_qmap(int width, int height, const int32_t* dqmap, const int32_t* sqmap,
                  int32_t* order, int32_t* group_q, int32_t* group_s,
                  int32_t* group_begin, int32_t* group_end, int32_t* q_first_group,
                  int32_t* pixels_per_bin, int group_cap, int q_cap, int s_cap,
                  int32_t* counts);

/* ---- IMM: src/xpcs/io/imm.h:63-144, src/xpcs/io/imm.cpp:59-129 --------- */
typedef struct {
    int64_t payload_off;   /* byte offset of the frame payload */
    uint32_t dlen;         /* entries in the payload */
    int32_t compressed;    /* header.compression != 0 */
    double elapsed;        /* header.elapsed  (clock) */
    int32_t corecotick;    /* header.corecotick (ticks) */
    int32_t pad;
} xo_imm_frame;
/* Returns the number of whole frames indexed, or -1 if a payload overruns. */
int64_t xo_imm_index(const uint8_t* bytes, int64_t nbytes, xo_imm_frame* frames, int64_t cap);

/*
 * Raw frame selection of main.cpp:222-268: dark frames are the first `darks`
 * frames of the file (dense files only), then frames are skipped up to
 * frame_from, then `frames` output frames are built from frames*stride*avg raw
 * frames, using raw frame block*R + j*stride (j < avg, R = stride*avg).
 * sel[k] = raw frame index, out_frame[k] = output frame.  Returns count.
 */
int64_t xo_select_frames(int64_t n_raw, int darks, int frame_from, int frames,
                         int stride, int avg, int64_t* sel, int32_t* out_frame, int64_t cap);

/* ---- filters: per-pixel event lists in compact pixel order -------------- */
/*
 * Sparse transpose, src/xpcs/filter/sparse_filter.cpp:115-193.
 * pix2c[W*H] : compact pixel id or -1.  Produces a CSR over compact pixels with
 * frames ascending: off[P+1], frame[], cnt[] (integer count sum), val[] (the
 * reference's float value sum(count*ff)/avg).  Entries are made for every
 * masked pixel touched in an output frame, including zero counts (as the
 * reference does).  Returns the number of entries, -1 if a pixel index is out
 * of range, -2 if cap is too small.  With frame == NULL only counts (off).
 */
int64_t xo_sparse_transpose(const uint8_t* bytes, const xo_imm_frame* frames,
                            const int64_t* sel, const int32_t* out_frame, int64_t nsel,
                            int npix_det, const int32_t* pix2c, int32_t P,
                            const double* flatfield, int avg,
                            int64_t* off, int32_t* frame, int64_t* cnt, float* val, int64_t cap);

/* Dark image, src/xpcs/data_structure/dark_image.cpp:81-106 (Welford, double). */
void xo_dark_image(const uint8_t* bytes, const xo_imm_frame* frames, int darks,
                   int npix_det, const double* flatfield, double* dark_avg, double* dark_std,
                   int64_t* dark_sum);

/*
 * Dense filter, src/xpcs/filter/dense_filter.cpp:121-210.
 * Exact mode: values are integers scaled by n = max(darks,1):
 *   y = n*raw - dark_sum, kept iff y > floor(n * thresh_f) where
 *   thresh_f = float(lld + sigma*dark_std) (0 without darks); cnt = max(y,0).
 * val[] holds the reference's float value.  Same return convention as above.
 */
int64_t xo_dense_transpose(const uint8_t* bytes, const xo_imm_frame* frames,
                           const int64_t* sel, const int32_t* out_frame, int64_t nsel,
                           int npix_det, const int32_t* pix2c, int32_t P,
                           const double* flatfield, int avg, int darks,
                           const double* dark_avg, const double* dark_std, const int64_t* dark_sum,
                           float lld, float sigma,
                           int64_t* off, int32_t* frame, int64_t* cnt, float* val, int64_t cap);

/* ---- correlator -------------------------------------------------------- */
/*
 * Exact closed form of Corr::multiTau2 (src/xpcs/corr.cpp:315-431), by dense
 * brute force per pixel: x_0[t] from the CSR, x_L[k] = x_{L-1}[2k]+x_{L-1}[2k+1]
 * for k < floor(lastframe_{L-1}/2), then per schedule entry
 *   G2 = sum_{t<N} x[t] x[t+tau'], IP = sum_{t<N} x[t], IF = sum_{tau'<=t<lf} x[t],
 * N = lf - tau'.  Outputs are [P][T] int64 (pixel-major).
 */
void xo_multitau_exact(int64_t P, const int64_t* off, const int32_t* frame, const int64_t* cnt,
                       int frames, int T, const int32_t* level, const int32_t* tau,
                       int64_t* G2, int64_t* IP, int64_t* IF, int nthreads);

/*
 * Float replica of Corr::multiTau2 (src/xpcs/corr.cpp:315-431): in-place
 * pairwise binning with float halving, float accumulation and, when
 * stale_tail != 0, the whole-vector lower_bound partner search.  Each pixel's
 * row is copied first (the input is not mutated).  Outputs are written in the
 * reference layout [T][npix_det] (index tauIndex*pixels + pixel) and are
 * accumulated into (callers zero them, as main.cpp:193-201 does).
 * c2pix maps compact id -> flat pixel.  OpenMP dynamic,20 over pixels.
 */
void xo_multitau_compat(int64_t P, const int64_t* off, const int32_t* frame, const float* val,
                        const int32_t* c2pix, int64_t npix_det,
                        int frames, int T, const int32_t* level, const int32_t* tau,
                        int stale_tail, float* G2, float* IP, float* IF, int nthreads);

/* The per-pixel body of the float replica, in place on one row. */
void xo_mt_compat_row(int32_t* idx, float* vp, int64_t n0, int64_t pix, int64_t npix_det,
                      int frames, int T, const int32_t* level, const int32_t* tau,
                      int stale_tail, float* G2s, float* IPs, float* IFs);

/* ---- normalisation ------------------------------------------------------ */
/*
 * FP64 normaliser of the exact sums (closed form of corr.cpp:927-1091):
 *   g2s(s,tau) = c_s * N * sum G2 / (sum IP * sum IF)         (weights w_p:
 *   sum over p of w_p^2 G2_p, w_p IP_p, w_p IF_p when a flat field is set)
 *   g2(q,tau)  = mean of non-NaN g2s over s in q, NaN if none
 *   stderr     = sequential Welford over p in q of x_p = N G2_p/(IP_p IF_p),
 *                0/0 -> 0, every pixel counted (App. quirk of corr.cpp:1069-1070).
 * Sums are indexed by compact id (first n_mapped ids follow the group table).
 * Outputs [T][nq_out] (rows of missing q are NaN).
 */
void xo_normalize_f64(int T, const int32_t* Nvec, int64_t P,
                      const int64_t* G2, const int64_t* IP, const int64_t* IF,
                      int ngroups, const int32_t* group_q, const int32_t* group_begin,
                      const int32_t* group_end, int nq_out, const double* weight,
                      double* g2, double* stderr_);

/*
 * Float32 replica of Corr::normalizeG2s (src/xpcs/corr.cpp:927-1091) on the
 * reference-layout float arrays [T][npix_det]; the group table lists flat
 * pixel ids through order[].  Outputs [T][nq_out] float.
 */
void xo_normalize_compat(int T, int64_t npix_det, const float* G2, const float* IP, const float* IF,
                         const int32_t* order, int ngroups, const int32_t* group_q,
                         const int32_t* group_s, const int32_t* group_begin,
                         const int32_t* group_end, int nq_out, int static_out,
                         float* g2, float* stderr_);

/* Whole CPU baseline in the reference's decomposition (serial full-detector
 * transpose, OpenMP multitau, serial normalise); returns seconds per stage. */
int xo_cpu_pipeline(const uint8_t* bytes, int64_t nbytes, int width, int height,
                    const int32_t* dqmap, const int32_t* sqmap, int frames, int dpl,
                    int nthreads, float* g2_out, float* stderr_out, double* stage_seconds);

/* ---- two-time (Corr::twotimeQBinThreading, corr.cpp:781-840) ------------ */
/* C (F x F, FP64) of the compact pixels [c0, c1): every event pair j <= k of a
 * row adds val_j val_k at C[f_j][f_k], then C /= (c1 - c0).  Rows hold one
 * value per frame, frames ascending. */
void xo_twotime_pairs(int64_t c0, int64_t c1, const int64_t* off, const int32_t* frame, const double* val,
                      int frames, double* C);
/* nbins q-bins (compact ranges) at once into C[nbins][F][F], OpenMP over bins. */
void xo_twotime_bins(int nbins, const int64_t* bin_begin, const int64_t* bin_end, const int64_t* off,
                     const int32_t* frame, const double* val, int frames, double* C, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
