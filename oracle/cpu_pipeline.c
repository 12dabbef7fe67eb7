/*
 * Oracle part 4: the whole CPU path in the reference's own decomposition, used
 * as the timed CPU baseline.  TEST INFRASTRUCTURE ONLY (see xpcs_oracle.h).
 *
 *   load/transpose  serial, full-detector clear + scan per frame
 *                   (src/xpcs/filter/sparse_filter.cpp:115-193) into one
 *                   growable row per detector pixel reserved at 100 entries
 *                   (src/xpcs/data_structure/sparse_data.cpp:59-69)
 *   multi-tau       OpenMP dynamic,20 over valid pixels, in-place float
 *                   binning and binary search (src/xpcs/corr.cpp:315-431)
 *   normalise       serial float partition reduction (src/xpcs/corr.cpp:927-1091)
 *
 * Compressed IMM only, stride = avg = 1, no flat field, data from frame 0.
 */
#include "xpcs_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void xo_mt_compat_row(int32_t* idx, float* vp, int64_t n0, int64_t pix, int64_t npix_det,
                      int frames, int T, const int32_t* level, const int32_t* tau,
                      int stale_tail, float* G2s, float* IPs, float* IFs);

typedef struct { int32_t* idx; float* val; int64_t n, cap; } row_t;

static void row_push(row_t* r, int32_t f, float v)
{
    if (r->n == r->cap) {
        int64_t nc = r->cap ? r->cap * 2 : 1;
        r->idx = (int32_t*)realloc(r->idx, sizeof(int32_t) * (size_t)nc);
        r->val = (float*)realloc(r->val, sizeof(float) * (size_t)nc);
        r->cap = nc;
    }
    r->idx[r->n] = f; r->val[r->n] = v; r->n++;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int xo_cpu_pipeline(const uint8_t* bytes, int64_t nbytes, int width, int height,
                    const int32_t* dqmap, const int32_t* sqmap, int frames, int dpl,
                    int nthreads, float* g2_out, float* stderr_out, double* stage_seconds)
{
    int64_t npix_det = (int64_t)width * height;
    double t0 = now_s();

    /* configuration: q map and schedule */
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)npix_det);
    int gcap = 1 << 16;
    int32_t* gq = (int32_t*)malloc(sizeof(int32_t) * (size_t)gcap);
    int32_t* gs = (int32_t*)malloc(sizeof(int32_t) * (size_t)gcap);
    int32_t* gb = (int32_t*)malloc(sizeof(int32_t) * (size_t)gcap);
    int32_t* ge = (int32_t*)malloc(sizeof(int32_t) * (size_t)gcap);
    int32_t counts[6];
    if (xo_build_qmap(width, height, dqmap, sqmap, order, gq, gs, gb, ge, NULL, NULL,
                      gcap, 0, 0, counts) != 0) return -1;
    int max_level = xo_level_max(frames, dpl);
    int32_t lev[512], tau[512];
    int T = xo_delays_per_level(frames, dpl, max_level, lev, tau, 512);
    short* mask = (short*)calloc((size_t)npix_det, sizeof(short));
    for (int64_t i = 0; i < npix_det; i++) mask[i] = (dqmap[i] >= 1 && sqmap[i] >= 1);

    /* load + transpose */
    double t1 = now_s();
    int64_t nfr_cap = 1 << 20;
    xo_imm_frame* fr = (xo_imm_frame*)malloc(sizeof(xo_imm_frame) * (size_t)nfr_cap);
    int64_t nfr = xo_imm_index(bytes, nbytes, fr, nfr_cap);
    if (nfr < frames) return -2;
    row_t* rows = (row_t*)calloc((size_t)npix_det, sizeof(row_t));
    short* valid = (short*)calloc((size_t)npix_det, sizeof(short));
    for (int64_t i = 0; i < npix_det; i++) {
        rows[i].cap = 100;
        rows[i].idx = (int32_t*)malloc(sizeof(int32_t) * 100);
        rows[i].val = (float*)malloc(sizeof(float) * 100);
    }
    float* pv = (float*)malloc(sizeof(float) * (size_t)npix_det);
    short* sm = (short*)malloc(sizeof(short) * (size_t)npix_det);
    float* pixels_sum = (float*)calloc((size_t)npix_det, sizeof(float));
    for (int f = 0; f < frames; f++) {
        for (int64_t j = 0; j < npix_det; j++) { pv[j] = 0.0f; sm[j] = 0; }
        const uint8_t* ip = bytes + fr[f].payload_off;
        const uint8_t* vp = ip + 4 * (int64_t)fr[f].dlen;
        for (uint32_t j = 0; j < fr[f].dlen; j++) {
            int32_t pix; int16_t raw;
            memcpy(&pix, ip + 4 * (int64_t)j, 4);
            memcpy(&raw, vp + 2 * (int64_t)j, 2);
            if (pix < 0 || pix >= npix_det) return -3;
            if (mask[pix]) { pv[pix] += (float)raw; sm[pix] = 1; }
        }
        for (int64_t i = 0; i < npix_det; i++) {
            if (!sm[i]) continue;
            pixels_sum[i] += pv[i];
            valid[i] = 1;
            row_push(&rows[i], f, pv[i]);
        }
    }

    /* multi-tau */
    double t2 = now_s();
    float* G2 = (float*)calloc((size_t)npix_det * T, sizeof(float));
    float* IP = (float*)calloc((size_t)npix_det * T, sizeof(float));
    float* IF = (float*)calloc((size_t)npix_det * T, sizeof(float));
    int64_t nvalid = 0;
    int64_t* vlist = (int64_t*)malloc(sizeof(int64_t) * (size_t)npix_det);
    for (int64_t i = 0; i < npix_det; i++) if (valid[i]) vlist[nvalid++] = i;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 20)
    for (int64_t k = 0; k < nvalid; k++) {
        row_t* r = &rows[vlist[k]];
        xo_mt_compat_row(r->idx, r->val, r->n, vlist[k], npix_det, frames, T, lev, tau, 1, G2, IP, IF);
    }

    /* normalise */
    double t3 = now_s();
    xo_normalize_compat(T, npix_det, G2, IP, IF, order, counts[2], gq, gs, gb, ge,
                        counts[3], counts[4], g2_out, stderr_out);
    double t4 = now_s();
    if (stage_seconds) {
        stage_seconds[0] = t1 - t0;
        stage_seconds[1] = t2 - t1;
        stage_seconds[2] = t3 - t2;
        stage_seconds[3] = t4 - t3;
    }
    for (int64_t i = 0; i < npix_det; i++) { free(rows[i].idx); free(rows[i].val); }
    free(rows); free(valid); free(pv); free(sm); free(pixels_sum); free(G2); free(IP); free(IF);
    free(vlist); free(fr); free(mask); free(order); free(gq); free(gs); free(gb); free(ge);
    return T;
}
