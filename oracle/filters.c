/*
 * Oracle part 2: frame -> pixel transposes (sparse and dense filters) and the
 * dark image.  TEST INFRASTRUCTURE ONLY (see xpcs_oracle.h).
 */
#include "xpcs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint16_t rd_u16(const uint8_t* p) { uint16_t x; memcpy(&x, p, 2); return x; }
static int32_t rd_i32(const uint8_t* p) { int32_t x; memcpy(&x, p, 4); return x; }

/*
 * src/xpcs/filter/sparse_filter.cpp:115-193.  Per output frame, masked events
 * of its raw frames are accumulated into a per-pixel scratch
 * (value*flatfield, :148-157); every touched pixel then gets one entry
 * (frame, sum/avg) (:168-179).  Values are read as int16 by the reference
 * (src/xpcs/io/imm.cpp:87-99); the float value replicates that, the integer
 * count uses uint16.
 */
int64_t xo_sparse_transpose(const uint8_t* bytes, const xo_imm_frame* frames,
                            const int64_t* sel, const int32_t* out_frame, int64_t nsel,
                            int npix_det, const int32_t* pix2c, int32_t P,
                            const double* flatfield, int avg,
                            int64_t* off, int32_t* frame, int64_t* cnt, float* val, int64_t cap)
{
    int64_t* acc_c = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    float* acc_v = (float*)calloc((size_t)P + 1, sizeof(float));
    int32_t* stamp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)P + 1));
    int32_t* touched = (int32_t*)malloc(sizeof(int32_t) * ((size_t)P + 1));
    int64_t* fill = NULL;
    for (int32_t c = 0; c < P; c++) stamp[c] = -1;
    int pass_fill = frame != NULL;
    int64_t rc = 0;
    for (int pass = 0; pass < (pass_fill ? 2 : 1); pass++) {
        if (pass == 0) { for (int32_t c = 0; c <= P; c++) off[c] = 0; }
        else {
            fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)P + 1));
            for (int32_t c = 0; c < P; c++) { fill[c] = off[c]; stamp[c] = -1; }
        }
        int64_t k = 0;
        while (k < nsel) {
            int32_t f = out_frame[k];
            int ntouch = 0;
            for (; k < nsel && out_frame[k] == f; k++) {
                const xo_imm_frame* fr = &frames[sel[k]];
                const uint8_t* idxp = bytes + fr->payload_off;
                const uint8_t* valp = idxp + 4 * (int64_t)fr->dlen;
                for (uint32_t j = 0; j < fr->dlen; j++) {
                    int32_t pix = rd_i32(idxp + 4 * (int64_t)j);
                    if (pix < 0 || pix >= npix_det) { rc = -1; goto done; }
                    int32_t c = pix2c[pix];
                    if (c < 0) continue;
                    uint16_t raw = rd_u16(valp + 2 * (int64_t)j);
                    if (stamp[c] != f) { stamp[c] = f; touched[ntouch++] = c; acc_c[c] = 0; acc_v[c] = 0.0f; }
                    acc_c[c] += raw;
                    float rv = (float)(int16_t)raw;
                    double ff = flatfield ? flatfield[pix] : 1.0;
                    float v = (float)((double)rv * ff);
                    acc_v[c] += v;
                }
            }
            for (int t = 0; t < ntouch; t++) {
                int32_t c = touched[t];
                if (pass == 0) off[c + 1]++;
                else {
                    int64_t p = fill[c]++;
                    if (p >= cap) { rc = -2; goto done; }
                    frame[p] = f;
                    cnt[p] = acc_c[c];
                    val[p] = acc_v[c] / (float)avg;
                }
            }
        }
        if (pass == 0) for (int32_t c = 0; c < P; c++) off[c + 1] += off[c];
    }
    rc = off[P];
done:
    free(acc_c); free(acc_v); free(stamp); free(touched); free(fill);
    return rc;
}

/* src/xpcs/data_structure/dark_image.cpp:81-106 (Welford in double).  Dark
 * values are uint16, as every count here (the reference reads them as int16,
 * src/xpcs/io/imm.cpp:87; the two agree below 32768), so the mean and standard
 * deviation that set the threshold describe the same values as the integer
 * dark sums that are subtracted. */
void xo_dark_image(const uint8_t* bytes, const xo_imm_frame* frames, int darks,
                   int npix_det, const double* flatfield, double* dark_avg, double* dark_std,
                   int64_t* dark_sum)
{
    for (int j = 0; j < npix_det; j++) { dark_avg[j] = 0.0; dark_std[j] = 0.0; dark_sum[j] = 0; }
    for (int i = 0; i < darks; i++) {
        const uint8_t* vp = bytes + frames[i].payload_off;
        for (int j = 0; j < npix_det; j++) {
            uint16_t raw = rd_u16(vp + 2 * (int64_t)j);
            double tmp = dark_avg[j];
            double pix = (double)raw * (flatfield ? flatfield[j] : 1.0);
            dark_avg[j] += ((pix - dark_avg[j]) / (double)(i + 1));
            dark_std[j] += ((pix - tmp) * (pix - dark_avg[j]));
            dark_sum[j] += raw;
        }
    }
    for (int j = 0; j < npix_det; j++) dark_std[j] = sqrt(dark_std[j] / darks);
}

/* src/xpcs/filter/dense_filter.cpp:121-210. */
int64_t xo_dense_transpose(const uint8_t* bytes, const xo_imm_frame* frames,
                           const int64_t* sel, const int32_t* out_frame, int64_t nsel,
                           int npix_det, const int32_t* pix2c, int32_t P,
                           const double* flatfield, int avg, int darks,
                           const double* dark_avg, const double* dark_std, const int64_t* dark_sum,
                           float lld, float sigma,
                           int64_t* off, int32_t* frame, int64_t* cnt, float* val, int64_t cap)
{
    int64_t n = darks > 0 ? darks : 1;
    int64_t* acc_c = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    float* acc_v = (float*)calloc((size_t)P + 1, sizeof(float));
    char* hit = (char*)calloc((size_t)P + 1, 1);
    int64_t* thr = (int64_t*)malloc(sizeof(int64_t) * (size_t)npix_det);
    float* thrf = (float*)malloc(sizeof(float) * (size_t)npix_det);
    for (int j = 0; j < npix_det; j++) {
        float t = 0.0f;
        if (darks > 0) t = (float)((double)lld + (double)sigma * dark_std[j]);
        thrf[j] = t;
        thr[j] = (int64_t)floor((double)n * (double)t);
    }
    int64_t* fill = NULL;
    int64_t rc = 0;
    for (int pass = 0; pass < (frame ? 2 : 1); pass++) {
        if (pass == 0) { for (int32_t c = 0; c <= P; c++) off[c] = 0; }
        else {
            fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)P + 1));
            for (int32_t c = 0; c < P; c++) fill[c] = off[c];
        }
        int64_t k = 0;
        while (k < nsel) {
            int32_t f = out_frame[k];
            for (int32_t c = 0; c < P; c++) { acc_c[c] = 0; acc_v[c] = 0.0f; hit[c] = 0; }
            for (; k < nsel && out_frame[k] == f; k++) {
                const xo_imm_frame* fr = &frames[sel[k]];
                if ((int64_t)fr->dlen != npix_det) { rc = -1; goto done; }
                const uint8_t* vp = bytes + fr->payload_off;
                for (int j = 0; j < npix_det; j++) {
                    int32_t c = pix2c[j];
                    if (c < 0) continue;
                    uint16_t raw = rd_u16(vp + 2 * (int64_t)j);
                    /* exact, n-scaled integer value */
                    int64_t y = n * (int64_t)raw - (darks > 0 ? dark_sum[j] : 0);
                    if (y > thr[j]) { hit[c] = 1; acc_c[c] += y > 0 ? y : 0; }
                    /* reference float value (its int16 read; the dark mean is
                     * the uint16 one, so this matches the reference below 32768) */
                    float v = (float)(int16_t)raw;
                    float thresh = 0.0f;
                    if (darks > 0) {
                        v = (float)((double)v - dark_avg[j]);
                        v = v > 0.0f ? v : 0.0f;
                        thresh = thrf[j];
                    }
                    if (v <= thresh) continue;
                    v = (float)((double)v * (flatfield ? flatfield[j] : 1.0));
                    acc_v[c] += v;
                }
            }
            for (int32_t c = 0; c < P; c++) {
                if (!hit[c]) continue;
                if (pass == 0) off[c + 1]++;
                else {
                    int64_t p = fill[c]++;
                    if (p >= cap) { rc = -2; goto done; }
                    frame[p] = f; cnt[p] = acc_c[c]; val[p] = acc_v[c] / (float)avg;
                }
            }
        }
        if (pass == 0) for (int32_t c = 0; c < P; c++) off[c + 1] += off[c];
    }
    rc = off[P];
done:
    free(acc_c); free(acc_v); free(hit); free(thr); free(thrf); free(fill);
    return rc;
}
