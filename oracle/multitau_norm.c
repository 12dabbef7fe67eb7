/*
 * Oracle part 3: the multi-tau correlator (exact int64 closed form and the
 * float replica of the reference) and the partition normalisation (FP64 and
 * the float replica).  TEST INFRASTRUCTURE ONLY (see xpcs_oracle.h).
 */
#include "xpcs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/*
 * Closed form of Corr::multiTau2 (src/xpcs/corr.cpp:315-431) by dense brute
 * force.  Level change (:349-390): lastframe = floor(lastframe/2) and
 * x[k] = x[2k] + x[2k+1] for k < lastframe (the reference stores x/2^L).
 * Per entry (:392-424) with tau' = tau/2^L and N = lastframe - tau':
 * G2 = sum_{t<N} x[t]x[t+tau'], IP = sum_{t<N} x[t], IF = sum_{tau'<=t<lastframe} x[t].
 */
void xo_multitau_exact(int64_t P, const int64_t* off, const int32_t* frame, const int64_t* cnt,
                       int frames, int T, const int32_t* level, const int32_t* tau,
                       int64_t* G2, int64_t* IP, int64_t* IF, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        int64_t* x = (int64_t*)malloc(sizeof(int64_t) * (size_t)(frames > 0 ? frames : 1));
#pragma omp for schedule(dynamic, 20)
        for (int64_t p = 0; p < P; p++) {
            int64_t* g2 = G2 + p * T;
            int64_t* ip = IP + p * T;
            int64_t* iff = IF + p * T;
            for (int i = 0; i < T; i++) { g2[i] = 0; ip[i] = 0; iff[i] = 0; }
            if (off[p + 1] == off[p]) continue;
            memset(x, 0, sizeof(int64_t) * (size_t)frames);
            for (int64_t e = off[p]; e < off[p + 1]; e++) x[frame[e]] += cnt[e];
            int lf = frames;
            int ll = 0;
            for (int i = 0; i < T; i++) {
                int L = level[i];
                if (L != ll) {
                    lf = lf / 2;
                    for (int k = 0; k < lf; k++) x[k] = x[2 * k] + x[2 * k + 1];
                    ll = L;
                }
                int tp = tau[i] >> L;
                int N = lf - tp;
                int64_t sg = 0, sp = 0, sf = 0;
                for (int t = 0; t < N; t++) { sg += x[t] * x[t + tp]; sp += x[t]; }
                for (int t = tp; t < lf; t++) sf += x[t];
                g2[i] = sg; ip[i] = sp; iff[i] = sf;
            }
        }
        free(x);
    }
}

/* libstdc++ std::lower_bound over [first, last) with operator<. */
static int64_t lower_bound_i32(const int32_t* a, int64_t first, int64_t last, int32_t val)
{
    int64_t len = last - first;
    while (len > 0) {
        int64_t half = len >> 1;
        int64_t mid = first + half;
        if (a[mid] < val) { first = mid + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}

/*
 * Float replica of the per-pixel body of src/xpcs/corr.cpp:330-430, statement
 * by statement, working in place on one row (idx/vp hold n0 entries).
 */
void xo_mt_compat_row(int32_t* idx, float* vp, int64_t n0, int64_t pix, int64_t npix_det,
                      int frames, int T, const int32_t* level, const int32_t* tau,
                      int stale_tail, float* G2s, float* IPs, float* IFs)
{
    int ll = 0;
    int lastframe = frames;
    int64_t lastIndex = n0;
    for (int it = 0; it < T; it++) {
        int lev = level[it];
        int tau_i = tau[it];
        if (ll != lev) {
            if (lastframe % 2) lastframe -= 1;
            lastframe = lastframe / 2;
            int64_t index = 0, cnt = 0;
            int i0, i1;
            i0 = (int)(idx[cnt] / 2.0);
            if (lastIndex > 0 && i0 < lastframe) { idx[index] = i0; cnt = 1; }
            while (cnt < lastIndex) {
                i1 = (int)(idx[cnt] / 2.0);
                if (i1 >= lastframe) break;
                if (i1 == i0) vp[index] += vp[cnt];
                else { idx[++index] = i1; vp[index] = vp[cnt]; }
                i0 = i1;
                cnt++;
            }
            lastIndex = index + 1;
            for (int64_t i = 0; i < lastIndex; i++) vp[i] /= 2.0f;
        }
        if (lev > 0) tau_i = (int)(tau_i / pow(2, lev));
        int64_t g = (int64_t)it * npix_det + pix;
        for (int64_t r = 0; r < lastIndex; r++) {
            int src = idx[r];
            if (src < (lastframe - tau_i)) {
                IPs[g] += vp[r];
                int64_t end = stale_tail ? n0 : lastIndex;
                int64_t pos = lower_bound_i32(idx, 0, end, src + tau_i);
                if (pos != end) {
                    if (pos < lastIndex && idx[pos] == (src + tau_i)) {
                        float prod = vp[r] * vp[pos];
                        G2s[g] += prod;
                    }
                }
            }
            if (src >= tau_i && src < lastframe) IFs[g] += vp[r];
        }
        if ((lastframe - tau_i) > 0) {
            G2s[g] /= (float)(lastframe - tau_i);
            IPs[g] /= (float)(lastframe - tau_i);
            IFs[g] /= (float)(lastframe - tau_i);
        }
        ll = lev;
    }
}

/* Float replica of src/xpcs/corr.cpp:315-431 over a CSR copy of the rows. */
void xo_multitau_compat(int64_t P, const int64_t* off, const int32_t* frame, const float* val,
                        const int32_t* c2pix, int64_t npix_det,
                        int frames, int T, const int32_t* level, const int32_t* tau,
                        int stale_tail, float* G2s, float* IPs, float* IFs, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t maxn = 1;
    for (int64_t p = 0; p < P; p++) if (off[p + 1] - off[p] > maxn) maxn = off[p + 1] - off[p];
#pragma omp parallel
    {
        int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxn);
        float* vp = (float*)malloc(sizeof(float) * (size_t)maxn);
#pragma omp for schedule(dynamic, 20)
        for (int64_t p = 0; p < P; p++) {
            int64_t n0 = off[p + 1] - off[p];
            if (n0 == 0) continue;      /* not a valid pixel (sparse_data.cpp:84-97) */
            memcpy(idx, frame + off[p], sizeof(int32_t) * (size_t)n0);
            memcpy(vp, val + off[p], sizeof(float) * (size_t)n0);
            xo_mt_compat_row(idx, vp, n0, c2pix[p], npix_det, frames, T, level, tau,
                             stale_tail, G2s, IPs, IFs);
        }
        free(idx); free(vp);
    }
}

/*
 * FP64 closed form of Corr::normalizeG2s (src/xpcs/corr.cpp:927-1091):
 * static sums (:966-991), g2 per static bin (:997-1014), NaN-skipping mean per
 * dynamic bin (:1021-1039), Welford stderr with every pixel counted
 * (:1042-1087).
 */
void xo_normalize_f64(int T, const int32_t* Nvec, int64_t P,
                      const int64_t* G2, const int64_t* IP, const int64_t* IF,
                      int ngroups, const int32_t* group_q, const int32_t* group_begin,
                      const int32_t* group_end, int nq_out, const double* weight,
                      double* g2, double* stderr_)
{
    (void)P;
    double* g2s = (double*)malloc(sizeof(double) * (size_t)(ngroups > 0 ? ngroups : 1) * (size_t)T);
    for (int g = 0; g < ngroups; g++) {
        for (int i = 0; i < T; i++) {
            double sg = 0, sp = 0, sf = 0;
            int64_t ig = 0, ipp = 0, iff = 0;
            for (int64_t c = group_begin[g]; c < group_end[g]; c++) {
                if (weight) {
                    double w = weight[c];
                    sg += w * w * (double)G2[c * T + i];
                    sp += w * (double)IP[c * T + i];
                    sf += w * (double)IF[c * T + i];
                } else {
                    ig += G2[c * T + i]; ipp += IP[c * T + i]; iff += IF[c * T + i];
                }
            }
            if (!weight) { sg = (double)ig; sp = (double)ipp; sf = (double)iff; }
            double cs = (double)(group_end[g] - group_begin[g]);
            g2s[(size_t)g * T + i] = cs * (double)Nvec[i] * sg / (sp * sf);
        }
    }
    for (int q = 1; q <= nq_out; q++) {
        for (int i = 0; i < T; i++) {
            double acc = 0.0, cntv = 0.0;
            for (int g = 0; g < ngroups; g++) {
                if (group_q[g] != q) continue;
                double v = g2s[(size_t)g * T + i];
                if (!isnan(v)) { acc += v; cntv += 1.0; }
            }
            g2[(size_t)i * nq_out + (q - 1)] = acc / cntv;
            /* Welford over pixels of q */
            double avg = 0.0, m2 = 0.0, n = 0.0;
            for (int g = 0; g < ngroups; g++) {
                if (group_q[g] != q) continue;
                for (int64_t c = group_begin[g]; c < group_end[g]; c++) {
                    double x = (double)Nvec[i] * (double)G2[c * T + i] /
                               ((double)IP[c * T + i] * (double)IF[c * T + i]);
                    if (weight && weight[c] == 0.0) x = 0.0 / 0.0;
                    if (isnan(x)) x = 0.0;
                    n += 1.0;
                    double prev = avg;
                    avg += (x - prev) / n;
                    m2 += (x - prev) * (x - avg);
                }
            }
            stderr_[(size_t)i * nq_out + (q - 1)] = sqrt(1.0 / n) * sqrt(m2 / n);
        }
    }
    free(g2s);
}

/* Float replica of src/xpcs/corr.cpp:927-1091 (Eigen float element ops). */
void xo_normalize_compat(int T, int64_t npix_det, const float* G2, const float* IP, const float* IF,
                         const int32_t* order, int ngroups, const int32_t* group_q,
                         const int32_t* group_s, const int32_t* group_begin,
                         const int32_t* group_end, int nq_out, int static_out,
                         float* g2, float* stderr_)
{
    float* gs = (float*)calloc((size_t)static_out * T, sizeof(float));
    float* is = (float*)calloc((size_t)static_out * T, sizeof(float));
    float* fs = (float*)calloc((size_t)static_out * T, sizeof(float));
    int* pc = (int*)calloc((size_t)static_out, sizeof(int));
    for (int g = 0; g < ngroups; g++) {
        int s = group_s[g] - 1;
        for (int64_t k = group_begin[g]; k < group_end[g]; k++) {
            int64_t p = order[k];
            for (int i = 0; i < T; i++) {
                gs[(size_t)s * T + i] += G2[(int64_t)i * npix_det + p];
                is[(size_t)s * T + i] += IP[(int64_t)i * npix_det + p];
                fs[(size_t)s * T + i] += IF[(int64_t)i * npix_det + p];
            }
            pc[s] += 1;
        }
    }
    for (int g = 0; g < ngroups; g++) {
        int s = group_s[g] - 1;
        for (int i = 0; i < T; i++) {
            float c = (float)pc[s];
            gs[(size_t)s * T + i] /= c;
            is[(size_t)s * T + i] /= c;
            fs[(size_t)s * T + i] /= c;
            float den = is[(size_t)s * T + i] * fs[(size_t)s * T + i];
            gs[(size_t)s * T + i] = gs[(size_t)s * T + i] / den;
        }
    }
    for (int q = 1; q <= nq_out; q++) {
        int any = 0;
        for (int g = 0; g < ngroups; g++) if (group_q[g] == q) any = 1;
        for (int i = 0; i < T; i++) {
            float acc = 0.0f, counts = 0.0f;
            for (int g = 0; g < ngroups; g++) {
                if (group_q[g] != q) continue;
                float v = gs[(size_t)(group_s[g] - 1) * T + i];
                acc += isnan(v) ? 0.0f : v;
                counts += isnan(v) ? 0.0f : 1.0f;
            }
            g2[(size_t)i * nq_out + (q - 1)] = acc / counts;
            float avg = 0.0f, stdv = 0.0f, samples = 0.0f;
            for (int g = 0; g < ngroups; g++) {
                if (group_q[g] != q) continue;
                for (int64_t k = group_begin[g]; k < group_end[g]; k++) {
                    int64_t p = order[k];
                    float tmpAvg = avg;
                    float den = IP[(int64_t)i * npix_det + p] * IF[(int64_t)i * npix_det + p];
                    float x = G2[(int64_t)i * npix_det + p] / den;
                    x = isnan(x) ? 0.0f : x;
                    samples += isnan(x) ? 0.0f : 1.0f;
                    float tmp0 = x - tmpAvg;
                    float tmp00 = tmp0 / samples;
                    tmp00 = isnan(tmp00) ? 0.0f : tmp00;
                    avg += tmp00;
                    float tmp1 = x - tmpAvg;
                    float tmp2 = x - avg;
                    float tmp3 = tmp1 * tmp2;
                    stdv += tmp3;
                }
            }
            float stdNorm = stdv / samples;
            float inv = 1.0f / samples;
            stderr_[(size_t)i * nq_out + (q - 1)] = sqrtf(inv) * sqrtf(stdNorm);
        }
        (void)any;
    }
    free(gs); free(is); free(fs); free(pc);
}
