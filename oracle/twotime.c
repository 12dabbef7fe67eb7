/* Two-time pair loop of the reference in FP64 (test infrastructure only).
 *
 * Restates Corr::twotimeQBinThreading (src/xpcs/corr.cpp:781-840): for every
 * pixel of a q-bin, every pair of its events j <= k (row order, frames
 * ascending, one value per frame) adds val_j * val_k at C[f_j][f_k]; the
 * matrix is then divided by the q-bin's pixel count.  Used as the CPU
 * baseline of the two-time measurement and checked against the BLAS Gram form
 * of the oracle. */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "xpcs_oracle.h"

void xo_twotime_pairs(int64_t c0, int64_t c1, const int64_t* off, const int32_t* frame, const double* val,
                      int frames, double* C)
{
    const size_t F = (size_t)frames;
    memset(C, 0, sizeof(double) * F * F);
    for (int64_t c = c0; c < c1; c++) {
        const int64_t b = off[c], e = off[c + 1];
        for (int64_t j = b; j < e; j++) {
            const double v0 = val[j];
            double* row = C + (size_t)frame[j] * F;
            for (int64_t k = j; k < e; k++) row[frame[k]] += v0 * val[k];
        }
    }
    const double n = (double)(c1 > c0 ? c1 - c0 : 1);
    for (size_t i = 0; i < F * F; i++) C[i] /= n;
}

/* Several q-bins at once, one OpenMP thread per bin as the reference's
 * schedule(dynamic) loop over bins (corr.cpp:799). */
void xo_twotime_bins(int nbins, const int64_t* bin_begin, const int64_t* bin_end, const int64_t* off,
                     const int32_t* frame, const double* val, int frames, double* C, int nthreads)
{
    const size_t FF = (size_t)frames * (size_t)frames;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nbins; i++)
        xo_twotime_pairs(bin_begin[i], bin_end[i], off, frame, val, frames, C + (size_t)i * FF);
}
