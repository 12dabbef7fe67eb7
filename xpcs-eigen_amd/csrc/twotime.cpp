// Two-time analysis over an ingested plan (Corr::twotime, src/xpcs/corr.cpp:562-924;
// smoothing Corr::Smoothing :546-560 with SmoothingStaticMap :433-494 /
// SmoothingSymmetric :496-544).  Per listed q-bin (Configuration::QbinPixelList,
// src/xpcs/configuration.cpp:359-380): the q-bin's pixels are a contiguous
// range of the plan's pixel order, their values become a dense FP64 matrix,
// optional smoothing divides by the bin's mean image, then the Gram kernel and
// the diagonal sums run on the device.  Results stay on the device until read.
#include <algorithm>
#include <vector>

#include "internal.h"

using namespace xpcs;

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(XPCS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace xpcs {

void twotime_release(xpcs_mt_plan* p)
{
    for (auto* v : {&p->tt_c2t, &p->tt_full, &p->tt_part, &p->tt_sg})
        for (double* d : *v)
            if (d) (void)hipFree(d);
    for (auto& b : p->tt_pool) (void)hipFree(b.first);
    p->tt_pool.clear();
    p->tt_c2t.clear();
    p->tt_full.clear();
    p->tt_part.clear();
    p->tt_sg.clear();
    p->tt_q.clear();
    p->tt_sg_rows.clear();
    p->tt_method.clear();
    if (p->tt_x) (void)hipFree(p->tt_x);
    p->tt_x = nullptr;
    p->tt_x_cap = 0;
    if (p->tt_dpart) (void)hipFree(p->tt_dpart);
    p->tt_dpart = nullptr;
    p->tt_dpart_cap = 0;
}

}  // namespace xpcs

namespace {

template <typename T>
xpcs_status talloc(T** d, size_t n)
{
    *d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(d), std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return fail(XPCS_ERR_NOMEM, "twotime: hipMalloc of %zu bytes failed", n * sizeof(T));
    return XPCS_OK;
}

// The previous call's results become the pool of this one.
void recycle(xpcs_mt_plan* p)
{
    auto keep = [&](std::vector<double*>& v, auto size_of) {
        for (size_t i = 0; i < v.size(); i++)
            if (v[i]) p->tt_pool.emplace_back(v[i], size_of(i));
        v.clear();
    };
    const size_t F = (size_t)p->h.F, W = (size_t)std::max(p->tt_wsize, 0), tot = (size_t)std::max(p->tt_total, 0);
    keep(p->tt_c2t, [&](size_t) { return F * F; });
    keep(p->tt_full, [&](size_t) { return F; });
    keep(p->tt_part, [&](size_t) { return std::max<size_t>(W * tot, 1); });
    keep(p->tt_sg, [&](size_t i) {
        return std::max<size_t>((size_t)p->tt_sg_rows[i] * (p->tt_sg_cols == 1 ? 1 : F), 1);
    });
    p->tt_q.clear();
    p->tt_sg_rows.clear();
    p->tt_method.clear();
}

// A pooled buffer of at least n doubles (the smallest that fits, if it wastes
// less than half), else a new one.
xpcs_status take(xpcs_mt_plan* p, double** d, size_t n)
{
    n = std::max<size_t>(n, 1);
    size_t best = p->tt_pool.size();
    for (size_t i = 0; i < p->tt_pool.size(); i++) {
        const size_t c = p->tt_pool[i].second;
        if (c >= n && c <= 2 * n && (best == p->tt_pool.size() || c < p->tt_pool[best].second)) best = i;
    }
    if (best < p->tt_pool.size()) {
        *d = p->tt_pool[best].first;
        p->tt_pool.erase(p->tt_pool.begin() + (ptrdiff_t)best);
        return XPCS_OK;
    }
    return talloc(d, n);
}

// A plan-owned scratch buffer of at least n doubles.
xpcs_status grow(double** d, size_t* cap, size_t n)
{
    if (*cap >= n && *d) return XPCS_OK;
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    *cap = 0;
    xpcs_status st = talloc(d, n);
    if (!st) *cap = n;
    return st;
}

// Device buffers that live for one xpcs_mt_twotime call: released once the
// stream has drained, on every return path.
struct Scratch {
    hipStream_t s;
    std::vector<void*> bufs;
    ~Scratch()
    {
        (void)hipStreamSynchronize(s);
        for (void* b : bufs) (void)hipFree(b);
    }
    template <typename T>
    xpcs_status alloc(T** d, size_t n)
    {
        xpcs_status st = talloc(d, n);
        if (!st) bufs.push_back(*d);
        return st;
    }
    template <typename T>
    xpcs_status upload(T** d, const std::vector<T>& v)
    {
        xpcs_status st = alloc(d, v.size());
        if (!st && !v.empty() && hipMemcpy(*d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
            st = fail(XPCS_ERR_HIP, "twotime: table upload failed");
        return st;
    }
};

constexpr size_t kTtDefaultChunkBytes = (size_t)4 << 30;
// Event pairs are taken when pairs * kPairVsGram < P_q F^2: one scattered
// FP64 atomic add costs about as much as this many matrix-core products.
constexpr double kPairVsGram = 256.0;

}  // namespace

extern "C" xpcs_status xpcs_mt_twotime(xpcs_mt_plan* p, const xpcs_tt_desc* d, void* stream)
{
    if (!p || !d || (d->nq > 0 && !d->qbins)) return fail(XPCS_ERR_INVALID_ARG, "twotime: NULL argument");
    if (p->mode == 0) return fail(XPCS_ERR_STATE, "twotime: nothing ingested");
    if (d->window < 1) return fail(XPCS_ERR_INVALID_ARG, "twotime: window size must be >= 1");
    if (d->smoothing < 0 || d->smoothing > 2) return fail(XPCS_ERR_INVALID_ARG, "twotime: smoothing must be 0, 1 or 2");
    if (d->chunk_bytes < 0) return fail(XPCS_ERR_INVALID_ARG, "twotime: chunk_bytes must be >= 0");
    if (d->method < 0 || d->method > 2) return fail(XPCS_ERR_INVALID_ARG, "twotime: method must be 0, 1 or 2");
    if (d->method == 2 && p->mode == 2)
        return fail(XPCS_ERR_UNSUPPORTED, "twotime: dense frames use the Gram product (method 0 or 1)");
    HostTables& h = p->h;
    const int32_t F = h.F;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->stream;
    // the previous results may still be read by work queued on another stream
    HIP_TRY(hipStreamSynchronize(p->stream));
    recycle(p);
    // pixel-major values of the ingested data
    if (p->mode == 1) {
        xpcs_status st = prepare_execute(p);
        if (st) return st;
        launch_fill_u32(p->d_counters, 8, s);
        launch_transpose(p, s);
    } else if (p->mode == 2) {
        const size_t need = (size_t)h.P * (size_t)F;
        if (need > p->rows_cap) {
            if (p->d_rows) (void)hipFree(p->d_rows);
            p->d_rows = nullptr;
            p->rows_cap = 0;
            xpcs_status st = talloc(&p->d_rows, need);
            if (st) return st;
            p->rows_cap = need;
        }
        launch_fill_u32(p->d_counters, 8, s);
        launch_dense_transpose(p, s);
    } else {
        launch_fill_u32(p->d_counters, 8, s);
    }
    if (p->fsnorm) {
        // normalize_by_framesum scales the values before two-time as well
        if (p->mode == 3)
            return fail(XPCS_ERR_UNSUPPORTED, "twotime: normalize_by_framesum needs frame sums (IMM or packed ingest)");
        launch_fill_u64(reinterpret_cast<uint64_t*>(p->d_fs), (size_t)h.F * h.ns, s);
        launch_side(p, s);
        launch_frame_weights(p, s);
    }
    // the reference's values: counts (summed over averaged frames) / average,
    // dense values per dark-frame scale; the flat field per pixel
    const double scale = 1.0 / ((double)h.avg * (double)p->nscale);
    // listed q-bins present in the map, ascending (std::map order)
    std::vector<int32_t> qs;
    for (int32_t i = 0; i < d->nq; i++) {
        const int32_t q = d->qbins[i];
        if (q >= 1 && q <= (int32_t)h.q_first_group.size() - 1 &&
            h.q_first_group[(size_t)q - 1] < h.q_first_group[(size_t)q])
            qs.push_back(q);
    }
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    const int32_t total = (F - d->window) / d->window;
    p->tt_wsize = d->window;
    p->tt_total = total;
    p->tt_sg_cols = d->average ? 1 : F;
    // method per q-bin: event pairs cost Sum_p n_p (n_p + 1) / 2 scattered
    // atomic adds, the Gram product P_q F^2 matrix-core products
    std::vector<int32_t> method(qs.size(), p->mode == 2 ? 1 : d->method);
    Scratch tmp{s, {}};
    if (p->mode != 2 && d->method == 0 && !qs.empty()) {
        std::vector<int32_t> cb(qs.size()), ce(qs.size());
        for (size_t k = 0; k < qs.size(); k++) {
            cb[k] = h.group_begin[(size_t)h.q_first_group[(size_t)qs[k] - 1]];
            ce[k] = h.group_end[(size_t)h.q_first_group[(size_t)qs[k]] - 1];
        }
        int32_t *dcb = nullptr, *dce = nullptr;
        double* dpairs = nullptr;
        xpcs_status st = tmp.upload(&dcb, cb);
        if (!st) st = tmp.upload(&dce, ce);
        if (!st) st = tmp.alloc(&dpairs, qs.size());
        if (st) return st;
        launch_tt_pair_count(p, dcb, dce, (int32_t)qs.size(), dpairs, s);
        std::vector<double> pairs(qs.size());
        HIP_TRY(hipMemcpyAsync(pairs.data(), dpairs, sizeof(double) * qs.size(), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (size_t k = 0; k < qs.size(); k++)
            method[k] = pairs[k] * kPairVsGram < (double)(ce[k] - cb[k]) * (double)F * (double)F ? 2 : 1;
    }
    int32_t maxpq = 0;
    for (size_t k = 0; k < qs.size(); k++) {
        const int g0 = h.q_first_group[(size_t)qs[k] - 1], g1 = h.q_first_group[(size_t)qs[k]];
        if (method[k] == 1) maxpq = std::max(maxpq, h.group_end[(size_t)g1 - 1] - h.group_begin[(size_t)g0]);
    }
    const size_t row_bytes = (size_t)F * sizeof(double);
    const size_t cb = d->chunk_bytes > 0 ? (size_t)d->chunk_bytes : kTtDefaultChunkBytes;
    const int32_t chunk = (int32_t)std::max<size_t>(1, std::min<size_t>((size_t)maxpq, cb / row_bytes));
    if (maxpq > 0) {
        xpcs_status st = grow(&p->tt_x, &p->tt_x_cap, (size_t)chunk * F);
        if (st) return st;
    }
    double* X = p->tt_x;
    if (!qs.empty()) {
        xpcs_status st = grow(&p->tt_dpart, &p->tt_dpart_cap, (size_t)tt_diag_segments(F) * F);
        if (st) return st;
    }
    double* dpart = p->tt_dpart;   // diagonal segment sums
    // an error leaves no partial results behind
    auto bail = [&](xpcs_status st) {
        (void)hipStreamSynchronize(s);
        twotime_release(p);
        return st;
    };
    for (size_t k = 0; k < qs.size(); k++) {
        const int32_t q = qs[k];
        const bool pairs = method[k] == 2;
        const int g0 = h.q_first_group[(size_t)q - 1], g1 = h.q_first_group[(size_t)q];
        const int32_t c0 = h.group_begin[(size_t)g0];
        const int32_t Pq = h.group_end[(size_t)g1 - 1] - c0;
        double *C = nullptr, *full = nullptr, *part = nullptr;
        xpcs_status st = take(p, &C, (size_t)F * F);
        if (!st) st = take(p, &full, (size_t)F);
        if (!st) st = take(p, &part, (size_t)d->window * (size_t)std::max(total, 0));
        p->tt_c2t.push_back(C);
        p->tt_full.push_back(full);
        p->tt_part.push_back(part);
        p->tt_q.push_back(q);
        p->tt_method.push_back(method[k]);
        if (st) return bail(st);
        launch_fill_u64(reinterpret_cast<uint64_t*>(C), (size_t)F * F, s);
        const bool one_chunk = Pq <= chunk;
        const double* div = nullptr;
        int32_t* dbor = nullptr;
        if (d->smoothing != 0) {
            // static map: one bin per (q, s) group; symmetric: the whole q-bin
            std::vector<int32_t> bb, be, bor((size_t)Pq);
            if (d->smoothing == 1) {
                for (int g = g0; g < g1; g++) {
                    bb.push_back(h.group_begin[(size_t)g] - c0);
                    be.push_back(h.group_end[(size_t)g] - c0);
                }
            } else {
                bb.push_back(0);
                be.push_back(Pq);
            }
            for (size_t b = 0; b < bb.size(); b++)
                for (int32_t i = bb[b]; i < be[b]; i++) bor[(size_t)i] = (int32_t)b;
            const int32_t nb = (int32_t)bb.size();
            int32_t *dbb = nullptr, *dbe = nullptr;
            double *sg = nullptr, *sga = nullptr;
            st = tmp.upload(&dbb, bb);
            if (!st) st = tmp.upload(&dbe, be);
            if (!st) st = tmp.upload(&dbor, bor);
            if (!st) st = take(p, &sg, (size_t)nb * F);
            if (!st && d->average) st = take(p, &sga, (size_t)nb);
            // the result keeps the mean image, or with average its frame mean
            if (d->average) {
                if (sg) tmp.bufs.push_back(sg);
                if (sga) p->tt_sg.push_back(sga);
            } else if (sg) {
                p->tt_sg.push_back(sg);
            }
            if (st) return bail(st);
            p->tt_sg_rows.push_back(nb);
            launch_fill_u64(reinterpret_cast<uint64_t*>(sg), (size_t)nb * F, s);
            if (pairs) {
                launch_tt_sg_sparse(p, c0, Pq, scale, dbor, sg, s);
            } else {
                for (int32_t lo = 0; lo < Pq; lo += chunk) {
                    const int32_t n = std::min(chunk, Pq - lo);
                    launch_tt_values(p, c0 + lo, n, scale, X, s);
                    launch_tt_colsum(p, X, F, lo, lo + n, dbb, dbe, nb, sg, s);
                }
            }
            launch_tt_sgfinal(sg, F, dbb, dbe, nb, d->average != 0, sga, s);
            div = d->average ? sga : sg;
        }
        if (pairs) {
            launch_tt_pairs(p, c0, Pq, scale, dbor, div, d->average == 0, C, s);
        } else {
            for (int32_t lo = 0; lo < Pq; lo += chunk) {
                const int32_t n = std::min(chunk, Pq - lo);
                // with smoothing and one chunk the values are already in place
                if (!(div && one_chunk)) launch_tt_values(p, c0 + lo, n, scale, X, s);
                if (div) launch_tt_divide(p, X, F, lo, n, dbor, div, d->average == 0, s);
                launch_tt_gram(p, X, n, F, 1.0 / (double)Pq, C, s);
            }
        }
        // the event-pair sums take their 1 / P_q in the diagonal pass
        launch_tt_diag(p, C, F, d->window, total, pairs ? 1.0 / (double)Pq : 1.0, dpart, full, part, s);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return bail(fail(XPCS_ERR_HIP, "twotime: launch failed: %s", hipGetErrorString(e)));
    }
    if (hipStreamSynchronize(s) != hipSuccess) return bail(fail(XPCS_ERR_HIP, "twotime: device error"));
    // pooled buffers this call did not reuse
    for (auto& b : p->tt_pool) (void)hipFree(b.first);
    p->tt_pool.clear();
    collect_timers(p);
    uint32_t cnt[2] = {0, 0};
    HIP_TRY(hipMemcpy(cnt, p->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[1] & 1u)
        return bail(fail(XPCS_ERR_RANGE, "twotime: an IMM pixel index lies outside the detector"));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_twotime_info(xpcs_mt_plan* p, int32_t* nq, int32_t* qbins, int32_t* methods,
                                            int32_t cap, int32_t* sg_rows, int32_t* sg_cols, int32_t* total_partials)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "twotime_info: NULL plan");
    if (nq) *nq = (int32_t)p->tt_q.size();
    for (int32_t i = 0; i < cap && i < (int32_t)p->tt_q.size(); i++) {
        if (qbins) qbins[i] = p->tt_q[(size_t)i];
        if (methods) methods[i] = p->tt_method[(size_t)i];
    }
    int32_t rows = 0;
    for (int32_t r : p->tt_sg_rows) rows += r;
    if (sg_rows) *sg_rows = rows;
    if (sg_cols) *sg_cols = rows ? p->tt_sg_cols : 0;
    if (total_partials) *total_partials = std::max(p->tt_total, 0);
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_twotime(xpcs_mt_plan* p, int32_t i, double* c2t)
{
    if (!p || !c2t) return fail(XPCS_ERR_INVALID_ARG, "read_twotime: NULL argument");
    if (i < 0 || i >= (int32_t)p->tt_c2t.size()) return fail(XPCS_ERR_RANGE, "read_twotime: no q-bin %d", i);
    const size_t F = (size_t)p->h.F;
    HIP_TRY(hipMemcpy(c2t, p->tt_c2t[(size_t)i], sizeof(double) * F * F, hipMemcpyDeviceToHost));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_twotime_1d(xpcs_mt_plan* p, double* g2full, double* g2partials, double* sg)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "read_twotime_1d: NULL plan");
    const size_t F = (size_t)p->h.F, nq = p->tt_q.size(), W = (size_t)p->tt_wsize;
    const size_t tot = (size_t)std::max(p->tt_total, 0);
    std::vector<double> tmp;
    for (size_t i = 0; i < nq; i++) {
        if (g2full) {
            tmp.resize(F);
            HIP_TRY(hipMemcpy(tmp.data(), p->tt_full[i], sizeof(double) * F, hipMemcpyDeviceToHost));
            for (size_t f = 0; f < F; f++) g2full[f * nq + i] = tmp[f];
        }
        if (g2partials && W * tot) {
            tmp.resize(W * tot);
            HIP_TRY(hipMemcpy(tmp.data(), p->tt_part[i], sizeof(double) * W * tot, hipMemcpyDeviceToHost));
            for (size_t a = 0; a < W; a++)
                for (size_t b = 0; b < tot; b++) g2partials[(a * tot + b) * nq + i] = tmp[a * tot + b];
        }
    }
    if (sg) {
        size_t row = 0;
        const size_t cols = (size_t)p->tt_sg_cols;
        for (size_t i = 0; i < p->tt_sg.size(); i++) {
            const size_t n = (size_t)p->tt_sg_rows[i] * cols;
            HIP_TRY(hipMemcpy(sg + row * cols, p->tt_sg[i], sizeof(double) * n, hipMemcpyDeviceToHost));
            row += (size_t)p->tt_sg_rows[i];
        }
    }
    return XPCS_OK;
}
