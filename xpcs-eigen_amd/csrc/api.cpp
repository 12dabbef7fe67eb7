// C ABI of the multi-tau library: plans, ingest, execute, results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "internal.h"

using xpcs::fail;

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(e_ == hipErrorOutOfMemory ? XPCS_ERR_NOMEM : XPCS_ERR_HIP,         \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                    \
    } while (0)

namespace xpcs {
int sparse_cap_long();
int reg_cap();

template <typename T>
static xpcs_status dalloc(T** p, size_t n)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(XPCS_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                    hipGetErrorString(e));
    }
    return XPCS_OK;
}

template <typename T>
static xpcs_status upload(T** p, const std::vector<T>& v, hipStream_t s)
{
    xpcs_status st = dalloc(p, v.size());
    if (st) return st;
    if (!v.empty()) {
        hipError_t e = hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail(XPCS_ERR_HIP, "upload: %s", hipGetErrorString(e));
    }
    return XPCS_OK;
}

template <typename T>
static void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void timer_begin(xpcs_mt_plan* p, hipStream_t s, const char* name)
{
    if (!(p->h.flags & XPCS_MT_TIMING) || p->capturing) return;
    // timers follow one another (none is nested), so one flag serves timer_end
    p->timer_skipped = !p->timer_only.empty() && p->timer_only != name;
    if (p->timer_skipped) return;
    KernelTimer t;
    t.name = name;
    if (p->event_pool.size() >= 2) {
        t.a = p->event_pool.back(); p->event_pool.pop_back();
        t.b = p->event_pool.back(); p->event_pool.pop_back();
    } else {
        (void)hipEventCreate(&t.a);
        (void)hipEventCreate(&t.b);
    }
    (void)hipEventRecord(t.a, s);
    p->pending.push_back(t);
}

void timer_end(xpcs_mt_plan* p, hipStream_t s)
{
    if (!(p->h.flags & XPCS_MT_TIMING) || p->capturing || p->timer_skipped) return;
    (void)hipEventRecord(p->pending.back().b, s);
}

void collect_timers(xpcs_mt_plan* p)
{
    for (auto& t : p->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            auto& e = p->times[t.name];
            e.first += ms;
            e.second += 1;
        }
        p->event_pool.push_back(t.a);
        p->event_pool.push_back(t.b);
    }
    p->pending.clear();
}

}  // namespace xpcs

using namespace xpcs;

extern "C" const char* xpcs_version(void) { return "xpcs-mt 0.1 (gfx950)"; }

extern "C" xpcs_status xpcs_mt_plan_create(const xpcs_mt_desc* desc, xpcs_mt_plan** out)
{
    if (!desc || !out) return fail(XPCS_ERR_INVALID_ARG, "plan_create: NULL argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(XPCS_ERR_NO_DEVICE, "plan_create: no HIP device visible");
    xpcs_mt_plan* p = new (std::nothrow) xpcs_mt_plan();
    if (!p) return fail(XPCS_ERR_NOMEM, "plan_create: host allocation failed");
    xpcs_status st = build_tables(*desc, p->h);
    if (st) {
        delete p;
        return st;
    }
    HostTables& h = p->h;
    (void)hipGetDevice(&p->device);
    st = noalloc_setup(p->device);
    if (st) {
        delete p;
        return st;
    }
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return fail(XPCS_ERR_HIP, "plan_create: stream creation failed");
    }
    p->own_stream = true;
    hipStream_t s = p->stream;
    std::vector<int32_t> gsize((size_t)h.ngroups);
    for (int g = 0; g < h.ngroups; g++) gsize[(size_t)g] = h.group_end[(size_t)g] - h.group_begin[(size_t)g];
    std::vector<int32_t> qchunk(h.q_chunk_first.begin(), h.q_chunk_first.end());
#define TRY(x)                        \
    do {                              \
        st = (x);                     \
        if (st) {                     \
            xpcs_mt_plan_destroy(p);  \
            return st;                \
        }                             \
    } while (0)
    TRY(upload(&p->d_pix2c, h.pix2c, s));
    TRY(upload(&p->d_taus, h.taus, s));
    TRY(upload(&p->d_levels, h.levels, s));
    {
        // distinct frame boundaries of the IP/IF prefix sums, (lf - tau') 2^L,
        // tau' 2^L and lf 2^L, each tau's indices among them, and per frame
        // the number of boundaries at or below it
        std::vector<int32_t> b;
        for (const TauParam& tp : h.taus) {
            b.push_back((tp.lf - tp.tp) << tp.level);
            b.push_back(tp.tp << tp.level);
            b.push_back(tp.lf << tp.level);
        }
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        auto index = [&](int32_t x) { return (int32_t)(std::lower_bound(b.begin(), b.end(), x) - b.begin()); };
        std::vector<int4> tj;
        for (const TauParam& tp : h.taus)
            tj.push_back(make_int4(index((tp.lf - tp.tp) << tp.level), index(tp.tp << tp.level),
                                   index(tp.lf << tp.level), 0));
        std::vector<uint16_t> rank((size_t)std::max(h.F, 1));
        size_t j = 0;
        for (int32_t f = 0; f < h.F; f++) {
            while (j < b.size() && b[j] <= f) j++;
            rank[(size_t)f] = (uint16_t)j;
        }
        p->nrank = (int32_t)b.size();
        // per 64 frames: the bits of the boundary frames in the word and the
        // rank before it; rank(f) = z + popcount(bits at or below f & 63)
        const int32_t nw = (h.F + 63) / 64;
        std::vector<uint4> words((size_t)std::max(nw, 1));
        for (int32_t w = 0; w < nw; w++) {
            uint64_t bits = 0;
            for (int32_t f = 64 * w; f < std::min(h.F, 64 * w + 64); f++)
                if (rank[(size_t)f] != (f > 0 ? rank[(size_t)f - 1] : 0)) bits |= 1ull << (f - 64 * w);
            words[(size_t)w] = make_uint4((uint32_t)bits, (uint32_t)(bits >> 32), w > 0 ? rank[(size_t)(64 * w - 1)] : 0u, 0u);
        }
        // ranks rise by one at each boundary (distinct, sorted), so the words
        // give back the table; checked here, and left unused otherwise
        bool words_ok = true;
        for (int32_t f = 0; f < h.F && words_ok; f++) {
            const uint4 q = words[(size_t)(f / 64)];
            const uint64_t bits = ((uint64_t)q.y << 32) | q.x;
            const uint32_t r = q.z + (uint32_t)__builtin_popcountll(bits & (~0ull >> (63 - (f & 63))));
            words_ok = r == rank[(size_t)f];
        }
        TRY(upload(&p->d_tau_bnd, tj, s));
        TRY(upload(&p->d_frame_rank, rank, s));
        if (words_ok && nw > 0) {
            TRY(upload(&p->d_rank_words, words, s));
            p->nrank_words = nw;
        }
    }
    TRY(upload(&p->d_chunks, h.chunks, s));
    TRY(upload(&p->d_group_chunk_first, h.group_chunk_first, s));
    TRY(upload(&p->d_group_q, h.group_q, s));
    TRY(upload(&p->d_group_size, gsize, s));
    TRY(upload(&p->d_q_chunk_first, qchunk, s));
    TRY(upload(&p->d_q_first_group, h.q_first_group, s));
    TRY(upload(&p->d_c2pix, h.order, s));
    TRY(upload(&p->d_sbin_of_c, h.sbin_of_c, s));
    if (!h.weight.empty()) TRY(upload(&p->d_weight, h.weight, s));
    const size_t PT = (size_t)h.P * (size_t)h.T;
    TRY(dalloc(&p->d_G2, PT));
    TRY(dalloc(&p->d_IP, PT));
    TRY(dalloc(&p->d_IF, PT));
    TRY(dalloc(&p->d_pix_total, (size_t)h.P));
    TRY(dalloc(&p->d_part, std::max<size_t>(1, h.chunks.size() * (size_t)h.T)));
    TRY(dalloc(&p->d_g2s, std::max<size_t>(1, (size_t)h.ngroups * h.T)));
    TRY(dalloc(&p->d_g2, std::max<size_t>(1, (size_t)h.nq * h.T)));
    TRY(dalloc(&p->d_se, std::max<size_t>(1, (size_t)h.nq * h.T)));
    TRY(dalloc(&p->d_off, (size_t)std::max<int64_t>(h.P, (int64_t)h.W * h.H) + 1));
    {
        const int64_t npd = (int64_t)h.W * h.H;
        std::vector<uint32_t> bits((size_t)(npd / 32 + 1), 0u);
        for (int64_t i = 0; i < npd; i++)
            if (h.pix2c[(size_t)i] >= 0) bits[(size_t)(i >> 5)] |= 1u << (i & 31);
        TRY(upload(&p->d_mapped, bits, s));
    }
    TRY(dalloc(&p->d_long_list, (size_t)h.P));
    TRY(dalloc(&p->d_mid_list, (size_t)h.P));
    TRY(dalloc(&p->d_counters, 8));
    {
        // level-0 prefix boundaries of the dense streaming path
        std::vector<int4> bnd;
        for (int i = 0; i < h.T; i++) {
            const TauParam& tp = h.taus[(size_t)i];
            bnd.push_back(make_int4((tp.lf - tp.tp) << tp.level, 0, i, 0));
            bnd.push_back(make_int4(tp.tp << tp.level, 1, i, 0));
            bnd.push_back(make_int4(tp.lf << tp.level, 2, i, 0));
        }
        std::stable_sort(bnd.begin(), bnd.end(), [](const int4& a, const int4& b) {
            return a.x != b.x ? a.x < b.x : a.y < b.y;
        });
        p->nbnd = (int32_t)bnd.size();
        TRY(upload(&p->d_bnd, bnd, s));
    }
    p->fsnorm = (h.flags & XPCS_MT_NORMALIZE_BY_FRAMESUM) != 0;
    if (h.flags & (XPCS_MT_SIDE_STATS | XPCS_MT_NORMALIZE_BY_FRAMESUM)) {
        if (h.ns > 4096 || (int64_t)h.F * h.ns > (int64_t(1) << 26)) {
            xpcs_mt_plan_destroy(p);
            return fail(XPCS_ERR_UNSUPPORTED, "plan: side statistics need S <= 4096 and F*S <= 2^26");
        }
        TRY(dalloc(&p->d_fs, (size_t)h.F * h.ns));
        // static bin (1..S, 0 unmapped) per raster pixel, one byte when S < 256:
        // one small gather per sparse event instead of the pixel's row id and
        // then its bin
        if (h.ns < 256) {
            std::vector<uint8_t> sbp((size_t)h.W * h.H, 0);
            for (int c = 0; c < h.P; c++) sbp[(size_t)h.order[(size_t)c]] = (uint8_t)h.sbin_of_c[(size_t)c];
            uint8_t* d = nullptr;
            TRY(upload(&d, sbp, s));
            p->d_sbin_pix = d;
            p->sbin_pix_bytes = 1;
        } else {
            std::vector<uint16_t> sbp((size_t)h.W * h.H, 0);
            for (int c = 0; c < h.P; c++) sbp[(size_t)h.order[(size_t)c]] = (uint16_t)h.sbin_of_c[(size_t)c];
            uint16_t* d = nullptr;
            TRY(upload(&d, sbp, s));
            p->d_sbin_pix = d;
            p->sbin_pix_bytes = 2;
        }
    }
    if (p->fsnorm) TRY(dalloc(&p->d_fw, (size_t)h.F));
    // dense-row scratch (allocated by the first execute that needs it): one row
    // of F 64-bit values per resident workgroup
    {
        size_t per = (size_t)h.F * sizeof(uint64_t);
        size_t budget = (size_t)1 << 30;
        p->long_blocks = (int)std::min<size_t>(512, std::max<size_t>(1, budget / per));
        // normalize_by_framesum runs every pixel through the series kernel,
        // one wavefront per pixel: more, smaller workgroups (4 GiB of rows)
        if (p->fsnorm) p->long_blocks = (int)std::min<size_t>(8192, std::max<size_t>(1, 4 * budget / per));
    }
    if (hipMemsetAsync(p->d_counters, 0, 32, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        xpcs_mt_plan_destroy(p);
        return fail(XPCS_ERR_HIP, "plan_create: initialisation failed");
    }
#undef TRY
    *out = p;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_plan_destroy(xpcs_mt_plan* p)
{
    if (!p) return XPCS_OK;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    dfree(p->d_pix2c); dfree(p->d_mapped); dfree(p->d_taus); dfree(p->d_levels); dfree(p->d_tau_bnd); dfree(p->d_frame_rank); dfree(p->d_rank_words); dfree(p->d_chunks);
    dfree(p->d_group_chunk_first); dfree(p->d_group_q); dfree(p->d_group_size);
    dfree(p->d_q_chunk_first); dfree(p->d_q_first_group); dfree(p->d_weight); dfree(p->d_c2pix);
    dfree(p->d_sbin_of_c); dfree(p->d_sbin_pix); dfree(p->d_imm_owned); dfree(p->d_sel_off); dfree(p->d_sel_dlen);
    dfree(p->d_sel_frame); dfree(p->d_dark_sum); dfree(p->d_thr); dfree(p->d_ev); dfree(p->d_off);
    dfree(p->d_b1); dfree(p->d_hist); dfree(p->d_hoff); dfree(p->d_bsum); dfree(p->d_rows); dfree(p->d_long_list); dfree(p->d_mid_list); dfree(p->d_counters);
    dfree(p->d_scratch); dfree(p->d_G2); dfree(p->d_IP); dfree(p->d_IF); dfree(p->d_pix_total); dfree(p->d_fw);
    dfree(p->d_deep);
    dfree(p->d_part); dfree(p->d_g2s); dfree(p->d_g2); dfree(p->d_se); dfree(p->d_fs); dfree(p->d_series_a); dfree(p->d_series_b); dfree(p->d_bnd);
    for (auto& t : p->pending) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : p->event_pool) (void)hipEventDestroy(e);
    twotime_release(p);
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_plan_info(const xpcs_mt_plan* p, xpcs_mt_info* info)
{
    if (!p || !info) return fail(XPCS_ERR_INVALID_ARG, "plan_info: NULL argument");
    const HostTables& h = p->h;
    info->taus = h.T;
    info->max_level = h.max_level;
    info->dynamic_bins = h.nq;
    info->static_bins = h.ns;
    info->pixels = h.P;
    info->mapped_pixels = h.n_mapped;
    info->groups = h.ngroups;
    info->events = p->events;
    uint32_t cnt[4] = {0, 0, 0, 0};
    if (p->executed && p->d_counters)
        (void)hipMemcpy(cnt, p->d_counters, sizeof(cnt), hipMemcpyDeviceToHost);
    info->exact_pixels = p->mode == 2 ? (int32_t)cnt[3] : 0;
    info->long_pixels = p->mode == 2 ? h.P : cnt[0];
    int stride = 0;
    info->deep_level = (p->mode == 1 || p->mode == 3) ? deep_level(p, &stride) : (int32_t)h.levels.size();
    info->deep_stride = stride;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_schedule(const xpcs_mt_plan* p, int32_t* level, int32_t* tau)
{
    if (!p || !level || !tau) return fail(XPCS_ERR_INVALID_ARG, "schedule: NULL argument");
    std::memcpy(level, p->h.level.data(), sizeof(int32_t) * (size_t)p->h.T);
    std::memcpy(tau, p->h.tau.data(), sizeof(int32_t) * (size_t)p->h.T);
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_pixel_order(const xpcs_mt_plan* p, int32_t* flat)
{
    if (!p || !flat) return fail(XPCS_ERR_INVALID_ARG, "pixel_order: NULL argument");
    std::memcpy(flat, p->h.order.data(), sizeof(int32_t) * (size_t)p->h.P);
    return XPCS_OK;
}

// ---- IMM -------------------------------------------------------------------------
// io::Header offsets (src/xpcs/io/imm.h:63-144): compression @4, elapsed @128,
// dlen @152, corecotick @620; the first header decides compression for the
// whole file (Imm::Imm, src/xpcs/io/imm.cpp:59-68).
extern "C" xpcs_status xpcs_imm_index(const void* imm, size_t nbytes, int64_t* payload_off,
                                      uint32_t* dlen, double* elapsed, int32_t* corecotick,
                                      int64_t cap, int64_t* frames, int32_t* compressed)
{
    if (!imm || !frames) return fail(XPCS_ERR_INVALID_ARG, "imm_index: NULL argument");
    const uint8_t* b = static_cast<const uint8_t*>(imm);
    *frames = 0;
    if (nbytes < 1024) return fail(XPCS_ERR_FORMAT, "imm_index: file shorter than one header");
    int32_t comp;
    std::memcpy(&comp, b + 4, 4);
    if (compressed) *compressed = comp != 0;
    const size_t per = comp != 0 ? 6 : 2;
    size_t off = 0;
    int64_t n = 0;
    while (off + 1024 <= nbytes) {
        uint32_t dl;
        std::memcpy(&dl, b + off + 152, 4);
        size_t pay = off + 1024;
        if (pay + (size_t)dl * per > nbytes)
            return fail(XPCS_ERR_FORMAT, "imm_index: frame %lld payload (%u entries) overruns the file",
                        (long long)n, dl);
        if (n < cap) {
            if (payload_off) payload_off[n] = (int64_t)pay;
            if (dlen) dlen[n] = dl;
            if (elapsed) std::memcpy(&elapsed[n], b + off + 128, 8);
            if (corecotick) std::memcpy(&corecotick[n], b + off + 620, 4);
        }
        n++;
        off = pay + (size_t)dl * per;
    }
    if (off != nbytes && nbytes - off < 1024 && nbytes - off > 0)
        return fail(XPCS_ERR_FORMAT, "imm_index: %zu trailing bytes after frame %lld",
                    nbytes - off, (long long)n);
    *frames = n;
    return XPCS_OK;
}

namespace {

// Raw frames used by the load loop of src/xpcs/main.cpp:222-268: darks are the
// first frames of the file (uncompressed files only), then skip to frame_from,
// then for output frame f the raw frames r0 + f*R + j*stride, j < avg,
// R = stride*avg (SparseFilter/DenseFilter sum them and divide by avg).
xpcs_status select_frames(const HostTables& h, int64_t n_raw, bool compressed, int64_t* first_data,
                          std::vector<int64_t>& sel, std::vector<int32_t>& out_frame)
{
    int64_t r = 0;
    if (!compressed && h.darks > 0) r += h.darks;
    if (h.frame_from > 0 && r < h.frame_from) r = h.frame_from;
    *first_data = r;
    const int64_t R = (int64_t)h.stride * h.avg;
    sel.clear();
    out_frame.clear();
    sel.reserve((size_t)h.F * h.avg);
    for (int64_t f = 0; f < h.F; f++)
        for (int64_t j = 0; j < h.avg; j++) {
            const int64_t raw = r + f * R + j * h.stride;
            if (raw >= n_raw)
                return fail(XPCS_ERR_FORMAT, "ingest: IMM holds %lld frames, output frame %lld needs raw frame %lld",
                            (long long)n_raw, (long long)f, (long long)raw);
            sel.push_back(raw);
            out_frame.push_back((int32_t)f);
        }
    return XPCS_OK;
}

// Dark image of the first `darks` frames (DarkImage::Compute,
// src/xpcs/data_structure/dark_image.cpp:81-106, Welford in double) plus the
// exact integer sums used for the n-scaled threshold of the dense path.  Dark
// values are uint16 like every count (the reference reads int16,
// src/xpcs/io/imm.cpp:87; equal below 32768), so the threshold's standard
// deviation and the subtracted sums describe the same values.
xpcs_status dense_prepare(xpcs_mt_plan* p, const uint8_t* const* dark_frames)
{
    HostTables& h = p->h;
    const int64_t npd = (int64_t)h.W * h.H;
    p->nscale = h.darks > 0 ? h.darks : 1;
    std::vector<int64_t> dsum((size_t)h.P, 0), thr((size_t)h.P, 0);
    p->dark_avg.assign((size_t)npd, 0.0);
    p->dark_std.assign((size_t)npd, 0.0);
    if (h.darks > 0) {
        if (!h.ff_det.empty())
            return fail(XPCS_ERR_UNSUPPORTED,
                        "dense path: a flat field together with dark frames has no exact integer form");
        std::vector<int64_t> isum((size_t)npd, 0);
        for (int i = 0; i < h.darks; i++) {
            const uint16_t* v = reinterpret_cast<const uint16_t*>(dark_frames[i]);
            for (int64_t j = 0; j < npd; j++) {
                const double x = (double)v[j];
                const double prev = p->dark_avg[(size_t)j];
                p->dark_avg[(size_t)j] += (x - p->dark_avg[(size_t)j]) / (double)(i + 1);
                p->dark_std[(size_t)j] += (x - prev) * (x - p->dark_avg[(size_t)j]);
                isum[(size_t)j] += v[j];
            }
        }
        for (int64_t j = 0; j < npd; j++) p->dark_std[(size_t)j] = std::sqrt(p->dark_std[(size_t)j] / h.darks);
        for (int c = 0; c < h.P; c++) {
            const int64_t j = h.order[(size_t)c];
            const float t = (float)((double)h.lld + (double)h.sigma * p->dark_std[(size_t)j]);
            dsum[(size_t)c] = isum[(size_t)j];
            thr[(size_t)c] = (int64_t)std::floor((double)p->nscale * (double)t);
        }
    }
    xpcs_status st = upload(&p->d_dark_sum, dsum, p->stream);
    if (st) return st;
    return upload(&p->d_thr, thr, p->stream);
}

xpcs_status finish_selection(xpcs_mt_plan* p, const std::vector<int64_t>& sel_off,
                             const std::vector<uint32_t>& sel_dlen, const std::vector<int32_t>& sel_frame)
{
    p->nsel = (int64_t)sel_off.size();
    xpcs_status st = upload(&p->d_sel_off, sel_off, p->stream);
    if (st) return st;
    st = upload(&p->d_sel_dlen, sel_dlen, p->stream);
    if (st) return st;
    st = upload(&p->d_sel_frame, sel_frame, p->stream);
    if (st) return st;
    int64_t raw = 0;
    for (uint32_t d : sel_dlen) raw += d;
    p->raw_events = raw;
    return XPCS_OK;
}

// Host bytes to device memory through two pinned staging buffers: while the
// DMA engine drains one, host threads fill the other (pageable memory cannot
// be handed to the DMA engine directly, and one thread's memcpy is slower than
// the PCIe link).  The buffers are allocated once per process and shared by
// all plans (one staged copy at a time).  Returns when every piece has been
// copied; the caller's buffer is not read afterwards.
struct StagePool {
    std::mutex mu;
    uint8_t* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    size_t bytes = 0;
};

StagePool& stage_pool()
{
    static StagePool pool;   // lives until the process ends
    return pool;
}

xpcs_status h2d_staged(uint8_t* dst, const uint8_t* src, size_t len, hipStream_t s)
{
    constexpr size_t kStage = size_t(32) << 20;
    if (len == 0) return XPCS_OK;
    StagePool& sp = stage_pool();
    std::lock_guard<std::mutex> lk(sp.mu);
    if (!sp.buf[0]) {
        for (int i = 0; i < 2; i++) {
            if (hipHostMalloc(reinterpret_cast<void**>(&sp.buf[i]), kStage, hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&sp.done[i], hipEventDisableTiming) != hipSuccess) {
                for (int j = 0; j <= i; j++) {
                    if (sp.buf[j]) (void)hipHostFree(sp.buf[j]);
                    sp.buf[j] = nullptr;
                }
                return fail(XPCS_ERR_NOMEM, "ingest: pinned staging of %zu bytes failed", kStage);
            }
        }
        sp.bytes = kStage;
    }
    const unsigned nthreads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    int b = 0;
    xpcs_status st = XPCS_OK;
    bool issued[2] = {false, false};
    for (size_t off = 0; off < len && !st; off += sp.bytes, b ^= 1) {
        const size_t n = std::min(sp.bytes, len - off);
        if (issued[b] && hipEventSynchronize(sp.done[b]) != hipSuccess) {
            st = fail(XPCS_ERR_HIP, "ingest: staging copy failed");
            break;
        }
        // parallel host copy into the pinned buffer
        uint8_t* stage = sp.buf[b];
        const size_t per = (n + nthreads - 1) / nthreads;
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nthreads && t * per < n; t++)
            pool.emplace_back([=] { std::memcpy(stage + t * per, src + off + t * per, std::min(per, n - t * per)); });
        std::memcpy(stage, src + off, std::min(per, n));
        for (auto& th : pool) th.join();
        if (hipMemcpyAsync(dst + off, stage, n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipEventRecord(sp.done[b], s) != hipSuccess)
            st = fail(XPCS_ERR_HIP, "ingest: host-to-device copy failed");
        issued[b] = true;
    }
    // the pool is free for the next caller only when its copies have finished
    for (int i = 0; i < 2; i++)
        if (issued[i]) (void)hipEventSynchronize(sp.done[i]);
    return st;
}

// Tile loads of the dense streaming kernel: one output frame per raw frame,
// frames evenly spaced and 16-byte aligned, and every frame readable up to the
// 64-pixel boundary after the detector's last pixel.
void dense_tiling(xpcs_mt_plan* p, const std::vector<int64_t>& so, const uint8_t* base, size_t buf_bytes)
{
    p->tile_off0 = -1;
    p->tile_step = 0;
    const HostTables& h = p->h;
    const int64_t npd = (int64_t)h.W * h.H;
    // 16-byte loads of 8 pixels in the frame-sum kernel: every selected frame
    // 16-byte aligned and the pixel count a multiple of 8
    p->side_vec = npd % 8 == 0;
    for (int64_t o : so)
        if (((uintptr_t)base + (uintptr_t)o) % 16 != 0) p->side_vec = false;
    if (h.avg != 1 || so.empty() || npd % 8 != 0) return;
    const int64_t step = so.size() > 1 ? so[1] - so[0] : 16;
    if (step <= 0 || step % 16 != 0 || ((uintptr_t)base + (uintptr_t)so[0]) % 16 != 0) return;
    for (size_t k = 0; k < so.size(); k++)
        if (so[k] != so[0] + (int64_t)k * step) return;
    if (so.back() + 2 * ((npd + 63) / 64 * 64) > (int64_t)buf_bytes) return;
    p->tile_off0 = so[0];
    p->tile_step = step;
}

}  // namespace

extern "C" xpcs_status xpcs_mt_ingest_imm(xpcs_mt_plan* p, const void* imm, size_t nbytes)
{
    if (!p || !imm) return fail(XPCS_ERR_INVALID_ARG, "ingest_imm: NULL argument");
    HostTables& h = p->h;
    int64_t n = 0;
    int32_t comp = 0;
    xpcs_status st = xpcs_imm_index(imm, nbytes, nullptr, nullptr, nullptr, nullptr, 0, &n, &comp);
    if (st) return st;
    std::vector<int64_t> off((size_t)n);
    std::vector<uint32_t> dl((size_t)n);
    std::vector<double> el((size_t)n);
    std::vector<int32_t> ct((size_t)n);
    st = xpcs_imm_index(imm, nbytes, off.data(), dl.data(), el.data(), ct.data(), n, &n, &comp);
    if (st) return st;
    std::vector<int64_t> sel;
    std::vector<int32_t> ofr;
    int64_t first = 0;
    st = select_frames(h, n, comp != 0, &first, sel, ofr);
    if (st) return st;
    const uint8_t* b = static_cast<const uint8_t*>(imm);
    if (!comp) {
        for (int64_t r : sel)
            if ((int64_t)dl[(size_t)r] != (int64_t)h.W * h.H)
                return fail(XPCS_ERR_FORMAT, "ingest_imm: uncompressed frame %lld has %u values, detector has %lld",
                            (long long)r, dl[(size_t)r], (long long)h.W * h.H);
        std::vector<const uint8_t*> darks;
        for (int i = 0; i < h.darks; i++) {
            if (i >= n) return fail(XPCS_ERR_FORMAT, "ingest_imm: fewer frames than dark frames");
            if ((int64_t)dl[(size_t)i] != (int64_t)h.W * h.H)
                return fail(XPCS_ERR_FORMAT, "ingest_imm: dark frame %d has the wrong size", i);
            darks.push_back(b + off[(size_t)i]);
        }
        st = dense_prepare(p, darks.data());
        if (st) return st;
    }
    // timestamps of every raw frame the load loop reads (SparseFilter::Apply
    // records clock/ticks for each frame of each block, :134-140)
    p->clock.clear();
    p->ticks.clear();
    const int64_t R = (int64_t)h.stride * h.avg;
    for (int64_t k = 0; k < (int64_t)h.F * R && first + k < n; k++) {
        p->clock.push_back(el[(size_t)(first + k)]);
        p->ticks.push_back((double)ct[(size_t)(first + k)]);
    }
    // copy the byte range covering the selected frames
    const int64_t lo_frame = sel.front(), hi_frame = sel.back();
    const size_t lo = (size_t)off[(size_t)lo_frame] - 1024;
    const size_t hi = (size_t)off[(size_t)hi_frame] + (size_t)dl[(size_t)hi_frame] * (comp ? 6 : 2);
    const size_t len = hi - lo;
    // the tail pad keeps tile loads of the last frame inside the allocation
    const size_t pad = 256;
    if (len + pad > p->imm_cap) {
        st = dalloc(&p->d_imm_owned, len + pad);
        if (st) return st;
        p->imm_cap = len + pad;
    }
    st = h2d_staged(p->d_imm_owned, b + lo, len, p->stream);
    if (st) return st;
    p->d_imm = p->d_imm_owned;
    std::vector<int64_t> so(sel.size());
    std::vector<uint32_t> sd(sel.size());
    for (size_t k = 0; k < sel.size(); k++) {
        so[k] = off[(size_t)sel[k]] - (int64_t)lo;
        sd[k] = dl[(size_t)sel[k]];
    }
    if (!comp) dense_tiling(p, so, p->d_imm, p->imm_cap);
    st = finish_selection(p, so, sd, ofr);
    if (st) return st;
    p->mode = comp ? 1 : 2;
    p->fmt = kFmtImm;
    if (comp) p->nscale = 1;
    p->nrows = (int64_t)h.W * h.H;
    p->d_rowmap = p->d_c2pix;
    p->executed = false;
    p->generation++;
    HIP_TRY(hipStreamSynchronize(p->stream));
    return XPCS_OK;
}

// Frame-major events from a caller's own decoder: the selected raw frames are
// laid out as compressed-IMM payload records (int32 pixel[n], u16 count[n],
// padded to 4 bytes) and take the sparse path exactly as a compressed IMM
// file does (src/xpcs/io/imm.cpp:87-99, main.cpp:222-268).
extern "C" xpcs_status xpcs_mt_ingest_events(xpcs_mt_plan* p, const int32_t* pixel, const uint16_t* count,
                                             const int64_t* frame_off, int64_t nframes)
{
    if (!p || !frame_off) return fail(XPCS_ERR_INVALID_ARG, "ingest_events: NULL argument");
    if (nframes < 0) return fail(XPCS_ERR_INVALID_ARG, "ingest_events: nframes < 0");
    if (frame_off[0] != 0) return fail(XPCS_ERR_INVALID_ARG, "ingest_events: frame_off[0] != 0");
    for (int64_t k = 0; k < nframes; k++) {
        if (frame_off[k + 1] < frame_off[k])
            return fail(XPCS_ERR_INVALID_ARG, "ingest_events: frame_off decreases at frame %lld", (long long)k);
        if (frame_off[k + 1] - frame_off[k] > (int64_t)UINT32_MAX)
            return fail(XPCS_ERR_RANGE, "ingest_events: frame %lld has more than 2^32-1 events", (long long)k);
    }
    if (frame_off[nframes] > 0 && (!pixel || !count))
        return fail(XPCS_ERR_INVALID_ARG, "ingest_events: NULL pixel/count arrays");
    HostTables& h = p->h;
    std::vector<int64_t> sel;
    std::vector<int32_t> ofr;
    int64_t first = 0;
    xpcs_status st = select_frames(h, nframes, true, &first, sel, ofr);
    if (st) return st;
    std::vector<int64_t> so(sel.size());
    std::vector<uint32_t> sd(sel.size());
    size_t bytes = 0;
    for (size_t k = 0; k < sel.size(); k++) {
        const int64_t n = frame_off[sel[k] + 1] - frame_off[sel[k]];
        so[k] = (int64_t)bytes;
        sd[k] = (uint32_t)n;
        bytes += ((size_t)n * 6 + 3) & ~(size_t)3;
    }
    std::vector<uint8_t> buf(bytes, 0);
    for (size_t k = 0; k < sel.size(); k++) {
        const int64_t e0 = frame_off[sel[k]];
        const size_t n = sd[k];
        uint8_t* rec = buf.data() + so[k];
        if (n) {
            memcpy(rec, pixel + e0, n * 4);
            memcpy(rec + n * 4, count + e0, n * 2);
        }
    }
    const size_t pad = 256;
    if (bytes + pad > p->imm_cap) {
        st = dalloc(&p->d_imm_owned, bytes + pad);
        if (st) return st;
        p->imm_cap = bytes + pad;
    }
    st = h2d_staged(p->d_imm_owned, buf.data(), bytes, p->stream);
    if (st) return st;
    p->d_imm = p->d_imm_owned;
    st = finish_selection(p, so, sd, ofr);
    if (st) return st;
    p->mode = 1;
    p->fmt = kFmtImm;
    p->nscale = 1;
    p->clock.clear();
    p->ticks.clear();
    p->nrows = (int64_t)h.W * h.H;
    p->d_rowmap = p->d_c2pix;
    p->executed = false;
    p->generation++;
    HIP_TRY(hipStreamSynchronize(p->stream));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_ingest_imm_device(xpcs_mt_plan* p, const void* dev_imm, size_t nbytes,
                                                 const int64_t* payload_off, const uint32_t* dlen,
                                                 int64_t raw_frames, int32_t compressed)
{
    if (!p || !dev_imm || !payload_off || !dlen || raw_frames <= 0)
        return fail(XPCS_ERR_INVALID_ARG, "ingest_imm_device: bad arguments");
    HostTables& h = p->h;
    std::vector<int64_t> sel;
    std::vector<int32_t> ofr;
    int64_t first = 0;
    xpcs_status st = select_frames(h, raw_frames, compressed != 0, &first, sel, ofr);
    if (st) return st;
    const size_t per = compressed ? 6 : 2;
    for (int64_t r : sel) {
        if (payload_off[r] < 0 || (size_t)payload_off[r] + (size_t)dlen[r] * per > nbytes)
            return fail(XPCS_ERR_FORMAT, "ingest_imm_device: frame %lld lies outside the buffer", (long long)r);
        if (!compressed && (int64_t)dlen[r] != (int64_t)h.W * h.H)
            return fail(XPCS_ERR_FORMAT, "ingest_imm_device: uncompressed frame %lld has the wrong size", (long long)r);
    }
    if (!compressed) {
        for (int64_t i = 0; i < h.darks; i++) {
            if (i >= raw_frames)
                return fail(XPCS_ERR_FORMAT, "ingest_imm_device: fewer frames than dark frames");
            if ((int64_t)dlen[i] != (int64_t)h.W * h.H)
                return fail(XPCS_ERR_FORMAT, "ingest_imm_device: dark frame %lld has the wrong size", (long long)i);
            if (payload_off[i] < 0 || (size_t)payload_off[i] + (size_t)dlen[i] * 2 > nbytes)
                return fail(XPCS_ERR_FORMAT, "ingest_imm_device: dark frame %lld lies outside the buffer",
                            (long long)i);
        }
        std::vector<std::vector<uint8_t>> dark_host((size_t)h.darks);
        std::vector<const uint8_t*> darks;
        for (int i = 0; i < h.darks; i++) {
            dark_host[(size_t)i].resize((size_t)h.W * h.H * 2);
            HIP_TRY(hipMemcpy(dark_host[(size_t)i].data(), static_cast<const uint8_t*>(dev_imm) + payload_off[i],
                              dark_host[(size_t)i].size(), hipMemcpyDeviceToHost));
            darks.push_back(dark_host[(size_t)i].data());
        }
        st = dense_prepare(p, darks.data());
        if (st) return st;
    }
    p->clock.clear();
    p->ticks.clear();
    p->d_imm = static_cast<const uint8_t*>(dev_imm);
    std::vector<int64_t> so(sel.size());
    std::vector<uint32_t> sd(sel.size());
    for (size_t k = 0; k < sel.size(); k++) {
        so[k] = payload_off[sel[k]];
        sd[k] = dlen[sel[k]];
    }
    if (!compressed) dense_tiling(p, so, p->d_imm, nbytes);
    st = finish_selection(p, so, sd, ofr);
    if (st) return st;
    p->mode = compressed ? 1 : 2;
    p->fmt = kFmtImm;
    if (compressed) p->nscale = 1;
    p->nrows = (int64_t)h.W * h.H;
    p->d_rowmap = p->d_c2pix;
    p->executed = false;
    p->generation++;
    HIP_TRY(hipStreamSynchronize(p->stream));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_ingest_packed(xpcs_mt_plan* p, const void* words, size_t nbytes, int32_t format)
{
    if (!p || (!words && nbytes)) return fail(XPCS_ERR_INVALID_ARG, "ingest_packed: NULL argument");
    HostTables& h = p->h;
    if (h.darks > 0) return fail(XPCS_ERR_INVALID_ARG, "ingest_packed: packed formats have no dark frames");
    std::vector<PackedRecord> rec;
    xpcs_status st = packed_records(words, nbytes, format, h, rec);
    if (st) return st;
    const int64_t wb = format == XPCS_FMT_UFXC ? 4 : 8;
    // copy the word range the records cover
    int64_t lo = 0, hi = 0;
    if (!rec.empty()) {
        lo = rec.front().word;
        for (const auto& r : rec) {
            lo = std::min(lo, r.word);
            hi = std::max(hi, r.word + (int64_t)r.count);
        }
    }
    const size_t len = (size_t)(hi - lo) * (size_t)wb;
    if (len + 256 > p->imm_cap) {
        st = dalloc(&p->d_imm_owned, len + 256);
        if (st) return st;
        p->imm_cap = len + 256;
    }
    st = h2d_staged(p->d_imm_owned, static_cast<const uint8_t*>(words) + (size_t)lo * (size_t)wb, len, p->stream);
    if (st) return st;
    p->d_imm = p->d_imm_owned;
    std::vector<int64_t> so(rec.size());
    std::vector<uint32_t> sd(rec.size());
    std::vector<int32_t> ofr(rec.size());
    for (size_t k = 0; k < rec.size(); k++) {
        so[k] = (rec[k].word - lo) * wb;
        sd[k] = rec[k].count;
        ofr[k] = rec[k].frame;
    }
    st = finish_selection(p, so, sd, ofr);
    if (st) return st;
    p->clock.clear();
    p->ticks.clear();
    p->nscale = 1;
    p->mode = 1;
    p->fmt = format;
    p->nrows = (int64_t)h.W * h.H;
    p->d_rowmap = p->d_c2pix;
    p->executed = false;
    p->generation++;
    HIP_TRY(hipStreamSynchronize(p->stream));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_ingest_frames(xpcs_mt_plan* p, const void* frames, int64_t nframes,
                                             int32_t bytes_per_value)
{
    if (!p || !frames || nframes <= 0) return fail(XPCS_ERR_INVALID_ARG, "ingest_frames: bad arguments");
    if (bytes_per_value != 2 && bytes_per_value != 4)
        return fail(XPCS_ERR_INVALID_ARG, "ingest_frames: values are 2 or 4 bytes, not %d", bytes_per_value);
    HostTables& h = p->h;
    if (h.darks > 0) return fail(XPCS_ERR_INVALID_ARG, "ingest_frames: raw frames have no dark frames");
    std::vector<int64_t> sel;
    std::vector<int32_t> ofr;
    int64_t first = 0;
    // the HDF5 reader skips to frame_from and has no darks (it reports compression)
    xpcs_status st = select_frames(h, nframes, true, &first, sel, ofr);
    if (st) return st;
    const int64_t npd = (int64_t)h.W * h.H;
    const int64_t f_lo = sel.front(), f_hi = sel.back() + 1;
    const size_t len = (size_t)(f_hi - f_lo) * (size_t)npd * 2;
    if (len + 256 > p->imm_cap) {
        st = dalloc(&p->d_imm_owned, len + 256);
        if (st) return st;
        p->imm_cap = len + 256;
    }
    if (bytes_per_value == 2) {
        st = h2d_staged(p->d_imm_owned, static_cast<const uint8_t*>(frames) + (size_t)f_lo * npd * 2, len, p->stream);
    } else {
        // 32-bit values narrowed on the host (the kernels read 16-bit values)
        const uint32_t* src = static_cast<const uint32_t*>(frames) + (size_t)f_lo * npd;
        std::vector<uint16_t> nar((size_t)(f_hi - f_lo) * (size_t)npd);
        for (size_t i = 0; i < nar.size(); i++) {
            if (src[i] > 65535u)
                return fail(XPCS_ERR_UNSUPPORTED, "ingest_frames: value %u at frame %lld exceeds 65535", src[i],
                            (long long)(f_lo + (int64_t)(i / (size_t)npd)));
            nar[i] = (uint16_t)src[i];
        }
        st = h2d_staged(p->d_imm_owned, reinterpret_cast<const uint8_t*>(nar.data()), len, p->stream);
    }
    if (st) return st;
    p->d_imm = p->d_imm_owned;
    st = dense_prepare(p, nullptr);
    if (st) return st;
    std::vector<int64_t> so(sel.size());
    std::vector<uint32_t> sd(sel.size(), (uint32_t)npd);
    for (size_t k = 0; k < sel.size(); k++) so[k] = (sel[k] - f_lo) * npd * 2;
    dense_tiling(p, so, p->d_imm, p->imm_cap);
    st = finish_selection(p, so, sd, ofr);
    if (st) return st;
    p->clock.clear();
    p->ticks.clear();
    p->mode = 2;
    p->fmt = kFmtImm;
    p->nrows = npd;
    p->d_rowmap = p->d_c2pix;
    p->executed = false;
    p->generation++;
    HIP_TRY(hipStreamSynchronize(p->stream));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_ingest_rows(xpcs_mt_plan* p, const int64_t* row_off, const int32_t* frame,
                                           const uint32_t* count)
{
    if (!p || !row_off) return fail(XPCS_ERR_INVALID_ARG, "ingest_rows: NULL argument");
    if ((!frame || !count) && row_off[(int64_t)p->h.W * p->h.H] > 0)
        return fail(XPCS_ERR_INVALID_ARG, "ingest_rows: NULL frame/count arrays");
    HostTables& h = p->h;
    const int64_t npd = (int64_t)h.W * h.H;
    if (row_off[0] < 0) return fail(XPCS_ERR_INVALID_ARG, "ingest_rows: row_off[0] < 0");
    for (int64_t j = 0; j < npd; j++)
        if (row_off[j + 1] < row_off[j])
            return fail(XPCS_ERR_INVALID_ARG, "ingest_rows: row_off decreases at pixel %lld", (long long)j);
    std::vector<int64_t> off((size_t)h.P + 1, 0);
    for (int c = 0; c < h.P; c++) {
        const int64_t j = h.order[(size_t)c];
        off[(size_t)c + 1] = off[(size_t)c] + (row_off[j + 1] - row_off[j]);
    }
    std::vector<uint64_t> ev;
    try {
        ev.resize((size_t)off[(size_t)h.P]);
    } catch (const std::bad_alloc&) {
        return fail(XPCS_ERR_NOMEM, "ingest_rows: %lld events do not fit host memory", (long long)off[(size_t)h.P]);
    }
    bool wide = false;   // a row short enough for the register kernels totals 2^32 or more
    for (int c = 0; c < h.P; c++) {
        const int64_t j = h.order[(size_t)c];
        int64_t o = off[(size_t)c];
        uint64_t tot = 0;
        for (int64_t e = row_off[j]; e < row_off[j + 1]; e++) {
            if (frame[e] < 0 || frame[e] >= h.F)
                return fail(XPCS_ERR_RANGE, "ingest_rows: frame %d of pixel %lld outside [0, %d)", frame[e],
                            (long long)j, h.F);
            ev[(size_t)o++] = ((uint64_t)(uint32_t)frame[e] << 32) | count[e];
            tot += count[e];
        }
        if (row_off[j + 1] - row_off[j] <= reg_cap() && tot > 0xFFFFFFFFull) wide = true;
    }
    xpcs_status st = upload(&p->d_ev, ev, p->stream);
    if (st) return st;
    p->ev_cap = (int64_t)ev.size();
    HIP_TRY(hipMemcpyAsync(p->d_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, p->stream));
    p->events = off[(size_t)h.P];
    p->nscale = 1;
    p->fmt = kFmtImm;
    p->mode = 3;
    p->wide_rows = wide;
    p->nrows = h.P;
    p->d_rowmap = nullptr;
    p->executed = false;
    p->generation++;
    HIP_TRY(hipStreamSynchronize(p->stream));
    return XPCS_OK;
}

namespace xpcs {

// Dense frames go through the thread-per-pixel streaming kernels when level
// values and pixel totals fit 32 bits and n-scaled raw values fit signed 32 bits.
static bool dense_stream_ok(const xpcs_mt_plan* p)
{
    const HostTables& h = p->h;
    return dense_stream_supported(h) &&
           (double)h.F * 65535.0 * (double)p->nscale * (double)h.avg < 4294967295.0 &&
           65535.0 * (double)p->nscale < 2147483647.0;
}

// Device buffers that depend on the ingested data and the tuning; no launches.
xpcs_status prepare_execute(xpcs_mt_plan* p)
{
    if (p->mode == 0) return fail(XPCS_ERR_STATE, "execute: nothing ingested");
    if (p->fsnorm && p->mode == 3)
        return fail(XPCS_ERR_UNSUPPORTED, "execute: normalize_by_framesum needs frame sums (IMM or packed ingest)");
    HostTables& h = p->h;
    const bool dense_stream = p->mode == 2 && dense_stream_ok(p);
    const bool long_rows = !dense_stream;
    if (long_rows && !p->d_scratch) {
        xpcs_status st = dalloc(&p->d_scratch, (size_t)p->long_blocks * h.F);
        if (st) return st;
    }
    if (p->mode == 1 || p->mode == 3) {
        rank_lds_setup(p);
        int stride = 0;
        if (deep_level(p, &stride) < (int)h.levels.size()) {
            const size_t need = (size_t)h.P * (size_t)stride;
            if (need > p->deep_cap) {
                xpcs_status st = dalloc(&p->d_deep, need);
                if (st) return st;
                p->deep_cap = need;
            }
        }
    }
    if (p->mode == 1) {
        if (p->raw_events > p->ev_cap) {
            xpcs_status st = dalloc(&p->d_ev, (size_t)p->raw_events);
            if (!st) st = dalloc(&p->d_b1, (size_t)p->raw_events);
            if (st) return st;
            p->ev_cap = p->raw_events;
        }
        p->tshape = transpose_shape(p->nrows, p->nsel, p->tune.tr_bucket_shift, p->tune.tr_chunk_frames);
        // the row placement holds one bucket's cursors in LDS: 4096 rows at most
        if (p->tshape.bshift > 12)
            return fail(XPCS_ERR_UNSUPPORTED,
                        "execute: the sparse transpose takes at most 4096 x 4096 detector pixels, not %lld",
                        (long long)p->nrows);
        const int64_t nh = (int64_t)p->tshape.nb * p->tshape.nblk;
        if (nh + 1 > p->hist_cap) {
            xpcs_status st = dalloc(&p->d_hist, (size_t)nh);
            if (!st) st = dalloc(&p->d_hoff, (size_t)nh + 1);
            if (st) return st;
            p->hist_cap = nh + 1;
        }
        const int64_t nbs = bucket_scan_scratch(p->tshape.nb, p->tshape.nblk);
        if (nbs > p->bsum_cap) {
            xpcs_status st = dalloc(&p->d_bsum, (size_t)nbs);
            if (st) return st;
            p->bsum_cap = nbs;
        }
    } else if (dense_stream) {
        const size_t need = dense_series_elems(p);
        if (need > p->series_cap) {
            xpcs_status st = dalloc(&p->d_series_a, need);
            if (!st) st = dalloc(&p->d_series_b, need);
            if (st) return st;
            p->series_cap = need;
        }
    } else if (p->mode == 2) {
        const size_t need = (size_t)h.P * (size_t)h.F;
        if (need > p->rows_cap) {
            xpcs_status st = dalloc(&p->d_rows, need);
            if (st) return st;
            p->rows_cap = need;
        }
    }
    return XPCS_OK;
}

// Every launch of one job on stream s: no allocation, no host synchronisation
// and no memset nodes (counters and side sums are cleared by a fill kernel),
// so the sequence can be captured into a graph.
void enqueue_execute(xpcs_mt_plan* p, hipStream_t s)
{
    HostTables& h = p->h;
    // integer IP/IF in the 32-bit layout (unless tuned off; FP64 sums stay 8
    // bytes), recorded with the launches that write them, so the readers follow
    // the last multi-tau run whatever ran on the plan since
    p->ipif32 = !p->fsnorm && p->tune.sums64 == 0;
    launch_fill_u32(p->d_counters, 8, s);
    if (p->d_fs) launch_fill_u64(reinterpret_cast<uint64_t*>(p->d_fs), (size_t)h.F * h.ns, s);
    if (p->fsnorm && p->mode == 2 && dense_stream_ok(p)) {
        // normalize_by_framesum on dense frames: frame sums from the raw
        // frames, then the streaming kernels over frame-weighted values
        launch_side_dense_raw(p, s);
        launch_frame_weights(p, s);
        launch_mt_dense(p, s, true);
    } else if (p->fsnorm) {
        // normalize_by_framesum: frame sums first, then every pixel in FP64
        const bool fused = side_fused(p);   // frame sums inside the transpose's count pass
        if (p->mode == 1) launch_transpose(p, s, fused);
        else launch_dense_transpose(p, s);
        if (!fused) launch_side(p, s);
        launch_frame_weights(p, s);
        launch_mt_fsnorm(p, s, p->mode == 2);
    } else if (p->mode == 1) {
        const bool fused = side_fused(p);
        launch_transpose(p, s, fused);
        if (p->d_fs && !fused) launch_side(p, s);
        launch_mt_sparse(p, s);
    } else if (p->mode == 2 && dense_stream_ok(p)) {
        // thread-per-pixel streaming multi-tau over the raw frames
        if (p->d_fs) launch_side_dense_raw(p, s);
        launch_mt_dense(p, s, false);
    } else if (p->mode == 2) {
        launch_dense_transpose(p, s);
        if (p->d_fs) launch_side(p, s);
        launch_mt_long(p, s, true);
    } else {
        launch_mt_sparse(p, s);
    }
    launch_normalize(p, s);
}

}  // namespace xpcs

extern "C" xpcs_status xpcs_mt_execute(xpcs_mt_plan* p, void* stream)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "execute: NULL plan");
    xpcs_status st = prepare_execute(p);
    if (st) return st;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : p->stream;
    enqueue_execute(p, s);
    HIP_TRY(hipGetLastError());
    p->stream_last = s;
    p->executed = true;
    return XPCS_OK;
}

static xpcs_status sync_and_check(xpcs_mt_plan* p)
{
    hipStream_t s = p->stream_last ? p->stream_last : p->stream;
    HIP_TRY(hipStreamSynchronize(s));
    collect_timers(p);
    uint32_t cnt[2] = {0, 0};
    HIP_TRY(hipMemcpy(cnt, p->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[1] & 1u) return fail(XPCS_ERR_RANGE, "execute: an IMM pixel index lies outside the detector");
    if (p->mode == 1) {
        int64_t e = 0;
        HIP_TRY(hipMemcpy(&e, p->d_off + p->nrows, sizeof(e), hipMemcpyDeviceToHost));
        p->events = e;
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_synchronize(xpcs_mt_plan* p)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "synchronize: NULL plan");
    if (!p->executed) return XPCS_OK;
    return sync_and_check(p);
}

// IP or IF of an integer plan as int64 [P][T]: the 32-bit layout widened with
// the high words of the pixels whose total passes 2^32 - 1.
static xpcs_status read_ipif(xpcs_mt_plan* p, const int64_t* d, int64_t* out)
{
    const size_t n = (size_t)p->h.P * p->h.T;
    if (!p->ipif32) {
        HIP_TRY(hipMemcpy(out, d, n * sizeof(int64_t), hipMemcpyDeviceToHost));
        return XPCS_OK;
    }
    std::vector<uint32_t> w(2 * n);
    std::vector<uint64_t> tot((size_t)p->h.P);
    HIP_TRY(hipMemcpy(w.data(), d, 2 * n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(tot.data(), p->d_pix_total, tot.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    const size_t T = (size_t)p->h.T;
    for (size_t i = 0; i < n; i++) {
        const uint64_t hi = tot[i / T] > 0xFFFFFFFFull ? (uint64_t)w[n + i] << 32 : 0;
        out[i] = (int64_t)(hi | w[i]);
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_sums_f64(xpcs_mt_plan* p, double* G2, double* IP, double* IF)
{
    if (!p || !G2 || !IP || !IF) return fail(XPCS_ERR_INVALID_ARG, "read_sums_f64: NULL argument");
    if (!p->executed) return fail(XPCS_ERR_STATE, "read_sums_f64: execute first");
    xpcs_status st = sync_and_check(p);
    if (st) return st;
    const size_t n = (size_t)p->h.P * p->h.T;
    std::vector<int64_t> tmp(n);
    double* out[3] = {G2, IP, IF};
    const int64_t* src[3] = {p->d_G2, p->d_IP, p->d_IF};
    for (int k = 0; k < 3; k++) {
        if (k > 0 && !p->fsnorm) {
            st = read_ipif(p, src[k], tmp.data());
            if (st) return st;
        } else {
            HIP_TRY(hipMemcpy(tmp.data(), src[k], n * sizeof(int64_t), hipMemcpyDeviceToHost));
        }
        for (size_t i = 0; i < n; i++) {
            double v;
            if (p->fsnorm) std::memcpy(&v, &tmp[i], sizeof(v));
            else v = (double)tmp[i];
            out[k][i] = v;
        }
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_sums_i64(xpcs_mt_plan* p, int64_t* G2, int64_t* IP, int64_t* IF)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "read_sums: NULL plan");
    if (!p->executed) return fail(XPCS_ERR_STATE, "read_sums: execute first");
    if (p->fsnorm) return fail(XPCS_ERR_STATE, "read_sums: normalize_by_framesum sums are FP64 (read_sums_f64)");
    xpcs_status st = sync_and_check(p);
    if (st) return st;
    const size_t n = (size_t)p->h.P * p->h.T * sizeof(int64_t);
    if (G2) HIP_TRY(hipMemcpy(G2, p->d_G2, n, hipMemcpyDeviceToHost));
    if (IP && (st = read_ipif(p, p->d_IP, IP))) return st;
    if (IF && (st = read_ipif(p, p->d_IF, IF))) return st;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_ref_f32(xpcs_mt_plan* p, float* G2, float* IP, float* IF)
{
    if (!p || !G2 || !IP || !IF) return fail(XPCS_ERR_INVALID_ARG, "read_ref_f32: NULL argument");
    if (!p->executed) return fail(XPCS_ERR_STATE, "read_ref_f32: execute first");
    xpcs_status st = sync_and_check(p);
    if (st) return st;
    const size_t n = (size_t)p->h.W * p->h.H * p->h.T;
    float* d = nullptr;
    HIP_TRY(hipMalloc(&d, 3 * n * sizeof(float)));
    hipError_t e = hipMemsetAsync(d, 0, 3 * n * sizeof(float), p->stream);
    if (e == hipSuccess) {
        launch_ref_f32(p, p->stream, d, d + n, d + 2 * n);
        e = hipStreamSynchronize(p->stream);
    }
    if (e == hipSuccess) e = hipMemcpy(G2, d, n * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(IP, d + n, n * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(IF, d + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(XPCS_ERR_HIP, "read_ref_f32: %s", hipGetErrorString(e));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_g2(xpcs_mt_plan* p, double* g2, double* se)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "read_g2: NULL plan");
    if (!p->executed && !p->normalized) return fail(XPCS_ERR_STATE, "read_g2: execute first");
    if (p->executed) {
        xpcs_status st = sync_and_check(p);
        if (st) return st;
    }
    const size_t n = (size_t)p->h.nq * p->h.T * sizeof(double);
    if (g2) HIP_TRY(hipMemcpy(g2, p->d_g2, n, hipMemcpyDeviceToHost));
    if (se) HIP_TRY(hipMemcpy(se, p->d_se, n, hipMemcpyDeviceToHost));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_normalize_ref_f32(xpcs_mt_plan* p, const float* G2, const float* IP,
                                                 const float* IF, int64_t pixels, int32_t taus, double* g2,
                                                 double* se)
{
    if (!p || !G2 || !IP || !IF) return fail(XPCS_ERR_INVALID_ARG, "normalize_ref_f32: NULL argument");
    const HostTables& h = p->h;
    const int64_t npd = (int64_t)h.W * h.H;
    if (pixels != npd || taus != h.T)
        return fail(XPCS_ERR_INVALID_ARG, "normalize_ref_f32: arrays are %d x %lld, the plan needs %d taus x %lld pixels",
                    taus, (long long)pixels, h.T, (long long)npd);
    // an execute still in flight writes the same partial and result buffers
    if (p->executed) HIP_TRY(hipStreamSynchronize(p->stream_last ? p->stream_last : p->stream));
    const size_t n_in = (size_t)npd * (size_t)h.T;
    const size_t n_out = std::max<size_t>(1, (size_t)h.P * (size_t)h.T);
    float* d_in = nullptr;
    int64_t* d_sum = nullptr;
    hipError_t e = hipMalloc(&d_in, 3 * n_in * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_sum, 3 * n_out * sizeof(int64_t));
    hipStream_t s = p->stream;
    if (e == hipSuccess) e = hipMemcpyAsync(d_in, G2, n_in * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in + n_in, IP, n_in * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in + 2 * n_in, IF, n_in * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        launch_from_ref_f32(p, s, d_in, d_in + n_in, d_in + 2 * n_in, d_sum, d_sum + n_out, d_sum + 2 * n_out);
        // FP64 sums, no flat-field weights (the values carry them)
        launch_normalize_sums(p, s, 0, h.T, h.T, p->d_taus, d_sum, d_sum + n_out, d_sum + 2 * n_out, true, false,
                              nullptr, plan_norm_out(p));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const size_t nr = (size_t)h.nq * h.T * sizeof(double);
    if (e == hipSuccess && g2) e = hipMemcpy(g2, p->d_g2, nr, hipMemcpyDeviceToHost);
    if (e == hipSuccess && se) e = hipMemcpy(se, p->d_se, nr, hipMemcpyDeviceToHost);
    if (d_in) (void)hipFree(d_in);
    if (d_sum) (void)hipFree(d_sum);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? XPCS_ERR_NOMEM : XPCS_ERR_HIP, "normalize_ref_f32: %s",
                    hipGetErrorString(e));
    collect_timers(p);
    p->normalized = true;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_device_results(xpcs_mt_plan* p, const double** g2, const double** se)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "device_results: NULL plan");
    if (g2) *g2 = p->d_g2;
    if (se) *se = p->d_se;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_side(xpcs_mt_plan* p, int32_t window, float* pixel_sum,
                                         float* frame_sum, float* partition_mean, float* partition_partial)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "read_side: NULL plan");
    if (!p->executed) return fail(XPCS_ERR_STATE, "read_side: execute first");
    (void)window;
    xpcs_status st = sync_and_check(p);
    if (st) return st;
    const HostTables& h = p->h;
    if ((frame_sum || partition_partial) && !p->d_fs)
        return fail(XPCS_ERR_STATE, "read_side: frame sums need a plan created with XPCS_MT_SIDE_STATS");
    if (partition_partial && window <= 0)
        return fail(XPCS_ERR_INVALID_ARG, "read_side: static window must be >= 1");
    const double vscale = 1.0 / ((double)h.avg * (double)p->nscale);
    if (frame_sum || partition_partial) {
        std::vector<double> fs((size_t)h.F * h.ns);
        HIP_TRY(hipMemcpy(fs.data(), p->d_fs, fs.size() * sizeof(double), hipMemcpyDeviceToHost));
        if (frame_sum) {
            // [2][F]: frame number (1-based), sum over masked pixels / (W*H)
            for (int f = 0; f < h.F; f++) {
                double t = 0.0;
                for (int b = 0; b < h.ns; b++) t += fs[(size_t)f * h.ns + b];
                frame_sum[f] = (float)(f + 1);
                frame_sum[h.F + f] = (float)(t * vscale / ((double)h.W * h.H));
            }
        }
        if (partition_partial) {
            // [floor(F/window)][S] means over each window of `window` frames
            const int nw = h.F / window;
            for (int w = 0; w < nw; w++)
                for (int b = 0; b < h.ns; b++) {
                    double t = 0.0;
                    for (int f = w * window; f < (w + 1) * window; f++) t += fs[(size_t)f * h.ns + b];
                    partition_partial[(size_t)w * h.ns + b] =
                        (float)(t * vscale / ((double)h.pixels_per_bin[(size_t)b] * window));
                }
        }
    }
    std::vector<uint64_t> tot((size_t)h.P);
    HIP_TRY(hipMemcpy(tot.data(), p->d_pix_total, tot.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    const double inv = 1.0 / ((double)h.avg * (double)p->nscale);
    if (pixel_sum) {
        std::fill(pixel_sum, pixel_sum + (size_t)h.W * h.H, 0.0f);
        for (int c = 0; c < h.P; c++) {
            const double w = h.weight.empty() ? 1.0 : h.weight[(size_t)c];
            pixel_sum[h.order[(size_t)c]] = (float)((double)tot[(size_t)c] * w * inv / (double)h.F);
        }
    }
    if (partition_mean) {
        std::vector<double> acc((size_t)h.ns, 0.0);
        for (int c = 0; c < h.P; c++) {
            const double w = h.weight.empty() ? 1.0 : h.weight[(size_t)c];
            acc[(size_t)h.sbin_of_c[(size_t)c] - 1] += (double)tot[(size_t)c] * w * inv;
        }
        for (int s = 0; s < h.ns; s++)
            partition_mean[s] = (float)(acc[(size_t)s] / ((double)h.pixels_per_bin[(size_t)s] * (double)h.F));
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_row_stats(xpcs_mt_plan* p, int64_t* pixels, int64_t* events)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "row_stats: NULL plan");
    if (!p->executed) return fail(XPCS_ERR_STATE, "row_stats: execute first");
    if (p->mode == 2) return fail(XPCS_ERR_STATE, "row_stats: dense ingests form no rows");
    xpcs_status st = sync_and_check(p);
    if (st) return st;
    const HostTables& h = p->h;
    std::vector<int64_t> off((size_t)p->nrows + 1);
    HIP_TRY(hipMemcpy(off.data(), p->d_off, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t np[6] = {0, 0, 0, 0, 0, 0}, ne[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < h.P; c++) {
        const int64_t row = p->d_rowmap ? (int64_t)h.order[(size_t)c] : c;
        const int64_t n = off[(size_t)row + 1] - off[(size_t)row];
        const int k = n == 0 ? 0 : n <= 64 ? 1 : n <= 128 ? 2 : n <= 512 ? 3 : n <= sparse_cap_long() ? 4 : 5;
        np[k]++;
        ne[k] += n;
    }
    for (int k = 0; k < 6; k++) {
        if (pixels) pixels[k] = np[k];
        if (events) events[k] = ne[k];
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_kernel_times(xpcs_mt_plan* p, char* names, size_t cap, double* ms,
                                            int64_t* launches, int32_t max_kernels, int32_t* n_kernels)
{
    if (!p || !n_kernels) return fail(XPCS_ERR_INVALID_ARG, "kernel_times: NULL argument");
    if (p->executed) {
        xpcs_status st = sync_and_check(p);
        if (st) return st;
    }
    std::string all;
    int k = 0;
    for (auto& kv : p->times) {
        if (k < max_kernels) {
            if (ms) ms[k] = kv.second.first;
            if (launches) launches[k] = kv.second.second;
        }
        all += kv.first;
        all += '\n';
        k++;
    }
    *n_kernels = k;
    if (names && cap > 0) {
        size_t n = std::min(cap - 1, all.size());
        std::memcpy(names, all.data(), n);
        names[n] = 0;
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_kernel_times_reset(xpcs_mt_plan* p)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "kernel_times_reset: NULL plan");
    if (p->executed) (void)hipStreamSynchronize(p->stream_last ? p->stream_last : p->stream);
    collect_timers(p);
    p->times.clear();
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_set_timing_filter(xpcs_mt_plan* p, const char* name)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "set_timing_filter: NULL plan");
    p->timer_only = name ? name : "";
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_read_dark(xpcs_mt_plan* p, double* dark_avg, double* dark_std)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "read_dark: NULL plan");
    const size_t n = (size_t)p->h.W * p->h.H;
    for (size_t i = 0; i < n; i++) {
        if (dark_avg) dark_avg[i] = p->dark_avg.size() == n ? p->dark_avg[i] : 0.0;
        if (dark_std) dark_std[i] = p->dark_std.size() == n ? p->dark_std[i] : 0.0;
    }
    return XPCS_OK;
}
