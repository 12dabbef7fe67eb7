// Streams: online multi-tau over frames that arrive in pieces, with the level
// state resident in HBM (k_stream.hip) and g2 snapshots at any frame count.
// No reference counterpart: the reference correlates a finished IMM file
// (src/xpcs/main.cpp:222-268 load loop, then Corr::multiTau2 and
// Corr::normalizeG2s).  A stream wraps a plan built for its maximum length; a
// snapshot after n frames gives exactly what a plan of n frames gives.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace xpcs;

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(XPCS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct xpcs_mt_stream {
    xpcs_mt_plan* plan = nullptr;
    int32_t max_frames = 0;
    int32_t n = 0;               // frames pushed
    int32_t nlev = 0;            // levels with state
    int32_t d2 = 0;              // 2 * dpl
    uint32_t* win = nullptr;
    uint32_t* head = nullptr;
    uint64_t* acc = nullptr;
    uint32_t* pend = nullptr;
    uint64_t* total = nullptr;
    uint32_t* tile = nullptr;
    uint8_t* d_imm = nullptr;    // copies of host-pushed frames
    size_t imm_cap = 0;
    uint8_t* h_imm = nullptr;    // pinned staging of host-pushed frames
    size_t h_cap = 0;
    hipEvent_t copied = nullptr; // the last staging copy has been read
    bool copy_pending = false;
    TauParam* d_taus_snap = nullptr;
    hipEvent_t done = nullptr;
    int32_t snap_n = 0;          // frames of the newest snapshot enqueued
    int32_t snap_T = 0;
    int32_t done_n = 0;          // frames of the newest snapshot seen finished
};

namespace {

template <typename T>
xpcs_status salloc(T** p, size_t n)
{
    *p = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return fail(XPCS_ERR_NOMEM, "stream: hipMalloc of %zu bytes failed", n * sizeof(T));
    return XPCS_OK;
}

bool stream_desc_ok(const xpcs_mt_desc& d, int32_t max_frames)
{
    return max_frames >= 1 && (d.delays_per_level == 4 || d.delays_per_level == 8) &&
           (d.stride == 0 || d.stride == 1) && (d.average == 0 || d.average == 1) && d.frame_from == 0 &&
           d.dark_frames == 0 && d.flags == 0;
}

size_t state_bytes(int32_t P, int32_t nlev, int32_t d2)
{
    const size_t per_level = (size_t)d2 * (4 + 4 + 8) + 4;
    return (size_t)P * ((size_t)nlev * per_level + 8 + 4 * (size_t)kStreamChunk);
}

StreamLaunch launch_args(xpcs_mt_stream* s)
{
    StreamLaunch a{};
    a.pix2c = s->plan->d_pix2c;
    a.npix_det = (uint32_t)((int64_t)s->plan->h.W * s->plan->h.H);
    a.P = s->plan->h.P;
    a.tile = s->tile;
    a.counters = s->plan->d_counters;
    a.d2 = s->d2;
    a.nlev = s->nlev;
    a.win = s->win;
    a.head = s->head;
    a.acc = s->acc;
    a.pend = s->pend;
    a.total = s->total;
    return a;
}

// Frames [n, n + nf) with payloads at imm + off[i]: updates of at most
// kStreamChunk frames each, their frame tables passed as kernel arguments.
void enqueue_updates(xpcs_mt_stream* s, const uint8_t* imm, const int64_t* off, const uint32_t* dlen, int32_t nf)
{
    StreamLaunch a = launch_args(s);
    a.imm = imm;
    for (int32_t k = 0; k < nf; k += kStreamChunk) {
        a.nb = std::min(kStreamChunk, nf - k);
        for (int32_t i = 0; i < a.nb; i++) {
            a.fr.off[i] = off[k + i];
            a.fr.dlen[i] = dlen[k + i];
        }
        a.n0 = s->n;
        launch_stream_update(a, s->plan->stream);
        s->n += a.nb;
    }
}

}  // namespace

extern "C" xpcs_status xpcs_mt_stream_bytes(const xpcs_mt_desc* desc, int32_t max_frames, uint64_t* bytes)
{
    if (!desc || !bytes) return fail(XPCS_ERR_INVALID_ARG, "stream_bytes: NULL argument");
    if (!stream_desc_ok(*desc, max_frames))
        return fail(XPCS_ERR_UNSUPPORTED, "stream: dpl 4 or 8, no stride/average/skip/darks/flags");
    xpcs_mt_desc d = *desc;
    d.frames = max_frames;
    HostTables h;
    xpcs_status st = build_tables(d, h);
    if (st) return st;
    const size_t P = (size_t)h.P, T = (size_t)h.T;
    size_t b = state_bytes(h.P, (int32_t)h.levels.size(), 2 * h.dpl);
    b += P * T * 24 + P * 8 + h.chunks.size() * T * sizeof(NormPartial) + (size_t)h.ngroups * T * 8 +
         2 * (size_t)h.nq * T * 8 + (size_t)h.W * h.H * 4 + P * 24;
    *bytes = b;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_open(const xpcs_mt_desc* desc, int32_t max_frames, xpcs_mt_stream** out)
{
    if (!desc || !out) return fail(XPCS_ERR_INVALID_ARG, "stream_open: NULL argument");
    *out = nullptr;
    if (!stream_desc_ok(*desc, max_frames))
        return fail(XPCS_ERR_UNSUPPORTED, "stream: dpl 4 or 8, no stride/average/skip/darks/flags");
    xpcs_mt_desc d = *desc;
    d.frames = max_frames;
    auto* s = new xpcs_mt_stream();
    xpcs_status st = xpcs_mt_plan_create(&d, &s->plan);
    if (st) {
        delete s;
        return st;
    }
    const HostTables& h = s->plan->h;
    s->max_frames = max_frames;
    s->d2 = 2 * h.dpl;
    s->nlev = (int32_t)h.levels.size();
    const size_t P = (size_t)h.P, L = (size_t)s->nlev, D2 = (size_t)s->d2;
    if (!st) st = salloc(&s->win, L * D2 * P);
    if (!st) st = salloc(&s->head, L * D2 * P);
    if (!st) st = salloc(&s->acc, L * D2 * P);
    if (!st) st = salloc(&s->pend, L * P);
    if (!st) st = salloc(&s->total, P);
    if (!st) st = salloc(&s->tile, (size_t)kStreamChunk * P);
    if (!st) st = salloc(&s->d_taus_snap, (size_t)h.T);
    hipStream_t q = s->plan->stream;
    if (!st) {
        launch_fill_u32(s->win, L * D2 * P, q);
        launch_fill_u32(s->head, L * D2 * P, q);
        launch_fill_u64(s->acc, L * D2 * P, q);
        launch_fill_u32(s->pend, L * P, q);
        launch_fill_u64(s->total, P, q);
        launch_fill_u32(s->tile, (size_t)kStreamChunk * P, q);
        launch_fill_u32(s->plan->d_counters, 8, q);
        if (hipEventCreateWithFlags(&s->done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->copied, hipEventDisableTiming) != hipSuccess ||
            hipGetLastError() != hipSuccess || hipStreamSynchronize(q) != hipSuccess)
            st = fail(XPCS_ERR_HIP, "stream_open: initialisation failed");
    }
    if (st) {
        xpcs_mt_stream_close(s);
        return st;
    }
    *out = s;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_push_imm_device(xpcs_mt_stream* s, const void* dev_imm, size_t nbytes,
                                                      const int64_t* payload_off, const uint32_t* dlen,
                                                      int32_t nframes)
{
    if (!s || (!dev_imm && nframes > 0) || (nframes > 0 && (!payload_off || !dlen)) || nframes < 0)
        return fail(XPCS_ERR_INVALID_ARG, "stream_push: bad arguments");
    if (nframes == 0) return XPCS_OK;
    if (nframes > s->max_frames - s->n)
        return fail(XPCS_ERR_RANGE, "stream_push: %d frames exceed the stream's %d", s->n + nframes,
                    s->max_frames);
    for (int32_t f = 0; f < nframes; f++)
        if (payload_off[f] < 0 || (size_t)payload_off[f] + (size_t)dlen[f] * 6 > nbytes)
            return fail(XPCS_ERR_FORMAT, "stream_push: frame %d lies outside the buffer", f);
    enqueue_updates(s, static_cast<const uint8_t*>(dev_imm), payload_off, dlen, nframes);
    HIP_TRY(hipGetLastError());
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_push_imm(xpcs_mt_stream* s, const void* imm, size_t nbytes)
{
    if (!s || !imm) return fail(XPCS_ERR_INVALID_ARG, "stream_push_imm: NULL argument");
    int64_t n = 0;
    int32_t comp = 0;
    xpcs_status st = xpcs_imm_index(imm, nbytes, nullptr, nullptr, nullptr, nullptr, 0, &n, &comp);
    if (st) return st;
    if (n == 0) return XPCS_OK;
    if (!comp) return fail(XPCS_ERR_UNSUPPORTED, "stream_push_imm: uncompressed frames");
    std::vector<int64_t> off((size_t)n);
    std::vector<uint32_t> dl((size_t)n);
    st = xpcs_imm_index(imm, nbytes, off.data(), dl.data(), nullptr, nullptr, n, &n, &comp);
    if (st) return st;
    hipStream_t q = s->plan->stream;
    if (nbytes > s->imm_cap || nbytes > s->h_cap) {
        // earlier pushes may still read the old buffers
        HIP_TRY(hipStreamSynchronize(q));
        s->copy_pending = false;
        if (s->d_imm) (void)hipFree(s->d_imm);
        if (s->h_imm) (void)hipHostFree(s->h_imm);
        s->d_imm = nullptr;
        s->h_imm = nullptr;
        s->imm_cap = s->h_cap = 0;
        st = salloc(&s->d_imm, nbytes);
        if (st) return st;
        if (hipHostMalloc(reinterpret_cast<void**>(&s->h_imm), nbytes, hipHostMallocDefault) != hipSuccess)
            return fail(XPCS_ERR_NOMEM, "stream_push_imm: pinned allocation of %zu bytes failed", nbytes);
        s->imm_cap = s->h_cap = nbytes;
    } else if (s->copy_pending) {
        HIP_TRY(hipEventSynchronize(s->copied));   // the staging buffer has been read
    }
    std::memcpy(s->h_imm, imm, nbytes);
    // ordered after the updates of earlier pushes on the same stream
    HIP_TRY(hipMemcpyAsync(s->d_imm, s->h_imm, nbytes, hipMemcpyHostToDevice, q));
    HIP_TRY(hipEventRecord(s->copied, q));
    s->copy_pending = true;
    return xpcs_mt_stream_push_imm_device(s, s->d_imm, nbytes, off.data(), dl.data(), (int32_t)n);
}

extern "C" xpcs_status xpcs_mt_stream_frames(xpcs_mt_stream* s, int32_t* frames, int32_t* taus)
{
    if (!s) return fail(XPCS_ERR_INVALID_ARG, "stream_frames: NULL stream");
    if (frames) *frames = s->n;
    if (taus) {
        *taus = 0;
        if (s->n >= 1) {
            const int32_t lm = xpcs_level_max(s->n, s->plan->h.dpl);
            int32_t cnt = 0;
            if (lm >= 0 && xpcs_delays_per_level(s->n, s->plan->h.dpl, lm, nullptr, nullptr, 0, &cnt) == XPCS_OK)
                *taus = cnt;
        }
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_snapshot(xpcs_mt_stream* s)
{
    if (!s) return fail(XPCS_ERR_INVALID_ARG, "stream_snapshot: NULL stream");
    int32_t T = 0;
    xpcs_mt_stream_frames(s, nullptr, &T);
    if (T <= 0) return fail(XPCS_ERR_STATE, "stream_snapshot: %d frames give no delay", s->n);
    // the schedule of n frames is a prefix of the schedule of the maximum length
    std::vector<int32_t> lev((size_t)T), tau((size_t)T);
    int32_t cnt = 0;
    xpcs_status st = xpcs_delays_per_level(s->n, s->plan->h.dpl, xpcs_level_max(s->n, s->plan->h.dpl), lev.data(),
                                           tau.data(), T, &cnt);
    if (st) return st;
    const HostTables& h = s->plan->h;
    if (T > h.T) return fail(XPCS_ERR_STATE, "stream_snapshot: schedule longer than the maximum length's");
    for (int32_t i = 0; i < T; i++)
        if (h.taus[(size_t)i].level != lev[(size_t)i] || (h.taus[(size_t)i].tp << h.taus[(size_t)i].level) != tau[(size_t)i])
            return fail(XPCS_ERR_STATE, "stream_snapshot: schedule of %d frames is not a prefix", s->n);
    hipStream_t q = s->plan->stream;
    const StreamLaunch a = launch_args(s);
    // Only what changed since the previous snapshot is recomputed.  Level L of
    // n frames differs from level L of the previous count n' only when
    // n >> L != n' >> L (the levels up to the highest bit of n ^ n'), and the
    // taus beyond the previous schedule are new; the sums, chunk partials and
    // g2 rows of every other tau are the previous snapshot's.  Sums and
    // partials use the maximum schedule's row stride, so they stay in place.
    const int32_t S = h.T;
    int32_t rng[2][2];
    int nr = 0;
    if (s->snap_T <= 0) {
        rng[nr][0] = 0;
        rng[nr++][1] = T;
    } else {
        const uint32_t x = (uint32_t)(s->n ^ s->snap_n);
        int32_t e = 0;   // taus [0, e) belong to changed levels
        if (x) {
            const int lc = 31 - __builtin_clz(x);
            e = lc + 1 < (int)h.levels.size() ? h.levels[(size_t)lc + 1].first : T;
            e = std::min(e, T);
        }
        const int32_t tp = s->snap_T;
        if (e >= tp) {
            rng[nr][0] = 0;
            rng[nr++][1] = T;
        } else {
            if (e > 0) {
                rng[nr][0] = 0;
                rng[nr++][1] = e;
            }
            if (T > tp) {
                rng[nr][0] = tp;
                rng[nr++][1] = T;
            }
        }
    }
    for (int r = 0; r < nr; r++) {
        launch_stream_snapshot(a, s->n, s->plan->d_levels, (int32_t)h.levels.size(), s->plan->d_taus, rng[r][0],
                               rng[r][1], S, s->d_taus_snap, s->plan->d_G2, s->plan->d_IP, s->plan->d_IF, q);
        launch_normalize_with(s->plan, q, rng[r][0], rng[r][1], S, s->d_taus_snap);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(s->done, q));
    s->snap_n = s->n;
    s->snap_T = T;
    s->plan->stream_last = q;
    s->plan->executed = true;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_poll(xpcs_mt_stream* s, int32_t* frames_done)
{
    if (!s || !frames_done) return fail(XPCS_ERR_INVALID_ARG, "stream_poll: NULL argument");
    if (s->snap_n > s->done_n) {
        const hipError_t e = hipEventQuery(s->done);
        if (e == hipSuccess) s->done_n = s->snap_n;
        else if (e != hipErrorNotReady) return fail(XPCS_ERR_HIP, "stream_poll: %s", hipGetErrorString(e));
    }
    *frames_done = s->done_n;
    return XPCS_OK;
}

static xpcs_status stream_wait(xpcs_mt_stream* s)
{
    if (s->snap_T <= 0) return fail(XPCS_ERR_STATE, "stream: no snapshot taken");
    HIP_TRY(hipStreamSynchronize(s->plan->stream));
    s->done_n = s->snap_n;
    uint32_t cnt[2] = {0, 0};
    HIP_TRY(hipMemcpy(cnt, s->plan->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[1] & 1u) return fail(XPCS_ERR_RANGE, "stream: an IMM pixel index lies outside the detector");
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_read_g2(xpcs_mt_stream* s, double* g2, double* se, int32_t* taus)
{
    if (!s) return fail(XPCS_ERR_INVALID_ARG, "stream_read_g2: NULL stream");
    xpcs_status st = stream_wait(s);
    if (st) return st;
    const size_t n = (size_t)s->snap_T * s->plan->h.nq * sizeof(double);
    if (g2) HIP_TRY(hipMemcpy(g2, s->plan->d_g2, n, hipMemcpyDeviceToHost));
    if (se) HIP_TRY(hipMemcpy(se, s->plan->d_se, n, hipMemcpyDeviceToHost));
    if (taus) *taus = s->snap_T;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_read_sums_i64(xpcs_mt_stream* s, int64_t* G2, int64_t* IP, int64_t* IF)
{
    if (!s) return fail(XPCS_ERR_INVALID_ARG, "stream_read_sums: NULL stream");
    xpcs_status st = stream_wait(s);
    if (st) return st;
    // device rows have the maximum schedule's stride
    const size_t w = (size_t)s->snap_T * sizeof(int64_t), pitch = (size_t)s->plan->h.T * sizeof(int64_t);
    const size_t rows = (size_t)s->plan->h.P;
    if (G2) HIP_TRY(hipMemcpy2D(G2, w, s->plan->d_G2, pitch, w, rows, hipMemcpyDeviceToHost));
    if (IP) HIP_TRY(hipMemcpy2D(IP, w, s->plan->d_IP, pitch, w, rows, hipMemcpyDeviceToHost));
    if (IF) HIP_TRY(hipMemcpy2D(IF, w, s->plan->d_IF, pitch, w, rows, hipMemcpyDeviceToHost));
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_stream_close(xpcs_mt_stream* s)
{
    if (!s) return XPCS_OK;
    if (s->plan && s->plan->stream) (void)hipStreamSynchronize(s->plan->stream);
    for (void* p : {(void*)s->win, (void*)s->head, (void*)s->acc, (void*)s->pend, (void*)s->total,
                    (void*)s->tile, (void*)s->d_imm, (void*)s->d_taus_snap})
        if (p) (void)hipFree(p);
    if (s->h_imm) (void)hipHostFree(s->h_imm);
    if (s->done) (void)hipEventDestroy(s->done);
    if (s->copied) (void)hipEventDestroy(s->copied);
    if (s->plan) xpcs_mt_plan_destroy(s->plan);
    delete s;
    return XPCS_OK;
}
