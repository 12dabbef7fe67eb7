// Launch-shape tuning, the occupancy-keyed autotuner and batched execution of
// many plans as one captured hipGraph.  None of this has a counterpart in the
// reference, which runs one job per process with one OpenMP loop
// (src/xpcs/corr.cpp:329); the per-job work is the plain path of api.cpp.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.h"

using namespace xpcs;

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(XPCS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

bool tuning_valid(const xpcs_mt_plan* p, const xpcs_mt_tuning& t)
{
    if (t.tr_bucket_shift != 0 && (t.tr_bucket_shift < 8 || t.tr_bucket_shift > 12)) return false;
    if (t.tr_chunk_frames < 0 || t.reg_blocks < 0 || t.reg_blocks > 65535) return false;
    if (t.lds_blocks < 0 || t.lds_blocks > 65535) return false;
    if (t.dense_levels != 0 && !dense_levels_allowed(p->h.dpl, t.dense_levels)) return false;
    if (t.deep_frames > 0 && (t.deep_frames < 16 || t.deep_frames > 96)) return false;
    if (t.sums64 != 0 && t.sums64 != 1) return false;
    return true;
}

// Occupancy key: ingest mode, dpl, events per pixel and frame (half octaves),
// log2 pixels, log2 frames.
using TuneKey = std::tuple<int, int, int, int, int>;

TuneKey key_of(const xpcs_mt_plan* p)
{
    const HostTables& h = p->h;
    int occ = 99;   // dense frames: every pixel of every frame
    if (p->mode == 1) {
        const double cells = (double)h.W * (double)h.H * (double)std::max<int64_t>(p->nsel, 1);
        const double x = (double)p->raw_events / cells;
        occ = x > 0 ? (int)std::floor(2.0 * std::log2(x)) : -99;
    } else if (p->mode == 3) {
        const double x = (double)p->events / ((double)std::max(h.P, 1) * (double)std::max(h.F, 1));
        occ = x > 0 ? (int)std::floor(2.0 * std::log2(x)) : -99;
    }
    return TuneKey{p->mode, h.dpl, occ, (int)std::floor(std::log2((double)std::max(h.P, 1))),
                   (int)std::floor(std::log2((double)std::max(h.F, 1)))};
}

std::mutex g_tune_mu;
std::map<TuneKey, xpcs_mt_tuning> g_tune_cache;

// Best of two runs of the whole job on s, in milliseconds (negative on error).
double time_job(xpcs_mt_plan* p, hipStream_t s, hipEvent_t a, hipEvent_t b)
{
    if (prepare_execute(p) != XPCS_OK) return -1.0;
    double best = 1e30;
    for (int r = 0; r < 2; r++) {
        if (hipEventRecord(a, s) != hipSuccess) return -1.0;
        enqueue_execute(p, s);
        if (hipGetLastError() != hipSuccess || hipEventRecord(b, s) != hipSuccess ||
            hipEventSynchronize(b) != hipSuccess)
            return -1.0;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.0;
        best = std::min(best, (double)ms);
    }
    return best;
}

}  // namespace

extern "C" xpcs_status xpcs_mt_set_tuning(xpcs_mt_plan* p, const xpcs_mt_tuning* t)
{
    if (!p || !t) return fail(XPCS_ERR_INVALID_ARG, "set_tuning: NULL argument");
    if (!tuning_valid(p, *t)) return fail(XPCS_ERR_INVALID_ARG, "set_tuning: value out of range");
    p->tune = *t;
    p->generation++;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_get_tuning(xpcs_mt_plan* p, xpcs_mt_tuning* t)
{
    if (!p || !t) return fail(XPCS_ERR_INVALID_ARG, "get_tuning: NULL argument");
    *t = xpcs_mt_tuning{};
    const TransposeShape sh = transpose_shape(std::max<int64_t>(p->nrows, 1), std::max<int64_t>(p->nsel, 1),
                                              p->tune.tr_bucket_shift, p->tune.tr_chunk_frames);
    t->tr_bucket_shift = sh.bshift;
    t->tr_chunk_frames = (int32_t)sh.cf;
    t->reg_blocks = reg_blocks(p);
    t->lds_blocks = lds_blocks(p);
    t->dense_levels = (p->h.dpl == 4 || p->h.dpl == 8) ? dense_levels(p) : 0;
    t->deep_frames = p->tune.deep_frames < 0 ? -1 : (p->tune.deep_frames > 0 ? p->tune.deep_frames : 96);
    t->sums64 = p->tune.sums64;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_autotune(xpcs_mt_plan* p, xpcs_mt_tuning* best_out)
{
    if (!p) return fail(XPCS_ERR_INVALID_ARG, "autotune: NULL plan");
    if (p->mode == 0) return fail(XPCS_ERR_STATE, "autotune: nothing ingested");
    const TuneKey key = key_of(p);
    // the sums layout is the caller's choice, not a launch shape: kept as set
    const int32_t sums64 = p->tune.sums64;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            p->tune = it->second;
            p->tune.sums64 = sums64;
            p->generation++;
            if (best_out) *best_out = p->tune;
            return XPCS_OK;
        }
    }
    hipStream_t s = p->stream;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(XPCS_ERR_HIP, "autotune: hipEventCreate failed");
    }
    const bool was = p->capturing;
    p->capturing = true;   // no timer events while tuning
    xpcs_mt_tuning cur{};
    cur.sums64 = sums64;
    p->tune = cur;
    double cur_ms = time_job(p, s, a, b);
    xpcs_status st = XPCS_OK;
    if (cur_ms < 0) st = fail(XPCS_ERR_HIP, "autotune: job failed");
    // coordinate descent over the shapes the job's kernels use
    const bool sparse_rows = p->mode == 1 || p->mode == 3;
    const bool dense_stream = p->mode == 2 && (p->h.dpl == 4 || p->h.dpl == 8);
    std::vector<std::pair<int32_t xpcs_mt_tuning::*, std::vector<int32_t>>> axes;
    if (sparse_rows) {
        if (p->mode == 1) {
            const TransposeShape sh = transpose_shape(std::max<int64_t>(p->nrows, 1), std::max<int64_t>(p->nsel, 1), 0, 0);
            const int32_t cf = (int32_t)sh.cf;
            axes.push_back({&xpcs_mt_tuning::tr_bucket_shift, {8, 9, 10}});
            axes.push_back({&xpcs_mt_tuning::tr_chunk_frames, {std::max(1, cf / 2), cf * 2, cf * 4}});
        }
        axes.push_back({&xpcs_mt_tuning::reg_blocks, {1024, 4096}});
        axes.push_back({&xpcs_mt_tuning::lds_blocks, {256, 2048, 4096, 8192}});
        // deep-level hand-over: none, or from the first level with <= 48 frames
        // (the default is <= 96)
        int dstride = 0;
        if (deep_level(p, &dstride) < (int)p->h.levels.size())
            axes.push_back({&xpcs_mt_tuning::deep_frames, {-1, 48}});
    } else if (dense_stream) {
        axes.push_back({&xpcs_mt_tuning::dense_levels,
                        p->h.dpl == 4 ? std::vector<int32_t>{3, 4, 5} : std::vector<int32_t>{2, 3}});
    }
    for (auto& ax : axes) {
        if (st) break;
        for (int32_t v : ax.second) {
            xpcs_mt_tuning t = cur;
            t.*(ax.first) = v;
            if (!tuning_valid(p, t)) continue;
            p->tune = t;
            const double ms = time_job(p, s, a, b);
            if (ms < 0) {
                st = fail(XPCS_ERR_HIP, "autotune: job failed");
                break;
            }
            if (ms < 0.97 * cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
    }
    p->tune = cur;
    p->capturing = was;
    p->generation++;
    // leave the results of the chosen shapes in place
    if (!st) {
        st = prepare_execute(p);
        if (!st) {
            enqueue_execute(p, s);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
                st = fail(XPCS_ERR_HIP, "autotune: final run failed");
            p->stream_last = s;
            p->executed = true;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st) return st;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = cur;
    }
    if (best_out) *best_out = cur;
    return XPCS_OK;
}

struct xpcs_mt_batch {
    std::vector<xpcs_mt_plan*> plans;
    std::vector<uint64_t> gens;          // plan generations at capture
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    hipStream_t own = nullptr;
    hipStream_t last = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void batch_release_graph(xpcs_mt_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Work estimate for the longest-first assignment of jobs to branches.
double job_work(const xpcs_mt_plan* p)
{
    if (p->mode == 1) return (double)p->raw_events + 8.0 * p->h.P;
    if (p->mode == 3) return (double)p->events + 8.0 * p->h.P;
    return (double)p->h.P * (double)p->h.F;
}

xpcs_status batch_capture(xpcs_mt_batch* b, hipStream_t s, uint32_t flags)
{
    batch_release_graph(b);
    for (xpcs_mt_plan* p : b->plans) {
        if (flags & XPCS_BATCH_AUTOTUNE) {
            xpcs_status st = xpcs_mt_autotune(p, nullptr);
            if (st) return st;
        }
        xpcs_status st = prepare_execute(p);
        if (st) return st;
    }
    // longest job first onto the least loaded branch
    const size_t nb = b->branch.size();
    std::vector<size_t> order(b->plans.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return job_work(b->plans[x]) > job_work(b->plans[y]); });
    std::vector<double> load(nb, 0.0);
    std::vector<size_t> where(b->plans.size(), 0);
    for (size_t i : order) {
        const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        where[i] = k;
        load[k] += job_work(b->plans[i]);
    }
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipEventRecord(b->fork, s);
    for (size_t k = 0; k < nb && e == hipSuccess; k++) e = hipStreamWaitEvent(b->branch[k], b->fork, 0);
    for (size_t i : order) {
        if (e != hipSuccess) break;
        xpcs_mt_plan* p = b->plans[i];
        p->capturing = true;
        enqueue_execute(p, b->branch[where[i]]);
        p->capturing = false;
        e = hipGetLastError();
    }
    for (size_t k = 0; k < nb && e == hipSuccess; k++) {
        e = hipEventRecord(b->join[k], b->branch[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, b->join[k], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(XPCS_ERR_HIP, "batch: graph capture failed: %s",
                    hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->gens.resize(b->plans.size());
    for (size_t i = 0; i < b->plans.size(); i++) b->gens[i] = b->plans[i]->generation;
    return XPCS_OK;
}

}  // namespace

extern "C" xpcs_status xpcs_mt_batch_create(int32_t branches, xpcs_mt_batch** out)
{
    if (!out || branches < 1 || branches > 64)
        return fail(XPCS_ERR_INVALID_ARG, "batch_create: branches must be in [1, 64]");
    *out = nullptr;
    auto* b = new xpcs_mt_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int32_t k = 0; k < branches && e == hipSuccess; k++) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) b->branch.push_back(st);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->join.push_back(ev);
    }
    if (e != hipSuccess) {
        xpcs_mt_batch_destroy(b);
        return fail(XPCS_ERR_HIP, "batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_batch_add(xpcs_mt_batch* b, xpcs_mt_plan* p, int32_t* index)
{
    if (!b || !p) return fail(XPCS_ERR_INVALID_ARG, "batch_add: NULL argument");
    if (std::find(b->plans.begin(), b->plans.end(), p) != b->plans.end())
        return fail(XPCS_ERR_INVALID_ARG, "batch_add: plan already in the batch");
    b->plans.push_back(p);
    batch_release_graph(b);
    if (index) *index = (int32_t)b->plans.size() - 1;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_batch_execute(xpcs_mt_batch* b, void* stream, uint32_t flags)
{
    if (!b) return fail(XPCS_ERR_INVALID_ARG, "batch_execute: NULL batch");
    if (b->plans.empty()) return fail(XPCS_ERR_STATE, "batch_execute: no plans");
    for (xpcs_mt_plan* p : b->plans)
        if (p->mode == 0) return fail(XPCS_ERR_STATE, "batch_execute: a plan has nothing ingested");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->own;
    bool stale = !b->exec || b->gens.size() != b->plans.size();
    for (size_t i = 0; !stale && i < b->plans.size(); i++) stale = b->gens[i] != b->plans[i]->generation;
    if (stale || (flags & XPCS_BATCH_AUTOTUNE)) {
        xpcs_status st = batch_capture(b, s, flags);
        if (st) return st;
    }
    HIP_TRY(hipGraphLaunch(b->exec, s));
    for (xpcs_mt_plan* p : b->plans) {
        p->stream_last = s;
        p->executed = true;
    }
    b->last = s;
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_batch_synchronize(xpcs_mt_batch* b)
{
    if (!b) return fail(XPCS_ERR_INVALID_ARG, "batch_synchronize: NULL batch");
    if (b->last) HIP_TRY(hipStreamSynchronize(b->last));
    for (xpcs_mt_plan* p : b->plans) {
        xpcs_status st = xpcs_mt_synchronize(p);
        if (st) return st;
    }
    return XPCS_OK;
}

extern "C" xpcs_status xpcs_mt_batch_destroy(xpcs_mt_batch* b)
{
    if (!b) return XPCS_OK;
    if (b->last) (void)hipStreamSynchronize(b->last);
    batch_release_graph(b);
    for (hipStream_t st : b->branch) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
    return XPCS_OK;
}
