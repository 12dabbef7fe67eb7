// Host-side tables of a plan: tau schedule, q-partition map, normalisation
// chunks.  Pure host code; the device copies are made by the plan.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <set>

#include "internal.h"

namespace {
thread_local char g_err[512] = "";
}

extern "C" const char* xpcs_last_error(void) { return g_err; }

namespace xpcs {

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

xpcs_status fail(xpcs_status st, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return st;
}

}  // namespace xpcs

// ---- tau schedule --------------------------------------------------------------
// Corr::calculateLevelMax, src/xpcs/corr.cpp:1156-1160.  The expression is kept
// in double exactly as written there, since the truncation of the last
// log2(dpl) term decides the level count for non-power-of-two dpl.
extern "C" int32_t xpcs_level_max(int32_t frames, int32_t dpl)
{
    if (dpl <= 0) return 0;
    if (frames < dpl * 2) return 0;
    return (int32_t)(std::floor(std::log2((double)frames) - std::log2(1.0 + 1.0 / (double)dpl)) -
                     std::log2((double)dpl));
}

// Corr::calculateDelayCount, src/xpcs/corr.cpp:1162-1164.
extern "C" int32_t xpcs_delay_count(int32_t dpl, int32_t level)
{
    return level == 0 ? (2 * dpl - 1) : dpl;
}

// Corr::delaysPerLevel, src/xpcs/corr.cpp:1133-1154: level 0 has 2*dpl taus at
// step 1, level i >= 1 has dpl taus at step 2^i; a tau is kept while
// previous_tau + 2*2^i <= frames.
extern "C" xpcs_status xpcs_delays_per_level(int32_t frames, int32_t dpl, int32_t max_level,
                                             int32_t* level, int32_t* tau, int32_t cap,
                                             int32_t* count)
{
    if (dpl <= 0 || frames < 0 || max_level < 0 || cap < 0 || !count ||
        (cap > 0 && (!level || !tau)))
        return xpcs::fail(XPCS_ERR_INVALID_ARG, "xpcs_delays_per_level: bad arguments");
    int64_t n = 0;
    int64_t prev = 0;
    for (int i = 0; i <= max_level && i < 62; i++) {
        int64_t step = int64_t(1) << i;
        int per = i == 0 ? 2 * dpl : dpl;
        for (int j = 0; j < per; j++) {
            if (prev + 2 * step > frames) break;
            if (n < cap) {
                level[n] = i;
                tau[n] = (int32_t)(prev + step);
            }
            n++;
            prev += step;
        }
    }
    *count = (int32_t)n;
    return XPCS_OK;
}

namespace xpcs {

// Configuration::BuildQMap, src/xpcs/configuration.cpp:244-381.
// m_mapping is {q -> {s -> [pixels]}} (std::map, so q ascending, s ascending,
// pixels in raster order).  When one static bin appears under several dynamic
// bins, the group with the most pixels (first q on ties) is kept and the other
// groups are erased; their pixels leave the mapping (the reference pushes them
// into a discarded copy, :327-328).  Compact pixel ids follow the mapping
// order; dropped pixels follow in raster order.
xpcs_status build_qmap(const int32_t* dq, const int32_t* sq, HostTables& h)
{
    const int64_t npd = (int64_t)h.W * h.H;
    std::map<int, std::map<int, std::vector<int32_t>>> mapping;
    std::map<int, std::set<int>> s2q;
    int max_q = 0, max_s = 0;
    for (int64_t i = 0; i < npd; i++) {
        if (dq[i] < 1 || sq[i] < 1) continue;
        mapping[dq[i]][sq[i]].push_back((int32_t)i);
        s2q[sq[i]].insert(dq[i]);
        if (dq[i] > max_q) max_q = dq[i];
        if (sq[i] > max_s) max_s = sq[i];
    }
    std::vector<int32_t> dropped;
    for (auto& kv : s2q) {
        if (kv.second.size() < 2) continue;
        int s = kv.first;
        int best_q = 0;
        size_t best_n = 0;
        for (int q : kv.second) {
            size_t n = mapping[q][s].size();
            if (n > best_n) { best_n = n; best_q = q; }
        }
        for (int q : kv.second) {
            if (q == best_q) continue;
            auto& m = mapping[q];
            auto it = m.find(s);
            dropped.insert(dropped.end(), it->second.begin(), it->second.end());
            m.erase(it);
        }
    }
    std::sort(dropped.begin(), dropped.end());

    h.order.clear();
    h.group_q.clear(); h.group_s.clear(); h.group_begin.clear(); h.group_end.clear();
    for (auto& qk : mapping) {
        for (auto& sk : qk.second) {
            h.group_q.push_back(qk.first);
            h.group_s.push_back(sk.first);
            h.group_begin.push_back((int32_t)h.order.size());
            h.order.insert(h.order.end(), sk.second.begin(), sk.second.end());
            h.group_end.push_back((int32_t)h.order.size());
        }
    }
    h.n_mapped = (int)h.order.size();
    h.order.insert(h.order.end(), dropped.begin(), dropped.end());
    h.P = (int)h.order.size();
    h.ngroups = (int)h.group_q.size();
    h.nq = max_q;
    h.ns = max_s;
    h.pix2c.assign((size_t)npd, -1);
    for (int c = 0; c < h.P; c++) h.pix2c[(size_t)h.order[c]] = c;
    h.sbin_of_c.resize((size_t)h.P);
    for (int c = 0; c < h.P; c++) h.sbin_of_c[(size_t)c] = sq[h.order[c]];
    h.pixels_per_bin.assign((size_t)max_s, 0);
    for (int64_t i = 0; i < npd; i++)
        if (dq[i] >= 1 && sq[i] >= 1) h.pixels_per_bin[(size_t)sq[i] - 1]++;
    h.q_first_group.assign((size_t)max_q + 1, h.ngroups);
    for (int g = h.ngroups - 1; g >= 0; g--) h.q_first_group[(size_t)h.group_q[g] - 1] = g;
    for (int q = max_q - 1; q >= 0; q--)
        if (h.q_first_group[(size_t)q] > h.q_first_group[(size_t)q + 1])
            h.q_first_group[(size_t)q] = h.q_first_group[(size_t)q + 1];
    return XPCS_OK;
}

xpcs_status build_tables(const xpcs_mt_desc& d, HostTables& h)
{
    if (d.width <= 0 || d.height <= 0 || d.frames <= 0 || d.delays_per_level <= 0 ||
        d.stride <= 0 || d.average <= 0 || d.frame_from < 0 || d.dark_frames < 0 || !d.dqmap ||
        !d.sqmap)
        return fail(XPCS_ERR_INVALID_ARG, "plan: invalid description (sizes, stride/average or maps)");
    if ((int64_t)d.width * d.height >= (int64_t(1) << 31))
        return fail(XPCS_ERR_INVALID_ARG, "plan: detector larger than 2^31 pixels");
    h.W = d.width; h.H = d.height; h.F = d.frames; h.dpl = d.delays_per_level;
    h.stride = d.stride; h.avg = d.average; h.frame_from = d.frame_from; h.darks = d.dark_frames;
    h.lld = d.lld; h.sigma = d.sigma; h.flags = d.flags;
    if (h.avg > 64)
        return fail(XPCS_ERR_UNSUPPORTED, "plan: avg_frames > 64 exceeds the 32-bit event value range");

    // schedule
    h.max_level = xpcs_level_max(h.F, h.dpl);
    int32_t T = 0;
    xpcs_delays_per_level(h.F, h.dpl, h.max_level, nullptr, nullptr, 0, &T);
    if (T <= 0) return fail(XPCS_ERR_INVALID_ARG, "plan: empty tau schedule (frames < 2)");
    if (T > kMaxTaus) return fail(XPCS_ERR_UNSUPPORTED, "plan: %d taus exceed %d", T, kMaxTaus);
    h.T = T;
    h.level.resize((size_t)T);
    h.tau.resize((size_t)T);
    xpcs_delays_per_level(h.F, h.dpl, h.max_level, h.level.data(), h.tau.data(), T, &T);
    h.taus.resize((size_t)T);
    h.levels.clear();
    for (int i = 0; i < T; i++) {
        TauParam& tp = h.taus[(size_t)i];
        tp.level = h.level[(size_t)i];
        tp.tp = h.tau[(size_t)i] >> tp.level;
        tp.lf = h.F >> tp.level;
        tp.N = tp.lf - tp.tp;
        if (h.levels.empty() || h.levels.back().shift != tp.level) {
            LevelParam lp;
            lp.first = i; lp.count = 0; lp.lo = tp.tp; lp.hi = tp.tp; lp.lf = tp.lf;
            lp.shift = tp.level;
            h.levels.push_back(lp);
        }
        LevelParam& lp = h.levels.back();
        lp.count++;
        lp.hi = tp.tp;
        if (lp.hi - lp.lo + 1 != lp.count)
            return fail(XPCS_ERR_UNSUPPORTED, "plan: non-contiguous lags in level %d", tp.level);
    }
    for (size_t k = 0; k < h.levels.size(); k++)
        if (h.levels[k].shift != (int)k)
            return fail(XPCS_ERR_UNSUPPORTED, "plan: schedule skips a level");

    xpcs_status st = build_qmap(d.dqmap, d.sqmap, h);
    if (st != XPCS_OK) return st;
    if (h.P == 0) return fail(XPCS_ERR_INVALID_ARG, "plan: no masked pixels (dqmap>=1 && sqmap>=1)");

    if (d.flatfield) {
        h.ff_det.assign(d.flatfield, d.flatfield + (size_t)h.W * h.H);
        h.weight.resize((size_t)h.P);
        for (int c = 0; c < h.P; c++) h.weight[(size_t)c] = h.ff_det[(size_t)h.order[c]];
    }

    // normalisation chunks: at most 256 pixels, never across a group
    const int CH = 256;
    h.chunks.clear();
    h.group_chunk_first.assign((size_t)h.ngroups + 1, 0);
    for (int g = 0; g < h.ngroups; g++) {
        h.group_chunk_first[(size_t)g] = (int32_t)h.chunks.size();
        for (int b = h.group_begin[(size_t)g]; b < h.group_end[(size_t)g]; b += CH) {
            NormChunk c;
            c.group = g; c.begin = b;
            c.end = std::min(b + CH, (int)h.group_end[(size_t)g]);
            c.q = h.group_q[(size_t)g];
            h.chunks.push_back(c);
        }
    }
    h.group_chunk_first[(size_t)h.ngroups] = (int32_t)h.chunks.size();
    h.q_chunk_first.assign((size_t)h.nq + 1, 0);
    for (int q = 0; q <= h.nq; q++) {
        int g = h.q_first_group[(size_t)q];
        h.q_chunk_first[(size_t)q] = h.group_chunk_first[(size_t)g];
    }
    return XPCS_OK;
}

}  // namespace xpcs

extern "C" xpcs_status xpcs_build_qmap(int32_t width, int32_t height, const int32_t* dqmap,
                                       const int32_t* sqmap, int32_t* order, int32_t* group_q,
                                       int32_t* group_s, int32_t* group_begin, int32_t* group_end,
                                       int32_t group_cap, int32_t* counts)
{
    if (width <= 0 || height <= 0 || !dqmap || !sqmap || !counts)
        return xpcs::fail(XPCS_ERR_INVALID_ARG, "build_qmap: bad arguments");
    xpcs::HostTables h;
    h.W = width;
    h.H = height;
    xpcs_status st = xpcs::build_qmap(dqmap, sqmap, h);
    if (st) return st;
    counts[0] = h.P;
    counts[1] = h.n_mapped;
    counts[2] = h.ngroups;
    counts[3] = h.nq;
    counts[4] = h.ns;
    if (order) std::copy(h.order.begin(), h.order.end(), order);
    const int ng = std::min(h.ngroups, group_cap);
    for (int g = 0; g < ng; g++) {
        if (group_q) group_q[g] = h.group_q[(size_t)g];
        if (group_s) group_s[g] = h.group_s[(size_t)g];
        if (group_begin) group_begin[g] = h.group_begin[(size_t)g];
        if (group_end) group_end[g] = h.group_end[(size_t)g];
    }
    return XPCS_OK;
}
