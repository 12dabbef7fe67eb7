/*
 * C++ facade with the names and argument meanings of the reference's in-process
 * API, over the C ABI of xpcs_mt.h (header only; link -lxpcs_mt):
 *
 *   xpcs::Corr                          src/xpcs/corr.h:66-126
 *   xpcs::data_structure::Row           src/xpcs/data_structure/row.h:55-77
 *   xpcs::data_structure::SparseData    src/xpcs/data_structure/sparse_data.h:57-79
 *
 * The reference's Corr reads its parameters from the Configuration singleton
 * (src/xpcs/corr.cpp:317-327, :932-939); here Corr::configure() takes them once
 * as an xpcs_mt_desc (the q-maps and flat field it points to must outlive the
 * calls). Errors the reference never reports (it has none) become
 * std::runtime_error carrying xpcs_last_error(); nothing throws across the C ABI.
 */
#ifndef XPCS_CORR_HPP
#define XPCS_CORR_HPP

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

extern "C" {
#include "xpcs_mt.h"
}

namespace xpcs {

namespace detail {
inline void check(xpcs_status st, const char* what)
{
    if (st != XPCS_OK) throw std::runtime_error(std::string(what) + ": " + xpcs_last_error());
}
}  // namespace detail

namespace data_structure {

// One pixel's events: ascending frame indices and their values (row.h:55-77).
class Row {
public:
    Row() {}
    explicit Row(int size)
    {
        indxPtr.reserve((size_t)size);
        valPtr.reserve((size_t)size);
    }
    std::vector<int> indxPtr;
    std::vector<float> valPtr;
};

// Pixel-major store of rows, one per detector pixel (sparse_data.h:57-79).
// Rows are created on first use (the reference pre-reserves one per pixel);
// Pixel(i) marks pixel i valid, ValidPixels() lists them in ascending order.
class SparseData {
public:
    explicit SparseData(int rows, int initialSize = 10) : m_rows(rows), m_initSize(initialSize), m_data((size_t)rows)
    {
    }

    Row* Pixel(int index)
    {
        std::unique_ptr<Row>& r = m_data.at((size_t)index);
        if (!r) r.reset(new Row(m_initSize));
        return r.get();
    }

    bool Exists(int index) const { return index >= 0 && index < m_rows && m_data[(size_t)index] != nullptr; }

    std::vector<int>& ValidPixels()
    {
        m_validPixels.clear();
        for (int i = 0; i < m_rows; i++)
            if (m_data[(size_t)i]) m_validPixels.push_back(i);
        return m_validPixels;
    }

    int Rows() const { return m_rows; }

private:
    int m_rows;
    int m_initSize;
    std::vector<std::unique_ptr<Row>> m_data;
    std::vector<int> m_validPixels;
};

}  // namespace data_structure

class Corr {
public:
    // Tau schedule (corr.cpp:1133-1164).
    static int calculateDelayCount(int dpl, int level) { return xpcs_delay_count(dpl, level); }

    static int calculateLevelMax(int frameCount, int dpl) { return xpcs_level_max(frameCount, dpl); }

    static std::vector<std::tuple<int, int>> delaysPerLevel(int frameCount, int dpl, int maxDelay)
    {
        int32_t n = 0;
        detail::check(xpcs_delays_per_level(frameCount, dpl, maxDelay, nullptr, nullptr, 0, &n), "delaysPerLevel");
        std::vector<int32_t> level((size_t)n), tau((size_t)n);
        detail::check(xpcs_delays_per_level(frameCount, dpl, maxDelay, level.data(), tau.data(), n, &n),
                      "delaysPerLevel");
        std::vector<std::tuple<int, int>> out;
        for (int32_t i = 0; i < n; i++) out.emplace_back(level[(size_t)i], tau[(size_t)i]);
        return out;
    }

    // Stands in for the Configuration singleton: width, height, frames,
    // delays per level, stride, average, q-maps and flat field of the job.
    static void configure(const xpcs_mt_desc& desc)
    {
        state().reset();
        state().desc = desc;
        state().configured = true;
    }

    // Corr::multiTau2 (corr.cpp:315-431): G2, IP and IF of every pixel of
    // `data` for every tau, ADDED into the caller's float arrays laid out
    // [tau][W*H] (the caller zeroes them first, as main.cpp:193-201 does).
    // Row values are the filter's values, sum(count * flatfield) / average;
    // they are turned back into integer counts (an error if one is not). The
    // input is not modified (the reference bins it in place).
    static void multiTau2(data_structure::SparseData* data, float* G2, float* IP, float* IF)
    {
        State& st = state();
        if (!st.configured) throw std::runtime_error("multiTau2: Corr::configure() not called");
        const xpcs_mt_desc& d = st.desc;
        const int npix = d.width * d.height;
        if (!data || data->Rows() != npix) throw std::runtime_error("multiTau2: SparseData rows != width*height");
        const double avg = d.average > 0 ? d.average : 1;
        std::vector<int64_t> off((size_t)npix + 1, 0);
        std::vector<int32_t> frame;
        std::vector<uint32_t> count;
        for (int i = 0; i < npix; i++) {
            if (data->Exists(i)) {
                const data_structure::Row* r = data->Pixel(i);
                if (r->indxPtr.size() != r->valPtr.size())
                    throw std::runtime_error("multiTau2: row with unequal index and value counts");
                const double ff = d.flatfield ? (double)d.flatfield[i] : 1.0;
                for (size_t k = 0; k < r->indxPtr.size(); k++) {
                    const double c = (double)r->valPtr[k] * avg / ff;
                    const double cr = std::nearbyint(c);
                    if (!(cr >= 0.0) || cr > 4294967295.0 || std::fabs(c - cr) > 1e-3 * std::fmax(1.0, cr))
                        throw std::runtime_error("multiTau2: pixel " + std::to_string(i) +
                                                 " holds a value that is not count * flatfield / average");
                    frame.push_back(r->indxPtr[k]);
                    count.push_back((uint32_t)cr);
                }
            }
            off[(size_t)i + 1] = (int64_t)frame.size();
        }
        st.reset();
        detail::check(xpcs_mt_plan_create(&d, &st.plan), "multiTau2");
        detail::check(xpcs_mt_ingest_rows(st.plan, off.data(), frame.data(), count.data()), "multiTau2");
        detail::check(xpcs_mt_execute(st.plan, nullptr), "multiTau2");
        xpcs_mt_info info;
        detail::check(xpcs_mt_plan_info(st.plan, &info), "multiTau2");
        const size_t n = (size_t)info.taus * (size_t)npix;
        std::vector<float> g((size_t)n), p((size_t)n), f((size_t)n);
        detail::check(xpcs_mt_read_ref_f32(st.plan, g.data(), p.data(), f.data()), "multiTau2");
        for (size_t k = 0; k < n; k++) {
            G2[k] += g[k];
            IP[k] += p[k];
            IF[k] += f[k];
        }
    }

    // Corr::normalizeG2s(G2, IP, IF) (corr.h:106-108, corr.cpp:927-1091) of
    // the caller's arrays, as main.cpp:441-445 calls it on the accumulated
    // G2s/IPs/IFs: each a pixels x taus matrix stored column-major, i.e.
    // [tau][W*H] floats (rows = W*H, cols = taus).  The reference writes the
    // result to HDF5 as norm-0-g2/norm-0-stderr; here g2 and stderr_ receive
    // it, [tau][q] doubles (taus x lastShape's qbins).  Uses the configured
    // q-maps and schedule; the arrays are whatever the caller holds (one
    // multiTau2, several accumulated, or edited), as in the reference.
    static void normalizeG2s(const float* G2, const float* IP, const float* IF, int rows, int cols, double* g2,
                             double* stderr_)
    {
        State& st = state();
        if (!st.configured) throw std::runtime_error("normalizeG2s: Corr::configure() not called");
        if (!st.plan) detail::check(xpcs_mt_plan_create(&st.desc, &st.plan), "normalizeG2s");
        detail::check(xpcs_mt_normalize_ref_f32(st.plan, G2, IP, IF, rows, cols, g2, stderr_), "normalizeG2s");
    }

    // Taus and dynamic q-bins of the configured job (the shapes of
    // normalizeG2s's outputs).
    static void lastShape(int* taus, int* qbins)
    {
        State& st = state();
        if (!st.configured) throw std::runtime_error("lastShape: Corr::configure() not called");
        if (!st.plan) detail::check(xpcs_mt_plan_create(&st.desc, &st.plan), "lastShape");
        xpcs_mt_info info;
        detail::check(xpcs_mt_plan_info(st.plan, &info), "lastShape");
        *taus = info.taus;
        *qbins = info.dynamic_bins;
    }

    // Frees the device state of the last multiTau2.
    static void release() { state().reset(); }

private:
    struct State {
        xpcs_mt_desc desc{};
        bool configured = false;
        xpcs_mt_plan* plan = nullptr;
        void reset()
        {
            if (plan) xpcs_mt_plan_destroy(plan);
            plan = nullptr;
        }
        ~State() { reset(); }
    };
    static State& state()
    {
        static State s;
        return s;
    }
};

}  // namespace xpcs

#endif
