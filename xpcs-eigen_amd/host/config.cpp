// HDF5 configuration reader of the drop-in `corr` tool (key set of
// Configuration::init, src/xpcs/configuration.cpp:80-242).
#include "config.h"

#include <algorithm>
#include <cctype>
#include <cstddef>
#include <cstring>

#include "h5lite.h"

namespace xpcs {

namespace {

struct Reader {
    h5::File& f;
    std::string* err;
    bool ok = true;

    bool need(const std::string& key)
    {
        if (!ok) return false;
        if (!f.exists(key)) {
            *err = "configuration key missing: " + key;
            ok = false;
        }
        return ok;
    }
    int64_t i64(const std::string& key)
    {
        int64_t v = 0;
        if (need(key) && !f.read_i64(key, &v)) {
            *err = "configuration key is not numeric: " + key;
            ok = false;
        }
        return v;
    }
    double f64(const std::string& key)
    {
        double v = 0.0;
        if (need(key) && !f.read_f64(key, &v)) {
            *err = "configuration key is not numeric: " + key;
            ok = false;
        }
        return v;
    }
    std::string str(const std::string& key)
    {
        std::string v;
        if (need(key) && !f.read_string(key, &v)) {
            *err = "configuration key is not a string: " + key;
            ok = false;
        }
        return v;
    }
    std::string str_opt(const std::string& key)
    {
        std::string v;
        if (ok && f.exists(key)) f.read_string(key, &v);
        return v;
    }
};

}  // namespace

static bool iequal(const std::string& a, const char* b)
{
    const size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; i++)
        if (std::toupper((unsigned char)a[i]) != std::toupper((unsigned char)b[i])) return false;
    return true;
}

bool load_config(const std::string& path, const std::string& entry, JobConfig* c, std::string* err)
{
    h5::File f;
    if (!f.open(path, false, err)) return false;
    Reader r{f, err};
    const std::string E = entry;
    const std::string det = "/measurement/instrument/detector/";
    c->path = path;
    c->entry = entry;
    c->compression = r.str(E + "/compression") == "ENABLED";
    c->output_group = r.str(E + "/output_data");
    c->width = (int32_t)r.i64(det + "x_dimension");
    c->height = (int32_t)r.i64(det + "y_dimension");
    if (r.need(E + "/dqmap") && r.need(E + "/sqmap")) {
        std::vector<h5::hsize_t> d1, d2;
        if (!f.read_i32(E + "/dqmap", &c->dqmap, &d1) || !f.read_i32(E + "/sqmap", &c->sqmap, &d2)) {
            *err = "cannot read dqmap/sqmap";
            return false;
        }
    }
    if (!r.ok) return false;
    const size_t npd = (size_t)c->width * (size_t)c->height;
    if (c->dqmap.size() != npd || c->sqmap.size() != npd) {
        *err = "dqmap/sqmap do not hold x_dimension*y_dimension values";
        return false;
    }
    // Optional transpose of the q-maps (configuration.cpp:100-134): element i
    // moves to (i % xdim) * ydim + i / xdim, then the dimensions swap.
    c->transposed = r.str_opt(E + "/transposed") == "ENABLED";
    if (c->transposed) {
        std::vector<int32_t> nd(npd), ns(npd);
        for (size_t i = 0; i < npd; i++) {
            const size_t row = i % (size_t)c->width, col = i / (size_t)c->width;
            nd[row * (size_t)c->height + col] = c->dqmap[i];
            ns[row * (size_t)c->height + col] = c->sqmap[i];
        }
        c->dqmap.swap(nd);
        c->sqmap.swap(ns);
        std::swap(c->width, c->height);
    }
    c->data_begin = r.i64(E + "/data_begin");
    c->data_end = r.i64(E + "/data_end");
    c->data_begin_todo = r.i64(E + "/data_begin_todo");
    c->data_end_todo = r.i64(E + "/data_end_todo");
    c->delays_per_level = (int32_t)r.i64(E + "/delays_per_level");
    c->dark_begin_todo = r.i64(E + "/dark_begin_todo");
    c->dark_end_todo = r.i64(E + "/dark_end_todo");
    c->stride = r.i64(E + "/stride_frames");
    c->average = r.i64(E + "/avg_frames");
    if (f.exists(E + "/normalize_by_framesum")) c->normalize_by_framesum = r.i64(E + "/normalize_by_framesum") != 0;
    if (!(c->dark_begin_todo == c->dark_end_todo || c->dark_end_todo == 0)) {
        c->dark_frames = (int32_t)(c->dark_end_todo - c->dark_begin_todo + 1);
        c->lld = (float)r.f64(E + "/lld");
        c->sigma = (float)r.f64(E + "/sigma");
    }
    // Norm factor in float, operation by operation (configuration.cpp:184-199).
    const float dpx = (float)r.f64(det + "x_pixel_size");
    const float dpy = (float)r.f64(det + "y_pixel_size");
    const float adu = (float)r.f64(det + "adu_per_photon");
    const float preset = (float)r.f64(det + "exposure_time");
    const float eff = (float)r.f64(det + "efficiency");
    const float dist = (float)r.f64(det + "distance");
    const float flux = (float)r.f64("/measurement/instrument/source_begin/beam_intensity_transmitted");
    const float thick = (float)r.f64("/measurement/sample/thickness");
    float nf = 1.0f;
    nf = nf / eff / adu / preset;
    nf = nf / (dpx / dist * dpy / dist);
    nf /= flux;
    nf /= thick;
    c->norm_factor = nf;
    c->static_window = (int32_t)r.i64(E + "/static_mean_window_size");
    c->flatfield_enabled = r.str(E + "/flatfield_enabled") == "ENABLED";
    if (c->flatfield_enabled && r.need(det + "flatfield")) {
        std::vector<h5::hsize_t> d;
        if (!f.read_f64v(det + "flatfield", &c->flatfield, &d) || c->flatfield.size() != npd) {
            *err = "flatfield does not hold x_dimension*y_dimension values";
            return false;
        }
    }
    if (f.exists(E + "/twotime2onetime_window_size"))
        c->twotime_window = (int32_t)r.i64(E + "/twotime2onetime_window_size");
    c->analysis_type = r.str_opt(E + "/analysis_type");
    // Two-time keys (configuration.cpp:216-232); string matches ignore case
    // (Configuration::SmoothingMethod/SmoothingFilter, :784-806).
    c->twotime = c->analysis_type == "Twotime";
    if (c->twotime) {
        const std::string m = r.str(E + "/smoothing_method"), flt = r.str(E + "/smoothing_filter");
        c->smoothing = iequal(m, "StaticMap") ? 1 : iequal(m, "symmetric") ? 2 : 0;
        c->smoothing_average = iequal(flt, "Average");
        if (r.need(E + "/qphi_bin_to_process")) {
            std::vector<h5::hsize_t> d;
            std::vector<int32_t> v;
            if (!f.read_i32(E + "/qphi_bin_to_process", &v, &d)) {
                *err = "cannot read qphi_bin_to_process";
                return false;
            }
            // the reference takes the first dimension's count of entries
            const size_t n = std::min(v.size(), d.empty() ? (size_t)1 : (size_t)d[0]);
            c->qphi_bins.assign(v.begin(), v.begin() + (ptrdiff_t)n);
        }
    }
    c->imm_path = r.str(E + "/input_file_local");
    if (!r.ok) return false;
    if (c->delays_per_level < 1 || c->stride < 1 || c->average < 1 || c->data_begin_todo < 1 ||
        c->data_end_todo < c->data_begin_todo) {
        *err = "invalid frame range, stride/average or delays_per_level in the configuration";
        return false;
    }
    return true;
}

}  // namespace xpcs
