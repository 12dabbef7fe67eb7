// HDF5 analysis configuration of the drop-in `corr` tool.
//
// Reads the keys Configuration::init reads (src/xpcs/configuration.cpp:80-242)
// with two deliberate differences: integers and floats are read through the
// native int64 / double memory types whatever their stored type (the reference
// reads them into 4-byte locals with the file type), and a missing key is an
// error naming the key rather than undefined behaviour.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace xpcs {

struct JobConfig {
    std::string path;                // the configuration file (results are written into it)
    std::string entry = "/xpcs";     // --entry
    bool compression = false;        // <entry>/compression == "ENABLED"
    std::string output_group;        // <entry>/output_data (or --exchange)
    std::string imm_path;            // <entry>/input_file_local (or --imm)
    int32_t width = 0, height = 0;   // /measurement/instrument/detector/{x,y}_dimension
    std::vector<int32_t> dqmap, sqmap;
    bool transposed = false;         // <entry>/transposed == "ENABLED" (maps transposed)
    int64_t data_begin = 0, data_end = 0, data_begin_todo = 0, data_end_todo = 0;
    int32_t delays_per_level = 0;
    int64_t dark_begin_todo = 0, dark_end_todo = 0;
    int32_t dark_frames = 0;         // 0 when dark_begin == dark_end or dark_end == 0
    float lld = 0.f, sigma = 0.f;    // only read when dark frames are set
    int64_t stride = 1, average = 1; // stride_frames, avg_frames
    bool normalize_by_framesum = false;
    int32_t static_window = 1;       // static_mean_window_size
    bool flatfield_enabled = false;
    std::vector<double> flatfield;   // /measurement/instrument/detector/flatfield
    float norm_factor = 1.f;         // partition_norm_factor (computed in float as the reference)
    std::string analysis_type;       // "Twotime" selects the two-time analysis
    int32_t twotime_window = 1;      // twotime2onetime_window_size (optional, default 1)
    bool twotime = false;            // analysis_type == "Twotime"
    int32_t smoothing = 0;           // smoothing_method: 1 "StaticMap", 2 "symmetric" (any case), 0 other
    bool smoothing_average = false;  // smoothing_filter == "Average" (any case)
    std::vector<int32_t> qphi_bins;  // qphi_bin_to_process

    int32_t frames() const { return (int32_t)((data_end_todo - data_begin_todo + 1) / (stride * average)); }
    int32_t real_frames() const { return (int32_t)(data_end_todo - data_begin_todo + 1); }
    int32_t frame_from() const { return (int32_t)(data_begin_todo - 1); }
};

// Loads `path`; returns false and sets *err (naming the key) on failure.
bool load_config(const std::string& path, const std::string& entry, JobConfig* cfg, std::string* err);

}  // namespace xpcs
