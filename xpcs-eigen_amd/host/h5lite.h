// Minimal HDF5 binding for the drop-in host layer.
//
// libhdf5 (C API, 1.10 ABI) is loaded at run time with dlopen, so the rest of
// the library builds and runs without it; only the functions below are bound.
// The library is looked up in $XPCS_HDF5_LIB, then libhdf5.so.103 /
// libhdf5.so on the loader path, then /opt/conda/lib.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace xpcs {
namespace h5 {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;

// Loads libhdf5 once; returns false (with a message) if it is unavailable.
bool available(std::string* why = nullptr);

// A file opened read-write (created when `create`), closed on destruction.
class File {
public:
    File() = default;
    ~File();
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    bool open(const std::string& path, bool create, std::string* err);
    void close();
    bool exists(const std::string& path) const;
    // Shape of a dataset (empty for a scalar); false if missing.
    bool dims(const std::string& path, std::vector<hsize_t>* d) const;
    // Stored element type: HDF5 class (0 integer, 1 float, 3 string), size in
    // bytes, and for integers the sign (0 unsigned, 1 signed; -1 unknown).
    bool type_of(const std::string& path, int* cls, size_t* size, int* sign) const;
    // Scalar readers converting from any stored numeric type (first element).
    bool read_i64(const std::string& path, int64_t* v) const;
    bool read_f64(const std::string& path, double* v) const;
    bool read_string(const std::string& path, std::string* v) const;
    // Whole-dataset readers converting to the requested native type.
    bool read_i32(const std::string& path, std::vector<int32_t>* v, std::vector<hsize_t>* d) const;
    bool read_f64v(const std::string& path, std::vector<double>* v, std::vector<hsize_t>* d) const;
    bool read_f32v(const std::string& path, std::vector<float>* v, std::vector<hsize_t>* d) const;
    // Unsigned 16- or 32-bit dataset (by its stored size) into v16 or v32.
    bool read_frames(const std::string& path, std::vector<uint16_t>* v16, std::vector<uint32_t>* v32,
                     std::vector<hsize_t>* d) const;
    // Writers: create or replace the dataset (and missing parent groups).
    bool write_f32(const std::string& path, const float* data, const std::vector<hsize_t>& d);
    bool write_f64(const std::string& path, const double* data, const std::vector<hsize_t>& d);
    bool write_i32(const std::string& path, const int32_t* data, const std::vector<hsize_t>& d);
    bool write_i64(const std::string& path, const int64_t* data, const std::vector<hsize_t>& d);
    bool write_u16(const std::string& path, const uint16_t* data, const std::vector<hsize_t>& d);
    bool write_u32(const std::string& path, const uint32_t* data, const std::vector<hsize_t>& d);
    bool write_string(const std::string& path, const std::string& s);
    // Chunked, deflate-compressed float dataset (plain when libhdf5 lacks the filter API).
    bool write_f32_deflate(const std::string& path, const float* data, const std::vector<hsize_t>& d, int level);
    hid_t id() const { return fid_; }

private:
    bool write_any(const std::string& path, hid_t memtype, hid_t filetype, const void* data,
                   const std::vector<hsize_t>& d, int deflate = -1);
    bool read_any(const std::string& path, hid_t memtype, void* buf, size_t elem, size_t cap,
                  std::vector<hsize_t>* d) const;
    hid_t fid_ = -1;
};

}  // namespace h5
}  // namespace xpcs
