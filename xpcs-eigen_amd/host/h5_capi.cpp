// C entry points of include/xpcs_h5.h over the dlopen HDF5 binding.
#include <cstring>
#include <string>
#include <vector>

#include "h5lite.h"
#include "xpcs_h5.h"

struct xpcs_h5_file {
    xpcs::h5::File f;
};

namespace {
thread_local std::string g_err;
int32_t err(const std::string& m)
{
    g_err = m;
    return 1;
}
}  // namespace

extern "C" const char* xpcs_h5_last_error(void) { return g_err.c_str(); }

extern "C" int32_t xpcs_h5_open(const char* path, int32_t create, xpcs_h5_file** out)
{
    if (!path || !out) return err("xpcs_h5_open: NULL argument");
    auto* h = new xpcs_h5_file();
    std::string why;
    if (!h->f.open(path, create != 0, &why)) {
        delete h;
        return err(why);
    }
    *out = h;
    return 0;
}

extern "C" int32_t xpcs_h5_close(xpcs_h5_file* f)
{
    delete f;
    return 0;
}

extern "C" int32_t xpcs_h5_dims(xpcs_h5_file* f, const char* path, uint64_t* dims, int32_t cap, int32_t* nd)
{
    std::vector<xpcs::h5::hsize_t> d;
    if (!f || !path || !nd || !f->f.dims(path, &d)) return err(std::string("no dataset ") + (path ? path : ""));
    *nd = (int32_t)d.size();
    for (size_t i = 0; i < d.size() && (int32_t)i < cap; i++) dims[i] = d[i];
    return 0;
}

extern "C" int32_t xpcs_h5_dtype(xpcs_h5_file* f, const char* path, int32_t* klass, int32_t* size,
                                 int32_t* is_signed)
{
    int cls = -1, sign = -1;
    size_t sz = 0;
    if (!f || !path || !klass || !size || !is_signed || !f->f.type_of(path, &cls, &sz, &sign))
        return err(std::string("no dataset ") + (path ? path : ""));
    *klass = cls;
    *size = (int32_t)sz;
    *is_signed = sign;
    return 0;
}

template <typename T, typename R>
static int32_t read_vec(xpcs_h5_file* f, const char* path, T* out, int64_t cap, R reader)
{
    std::vector<T> v;
    std::vector<xpcs::h5::hsize_t> d;
    if (!f || !path || !(f->f.*reader)(path, &v, &d)) return err(std::string("cannot read ") + (path ? path : ""));
    if ((int64_t)v.size() > cap) return err(std::string("buffer too small for ") + path);
    std::memcpy(out, v.data(), v.size() * sizeof(T));
    return 0;
}

extern "C" int32_t xpcs_h5_read_f64(xpcs_h5_file* f, const char* path, double* out, int64_t cap)
{
    return read_vec(f, path, out, cap, &xpcs::h5::File::read_f64v);
}
extern "C" int32_t xpcs_h5_read_f32(xpcs_h5_file* f, const char* path, float* out, int64_t cap)
{
    return read_vec(f, path, out, cap, &xpcs::h5::File::read_f32v);
}
extern "C" int32_t xpcs_h5_read_i32(xpcs_h5_file* f, const char* path, int32_t* out, int64_t cap)
{
    return read_vec(f, path, out, cap, &xpcs::h5::File::read_i32);
}

extern "C" int32_t xpcs_h5_read_string(xpcs_h5_file* f, const char* path, char* buf, int64_t cap)
{
    std::string s;
    if (!f || !path || !buf || cap <= 0 || !f->f.read_string(path, &s)) return err(std::string("cannot read string ") + (path ? path : ""));
    const size_t n = std::min((size_t)cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
    return 0;
}

extern "C" int32_t xpcs_h5_write(xpcs_h5_file* f, const char* path, int32_t dtype, const void* data,
                                 const uint64_t* dims, int32_t nd)
{
    if (!f || !path || !data || nd < 0) return err("xpcs_h5_write: bad arguments");
    std::vector<xpcs::h5::hsize_t> d(dims, dims + nd);
    bool ok = false;
    switch (dtype) {
    case 0: ok = f->f.write_i32(path, static_cast<const int32_t*>(data), d); break;
    case 1: ok = f->f.write_i64(path, static_cast<const int64_t*>(data), d); break;
    case 2: ok = f->f.write_f32(path, static_cast<const float*>(data), d); break;
    case 3: ok = f->f.write_f64(path, static_cast<const double*>(data), d); break;
    case 4: ok = f->f.write_u16(path, static_cast<const uint16_t*>(data), d); break;
    case 5: ok = f->f.write_u32(path, static_cast<const uint32_t*>(data), d); break;
    default: return err("xpcs_h5_write: unknown dtype");
    }
    return ok ? 0 : err(std::string("cannot write ") + path);
}

extern "C" int32_t xpcs_h5_write_string(xpcs_h5_file* f, const char* path, const char* s)
{
    if (!f || !path || !s) return err("xpcs_h5_write_string: bad arguments");
    return f->f.write_string(path, s) ? 0 : err(std::string("cannot write ") + path);
}
