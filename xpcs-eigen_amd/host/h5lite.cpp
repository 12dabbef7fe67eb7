// dlopen binding of the HDF5 C API (1.10) used by the drop-in host layer.
#include "h5lite.h"

#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <mutex>

namespace xpcs {
namespace h5 {

namespace {

using htri_t = int;
using H5E_auto2_t = herr_t (*)(hid_t, void*);

struct Api {
    void* so = nullptr;
    herr_t (*open)(void);
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    hid_t (*Dget_type)(hid_t);
    hid_t (*Dget_space)(hid_t);
    herr_t (*Dclose)(hid_t);
    herr_t (*Dvlen_reclaim)(hid_t, hid_t, hid_t, void*);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    long long (*Sget_simple_extent_npoints)(hid_t);
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
    hid_t (*Screate)(int);
    herr_t (*Sclose)(hid_t);
    htri_t (*Tis_variable_str)(hid_t);
    size_t (*Tget_size)(hid_t);
    int (*Tget_class)(hid_t);
    int (*Tget_sign)(hid_t) = nullptr;   // optional: integer signedness
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    herr_t (*Tclose)(hid_t);
    herr_t (*Eset_auto2)(hid_t, H5E_auto2_t, void*);
    htri_t (*Lexists)(hid_t, const char*, hid_t);
    herr_t (*Ldelete)(hid_t, const char*, hid_t);
    hid_t (*Pcreate)(hid_t);
    herr_t (*Pclose)(hid_t);
    herr_t (*Pset_create_intermediate_group)(hid_t, unsigned);
    // optional (deflate-compressed datasets): absent symbols mean plain datasets
    herr_t (*Pset_chunk)(hid_t, int, const hsize_t*) = nullptr;
    herr_t (*Pset_deflate)(hid_t, unsigned) = nullptr;
    hid_t* P_CLS_DATASET_CREATE = nullptr;
    hid_t* T_NATIVE_INT;
    hid_t* T_NATIVE_LLONG;
    hid_t* T_NATIVE_USHORT;
    hid_t* T_NATIVE_UINT;
    hid_t* T_NATIVE_FLOAT;
    hid_t* T_NATIVE_DOUBLE;
    hid_t* T_C_S1;
    hid_t* P_CLS_LINK_CREATE;
    std::string why;
    bool ok = false;
};

Api g;
std::once_flag g_once;

template <typename F>
bool bind(F& f, const char* name)
{
    f = reinterpret_cast<F>(dlsym(g.so, name));
    if (!f) g.why = std::string("libhdf5 lacks ") + name;
    return f != nullptr;
}

void load()
{
    const char* env = std::getenv("XPCS_HDF5_LIB");
    const char* cands[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                           "/opt/conda/lib/libhdf5.so"};
    for (const char* c : cands) {
        if (!c || !*c) continue;
        g.so = dlopen(c, RTLD_NOW | RTLD_LOCAL);
        if (g.so) break;
    }
    if (!g.so) {
        g.why = "libhdf5 not found (set XPCS_HDF5_LIB)";
        return;
    }
    bool ok = bind(g.open, "H5open") && bind(g.Fopen, "H5Fopen") && bind(g.Fcreate, "H5Fcreate") &&
              bind(g.Fclose, "H5Fclose") && bind(g.Dopen2, "H5Dopen2") && bind(g.Dcreate2, "H5Dcreate2") &&
              bind(g.Dread, "H5Dread") && bind(g.Dwrite, "H5Dwrite") && bind(g.Dget_type, "H5Dget_type") &&
              bind(g.Dget_space, "H5Dget_space") && bind(g.Dclose, "H5Dclose") &&
              bind(g.Dvlen_reclaim, "H5Dvlen_reclaim") &&
              bind(g.Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims") &&
              bind(g.Sget_simple_extent_dims, "H5Sget_simple_extent_dims") &&
              bind(g.Sget_simple_extent_npoints, "H5Sget_simple_extent_npoints") &&
              bind(g.Screate_simple, "H5Screate_simple") && bind(g.Screate, "H5Screate") &&
              bind(g.Sclose, "H5Sclose") && bind(g.Tis_variable_str, "H5Tis_variable_str") &&
              bind(g.Tget_size, "H5Tget_size") && bind(g.Tget_class, "H5Tget_class") &&
              bind(g.Tcopy, "H5Tcopy") && bind(g.Tset_size, "H5Tset_size") && bind(g.Tclose, "H5Tclose") &&
              bind(g.Eset_auto2, "H5Eset_auto2") && bind(g.Lexists, "H5Lexists") &&
              bind(g.Ldelete, "H5Ldelete") && bind(g.Pcreate, "H5Pcreate") && bind(g.Pclose, "H5Pclose") &&
              bind(g.Pset_create_intermediate_group, "H5Pset_create_intermediate_group") &&
              bind(g.T_NATIVE_INT, "H5T_NATIVE_INT_g") && bind(g.T_NATIVE_LLONG, "H5T_NATIVE_LLONG_g") &&
              bind(g.T_NATIVE_USHORT, "H5T_NATIVE_USHORT_g") && bind(g.T_NATIVE_UINT, "H5T_NATIVE_UINT_g") &&
              bind(g.T_NATIVE_FLOAT, "H5T_NATIVE_FLOAT_g") && bind(g.T_NATIVE_DOUBLE, "H5T_NATIVE_DOUBLE_g") &&
              bind(g.T_C_S1, "H5T_C_S1_g") && bind(g.P_CLS_LINK_CREATE, "H5P_CLS_LINK_CREATE_ID_g");
    if (!ok) return;
    g.Tget_sign = reinterpret_cast<decltype(g.Tget_sign)>(dlsym(g.so, "H5Tget_sign"));
    g.Pset_chunk = reinterpret_cast<decltype(g.Pset_chunk)>(dlsym(g.so, "H5Pset_chunk"));
    g.Pset_deflate = reinterpret_cast<decltype(g.Pset_deflate)>(dlsym(g.so, "H5Pset_deflate"));
    g.P_CLS_DATASET_CREATE = reinterpret_cast<hid_t*>(dlsym(g.so, "H5P_CLS_DATASET_CREATE_ID_g"));
    if (g.open() < 0) {
        g.why = "H5open failed";
        return;
    }
    g.Eset_auto2(0, nullptr, nullptr);   // H5E_DEFAULT: silence the error stack printer
    g.ok = true;
}

constexpr int kClassString = 3;   // H5T_STRING

// Does every component of an absolute path exist?
bool path_exists(hid_t fid, const std::string& path)
{
    if (path.empty() || path == "/") return true;
    size_t pos = 0;
    while (true) {
        pos = path.find('/', pos + 1);
        std::string part = path.substr(0, pos);
        if (!part.empty() && part != "/" && g.Lexists(fid, part.c_str(), 0) <= 0) return false;
        if (pos == std::string::npos) return true;
    }
}

}  // namespace

bool available(std::string* why)
{
    std::call_once(g_once, load);
    if (!g.ok && why) *why = g.why;
    return g.ok;
}

File::~File() { close(); }

void File::close()
{
    if (fid_ >= 0) g.Fclose(fid_);
    fid_ = -1;
}

bool File::open(const std::string& path, bool create, std::string* err)
{
    std::string why;
    if (!available(&why)) {
        if (err) *err = why;
        return false;
    }
    close();
    fid_ = create ? g.Fcreate(path.c_str(), 0x0002u /* TRUNC */, 0, 0) : g.Fopen(path.c_str(), 0x0001u /* RDWR */, 0);
    if (fid_ < 0 && !create) fid_ = g.Fopen(path.c_str(), 0x0000u /* RDONLY */, 0);
    if (fid_ < 0 && err) *err = "cannot open HDF5 file " + path;
    return fid_ >= 0;
}

bool File::exists(const std::string& path) const { return fid_ >= 0 && path_exists(fid_, path); }

bool File::dims(const std::string& path, std::vector<hsize_t>* d) const
{
    if (!exists(path)) return false;
    hid_t ds = g.Dopen2(fid_, path.c_str(), 0);
    if (ds < 0) return false;
    hid_t sp = g.Dget_space(ds);
    int nd = g.Sget_simple_extent_ndims(sp);
    d->assign(nd > 0 ? (size_t)nd : 0, 0);
    if (nd > 0) g.Sget_simple_extent_dims(sp, d->data(), nullptr);
    g.Sclose(sp);
    g.Dclose(ds);
    return nd >= 0;
}

bool File::type_of(const std::string& path, int* cls, size_t* size, int* sign) const
{
    if (!exists(path)) return false;
    hid_t ds = g.Dopen2(fid_, path.c_str(), 0);
    if (ds < 0) return false;
    hid_t ft = g.Dget_type(ds);
    *cls = g.Tget_class(ft);
    *size = g.Tget_size(ft);
    *sign = (*cls == 0 && g.Tget_sign) ? g.Tget_sign(ft) : -1;
    g.Tclose(ft);
    g.Dclose(ds);
    return *cls >= 0;
}

bool File::read_any(const std::string& path, hid_t memtype, void* buf, size_t elem, size_t cap,
                    std::vector<hsize_t>* d) const
{
    if (!exists(path)) return false;
    hid_t ds = g.Dopen2(fid_, path.c_str(), 0);
    if (ds < 0) return false;
    hid_t sp = g.Dget_space(ds);
    long long n = g.Sget_simple_extent_npoints(sp);
    if (d) {
        int nd = g.Sget_simple_extent_ndims(sp);
        d->assign(nd > 0 ? (size_t)nd : 0, 0);
        if (nd > 0) g.Sget_simple_extent_dims(sp, d->data(), nullptr);
    }
    bool ok = n >= 0 && (size_t)n <= cap;
    (void)elem;
    if (ok) ok = g.Dread(ds, memtype, 0, 0, 0, buf) >= 0;
    g.Sclose(sp);
    g.Dclose(ds);
    return ok;
}

static long long npoints(const File& f, const std::string& path)
{
    std::vector<hsize_t> d;
    if (!f.dims(path, &d)) return -1;
    long long n = 1;
    for (hsize_t x : d) n *= (long long)x;
    return n;
}

bool File::read_i64(const std::string& path, int64_t* v) const
{
    long long n = npoints(*this, path);
    if (n < 1) return false;
    std::vector<int64_t> b((size_t)n);
    if (!read_any(path, *g.T_NATIVE_LLONG, b.data(), 8, b.size(), nullptr)) return false;
    *v = b[0];
    return true;
}

bool File::read_f64(const std::string& path, double* v) const
{
    long long n = npoints(*this, path);
    if (n < 1) return false;
    std::vector<double> b((size_t)n);
    if (!read_any(path, *g.T_NATIVE_DOUBLE, b.data(), 8, b.size(), nullptr)) return false;
    *v = b[0];
    return true;
}

bool File::read_i32(const std::string& path, std::vector<int32_t>* v, std::vector<hsize_t>* d) const
{
    long long n = npoints(*this, path);
    if (n < 0) return false;
    v->assign((size_t)n, 0);
    return read_any(path, *g.T_NATIVE_INT, v->data(), 4, v->size(), d);
}

bool File::read_f64v(const std::string& path, std::vector<double>* v, std::vector<hsize_t>* d) const
{
    long long n = npoints(*this, path);
    if (n < 0) return false;
    v->assign((size_t)n, 0.0);
    return read_any(path, *g.T_NATIVE_DOUBLE, v->data(), 8, v->size(), d);
}

bool File::read_f32v(const std::string& path, std::vector<float>* v, std::vector<hsize_t>* d) const
{
    long long n = npoints(*this, path);
    if (n < 0) return false;
    v->assign((size_t)n, 0.0f);
    return read_any(path, *g.T_NATIVE_FLOAT, v->data(), 4, v->size(), d);
}

bool File::read_string(const std::string& path, std::string* v) const
{
    if (!exists(path)) return false;
    hid_t ds = g.Dopen2(fid_, path.c_str(), 0);
    if (ds < 0) return false;
    hid_t ft = g.Dget_type(ds);
    hid_t sp = g.Dget_space(ds);
    bool ok = g.Tget_class(ft) == kClassString;
    long long n = g.Sget_simple_extent_npoints(sp);
    if (ok && n >= 1) {
        if (g.Tis_variable_str(ft) > 0) {
            hid_t mt = g.Tcopy(*g.T_C_S1);
            g.Tset_size(mt, (size_t)-1);
            std::vector<char*> s((size_t)n, nullptr);
            ok = g.Dread(ds, mt, 0, 0, 0, s.data()) >= 0;
            if (ok) *v = s[0] ? std::string(s[0]) : std::string();
            if (ok) g.Dvlen_reclaim(mt, sp, 0, s.data());
            g.Tclose(mt);
        } else {
            size_t sz = g.Tget_size(ft);
            std::vector<char> b(sz * (size_t)n + 1, 0);
            ok = g.Dread(ds, ft, 0, 0, 0, b.data()) >= 0;
            if (ok) *v = std::string(b.data(), strnlen(b.data(), sz));
        }
    } else {
        ok = false;
    }
    g.Sclose(sp);
    g.Tclose(ft);
    g.Dclose(ds);
    return ok;
}

bool File::write_any(const std::string& path, hid_t memtype, hid_t filetype, const void* data,
                     const std::vector<hsize_t>& d, int deflate)
{
    if (fid_ < 0) return false;
    if (exists(path)) g.Ldelete(fid_, path.c_str(), 0);
    hid_t lcpl = g.Pcreate(*g.P_CLS_LINK_CREATE);
    g.Pset_create_intermediate_group(lcpl, 1);
    hid_t sp = d.empty() ? g.Screate(0 /* H5S_SCALAR */) : g.Screate_simple((int)d.size(), d.data(), nullptr);
    // chunked + deflate when asked for and available (chunks of at most 128 per dimension)
    hid_t dcpl = 0;
    bool nonzero = !d.empty();
    for (hsize_t x : d) nonzero = nonzero && x > 0;
    if (deflate >= 0 && nonzero && g.Pset_chunk && g.Pset_deflate && g.P_CLS_DATASET_CREATE) {
        std::vector<hsize_t> cd(d.size());
        for (size_t i = 0; i < d.size(); i++) cd[i] = d[i] < 128 ? d[i] : 128;
        dcpl = g.Pcreate(*g.P_CLS_DATASET_CREATE);
        if (dcpl < 0 || g.Pset_chunk(dcpl, (int)cd.size(), cd.data()) < 0 || g.Pset_deflate(dcpl, (unsigned)deflate) < 0) {
            if (dcpl > 0) g.Pclose(dcpl);
            dcpl = 0;
        }
    }
    hid_t ds = g.Dcreate2(fid_, path.c_str(), filetype, sp, lcpl, dcpl, 0);
    bool ok = ds >= 0 && g.Dwrite(ds, memtype, 0, 0, 0, data) >= 0;
    if (ds >= 0) g.Dclose(ds);
    if (dcpl > 0) g.Pclose(dcpl);
    g.Sclose(sp);
    g.Pclose(lcpl);
    return ok;
}

bool File::write_f32_deflate(const std::string& path, const float* data, const std::vector<hsize_t>& d, int level)
{
    return write_any(path, *g.T_NATIVE_FLOAT, *g.T_NATIVE_FLOAT, data, d, level);
}

bool File::write_u16(const std::string& path, const uint16_t* data, const std::vector<hsize_t>& d)
{
    return write_any(path, *g.T_NATIVE_USHORT, *g.T_NATIVE_USHORT, data, d);
}

bool File::write_u32(const std::string& path, const uint32_t* data, const std::vector<hsize_t>& d)
{
    return write_any(path, *g.T_NATIVE_UINT, *g.T_NATIVE_UINT, data, d);
}

bool File::read_frames(const std::string& path, std::vector<uint16_t>* v16, std::vector<uint32_t>* v32,
                       std::vector<hsize_t>* d) const
{
    if (!exists(path)) return false;
    hid_t ds = g.Dopen2(fid_, path.c_str(), 0);
    if (ds < 0) return false;
    hid_t ft = g.Dget_type(ds);
    const size_t sz = g.Tget_size(ft);
    g.Tclose(ft);
    hid_t sp = g.Dget_space(ds);
    const long long n = g.Sget_simple_extent_npoints(sp);
    const int nd = g.Sget_simple_extent_ndims(sp);
    d->assign(nd > 0 ? (size_t)nd : 0, 0);
    if (nd > 0) g.Sget_simple_extent_dims(sp, d->data(), nullptr);
    bool ok = n >= 0 && (sz == 2 || sz == 4);
    v16->clear();
    v32->clear();
    if (ok && sz == 2) {
        v16->resize((size_t)n);
        ok = g.Dread(ds, *g.T_NATIVE_USHORT, 0, 0, 0, v16->data()) >= 0;
    } else if (ok) {
        v32->resize((size_t)n);
        ok = g.Dread(ds, *g.T_NATIVE_UINT, 0, 0, 0, v32->data()) >= 0;
    }
    g.Sclose(sp);
    g.Dclose(ds);
    return ok;
}

bool File::write_f32(const std::string& path, const float* data, const std::vector<hsize_t>& d)
{
    return write_any(path, *g.T_NATIVE_FLOAT, *g.T_NATIVE_FLOAT, data, d);
}

bool File::write_f64(const std::string& path, const double* data, const std::vector<hsize_t>& d)
{
    return write_any(path, *g.T_NATIVE_DOUBLE, *g.T_NATIVE_DOUBLE, data, d);
}

bool File::write_i32(const std::string& path, const int32_t* data, const std::vector<hsize_t>& d)
{
    return write_any(path, *g.T_NATIVE_INT, *g.T_NATIVE_INT, data, d);
}

bool File::write_i64(const std::string& path, const int64_t* data, const std::vector<hsize_t>& d)
{
    return write_any(path, *g.T_NATIVE_LLONG, *g.T_NATIVE_LLONG, data, d);
}

bool File::write_string(const std::string& path, const std::string& s)
{
    hid_t t = g.Tcopy(*g.T_C_S1);
    g.Tset_size(t, s.size() + 1);
    std::vector<char> b(s.begin(), s.end());
    b.push_back(0);
    bool ok = write_any(path, t, t, b.data(), {1});
    g.Tclose(t);
    return ok;
}

}  // namespace h5
}  // namespace xpcs
