"""ctypes front end of libxpcs_host.so (include/xpcs_h5.h): HDF5 datasets for
configurations and results, without h5py."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libxpcs_host.so"
EXPORTS = ["xpcs_h5_last_error", "xpcs_h5_open", "xpcs_h5_close", "xpcs_h5_dims", "xpcs_h5_read_f64",
           "xpcs_h5_read_f32", "xpcs_h5_read_i32", "xpcs_h5_read_string", "xpcs_h5_write",
           "xpcs_h5_write_string", "xpcs_h5_dtype"]
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C {HERE}`")
        L = C.CDLL(str(LIB_PATH))
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.xpcs_h5_last_error.restype = C.c_char_p
        L.xpcs_h5_open.argtypes = [C.c_char_p, i32, vp]
        L.xpcs_h5_close.argtypes = [vp]
        L.xpcs_h5_dims.argtypes = [vp, C.c_char_p, vp, i32, vp]
        for n in ("xpcs_h5_read_f64", "xpcs_h5_read_f32", "xpcs_h5_read_i32"):
            getattr(L, n).argtypes = [vp, C.c_char_p, vp, i64]
        L.xpcs_h5_read_string.argtypes = [vp, C.c_char_p, C.c_char_p, i64]
        L.xpcs_h5_write.argtypes = [vp, C.c_char_p, i32, vp, vp, i32]
        L.xpcs_h5_write_string.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.xpcs_h5_dtype.argtypes = [vp, C.c_char_p, vp, vp, vp]
        for n in EXPORTS[1:]:
            getattr(L, n).restype = i32
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().xpcs_h5_last_error().decode())


class H5:
    def __init__(self, path, create=False):
        self._f = C.c_void_p()
        _check(lib().xpcs_h5_open(str(path).encode(), 1 if create else 0, C.byref(self._f)))

    def close(self):
        if self._f:
            lib().xpcs_h5_close(self._f)
            self._f = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def dims(self, path):
        d = (C.c_uint64 * 8)()
        n = C.c_int32(0)
        _check(lib().xpcs_h5_dims(self._f, path.encode(), d, 8, C.byref(n)))
        return tuple(int(d[i]) for i in range(n.value))

    def dtype(self, path):
        """The stored element type: a numpy dtype, or str for strings."""
        k, n, sg = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        _check(lib().xpcs_h5_dtype(self._f, path.encode(), C.byref(k), C.byref(n), C.byref(sg)))
        if k.value == 3:
            return str
        if k.value == 1:
            return np.dtype(f"f{n.value}")
        if k.value == 0:
            return np.dtype(f"{'u' if sg.value == 0 else 'i'}{n.value}")
        raise RuntimeError(f"{path}: HDF5 type class {k.value}")

    def read(self, path, dtype=np.float64):
        shape = self.dims(path)
        out = np.zeros(int(np.prod(shape)) if shape else 1, dtype)
        fn = {np.dtype(np.float64): lib().xpcs_h5_read_f64, np.dtype(np.float32): lib().xpcs_h5_read_f32,
              np.dtype(np.int32): lib().xpcs_h5_read_i32}[np.dtype(dtype)]
        _check(fn(self._f, path.encode(), out.ctypes.data_as(C.c_void_p), out.size))
        return out.reshape(shape) if shape else out[0]

    def read_string(self, path):
        b = C.create_string_buffer(4096)
        _check(lib().xpcs_h5_read_string(self._f, path.encode(), b, 4096))
        return b.value.decode()

    def write(self, path, arr):
        a = np.ascontiguousarray(arr)
        code = {np.dtype(np.int32): 0, np.dtype(np.int64): 1, np.dtype(np.float32): 2,
                np.dtype(np.float64): 3, np.dtype(np.uint16): 4, np.dtype(np.uint32): 5}[a.dtype]
        dims = (C.c_uint64 * max(a.ndim, 1))(*a.shape)
        _check(lib().xpcs_h5_write(self._f, path.encode(), code, a.ctypes.data_as(C.c_void_p), dims, a.ndim))

    def write_string(self, path, s):
        _check(lib().xpcs_h5_write_string(self._f, path.encode(), s.encode()))


def write_config(path, *, width, height, dqmap, sqmap, imm_path, data_begin_todo=1, data_end_todo,
                 dpl, entry="/xpcs", output="/exchange", compression=True, stride=1, avg=1,
                 dark_begin_todo=0, dark_end_todo=0, lld=0.0, sigma=0.0, static_window=100,
                 flatfield=None, int_dtype=np.int32, float_dtype=np.float32, analysis_type="Multitau",
                 smoothing_method="StaticMap", smoothing_filter="None", qphi_bins=(1,), twotime_window=1,
                 normalize_by_framesum=0):
    """Writes every key Configuration::init reads (src/xpcs/configuration.cpp:80-242);
    the two-time keys (smoothing_method, smoothing_filter, qphi_bin_to_process)
    only with analysis_type "Twotime"."""
    I = lambda v: np.array([v], int_dtype)
    Fv = lambda v: np.array([v], float_dtype)
    det = "/measurement/instrument/detector/"
    with H5(path, create=True) as h:
        h.write_string(entry + "/compression", "ENABLED" if compression else "DISABLED")
        h.write_string(entry + "/output_data", output)
        h.write(det + "x_dimension", I(width))
        h.write(det + "y_dimension", I(height))
        h.write(entry + "/dqmap", np.asarray(dqmap, np.int32).reshape(height, width))
        h.write(entry + "/sqmap", np.asarray(sqmap, np.int32).reshape(height, width))
        h.write(entry + "/data_begin", I(1))
        h.write(entry + "/data_end", I(data_end_todo))
        h.write(entry + "/data_begin_todo", I(data_begin_todo))
        h.write(entry + "/data_end_todo", I(data_end_todo))
        h.write(entry + "/delays_per_level", I(dpl))
        h.write(entry + "/dark_begin_todo", I(dark_begin_todo))
        h.write(entry + "/dark_end_todo", I(dark_end_todo))
        h.write(entry + "/twotime2onetime_window_size", I(twotime_window))
        h.write(entry + "/stride_frames", np.array([stride], np.int64))
        h.write(entry + "/avg_frames", np.array([avg], np.int64))
        h.write(entry + "/normalize_by_framesum", I(normalize_by_framesum))
        h.write(entry + "/lld", Fv(lld))
        h.write(entry + "/sigma", Fv(sigma))
        for k, v in (("x_pixel_size", 0.075), ("y_pixel_size", 0.075), ("adu_per_photon", 1.0),
                     ("exposure_time", 0.001), ("efficiency", 1.0), ("distance", 5000.0)):
            h.write(det + k, Fv(v))
        h.write("/measurement/instrument/source_begin/beam_intensity_transmitted", Fv(1.0))
        h.write("/measurement/sample/thickness", Fv(1.0))
        h.write(entry + "/static_mean_window_size", I(static_window))
        h.write_string(entry + "/flatfield_enabled", "ENABLED" if flatfield is not None else "DISABLED")
        if flatfield is not None:
            h.write(det + "flatfield", np.asarray(flatfield, np.float64).reshape(height, width))
        h.write_string(entry + "/analysis_type", analysis_type)
        if analysis_type == "Twotime":
            h.write_string(entry + "/smoothing_method", smoothing_method)
            h.write_string(entry + "/smoothing_filter", smoothing_filter)
            h.write(entry + "/qphi_bin_to_process", np.asarray(qphi_bins, np.int64).reshape(-1, 1))
        h.write_string(entry + "/input_file_local", str(imm_path))
