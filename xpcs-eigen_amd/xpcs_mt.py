"""Python front end of libxpcs_mt.so (the C ABI declared in include/xpcs_mt.h).

Mirrors the reference's multi-tau interface (xpcs::Corr, src/xpcs/corr.h:66-126)
for tests, benchmarks and scripting:

    plan = Plan(width, height, frames, dpl, dqmap, sqmap)
    plan.ingest_imm(imm_bytes)            # Reader + Filter (main.cpp:222-268)
    plan.execute()                        # multiTau2 + normalizeG2s
    g2, stderr = plan.g2()                # norm-0-g2 / norm-0-stderr, [T, Q]

There is no CPU fallback: loading fails loudly when the HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libxpcs_mt.so"

XPCS_MT_TIMING = 0x1
XPCS_MT_SIDE_STATS = 0x2
XPCS_MT_NORMALIZE_BY_FRAMESUM = 0x4
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "HIP", 3: "FORMAT", 4: "RANGE", 5: "UNSUPPORTED",
          6: "STATE", 7: "NOMEM", 8: "NO_DEVICE"}

EXPORTS = [
    "xpcs_version", "xpcs_last_error", "xpcs_level_max", "xpcs_delay_count",
    "xpcs_delays_per_level", "xpcs_mt_plan_create", "xpcs_mt_plan_destroy", "xpcs_mt_plan_info",
    "xpcs_mt_schedule", "xpcs_mt_pixel_order", "xpcs_mt_ingest_imm", "xpcs_mt_ingest_imm_device",
    "xpcs_imm_index", "xpcs_mt_ingest_rows", "xpcs_mt_ingest_events", "xpcs_mt_execute", "xpcs_mt_synchronize",
    "xpcs_mt_read_sums_i64", "xpcs_mt_read_ref_f32", "xpcs_mt_read_g2", "xpcs_mt_device_results",
    "xpcs_mt_read_side", "xpcs_mt_kernel_times", "xpcs_mt_kernel_times_reset", "xpcs_mt_set_timing_filter",
    "xpcs_build_qmap",
    "xpcs_mt_read_dark", "xpcs_mt_set_tuning", "xpcs_mt_get_tuning", "xpcs_mt_autotune",
    "xpcs_mt_batch_create", "xpcs_mt_batch_add", "xpcs_mt_batch_execute", "xpcs_mt_batch_synchronize",
    "xpcs_mt_batch_destroy", "xpcs_mt_stream_bytes", "xpcs_mt_stream_open", "xpcs_mt_stream_push_imm",
    "xpcs_mt_stream_push_imm_device", "xpcs_mt_stream_frames", "xpcs_mt_stream_snapshot",
    "xpcs_mt_stream_poll", "xpcs_mt_stream_read_g2", "xpcs_mt_stream_read_sums_i64", "xpcs_mt_stream_close",
    "xpcs_mt_twotime", "xpcs_mt_twotime_info", "xpcs_mt_read_twotime", "xpcs_mt_read_twotime_1d",
    "xpcs_mt_ingest_packed", "xpcs_mt_ingest_frames", "xpcs_packed_records", "xpcs_mt_read_sums_f64",
    "xpcs_mt_normalize_ref_f32", "xpcs_mt_row_stats", "xpcs_mt_normalize_ref_f32_noalloc",
]
FMT_UFXC = 1
FMT_RIGAKU = 2
XPCS_BATCH_AUTOTUNE = 0x1


class XpcsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Desc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("frames", C.c_int32),
                ("delays_per_level", C.c_int32), ("stride", C.c_int32), ("average", C.c_int32),
                ("frame_from", C.c_int32), ("dark_frames", C.c_int32), ("lld", C.c_float),
                ("sigma", C.c_float), ("dqmap", C.c_void_p), ("sqmap", C.c_void_p),
                ("flatfield", C.c_void_p), ("flags", C.c_uint32)]


class Info(C.Structure):
    _fields_ = [("taus", C.c_int32), ("max_level", C.c_int32), ("dynamic_bins", C.c_int32),
                ("static_bins", C.c_int32), ("pixels", C.c_int32), ("mapped_pixels", C.c_int32),
                ("groups", C.c_int32), ("exact_pixels", C.c_int32), ("events", C.c_int64),
                ("long_pixels", C.c_int64), ("deep_level", C.c_int32), ("deep_stride", C.c_int32)]


_LIB = None


def lib():
    """Load libxpcs_mt.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                                    "(or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        vp, i32, i64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32
        sig = {
            "xpcs_version": ([], C.c_char_p),
            "xpcs_last_error": ([], C.c_char_p),
            "xpcs_level_max": ([i32, i32], i32),
            "xpcs_delay_count": ([i32, i32], i32),
            "xpcs_delays_per_level": ([i32, i32, i32, vp, vp, i32, vp], i32),
            "xpcs_mt_plan_create": ([vp, vp], i32),
            "xpcs_mt_plan_destroy": ([vp], i32),
            "xpcs_mt_plan_info": ([vp, vp], i32),
            "xpcs_mt_schedule": ([vp, vp, vp], i32),
            "xpcs_mt_pixel_order": ([vp, vp], i32),
            "xpcs_mt_ingest_imm": ([vp, vp, C.c_size_t], i32),
            "xpcs_mt_ingest_imm_device": ([vp, vp, C.c_size_t, vp, vp, i64, i32], i32),
            "xpcs_imm_index": ([vp, C.c_size_t, vp, vp, vp, vp, i64, vp, vp], i32),
            "xpcs_mt_ingest_rows": ([vp, vp, vp, vp], i32),
            "xpcs_mt_ingest_events": ([vp, vp, vp, vp, i64], i32),
            "xpcs_mt_execute": ([vp, vp], i32),
            "xpcs_mt_synchronize": ([vp], i32),
            "xpcs_mt_read_sums_i64": ([vp, vp, vp, vp], i32),
            "xpcs_mt_read_ref_f32": ([vp, vp, vp, vp], i32),
            "xpcs_mt_read_g2": ([vp, vp, vp], i32),
            "xpcs_mt_device_results": ([vp, vp, vp], i32),
            "xpcs_mt_read_side": ([vp, i32, vp, vp, vp, vp], i32),
            "xpcs_mt_kernel_times": ([vp, vp, C.c_size_t, vp, vp, i32, vp], i32),
            "xpcs_mt_kernel_times_reset": ([vp], i32),
            "xpcs_mt_set_timing_filter": ([vp, C.c_char_p], i32),
            "xpcs_build_qmap": ([i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, vp], i32),
            "xpcs_mt_read_dark": ([vp, vp, vp], i32),
            "xpcs_mt_set_tuning": ([vp, vp], i32),
            "xpcs_mt_get_tuning": ([vp, vp], i32),
            "xpcs_mt_autotune": ([vp, vp], i32),
            "xpcs_mt_batch_create": ([i32, vp], i32),
            "xpcs_mt_batch_add": ([vp, vp, vp], i32),
            "xpcs_mt_batch_execute": ([vp, vp, u32], i32),
            "xpcs_mt_batch_synchronize": ([vp], i32),
            "xpcs_mt_batch_destroy": ([vp], i32),
            "xpcs_mt_stream_bytes": ([vp, i32, vp], i32),
            "xpcs_mt_stream_open": ([vp, i32, vp], i32),
            "xpcs_mt_stream_push_imm": ([vp, vp, C.c_size_t], i32),
            "xpcs_mt_stream_push_imm_device": ([vp, vp, C.c_size_t, vp, vp, i32], i32),
            "xpcs_mt_stream_frames": ([vp, vp, vp], i32),
            "xpcs_mt_stream_snapshot": ([vp], i32),
            "xpcs_mt_stream_poll": ([vp, vp], i32),
            "xpcs_mt_stream_read_g2": ([vp, vp, vp, vp], i32),
            "xpcs_mt_stream_read_sums_i64": ([vp, vp, vp, vp], i32),
            "xpcs_mt_stream_close": ([vp], i32),
            "xpcs_mt_twotime": ([vp, vp, vp], i32),
            "xpcs_mt_twotime_info": ([vp, vp, vp, vp, i32, vp, vp, vp], i32),
            "xpcs_mt_read_twotime": ([vp, i32, vp], i32),
            "xpcs_mt_read_twotime_1d": ([vp, vp, vp, vp], i32),
            "xpcs_mt_ingest_packed": ([vp, vp, C.c_size_t, i32], i32),
            "xpcs_mt_read_sums_f64": ([vp, vp, vp, vp], i32),
            "xpcs_mt_ingest_frames": ([vp, vp, C.c_int64, i32], i32),
            "xpcs_packed_records": ([vp, vp, C.c_size_t, i32, vp, vp, vp, C.c_int64, vp], i32),
            "xpcs_mt_normalize_ref_f32": ([vp, vp, vp, vp, i64, i32, vp, vp], i32),
            "xpcs_mt_row_stats": ([vp, vp, vp], i32),
            "xpcs_mt_normalize_ref_f32_noalloc": ([vp, vp, vp, vp, i64, i32, vp, vp, vp], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def check(st):
    if st != 0:
        raise XpcsError(st, lib().xpcs_last_error().decode())


def level_max(frames, dpl):
    return lib().xpcs_level_max(frames, dpl)


def delays_per_level(frames, dpl, max_level=None):
    if max_level is None:
        max_level = level_max(frames, dpl)
    n = C.c_int32(0)
    check(lib().xpcs_delays_per_level(frames, dpl, max_level, None, None, 0, C.byref(n)))
    lev = np.zeros(n.value, np.int32)
    tau = np.zeros(n.value, np.int32)
    check(lib().xpcs_delays_per_level(frames, dpl, max_level, _p(lev), _p(tau), n.value, C.byref(n)))
    return lev, tau


def imm_index(buf):
    b = np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else np.ascontiguousarray(buf)
    n = C.c_int64(0)
    comp = C.c_int32(0)
    check(lib().xpcs_imm_index(_p(b), b.size, None, None, None, None, 0, C.byref(n), C.byref(comp)))
    off = np.zeros(n.value, np.int64)
    dl = np.zeros(n.value, np.uint32)
    el = np.zeros(n.value, np.float64)
    ct = np.zeros(n.value, np.int32)
    check(lib().xpcs_imm_index(_p(b), b.size, _p(off), _p(dl), _p(el), _p(ct), n.value, C.byref(n),
                               C.byref(comp)))
    return dict(payload_off=off, dlen=dl, elapsed=el, corecotick=ct, compressed=bool(comp.value))


def build_qmap(dqmap, sqmap):
    """Configuration::BuildQMap tables (host only)."""
    dq = np.ascontiguousarray(dqmap, np.int32)
    sq = np.ascontiguousarray(sqmap, np.int32)
    h, w = dq.shape
    counts = np.zeros(5, np.int32)
    check(lib().xpcs_build_qmap(w, h, _p(dq), _p(sq), None, None, None, None, None, 0, _p(counts)))
    P, nm, ng = int(counts[0]), int(counts[1]), int(counts[2])
    order = np.zeros(max(P, 1), np.int32)
    g = [np.zeros(max(ng, 1), np.int32) for _ in range(4)]
    check(lib().xpcs_build_qmap(w, h, _p(dq), _p(sq), _p(order), _p(g[0]), _p(g[1]), _p(g[2]),
                                _p(g[3]), ng, _p(counts)))
    return dict(order=order[:P], n_mapped=nm, ngroups=ng, group_q=g[0][:ng], group_s=g[1][:ng],
                group_begin=g[2][:ng], group_end=g[3][:ng], nq=int(counts[3]), ns=int(counts[4]))


class TTDesc(C.Structure):
    _fields_ = [("qbins", C.c_void_p), ("nq", C.c_int32), ("smoothing", C.c_int32), ("average", C.c_int32),
                ("window", C.c_int32), ("method", C.c_int32), ("chunk_bytes", C.c_int64)]


class Tuning(C.Structure):
    """Launch shapes of a plan (include/xpcs_mt.h, 0 = automatic)."""
    _fields_ = [("tr_bucket_shift", C.c_int32), ("tr_chunk_frames", C.c_int32), ("reg_blocks", C.c_int32),
                ("lds_blocks", C.c_int32), ("dense_levels", C.c_int32), ("deep_frames", C.c_int32),
                ("sums64", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


def _desc(width, height, frames, dpl, dq, sq, stride=1, average=1, frame_from=0, dark_frames=0, lld=0.0,
          sigma=0.0, ff=None, flags=0):
    return Desc(width, height, frames, dpl, stride, average, frame_from, dark_frames, lld, sigma,
                _p(dq), _p(sq), _p(ff), flags)


def packed_records(words, fmt, width, height, frames, dqmap, sqmap, stride=1, average=1, frame_from=0):
    """Frame records of packed UFXC / Rigaku words (host only): arrays of
    (first word, word count, output frame)."""
    dq = np.ascontiguousarray(dqmap, np.int32).reshape(-1)
    sq = np.ascontiguousarray(sqmap, np.int32).reshape(-1)
    w = np.frombuffer(words, np.uint8) if not isinstance(words, np.ndarray) else np.ascontiguousarray(words)
    d = _desc(width, height, frames, 4, dq, sq, stride, average, frame_from)
    n = C.c_int64(0)
    check(lib().xpcs_packed_records(C.byref(d), _p(w), w.nbytes, fmt, None, None, None, 0, C.byref(n)))
    fw = np.zeros(max(n.value, 1), np.int64)
    cnt = np.zeros(max(n.value, 1), np.uint32)
    fr = np.zeros(max(n.value, 1), np.int32)
    check(lib().xpcs_packed_records(C.byref(d), _p(w), w.nbytes, fmt, _p(fw), _p(cnt), _p(fr), n.value,
                                    C.byref(n)))
    k = n.value
    return fw[:k], cnt[:k], fr[:k]


class Plan:
    """One multi-tau job on the current HIP device."""

    def __init__(self, width, height, frames, dpl, dqmap, sqmap, stride=1, average=1,
                 frame_from=0, dark_frames=0, lld=0.0, sigma=0.0, flatfield=None, timing=False,
                 side_stats=False, normalize_by_framesum=False):
        self._dq = np.ascontiguousarray(dqmap, np.int32).reshape(-1)
        self._sq = np.ascontiguousarray(sqmap, np.int32).reshape(-1)
        self._ff = None if flatfield is None else np.ascontiguousarray(flatfield, np.float64).reshape(-1)
        if self._dq.size != width * height or self._sq.size != width * height:
            raise ValueError("q-maps must have height*width entries")
        d = Desc(width, height, frames, dpl, stride, average, frame_from, dark_frames, lld, sigma,
                 _p(self._dq), _p(self._sq), _p(self._ff),
                 (XPCS_MT_TIMING if timing else 0) | (XPCS_MT_SIDE_STATS if side_stats else 0)
                 | (XPCS_MT_NORMALIZE_BY_FRAMESUM if normalize_by_framesum else 0))
        self._h = C.c_void_p()
        check(lib().xpcs_mt_plan_create(C.byref(d), C.byref(self._h)))
        info = self.info()
        self.T, self.Q, self.S, self.P = info.taus, info.dynamic_bins, info.static_bins, info.pixels
        self.width, self.height, self.frames = width, height, frames
        self.fsnorm = bool(normalize_by_framesum)

    def close(self):
        if self._h:
            lib().xpcs_mt_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        i = Info()
        check(lib().xpcs_mt_plan_info(self._h, C.byref(i)))
        return i

    def schedule(self):
        lev = np.zeros(self.T, np.int32)
        tau = np.zeros(self.T, np.int32)
        check(lib().xpcs_mt_schedule(self._h, _p(lev), _p(tau)))
        return lev, tau

    def pixel_order(self):
        o = np.zeros(self.P, np.int32)
        check(lib().xpcs_mt_pixel_order(self._h, _p(o)))
        return o

    def ingest_imm(self, buf):
        b = np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else np.ascontiguousarray(buf)
        check(lib().xpcs_mt_ingest_imm(self._h, _p(b), b.size))

    def ingest_imm_device(self, dev_ptr, nbytes, payload_off, dlen, compressed=True):
        off = np.ascontiguousarray(payload_off, np.int64)
        dl = np.ascontiguousarray(dlen, np.uint32)
        check(lib().xpcs_mt_ingest_imm_device(self._h, C.c_void_p(dev_ptr), nbytes, _p(off), _p(dl),
                                              off.size, 1 if compressed else 0))

    def ingest_packed(self, words, fmt):
        """UFXC (FMT_UFXC, u32) or Rigaku (FMT_RIGAKU, u64) words, as bytes or an array."""
        w = np.frombuffer(words, np.uint8) if not isinstance(words, np.ndarray) else np.ascontiguousarray(words)
        check(lib().xpcs_mt_ingest_packed(self._h, _p(w), w.nbytes, int(fmt)))

    def ingest_frames(self, frames):
        """Raw frames [n, height, width] of uint16 or uint32 (the HDF5 reader's dataset)."""
        a = np.ascontiguousarray(frames)
        if a.dtype not in (np.uint16, np.uint32):
            raise ValueError("frames must be uint16 or uint32")
        n = a.shape[0] if a.ndim else 0
        if a.size != n * self.width * self.height:
            raise ValueError("frames must hold n * height * width values")
        check(lib().xpcs_mt_ingest_frames(self._h, _p(a), n, a.dtype.itemsize))

    def ingest_rows(self, row_off, frame, count):
        ro = np.ascontiguousarray(row_off, np.int64).reshape(-1)
        fr = np.ascontiguousarray(frame, np.int32).reshape(-1)
        ct = np.ascontiguousarray(count, np.uint32).reshape(-1)
        if ro.size != self.width * self.height + 1:
            raise ValueError(f"row_off needs width*height+1 = {self.width * self.height + 1} entries, got {ro.size}")
        if fr.size < ro[-1] or ct.size < ro[-1]:
            raise ValueError(f"frame/count hold fewer than row_off[-1] = {ro[-1]} events")
        check(lib().xpcs_mt_ingest_rows(self._h, _p(ro), _p(fr), _p(ct)))

    def ingest_events(self, pixel, count, frame_off):
        """Frame-major events: raw frame k holds pixel[frame_off[k]:frame_off[k+1]]
        (raster indices) with count[...] (u16)."""
        fo = np.ascontiguousarray(frame_off, np.int64).reshape(-1)
        px = np.ascontiguousarray(pixel, np.int32).reshape(-1)
        ct = np.ascontiguousarray(count, np.uint16).reshape(-1)
        if fo.size < 1:
            raise ValueError("frame_off needs nframes+1 entries")
        if px.size < fo[-1] or ct.size < fo[-1]:
            raise ValueError(f"pixel/count hold fewer than frame_off[-1] = {fo[-1]} events")
        check(lib().xpcs_mt_ingest_events(self._h, _p(px), _p(ct), _p(fo), fo.size - 1))

    def execute(self, stream=None):
        check(lib().xpcs_mt_execute(self._h, None if stream is None else C.c_void_p(stream)))

    def synchronize(self):
        check(lib().xpcs_mt_synchronize(self._h))

    def twotime(self, qbins, smoothing=0, average=False, window=1, method=0, chunk_bytes=0):
        """Two-time analysis (xpcs_mt_twotime); returns dict(qbins, methods
        (1 Gram, 2 event pairs), c2t list, g2full [F, nq], g2partials
        [window, total, nq], sg)."""
        info = self.twotime_run(qbins, smoothing, average, window, method, chunk_bytes)
        rows, cols, tot = info["sg_rows"], info["sg_cols"], info["total_partials"]
        F, n = self.frames, len(info["qbins"])
        c2t = []
        for i in range(n):
            m = np.zeros((F, F), np.float64)
            check(lib().xpcs_mt_read_twotime(self._h, i, _p(m)))
            c2t.append(m)
        full = np.zeros((F, n), np.float64)
        part = np.zeros((int(window), tot, n), np.float64)
        sg = np.zeros((rows, cols), np.float64) if rows else None
        check(lib().xpcs_mt_read_twotime_1d(self._h, _p(full), _p(part) if part.size else None, _p(sg)))
        return dict(qbins=info["qbins"], methods=info["methods"], c2t=c2t, g2full=full, g2partials=part, sg=sg)

    def twotime_run(self, qbins, smoothing=0, average=False, window=1, method=0, chunk_bytes=0):
        """Runs xpcs_mt_twotime and returns its info (results stay on the device)."""
        qb = np.ascontiguousarray(qbins, np.int32).reshape(-1)
        d = TTDesc(_p(qb), qb.size, int(smoothing), 1 if average else 0, int(window), int(method), int(chunk_bytes))
        check(lib().xpcs_mt_twotime(self._h, C.byref(d), None))
        nq, rows, cols, tot = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        qs = np.zeros(max(qb.size, 1), np.int32)
        ms = np.zeros(max(qb.size, 1), np.int32)
        check(lib().xpcs_mt_twotime_info(self._h, C.byref(nq), _p(qs), _p(ms), qs.size, C.byref(rows),
                                         C.byref(cols), C.byref(tot)))
        n = nq.value
        return dict(qbins=qs[:n].tolist(), methods=ms[:n].tolist(), sg_rows=rows.value, sg_cols=cols.value,
                    total_partials=tot.value)

    def set_tuning(self, **shapes):
        t = Tuning()
        for k, v in shapes.items():
            setattr(t, k, int(v))
        check(lib().xpcs_mt_set_tuning(self._h, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        check(lib().xpcs_mt_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def autotune(self):
        t = Tuning()
        check(lib().xpcs_mt_autotune(self._h, C.byref(t)))
        return t.as_dict()

    def sums_f64(self):
        """G2/IP/IF as FP64 [P][T] (the normalize_by_framesum sums)."""
        G2 = np.zeros((self.P, self.T), np.float64)
        IP = np.zeros((self.P, self.T), np.float64)
        IF = np.zeros((self.P, self.T), np.float64)
        check(lib().xpcs_mt_read_sums_f64(self._h, _p(G2), _p(IP), _p(IF)))
        return G2, IP, IF

    def sums(self):
        G2 = np.zeros((self.P, self.T), np.int64)
        IP = np.zeros((self.P, self.T), np.int64)
        IF = np.zeros((self.P, self.T), np.int64)
        check(lib().xpcs_mt_read_sums_i64(self._h, _p(G2), _p(IP), _p(IF)))
        return G2, IP, IF

    def ref_f32(self):
        n = self.width * self.height
        G2 = np.zeros((self.T, n), np.float32)
        IP = np.zeros((self.T, n), np.float32)
        IF = np.zeros((self.T, n), np.float32)
        check(lib().xpcs_mt_read_ref_f32(self._h, _p(G2), _p(IP), _p(IF)))
        return G2, IP, IF

    def g2(self):
        g2 = np.zeros((self.T, self.Q), np.float64)
        se = np.zeros((self.T, self.Q), np.float64)
        check(lib().xpcs_mt_read_g2(self._h, _p(g2), _p(se)))
        return g2, se

    def row_stats(self):
        """Pixels and level-0 events per row class (empty, <=64, 65..128,
        129..512, 513..2048, longer) of the last sparse execute."""
        px = np.zeros(6, np.int64)
        ev = np.zeros(6, np.int64)
        check(lib().xpcs_mt_row_stats(self._h, _p(px), _p(ev)))
        return px, ev

    def normalize_ref_f32(self, G2, IP, IF):
        """Corr::normalizeG2s of caller arrays in the reference's units,
        [T][W*H] float32 each (src/xpcs/corr.cpp:927-1091): g2, stderr [T][Q]."""
        n = self.width * self.height
        arrs = [np.ascontiguousarray(a, np.float32).reshape(self.T, n) for a in (G2, IP, IF)]
        g2 = np.zeros((self.T, self.Q), np.float64)
        se = np.zeros((self.T, self.Q), np.float64)
        check(lib().xpcs_mt_normalize_ref_f32(self._h, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), n, self.T,
                                              _p(g2), _p(se)))
        return g2, se

    def device_results(self):
        a, b = C.c_void_p(), C.c_void_p()
        check(lib().xpcs_mt_device_results(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def side(self, window=0, frame_sum=False, partial=False):
        ps = np.zeros(self.width * self.height, np.float32)
        pm = np.zeros(self.S, np.float32)
        fs = np.zeros((2, self.frames), np.float32) if frame_sum else None
        pp = np.zeros((max(self.frames // max(window, 1), 0), self.S), np.float32) if partial else None
        check(lib().xpcs_mt_read_side(self._h, window, _p(ps), _p(fs), _p(pm), _p(pp)))
        return dict(pixel_sum=ps, partition_mean=pm, frame_sum=fs, partition_partial=pp)

    def kernel_times(self):
        names = C.create_string_buffer(4096)
        ms = np.zeros(64, np.float64)
        ln = np.zeros(64, np.int64)
        n = C.c_int32(0)
        check(lib().xpcs_mt_kernel_times(self._h, names, 4096, _p(ms), _p(ln), 64, C.byref(n)))
        keys = names.value.decode().split("\n")[: n.value]
        return {k: (float(ms[i]), int(ln[i])) for i, k in enumerate(keys)}

    def kernel_times_reset(self):
        check(lib().xpcs_mt_kernel_times_reset(self._h))

    def set_timing_filter(self, name=None):
        """Time only the kernel timer called `name` (None: every timer)."""
        check(lib().xpcs_mt_set_timing_filter(self._h, name.encode() if name else None))


class Batch:
    """Many plans run as one captured hipGraph (xpcs_mt_batch_*)."""

    def __init__(self, branches=8):
        self._h = C.c_void_p()
        self._plans = []
        check(lib().xpcs_mt_batch_create(branches, C.byref(self._h)))

    def add(self, plan):
        i = C.c_int32(0)
        check(lib().xpcs_mt_batch_add(self._h, plan._h, C.byref(i)))
        self._plans.append(plan)
        return i.value

    def execute(self, stream=None, autotune=False):
        check(lib().xpcs_mt_batch_execute(self._h, None if stream is None else C.c_void_p(stream),
                                          XPCS_BATCH_AUTOTUNE if autotune else 0))

    def synchronize(self):
        check(lib().xpcs_mt_batch_synchronize(self._h))

    def close(self):
        if self._h:
            lib().xpcs_mt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _stream_desc(width, height, dpl, dq, sq, flatfield):
    return Desc(width, height, 0, dpl, 1, 1, 0, 0, 0.0, 0.0, _p(dq), _p(sq), _p(flatfield), 0)


def stream_bytes(width, height, max_frames, dpl, dqmap, sqmap, flatfield=None):
    """Device bytes a stream of at most max_frames frames holds (host-only)."""
    dq = np.ascontiguousarray(dqmap, np.int32).reshape(-1)
    sq = np.ascontiguousarray(sqmap, np.int32).reshape(-1)
    ff = None if flatfield is None else np.ascontiguousarray(flatfield, np.float64).reshape(-1)
    b = C.c_uint64(0)
    check(lib().xpcs_mt_stream_bytes(C.byref(_stream_desc(width, height, dpl, dq, sq, ff)), max_frames,
                                     C.byref(b)))
    return b.value


class Stream:
    """Online multi-tau over frames pushed in pieces (xpcs_mt_stream_*)."""

    def __init__(self, width, height, max_frames, dpl, dqmap, sqmap, flatfield=None):
        self._dq = np.ascontiguousarray(dqmap, np.int32).reshape(-1)
        self._sq = np.ascontiguousarray(sqmap, np.int32).reshape(-1)
        self._ff = None if flatfield is None else np.ascontiguousarray(flatfield, np.float64).reshape(-1)
        self._h = C.c_void_p()
        check(lib().xpcs_mt_stream_open(C.byref(_stream_desc(width, height, dpl, self._dq, self._sq, self._ff)),
                                        max_frames, C.byref(self._h)))
        self.width, self.height, self.max_frames, self.dpl = width, height, max_frames, dpl
        self.Q = int(self._dq.max())
        self.P = int(np.count_nonzero((self._dq > 0) & (self._sq > 0)))

    def push_imm(self, buf):
        b = np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else np.ascontiguousarray(buf)
        check(lib().xpcs_mt_stream_push_imm(self._h, _p(b), b.size))

    def push_imm_device(self, ptr, nbytes, payload_off, dlen):
        off = np.ascontiguousarray(payload_off, np.int64)
        dl = np.ascontiguousarray(dlen, np.uint32)
        check(lib().xpcs_mt_stream_push_imm_device(self._h, C.c_void_p(ptr), nbytes, _p(off), _p(dl), off.size))

    def frames(self):
        n, t = C.c_int32(0), C.c_int32(0)
        check(lib().xpcs_mt_stream_frames(self._h, C.byref(n), C.byref(t)))
        return n.value, t.value

    def snapshot(self):
        check(lib().xpcs_mt_stream_snapshot(self._h))

    def poll(self):
        n = C.c_int32(0)
        check(lib().xpcs_mt_stream_poll(self._h, C.byref(n)))
        return n.value

    def g2(self):
        _, T = self.frames()
        T = max(T, 1)
        g2 = np.zeros((T, self.Q), np.float64)
        se = np.zeros((T, self.Q), np.float64)
        t = C.c_int32(0)
        check(lib().xpcs_mt_stream_read_g2(self._h, _p(g2), _p(se), C.byref(t)))
        return g2[: t.value], se[: t.value]

    def sums(self):
        _, T = self.frames()
        G2 = np.zeros((self.P, T), np.int64)
        IP = np.zeros((self.P, T), np.int64)
        IF = np.zeros((self.P, T), np.int64)
        check(lib().xpcs_mt_stream_read_sums_i64(self._h, _p(G2), _p(IP), _p(IF)))
        return G2, IP, IF

    def close(self):
        if self._h:
            lib().xpcs_mt_stream_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
