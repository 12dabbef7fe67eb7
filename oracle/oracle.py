"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class ImmFrame(C.Structure):
    _fields_ = [("payload_off", C.c_int64), ("dlen", C.c_uint32), ("compressed", C.c_int32),
                ("elapsed", C.c_double), ("corecotick", C.c_int32), ("pad", C.c_int32)]


FRAME_DTYPE = np.dtype([("payload_off", np.int64), ("dlen", np.uint32), ("compressed", np.int32),
                        ("elapsed", np.float64), ("corecotick", np.int32), ("pad", np.int32)])
framep = np.ctypeslib.ndpointer(FRAME_DTYPE, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        so = HERE / "liboracle.so"
        if not so.exists():
            build()
        L = C.CDLL(str(so))
        L.xo_level_max.argtypes = [C.c_int, C.c_int]
        L.xo_level_max.restype = C.c_int
        L.xo_delays_per_level.argtypes = [C.c_int, C.c_int, C.c_int, i32p, i32p, C.c_int]
        L.xo_delays_per_level.restype = C.c_int
        L.xo_build_qmap.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, i32p, i32p, i32p, i32p,
                                    i32p, i32p, C.c_int, C.c_int, C.c_int, i32p]
        L.xo_build_qmap.restype = C.c_int
        L.xo_imm_index.argtypes = [u8p, C.c_int64, framep, C.c_int64]
        L.xo_imm_index.restype = C.c_int64
        L.xo_select_frames.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       i64p, i32p, C.c_int64]
        L.xo_select_frames.restype = C.c_int64
        L.xo_sparse_transpose.argtypes = [u8p, framep, i64p, i32p, C.c_int64, C.c_int, i32p,
                                          C.c_int32, C.c_void_p, C.c_int, i64p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_int64]
        L.xo_sparse_transpose.restype = C.c_int64
        L.xo_dark_image.argtypes = [u8p, framep, C.c_int, C.c_int, C.c_void_p, f64p, f64p, i64p]
        L.xo_dark_image.restype = None
        L.xo_dense_transpose.argtypes = [u8p, framep, i64p, i32p, C.c_int64, C.c_int, i32p,
                                         C.c_int32, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_float, C.c_float, i64p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.xo_dense_transpose.restype = C.c_int64
        L.xo_multitau_exact.argtypes = [C.c_int64, i64p, i32p, i64p, C.c_int, C.c_int, i32p, i32p,
                                        i64p, i64p, i64p, C.c_int]
        L.xo_multitau_exact.restype = None
        L.xo_multitau_compat.argtypes = [C.c_int64, i64p, i32p, f32p, i32p, C.c_int64, C.c_int,
                                         C.c_int, i32p, i32p, C.c_int, f32p, f32p, f32p, C.c_int]
        L.xo_multitau_compat.restype = None
        L.xo_normalize_f64.argtypes = [C.c_int, i32p, C.c_int64, i64p, i64p, i64p, C.c_int, i32p,
                                       i32p, i32p, C.c_int, C.c_void_p, f64p, f64p]
        L.xo_normalize_f64.restype = None
        L.xo_normalize_compat.argtypes = [C.c_int, C.c_int64, f32p, f32p, f32p, i32p, C.c_int,
                                          i32p, i32p, i32p, i32p, C.c_int, C.c_int, f32p, f32p]
        L.xo_normalize_compat.restype = None
        L.xo_cpu_pipeline.argtypes = [u8p, C.c_int64, C.c_int, C.c_int, i32p, i32p, C.c_int,
                                      C.c_int, C.c_int, f32p, f32p, f64p]
        L.xo_cpu_pipeline.restype = C.c_int
        L.xo_twotime_pairs.argtypes = [C.c_int64, C.c_int64, i64p, i32p, f64p, C.c_int, f64p]
        L.xo_twotime_pairs.restype = None
        L.xo_twotime_bins.argtypes = [C.c_int, i64p, i64p, i64p, i32p, f64p, C.c_int, f64p, C.c_int]
        L.xo_twotime_bins.restype = None
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- schedule ---------------------------------------------------------------
def level_max(frames, dpl):
    return lib().xo_level_max(frames, dpl)


def schedule(frames, dpl):
    """(level[T], tau[T]) of Corr::delaysPerLevel for frames, dpl."""
    ml = level_max(frames, dpl)
    lev = np.zeros(4096, np.int32)
    tau = np.zeros(4096, np.int32)
    n = lib().xo_delays_per_level(frames, dpl, ml, lev, tau, 4096)
    return lev[:n].copy(), tau[:n].copy()


def schedule_N(frames, level, tau):
    """N = lastframe_L - tau' per schedule entry."""
    lf = np.array([frames >> int(L) for L in level], dtype=np.int64)
    return (lf - (tau >> level)).astype(np.int32)


# ---- q map --------------------------------------------------------------------
def build_qmap(dqmap, sqmap):
    dq = np.ascontiguousarray(dqmap, np.int32).reshape(-1)
    sq = np.ascontiguousarray(sqmap, np.int32).reshape(-1)
    h, w = np.asarray(dqmap).shape
    gcap = 1 << 16
    order = np.zeros(w * h, np.int32)
    gq, gs, gb, ge = (np.zeros(gcap, np.int32) for _ in range(4))
    qcap = int(max(dq.max(), 0)) + 2
    scap = int(max(sq.max(), 0)) + 2
    qfirst = np.zeros(qcap, np.int32)
    ppb = np.zeros(scap, np.int32)
    counts = np.zeros(6, np.int32)
    rc = lib().xo_build_qmap(w, h, dq, sq, order, gq, gs, gb, ge, qfirst, ppb, gcap, qcap, scap, counts)
    assert rc == 0
    npix, n_mapped, ng, nq, ns, nkeys = (int(x) for x in counts)
    pix2c = np.full(w * h, -1, np.int32)
    pix2c[order[:npix]] = np.arange(npix, dtype=np.int32)
    return dict(order=order[:npix].copy(), pix2c=pix2c, n_mapped=n_mapped, ngroups=ng,
                group_q=gq[:ng].copy(), group_s=gs[:ng].copy(), group_begin=gb[:ng].copy(),
                group_end=ge[:ng].copy(), q_first_group=qfirst[:nq + 1].copy(),
                pixels_per_bin=ppb[:ns].copy(), nq=nq, ns=ns, nkeys=nkeys, npix=npix,
                width=w, height=h)


# ---- IMM / filters --------------------------------------------------------------
def imm_index(buf):
    b = np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else buf
    b = np.ascontiguousarray(b)
    cap = max(16, b.size // 1024 + 1)
    fr = np.zeros(cap, FRAME_DTYPE)
    n = lib().xo_imm_index(b, b.size, fr, cap)
    if n < 0:
        raise ValueError("truncated IMM payload")
    return fr[:n].copy()


def select_frames(n_raw, frames, darks=0, frame_from=0, stride=1, avg=1):
    cap = frames * avg + 1
    sel = np.zeros(cap, np.int64)
    of = np.zeros(cap, np.int32)
    n = lib().xo_select_frames(n_raw, darks, frame_from, frames, stride, avg, sel, of, cap)
    if n < 0:
        raise ValueError("IMM has fewer frames than the selection needs")
    return sel[:n].copy(), of[:n].copy()


def sparse_transpose(buf, frames_idx, sel, out_frame, qm, flatfield=None, avg=1):
    b = np.ascontiguousarray(np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else buf)
    P = qm["npix"]
    off = np.zeros(P + 1, np.int64)
    ff = None if flatfield is None else np.ascontiguousarray(flatfield, np.float64).reshape(-1)
    npd = qm["width"] * qm["height"]
    n = lib().xo_sparse_transpose(b, frames_idx, sel, out_frame, sel.size, npd, qm["pix2c"], P,
                                  _ptr(ff), avg, off, None, None, None, 0)
    if n < 0:
        raise ValueError("pixel index out of range")
    fr = np.zeros(max(n, 1), np.int32)
    cnt = np.zeros(max(n, 1), np.int64)
    val = np.zeros(max(n, 1), np.float32)
    n2 = lib().xo_sparse_transpose(b, frames_idx, sel, out_frame, sel.size, npd, qm["pix2c"], P,
                                   _ptr(ff), avg, off, _ptr(fr), _ptr(cnt), _ptr(val), max(n, 1))
    assert n2 == n
    return dict(off=off, frame=fr[:n], cnt=cnt[:n], val=val[:n])


def dark_image(buf, frames_idx, darks, npix_det, flatfield=None):
    b = np.ascontiguousarray(np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else buf)
    avgd = np.zeros(npix_det, np.float64)
    std = np.zeros(npix_det, np.float64)
    s = np.zeros(npix_det, np.int64)
    ff = None if flatfield is None else np.ascontiguousarray(flatfield, np.float64).reshape(-1)
    lib().xo_dark_image(b, frames_idx, darks, npix_det, _ptr(ff), avgd, std, s)
    return avgd, std, s


def dense_transpose(buf, frames_idx, sel, out_frame, qm, darks=0, lld=0.0, sigma=0.0,
                    flatfield=None, avg=1):
    b = np.ascontiguousarray(np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else buf)
    P = qm["npix"]
    npd = qm["width"] * qm["height"]
    off = np.zeros(P + 1, np.int64)
    ff = None if flatfield is None else np.ascontiguousarray(flatfield, np.float64).reshape(-1)
    davg = dstd = dsum = None
    if darks > 0:
        davg, dstd, dsum = dark_image(b, frames_idx, darks, npd, ff)
    args = (b, frames_idx, sel, out_frame, sel.size, npd, qm["pix2c"], P, _ptr(ff), avg, darks,
            _ptr(davg), _ptr(dstd), _ptr(dsum), C.c_float(lld), C.c_float(sigma), off)
    n = lib().xo_dense_transpose(*args, None, None, None, 0)
    if n < 0:
        raise ValueError("dense frame size mismatch")
    fr = np.zeros(max(n, 1), np.int32)
    cnt = np.zeros(max(n, 1), np.int64)
    val = np.zeros(max(n, 1), np.float32)
    n2 = lib().xo_dense_transpose(*args, _ptr(fr), _ptr(cnt), _ptr(val), max(n, 1))
    assert n2 == n
    return dict(off=off, frame=fr[:n], cnt=cnt[:n], val=val[:n], dark_avg=davg, dark_std=dstd,
                dark_sum=dsum)


# ---- correlator ------------------------------------------------------------------
def multitau_exact(csr, frames, level, tau, nthreads=0):
    P = csr["off"].size - 1
    T = level.size
    G2 = np.zeros((P, T), np.int64)
    IP = np.zeros((P, T), np.int64)
    IF = np.zeros((P, T), np.int64)
    lib().xo_multitau_exact(P, csr["off"], np.ascontiguousarray(csr["frame"], np.int32),
                            np.ascontiguousarray(csr["cnt"], np.int64), frames, T,
                            level, tau, G2, IP, IF, nthreads)
    return G2, IP, IF


def multitau_compat(csr, c2pix, npix_det, frames, level, tau, stale_tail=True, nthreads=0):
    P = csr["off"].size - 1
    T = level.size
    G2 = np.zeros((T, npix_det), np.float32)
    IP = np.zeros((T, npix_det), np.float32)
    IF = np.zeros((T, npix_det), np.float32)
    lib().xo_multitau_compat(P, csr["off"], np.ascontiguousarray(csr["frame"], np.int32),
                             np.ascontiguousarray(csr["val"], np.float32),
                             np.ascontiguousarray(c2pix, np.int32), npix_det, frames, T, level,
                             tau, int(bool(stale_tail)), G2, IP, IF, nthreads)
    return G2, IP, IF


def normalize_f64(G2, IP, IF, Nvec, qm, weight=None):
    T = G2.shape[1]
    nq = qm["nq"]
    g2 = np.zeros((T, nq), np.float64)
    se = np.zeros((T, nq), np.float64)
    w = None if weight is None else np.ascontiguousarray(weight, np.float64)
    lib().xo_normalize_f64(T, np.ascontiguousarray(Nvec, np.int32), G2.shape[0],
                           np.ascontiguousarray(G2), np.ascontiguousarray(IP),
                           np.ascontiguousarray(IF), qm["ngroups"], qm["group_q"],
                           qm["group_begin"], qm["group_end"], nq, _ptr(w), g2, se)
    return g2, se


def normalize_compat(G2r, IPr, IFr, qm):
    T, npd = G2r.shape
    nq = qm["nq"]
    g2 = np.zeros((T, nq), np.float32)
    se = np.zeros((T, nq), np.float32)
    lib().xo_normalize_compat(T, npd, G2r, IPr, IFr, qm["order"], qm["ngroups"], qm["group_q"],
                              qm["group_s"], qm["group_begin"], qm["group_end"], nq, qm["ns"],
                              g2, se)
    return g2, se


def cpu_pipeline(buf, width, height, dqmap, sqmap, frames, dpl, nthreads=0):
    b = np.ascontiguousarray(np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else buf)
    dq = np.ascontiguousarray(dqmap, np.int32).reshape(-1)
    sq = np.ascontiguousarray(sqmap, np.int32).reshape(-1)
    lev, tau = schedule(frames, dpl)
    nq = int(dq.max())
    g2 = np.zeros((lev.size, nq), np.float32)
    se = np.zeros((lev.size, nq), np.float32)
    st = np.zeros(4, np.float64)
    rc = lib().xo_cpu_pipeline(b, b.size, width, height, dq, sq, frames, dpl, nthreads, g2, se, st)
    if rc < 0:
        raise RuntimeError(f"cpu pipeline failed rc={rc}")
    return g2, se, st


# ---- convenience: whole exact path from IMM bytes ------------------------------------
def run_exact(buf, dqmap, sqmap, frames, dpl, darks=0, frame_from=0, stride=1, avg=1,
              flatfield=None, lld=0.0, sigma=0.0, nthreads=0):
    qm = build_qmap(dqmap, sqmap)
    fr = imm_index(buf)
    sel, of = select_frames(fr.size, frames, darks if (fr.size and not fr[0]["compressed"]) else 0,
                            frame_from, stride, avg)
    if fr.size and fr[0]["compressed"]:
        csr = sparse_transpose(buf, fr, sel, of, qm, flatfield, avg)
    else:
        csr = dense_transpose(buf, fr, sel, of, qm, darks, lld, sigma, flatfield, avg)
    lev, tau = schedule(frames, dpl)
    G2, IP, IF = multitau_exact(csr, frames, lev, tau, nthreads)
    Nv = schedule_N(frames, lev, tau)
    weight = None
    if flatfield is not None:
        weight = np.asarray(flatfield, np.float64).reshape(-1)[qm["order"]]
    g2, se = normalize_f64(G2, IP, IF, Nv, qm, weight)
    return dict(qm=qm, csr=csr, level=lev, tau=tau, N=Nv, G2=G2, IP=IP, IF=IF, g2=g2, stderr=se)


# ---- two-time (Corr::twotime, src/xpcs/corr.cpp:562-924) -------------------------
def _twotime_bins(qm, qbins):
    """Listed q-bins present in m_mapping (ascending) and their groups, as
    Configuration::QbinPixelList builds them (src/xpcs/configuration.cpp:359-380)."""
    listed = {int(q) for q in qbins}
    out = []
    for q in sorted({int(x) for x in qm["group_q"]}):
        if q not in listed:
            continue
        gs = [g for g in range(qm["ngroups"]) if int(qm["group_q"][g]) == q]
        out.append((q, gs))
    return out


def _twotime_values(csr, qm, frames, weight=None, value_scale=1.0):
    """Dense pixel x frame values (float64) from the pixel-major rows."""
    P = qm["npix"]
    X = np.zeros((P, frames), np.float64)
    off = csr["off"]
    for c in range(P):
        b, e = int(off[c]), int(off[c + 1])
        if e > b:
            w = 1.0 if weight is None else float(weight[c])
            X[c, csr["frame"][b:e]] += csr["cnt"][b:e].astype(np.float64) * w * value_scale
    return X


def twotime_f64(csr, qm, frames, qbins, smoothing=0, average=False, wsize=1, weight=None, value_scale=1.0):
    """FP64 two-time: smoothing (0 none, 1 StaticMap corr.cpp:433-494 with
    ComputeSGStaticMap :1228-1305, 2 symmetric :496-544 with ComputeSGSymmetric
    :1166-1226; average = smoothing_filter "Average"), then per listed q-bin
    C2T[f0][f1] = sum_p x[p,f0] x[p,f1] / P_q for f0 <= f1 (0 below the
    diagonal, :800-840), g2full[ff] = mean of diagonal ff, g2partials[ff][w] =
    sum of diagonal ff over window w / wsize (:842-870).  The Gram matrix is a
    float64 BLAS product (the textbook form of the reference's pair loops)."""
    F = frames
    X = _twotime_values(csr, qm, F, weight, value_scale)
    bins = _twotime_bins(qm, qbins)
    gb, ge = qm["group_begin"], qm["group_end"]
    sg = None
    if smoothing == 1:
        groups = [g for _, gs in bins for g in gs]
        sg = np.zeros((len(groups), F))
        for i, g in enumerate(groups):
            sg[i] = X[gb[g]:ge[g]].sum(0) / (ge[g] - gb[g])
        if average:
            sg = sg.sum(1, keepdims=True) / F
        for i, g in enumerate(groups):
            blk = X[gb[g]:ge[g]]
            d = np.broadcast_to(sg[i], (blk.shape[0], F))
            X[gb[g]:ge[g]] = np.where(blk != 0, blk / np.where(d != 0, d, 1.0), 0.0)
    elif smoothing == 2:
        sg = np.zeros((len(bins), F))
        for i, (q, gs) in enumerate(bins):
            rows = np.concatenate([np.arange(gb[g], ge[g]) for g in gs])
            sg[i] = X[rows].sum(0) / rows.size
        if average:
            sg = sg.sum(1, keepdims=True) / F
        for i, (q, gs) in enumerate(bins):
            for g in gs:
                blk = X[gb[g]:ge[g]]
                d = np.broadcast_to(sg[i], (blk.shape[0], F))
                X[gb[g]:ge[g]] = np.where(blk != 0, blk / np.where(d != 0, d, 1.0), 0.0)
    total = (F - wsize) // wsize
    nqb = len(bins)
    c2t = []
    g2full = np.zeros((F, nqb))
    g2part = np.zeros((wsize, max(total, 0), nqb))
    for i, (q, gs) in enumerate(bins):
        rows = np.concatenate([np.arange(gb[g], ge[g]) for g in gs])
        Xq = X[rows]
        G = np.triu(Xq.T @ Xq) / rows.size
        c2t.append(G)
        for ff in range(F):
            d = np.diagonal(G, ff)
            g2full[ff, i] = d.mean()
            if ff < wsize:
                for w in range(total):
                    g2part[ff, w, i] = d[w * wsize:min((w + 1) * wsize, d.size)].sum() / wsize
    return dict(qbins=[q for q, _ in bins], c2t=c2t, g2full=g2full, g2partials=g2part, sg=sg)


def twotime_pairs(csr, qm, frames, q, value_scale=1.0):
    """C2T of one q-bin by the reference's own loop order (corr.cpp:816-835:
    for each pixel, every event pair j <= k adds val_j val_k at (f_j, f_k)),
    for small cases."""
    bins = dict(_twotime_bins(qm, [q]))
    gb, ge = qm["group_begin"], qm["group_end"]
    G = np.zeros((frames, frames))
    n = 0
    for g in bins.get(q, []):
        for c in range(gb[g], ge[g]):
            b, e = int(csr["off"][c]), int(csr["off"][c + 1])
            fr, v = csr["frame"][b:e], csr["cnt"][b:e].astype(np.float64) * value_scale
            for j in range(e - b):
                for k in range(j, e - b):
                    G[fr[j], fr[k]] += v[j] * v[k]
            n += 1
    return G / max(n, 1)


def twotime_pairs_c(csr, qm, frames, qbins, value_scale=1.0, weight=None, nthreads=0):
    """C2T of the listed q-bins by the reference's pair loop in C (FP64);
    rows must hold one value per frame, frames ascending (the oracle's
    transposes produce such rows)."""
    bins = _twotime_bins(qm, qbins)
    gb, ge = qm["group_begin"], qm["group_end"]
    bb = np.array([gb[gs[0]] for _, gs in bins], np.int64)
    be = np.array([ge[gs[-1]] for _, gs in bins], np.int64)
    val = csr["cnt"].astype(np.float64) * value_scale
    if weight is not None:
        val *= np.repeat(np.asarray(weight, np.float64), np.diff(csr["off"]))
    C = np.zeros((max(len(bins), 1), frames, frames))
    lib().xo_twotime_bins(len(bins), bb, be, np.ascontiguousarray(csr["off"], np.int64),
                          np.ascontiguousarray(csr["frame"], np.int32), np.ascontiguousarray(val), frames, C,
                          nthreads)
    return [q for q, _ in bins], C[:len(bins)]


# ---- packed detector formats (src/xpcs/io/ufxc.cpp, src/xpcs/io/rigaku.cpp) ---------
def _raster(col_major, width, height):
    """Detector pixel of a column-major id (ufxc.cpp:142-146, rigaku.cpp:201-203)."""
    col_major = np.asarray(col_major, np.int64)
    return (col_major % height) * width + col_major // height


def ufxc_decode(words, width, height):
    """Raw frame (unwrapped counter, ufxc.cpp:75-97), raster pixel and value
    ((w >> 15) & 3, ufxc.cpp:140-148) of every UFXC word."""
    w = np.asarray(words, np.uint32)
    n = w.size
    ff = np.zeros(n, np.int64)
    if n:
        value = int(w[0] >> 21)
        f0 = value
        bf = 0
        for i in range(1, n):
            c = int(w[i] >> 21)
            diff = c - value
            if diff < -2000:
                bf += 2048
            elif diff > 2000:
                bf -= 2048
            ff[i] = c + bf - f0
            value = c
    return ff, _raster(w & 0x7FFF, width, height), ((w >> 15) & 3).astype(np.int64)


def ufxc_imm(words, width, height, raw_frames):
    """The UFXC words as an equivalent compressed IMM of raw_frames frames (the
    reader's frames 0, 1, ...; words of a frame in file order; empty frames
    where the file has none)."""
    ff, pix, val = ufxc_decode(words, width, height)
    ev = []
    for f in range(raw_frames):
        m = ff == f
        ev.append((pix[m].astype(np.int32), val[m].astype(np.uint16)))
    import synth
    return synth.imm_bytes_events(ev, width, height)


def rigaku_assign(words, width, height, mapped, start, frames, stride=1, avg=1):
    """Output frame of every Rigaku word (-1: not used) by the reader's state
    machine (rigaku.cpp:128-255): skip frames <= start and, with stride > 1,
    frames not divisible by it; flush the pending frame when a word's frame
    differs from the previous flush's (avg 1) or passes the next expected frame
    (avg > 1; it then advances by the read-in count); stop at `frames` output
    frames; keep the last pending frame only with more than one pixel.  Masked
    pixels are not pending pixels."""
    w = np.asarray(words, np.uint64)
    read_in = stride if stride > 1 else avg
    if stride > 1 and avg > 1:
        read_in = stride * avg
    real = 0
    next_expected = start + read_in
    previous = start + 1
    out = np.full(w.size, -1, np.int64)
    pending = {}
    for i in range(w.size):
        frame = int(w[i] >> 40)
        if frame <= start:
            continue
        if real >= frames:
            break
        if stride > 1 and frame != 0 and frame % stride != 0:
            continue
        if (avg > 1 and frame > next_expected) or (avg == 1 and frame != previous):
            for k in pending.values():
                out[k] = real
            pending = {}
            previous = frame
            real += 1
            next_expected += read_in
        pix = int(_raster([(int(w[i]) >> 16) & 0xFFFFF], width, height)[0])
        if not mapped[pix]:
            continue
        pending.setdefault(pix, []).append(i)
    if len(pending) > 1 and real < frames:
        for k in pending.values():
            out[k] = real
    return out


def rigaku_imm(words, width, height, mapped, start, frames, stride=1, avg=1):
    """The Rigaku words as an equivalent compressed IMM of `frames` output
    frames (summed counts of each output frame's words)."""
    w = np.asarray(words, np.uint64)
    of = rigaku_assign(w, width, height, mapped, start, frames, stride, avg)
    pix = _raster((w >> np.uint64(16)) & np.uint64(0xFFFFF), width, height)
    val = (w & np.uint64(0x7FF)).astype(np.int64)
    ev = []
    for f in range(frames):
        m = of == f
        acc = np.zeros(width * height, np.int64)
        np.add.at(acc, pix[m], val[m])
        idx = np.flatnonzero(acc)
        ev.append((idx.astype(np.int32), acc[idx].astype(np.uint16)))
    import synth
    return synth.imm_bytes_events(ev, width, height)


# ---- normalize_by_framesum (src/xpcs/main.cpp:313-336) ------------------------------
def frame_weights(csr, frames, weight=None):
    """mean_f(S_f) / S_f with S_f the frame's summed values over mapped pixels
    (flat-field weighted), 0 for a frame without counts; the reference's
    1/(avg * pixels) factors of frameSum cancel in the ratio."""
    S = np.zeros(frames)
    cnt = csr["cnt"].astype(np.float64)
    if weight is not None:
        cnt = cnt * np.repeat(np.asarray(weight, np.float64), np.diff(csr["off"]))
    np.add.at(S, csr["frame"], cnt)
    m = S.mean()
    return np.where(S > 0, m / np.where(S > 0, S, 1.0), 0.0)


def multitau_f64(csr, frames, level, tau, frame_weight=None):
    """FP64 G2/IP/IF [P][T] of the values count x frame_weight[frame]
    (corr.cpp:315-431 in closed form: binned series x_L, N = lf_L - tau',
    G2 = sum_{t<N} x_L[t] x_L[t+tau'], IP = sum_{t<N} x_L[t],
    IF = sum_{tau'<=t<lf_L} x_L[t])."""
    P = csr["off"].size - 1
    X = np.zeros((P, frames))
    rows = np.repeat(np.arange(P), np.diff(csr["off"]))
    v = csr["cnt"].astype(np.float64)
    if frame_weight is not None:
        v = v * np.asarray(frame_weight)[csr["frame"]]
    np.add.at(X, (rows, csr["frame"]), v)
    T = level.size
    G2, IP, IF = (np.zeros((P, T)) for _ in range(3))
    series = {0: X}
    for L in range(1, int(level.max()) + 1 if T else 1):
        prev = series[L - 1]
        lf = frames >> L
        series[L] = prev[:, 0:2 * lf:2] + prev[:, 1:2 * lf:2]
    for i in range(T):
        L = int(level[i])
        x = series[L]
        lf = frames >> L
        tp = int(tau[i]) >> L
        N = lf - tp
        G2[:, i] = (x[:, :N] * x[:, tp:tp + N]).sum(1)
        IP[:, i] = x[:, :N].sum(1)
        IF[:, i] = x[:, tp:lf].sum(1)
    return G2, IP, IF


def normalize_f64_sums(G2, IP, IF, Nvec, qm, weight=None):
    """normalizeG2s (corr.cpp:927-1091) in FP64 over FP64 sums: the numpy form
    of xo_normalize_f64 (static-bin g2s, NaN-skipping mean per q, stderr of the
    per-pixel ratios with population variance)."""
    G2, IP, IF = (np.asarray(a, np.float64) for a in (G2, IP, IF))
    T = G2.shape[1]
    nq = qm["nq"]
    w = np.ones(G2.shape[0]) if weight is None else np.asarray(weight, np.float64)
    N = np.asarray(Nvec, np.float64)
    ng = qm["ngroups"]
    g2s = np.zeros((ng, T))
    with np.errstate(divide="ignore", invalid="ignore"):
        for g in range(ng):
            b, e = qm["group_begin"][g], qm["group_end"][g]
            sg = (w[b:e, None] ** 2 * G2[b:e]).sum(0)
            sp = (w[b:e, None] * IP[b:e]).sum(0)
            sf = (w[b:e, None] * IF[b:e]).sum(0)
            g2s[g] = (e - b) * N * sg / (sp * sf)
        g2 = np.zeros((T, nq))
        se = np.zeros((T, nq))
        for q in range(1, nq + 1):
            gs = [g for g in range(ng) if qm["group_q"][g] == q]
            v = g2s[gs]
            ok = ~np.isnan(v)
            g2[:, q - 1] = np.where(ok, v, 0.0).sum(0) / ok.sum(0)
            rows = np.concatenate([np.arange(qm["group_begin"][g], qm["group_end"][g]) for g in gs]) if gs else \
                np.zeros(0, np.int64)
            x = N[None, :] * G2[rows] / (IP[rows] * IF[rows])
            x[w[rows] == 0.0] = np.nan
            x = np.nan_to_num(x, nan=0.0, posinf=np.inf)
            n = rows.size
            se[:, q - 1] = np.sqrt(1.0 / n) * np.sqrt(x.var(0)) if n else np.nan
    return g2, se


def run_fsnorm(buf, dqmap, sqmap, frames, dpl, darks=0, frame_from=0, stride=1, avg=1, flatfield=None, lld=0.0,
               sigma=0.0):
    """Whole normalize_by_framesum path: the exact path's rows, frame weights,
    FP64 multi-tau and normalisation."""
    r = run_exact(buf, dqmap, sqmap, frames, dpl, darks, frame_from, stride, avg, flatfield, lld, sigma)
    qm, csr = r["qm"], r["csr"]
    weight = None
    if flatfield is not None:
        weight = np.asarray(flatfield, np.float64).reshape(-1)[qm["order"]]
    fw = frame_weights(csr, frames, weight)
    G2, IP, IF = multitau_f64(csr, frames, r["level"], r["tau"], fw)
    g2, se = normalize_f64_sums(G2, IP, IF, r["N"], qm, weight)
    return dict(r, fw=fw, G2f=G2, IPf=IP, IFf=IF, g2f=g2, stderrf=se)
