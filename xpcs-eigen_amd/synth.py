"""Synthetic XPCS inputs: OU speckle photon counts, annular q-maps, IMM files.

Speckle model (per pixel): a complex Ornstein-Uhlenbeck field
``E[t+1] = a*E[t] + sqrt(1-a^2)*xi[t]`` with ``xi ~ CN(0,1)``, ``a = exp(-1/tau_c)``
and ``E[0] ~ CN(0,1)``; intensity ``I = mu*|E|^2``; counts ``Poisson(I)`` as uint16.
For such a field the level-0 intensity autocorrelation of a single pixel tends
to ``1 + exp(-2*tau/tau_c)``.

Q-maps are concentric annuli about the detector centre with ``R_max = W/2``:
``sbin = 1 + floor(S*r/R_max)`` and ``qbin = 1 + floor((sbin-1)/(S/Q))``, so
static rings nest inside dynamic rings; pixels outside ``R_max`` are 0 (masked).

IMM layout (1024-byte header per frame, then the payload) follows the header
of the reference reader (src/xpcs/io/imm.h:63-144): compression @4, rows @108,
cols @112, bytes @116, elapsed @128 (f64), dlen @152, buffer_number @160,
immversion @616, corecotick @620.  Compressed payload: int32 index[dlen] then
uint16 value[dlen]; uncompressed payload: uint16 value[W*H].
"""
from __future__ import annotations

import numpy as np

HEADER_BYTES = 1024


def annular_qmaps(width: int, height: int, n_static: int, n_dynamic: int):
    """Return (dqmap, sqmap) int32 [height, width] nested annuli (R_max = W/2)."""
    if n_static % n_dynamic:
        raise ValueError("n_static must be a multiple of n_dynamic")
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    cx, cy = width / 2.0 - 0.5, height / 2.0 - 0.5
    r = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2)
    rmax = width / 2.0
    sbin = 1 + np.floor(n_static * r / rmax).astype(np.int64)
    sbin[r >= rmax] = 0
    per = n_static // n_dynamic
    qbin = np.where(sbin > 0, 1 + (sbin - 1) // per, 0)
    return qbin.astype(np.int32), sbin.astype(np.int32)


def ou_counts(width, height, frames, mu, tau_c, seed, max_count=65535):
    """Dense counts [frames, height*width] uint16 from the OU speckle model.

    ``tau_c`` is a scalar or a per-pixel array (length height*width).
    Draw order per frame: real normals, imaginary normals, Poisson.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    npix = width * height
    tau = np.broadcast_to(np.asarray(tau_c, dtype=np.float64), (npix,))
    a = np.exp(-1.0 / tau)
    b = np.sqrt(1.0 - a * a)
    s = np.sqrt(0.5)
    e = (rng.standard_normal(npix) + 1j * rng.standard_normal(npix)) * s
    out = np.empty((frames, npix), dtype=np.uint16)
    for t in range(frames):
        if t > 0:
            xi = (rng.standard_normal(npix) + 1j * rng.standard_normal(npix)) * s
            e = a * e + b * xi
        lam = mu * (e.real * e.real + e.imag * e.imag)
        c = rng.poisson(lam)
        np.minimum(c, max_count, out=c)
        out[t] = c
    return out


def _header(compression, rows, cols, dlen, t, exposure=1.0e-3):
    h = np.zeros(HEADER_BYTES, dtype=np.uint8)
    def put(off, val, dt):
        h[off:off + np.dtype(dt).itemsize] = np.frombuffer(np.array(val, dtype=dt).tobytes(), np.uint8)
    put(0, 2 if compression else 0, np.int32)        # mode
    put(4, compression, np.int32)
    put(108, rows, np.int32)
    put(112, cols, np.int32)
    put(116, 2, np.int32)                             # bytes per value
    put(128, t * exposure, np.float64)                # elapsed
    put(136, exposure, np.float64)                    # preset
    put(152, dlen, np.uint32)
    put(156, 1, np.int32)                             # roi_number
    put(160, t, np.int32)                             # buffer_number
    put(616, 12, np.int32)                            # immversion
    put(620, t, np.int32)                             # corecotick
    return h


def imm_bytes_sparse(counts: np.ndarray, width: int, height: int) -> bytes:
    """Compressed IMM (compression=6): per frame the non-zero pixels."""
    parts = []
    for t in range(counts.shape[0]):
        idx = np.flatnonzero(counts[t]).astype(np.int32)
        val = counts[t][idx].astype(np.uint16)
        parts.append(_header(6, height, width, idx.size, t).tobytes())
        parts.append(idx.tobytes())
        parts.append(val.tobytes())
    return b"".join(parts)


def imm_bytes_events(frames_events, width: int, height: int) -> bytes:
    """Compressed IMM from a list of (index int32[], value uint16[]) per frame."""
    parts = []
    for t, (idx, val) in enumerate(frames_events):
        idx = np.asarray(idx, dtype=np.int32)
        val = np.asarray(val, dtype=np.uint16)
        parts.append(_header(6, height, width, idx.size, t).tobytes())
        parts.append(idx.tobytes())
        parts.append(val.tobytes())
    return b"".join(parts)


def imm_bytes_dense(counts: np.ndarray, width: int, height: int) -> bytes:
    """Uncompressed IMM (compression=0): every pixel of every frame."""
    parts = []
    for t in range(counts.shape[0]):
        parts.append(_header(0, height, width, width * height, t).tobytes())
        parts.append(np.ascontiguousarray(counts[t], dtype=np.uint16).tobytes())
    return b"".join(parts)


def _col_major(idx, width, height):
    """Column-major pixel id of raster pixel idx (row * width + col): col * height + row."""
    idx = np.asarray(idx, np.int64)
    return (idx % width) * height + idx // width


def ufxc_words(counts: np.ndarray, width: int, height: int, counter0: int = 0, split=()) -> np.ndarray:
    """Packed UFXC u32 words of counts [frames, width*height] (values clipped to
    the 2-bit range): frame counter (counter0 + f) mod 2048 << 21 | value << 15
    | column-major pixel.  Frames in `split` are written in two halves with the
    next frame's words between them (non-contiguous frames)."""
    if width * height > (1 << 15):
        raise ValueError("UFXC pixel ids are 15 bits")
    per = []
    for f in range(counts.shape[0]):
        idx = np.flatnonzero(counts[f])
        v = np.minimum(counts[f, idx], 3).astype(np.uint32)
        w = ((np.uint32((counter0 + f) % 2048) << np.uint32(21)) | (v << np.uint32(15))
             | _col_major(idx, width, height).astype(np.uint32))
        per.append(w.astype(np.uint32))
    out = []
    f = 0
    while f < len(per):
        if f in split and f + 1 < len(per):
            h = per[f].size // 2
            out += [per[f][:h], per[f + 1], per[f][h:]]
            f += 2
        else:
            out.append(per[f])
            f += 1
    return np.concatenate(out) if out else np.zeros(0, np.uint32)


def rigaku_words(counts: np.ndarray, width: int, height: int, first_frame: int = 1, frame_numbers=None,
                 extra=()) -> np.ndarray:
    """Packed Rigaku u64 words of counts [frames, width*height] (values clipped to
    11 bits): file frame << 40 | column-major pixel << 16 | value, file frames
    first_frame + f (or frame_numbers[f]); frames without counts have no words.
    `extra` appends (file frame, raster pixel, value) words at the end."""
    if width * height > (1 << 20):
        raise ValueError("Rigaku pixel ids are 20 bits")
    out = []
    for f in range(counts.shape[0]):
        idx = np.flatnonzero(counts[f])
        v = np.minimum(counts[f, idx], 2047).astype(np.uint64)
        fn = np.uint64(frame_numbers[f] if frame_numbers is not None else first_frame + f)
        out.append((fn << np.uint64(40)) | (_col_major(idx, width, height).astype(np.uint64) << np.uint64(16)) | v)
    for fn, idx, v in extra:
        out.append(np.array([(np.uint64(fn) << np.uint64(40))
                             | (np.uint64(_col_major([idx], width, height)[0]) << np.uint64(16))
                             | np.uint64(v)], np.uint64))
    return np.concatenate(out).astype(np.uint64) if out else np.zeros(0, np.uint64)


def dark_frames(width, height, n, mean=2.0, sd=0.5, seed=0):
    """n identical integer dark frames D = clip(round(N(mean, sd)), 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = np.clip(np.rint(rng.normal(mean, sd, width * height)), 0, None).astype(np.uint16)
    return np.repeat(d[None, :], n, axis=0)


def ring_tau(dqmap: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Per-pixel correlation time, log-spaced over the dynamic rings in [lo, hi]."""
    q = dqmap.reshape(-1)
    nq = max(int(q.max()), 1)
    taus = np.geomspace(lo, hi, nq)
    out = np.full(q.shape, lo, dtype=np.float64)
    m = q > 0
    out[m] = taus[q[m] - 1]
    return out


# ---------------------------------------------------------------------------
# Device-side generator for the large workloads (torch, any device).  Same
# model, different random stream from the numpy generator.
# ---------------------------------------------------------------------------
def _headers(compression, rows, cols, dlens, ts, exposure=1.0e-3):
    """IMM headers of many frames at once: uint8 [n, 1024] (fields as _header)."""
    ts = np.asarray(ts, np.int64)
    n = ts.size
    h = np.zeros((n, HEADER_BYTES), dtype=np.uint8)

    def put(off, vals, dt):
        a = np.empty((n, 1), dtype=dt)
        a[:, 0] = vals
        h[:, off:off + a.itemsize] = a.view(np.uint8).reshape(n, a.itemsize)
    put(0, 2 if compression else 0, np.int32)
    put(4, compression, np.int32)
    put(108, rows, np.int32)
    put(112, cols, np.int32)
    put(116, 2, np.int32)
    put(128, ts * exposure, np.float64)
    put(136, exposure, np.float64)
    put(152, np.asarray(dlens, np.int64), np.uint32)
    put(156, 1, np.int32)
    put(160, ts, np.int32)
    put(616, 12, np.int32)
    put(620, ts, np.int32)
    return h


def ou_events_torch(width, height, frames, mu, tau_c, seed, device="cuda", chunk=256):
    """Yield (t0, counts[n, npix] int32 tensor) blocks of up to `chunk` frames.

    Within a block the OU recursion is one lower-triangular matrix product per
    distinct correlation time: E[t0+k] = a^(k+1) E[t0-1] + sum_j a^(k-j) c_j xi[t0+j]
    with c_j = sqrt(1-a^2) (c = 1 for the very first frame, so E[0] = xi[0])."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    npix = width * height
    tau = np.broadcast_to(np.asarray(tau_c, np.float64), (npix,))
    uniq, inv = np.unique(tau, return_inverse=True)
    groups = []
    k = torch.arange(chunk, device=device, dtype=torch.float64)
    for gi, tc in enumerate(uniq):
        a = float(np.exp(-1.0 / tc))
        b = float(np.sqrt(1.0 - a * a))
        d = (k[:, None] - k[None, :])
        M = torch.where(d >= 0, torch.pow(torch.tensor(a, device=device, dtype=torch.float64), d.clamp(min=0)),
                        torch.zeros((), device=device, dtype=torch.float64)).to(torch.float32)
        carry = torch.pow(torch.tensor(a, device=device, dtype=torch.float64), k + 1).to(torch.float32)
        idx = None if uniq.size == 1 else torch.as_tensor(np.flatnonzero(inv == gi), device=device)
        groups.append((M, carry, b, idx))
    s = float(np.sqrt(0.5))
    er = torch.zeros(npix, device=device)
    ei = torch.zeros(npix, device=device)
    t = 0
    while t < frames:
        n = min(chunk, frames - t)
        xr = torch.randn((n, npix), generator=g, device=device) * s
        xi = torch.randn((n, npix), generator=g, device=device) * s
        nr = torch.empty_like(xr)
        ni = torch.empty_like(xi)
        for M, carry, b, idx in groups:
            cr = (xr if idx is None else xr[:, idx]) * b
            ci = (xi if idx is None else xi[:, idx]) * b
            if t == 0:
                cr[0] /= b
                ci[0] /= b
            pr = er if idx is None else er[idx]
            pi = ei if idx is None else ei[idx]
            vr = M[:n, :n] @ cr + carry[:n, None] * pr[None, :]
            vi = M[:n, :n] @ ci + carry[:n, None] * pi[None, :]
            if idx is None:
                nr, ni = vr, vi
            else:
                nr[:, idx] = vr
                ni[:, idx] = vi
        er, ei = nr[-1].clone(), ni[-1].clone()
        lam = mu * (nr * nr + ni * ni)
        counts = torch.poisson(lam, generator=g).clamp_(max=65535).to(torch.int32)
        yield t, counts
        t += n


def imm_device_sparse(width, height, frames, mu, tau_c, seed, device="cuda"):
    """Build a compressed IMM image directly in device memory.

    Returns (imm uint8 tensor on device, payload_off int64[F], dlen uint32[F]).
    Headers carry compression=6, dlen, buffer_number/corecotick = t and
    elapsed = t*1e-3.
    """
    import torch
    parts = []
    offs, dls = [], []
    pos = 0
    for t0, c in ou_events_torch(width, height, frames, mu, tau_c, seed, device):
        n = c.shape[0]
        nz = torch.nonzero(c, as_tuple=False)           # row-major: frame, then pixel
        fr, px = nz[:, 0], nz[:, 1]
        val = c[fr, px].to(torch.int16)
        dl = torch.bincount(fr, minlength=n)
        dl_h = dl.cpu().numpy().astype(np.int64)
        rec = HEADER_BYTES + 6 * dl_h
        start = np.concatenate([[0], np.cumsum(rec)[:-1]])          # frame record starts in the block
        total = int(rec.sum())
        buf = torch.empty(total, dtype=torch.uint8, device=device)
        hdr = torch.from_numpy(_headers(6, height, width, dl_h, np.arange(t0, t0 + n))).to(device)
        st = torch.as_tensor(start, device=device)
        buf[(st[:, None] + torch.arange(HEADER_BYTES, device=device)[None, :]).reshape(-1)] = hdr.reshape(-1)
        first = torch.cumsum(dl, 0) - dl                                 # first event of each frame
        rank = torch.arange(fr.numel(), device=device) - first[fr]
        ib = st[fr] + HEADER_BYTES + 4 * rank
        buf[(ib[:, None] + torch.arange(4, device=device)[None, :]).reshape(-1)] = \
            px.to(torch.int32).contiguous().view(torch.uint8)
        vb = st[fr] + HEADER_BYTES + 4 * dl[fr] + 2 * rank
        buf[(vb[:, None] + torch.arange(2, device=device)[None, :]).reshape(-1)] = val.contiguous().view(torch.uint8)
        parts.append(buf)
        offs.append(pos + start + HEADER_BYTES)
        dls.append(dl_h)
        pos += total
    return torch.cat(parts), np.concatenate(offs).astype(np.int64), np.concatenate(dls).astype(np.uint32)


def imm_device_dense(width, height, frames, mu, tau_c, seed, darks=2, dark_mean=2.0, dark_sd=0.5,
                     device="cuda"):
    """Uncompressed IMM image in device memory: `darks` identical integer dark
    frames D = clip(round(N(dark_mean, dark_sd)), 0), then `frames` frames of
    OU-speckle counts + D.  Returns (imm uint8 tensor, payload_off, dlen)."""
    import torch
    npix = width * height
    rec = HEADER_BYTES + 2 * npix
    total = darks + frames
    imm = torch.empty(total * rec, dtype=torch.uint8, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(seed + 7)
    d = torch.clamp(torch.round(torch.randn(npix, generator=g, device=device) * dark_sd + dark_mean), min=0)
    d16 = d.to(torch.int32)
    view = imm.view(total, rec)
    for t in range(darks):
        view[t, :HEADER_BYTES] = torch.from_numpy(_header(0, height, width, npix, t)).to(device)
        view[t, HEADER_BYTES:] = d16.to(torch.int16).view(torch.uint8)
    for t0, c in ou_events_torch(width, height, frames, mu, tau_c, seed, device):
        n = c.shape[0]
        t = darks + t0
        hdr = _headers(0, height, width, np.full(n, npix), np.arange(t, t + n))
        view[t:t + n, :HEADER_BYTES] = torch.from_numpy(hdr).to(device)
        view[t:t + n, HEADER_BYTES:] = (c + d16[None, :]).clamp_(max=65535).to(torch.int16).view(torch.uint8)
    offs = np.arange(total, dtype=np.int64) * rec + HEADER_BYTES
    return imm, offs, np.full(total, npix, np.uint32)


def batched_jobs(n_jobs=64):
    """The batched workload: job j draws (seed 4000 + j) frames U[500, 20000], a
    square crop U[128, 1024] px, occupancy log-U[0.1 %, 5 %] (mu = occ / (1 - occ)),
    OU correlation time log-U[1, 1000] frames, dpl 4 or 8, Q ~ U[2, 12], S = 3Q."""
    jobs = []
    for j in range(n_jobs):
        r = np.random.default_rng(4000 + j)
        F = int(r.integers(500, 20001))
        W = int(r.integers(128, 1025))
        occ = float(np.exp(r.uniform(np.log(0.001), np.log(0.05))))
        tau = float(np.exp(r.uniform(0.0, np.log(1000.0))))
        dpl = int(r.choice([4, 8]))
        Q = int(r.integers(2, 13))
        jobs.append(dict(F=F, W=W, mu=occ / (1.0 - occ), tau=tau, dpl=dpl, Q=Q, S=3 * Q, seed=4000 + j))
    return jobs
