"""Shared synthetic cases for the parity tests (seeded, small)."""
import numpy as np

import synth


def sparse_case(W=32, H=32, F=500, dpl=4, mu=0.05, tau_c=10.0, seed=7, S=6, Q=3):
    dq, sq = synth.annular_qmaps(W, H, S, Q)
    counts = synth.ou_counts(W, H, F, mu, tau_c, seed)
    return dict(W=W, H=H, F=F, dpl=dpl, dq=dq, sq=sq, counts=counts,
                imm=synth.imm_bytes_sparse(counts, W, H))


def counterexample_case():
    """Level-0 frames [0..15, 100, 101, 110, 111], counts 1, F=256, dpl=4, one pixel."""
    F = 256
    counts = np.zeros((F, 1), np.uint16)
    counts[list(range(16)) + [100, 101, 110, 111], 0] = 1
    dq = np.ones((1, 1), np.int32)
    sq = np.ones((1, 1), np.int32)
    return dict(W=1, H=1, F=F, dpl=4, dq=dq, sq=sq, counts=counts,
                imm=synth.imm_bytes_sparse(counts, 1, 1))


def twotime_inputs(orc, c, darks=0, avg=1, stride=1, frame_from=0, flatfield=None):
    """Oracle rows, q map, value scale and flat-field weights of a case, in the
    units the two-time analysis uses (counts per output frame / average, dense
    values per dark-frame scale)."""
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], 2, darks=darks, frame_from=frame_from,
                      stride=stride, avg=avg, flatfield=flatfield)
    qm = r["qm"]
    dense = "dark_sum" in r["csr"]
    nscale = darks if (dense and darks > 0) else 1
    weight = None
    if flatfield is not None:
        weight = np.asarray(flatfield, np.float64).reshape(-1)[qm["order"]]
    return r["csr"], qm, 1.0 / (avg * nscale), weight


def geometric_case(F=16):
    """One pixel, count 2^f in frame f (closed forms for every two-time output)."""
    counts = (2 ** np.arange(F, dtype=np.int64)).astype(np.uint16).reshape(F, 1)
    dq = np.ones((1, 1), np.int32)
    sq = np.ones((1, 1), np.int32)
    return dict(W=1, H=1, F=F, dpl=4, dq=dq, sq=sq, counts=counts, imm=synth.imm_bytes_sparse(counts, 1, 1))


def geometric_expect(F, wsize):
    """Closed forms for geometric_case: C[i][j] = 2^(i+j) (j >= i),
    g2full[d] = 2^d (4^(F-d) - 1) / 3 / (F-d),
    g2partials[d][w] = 2^d 4^(w*wsize) (4^wsize - 1) / 3 / wsize."""
    i, j = np.meshgrid(np.arange(F), np.arange(F), indexing="ij")
    C = np.where(j >= i, 2.0 ** (i + j), 0.0)
    full = np.array([2.0 ** d * (4.0 ** (F - d) - 1) / 3 / (F - d) for d in range(F)])
    total = (F - wsize) // wsize
    part = np.array([[2.0 ** d * 4.0 ** (w * wsize) * (4.0 ** wsize - 1) / 3 / wsize for w in range(total)]
                     for d in range(wsize)]).reshape(wsize, total)
    return C, full, part
