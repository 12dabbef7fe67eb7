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
