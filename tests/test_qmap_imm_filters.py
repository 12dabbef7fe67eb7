"""Host logic: q-map tables (product and oracle), IMM indexing, frame selection,
and the oracle's frame->pixel transposes."""
import numpy as np
import pytest

import cases
import synth


def test_qmap_dedupe_known_answer(xmt, orc):
    """sbin 2 spans q=1 (2 pixels) and q=2 (3 pixels): the q=2 group is kept and
    the q=1 group's pixels leave the mapping (src/xpcs/configuration.cpp:307-345)."""
    dq = np.array([[1, 1, 1, 2], [2, 2, 2, 0]], np.int32)
    sq = np.array([[1, 2, 2, 2], [2, 2, 3, 4]], np.int32)
    for qm in (xmt.build_qmap(dq, sq), orc.build_qmap(dq, sq)):
        assert qm["group_q"].tolist() == [1, 2, 2]
        assert qm["group_s"].tolist() == [1, 2, 3]
        assert qm["order"].tolist() == [0, 3, 4, 5, 6, 1, 2]       # flat ids; 1, 2 dropped
        assert qm["n_mapped"] == 5


def test_qmap_dedupe_tie_keeps_first_q(xmt, orc):
    """Equal group sizes: the first (smallest) q keeps the static bin (strict >)."""
    dq = np.array([[1, 1, 1, 2], [2, 2, 2, 0]], np.int32)
    sq = np.array([[1, 2, 2, 2], [2, 3, 3, 4]], np.int32)
    for qm in (xmt.build_qmap(dq, sq), orc.build_qmap(dq, sq)):
        assert qm["group_q"].tolist() == [1, 1, 2]
        assert qm["group_s"].tolist() == [1, 2, 3]
        assert qm["order"].tolist() == [0, 1, 2, 5, 6, 3, 4]


@pytest.mark.parametrize("seed", range(4))
def test_qmap_product_equals_oracle(xmt, orc, seed):
    rng = np.random.default_rng(seed)
    dq = rng.integers(0, 5, (23, 17)).astype(np.int32)
    sq = rng.integers(0, 9, (23, 17)).astype(np.int32)
    a, b = xmt.build_qmap(dq, sq), orc.build_qmap(dq, sq)
    for k in ("order", "group_q", "group_s", "group_begin", "group_end"):
        assert np.array_equal(a[k], b[k]), k
    assert a["n_mapped"] == b["n_mapped"] and a["nq"] == b["nq"] and a["ns"] == b["ns"]


def test_annular_nesting():
    dq, sq = synth.annular_qmaps(64, 64, 18, 6)
    m = sq > 0
    assert np.array_equal(dq[m], 1 + (sq[m] - 1) // 3)
    assert dq.max() == 6 and sq.max() == 18


def test_imm_index_roundtrip(xmt, orc):
    counts = synth.ou_counts(10, 7, 25, 0.3, 4.0, seed=1)
    imm = synth.imm_bytes_sparse(counts, 10, 7)
    a = xmt.imm_index(imm)
    b = orc.imm_index(imm)
    assert a["compressed"] and np.array_equal(a["payload_off"], b["payload_off"])
    assert np.array_equal(a["dlen"], (counts > 0).sum(1))
    assert np.array_equal(a["corecotick"], np.arange(25))
    assert np.allclose(a["elapsed"], np.arange(25) * 1e-3)
    with pytest.raises(xmt.XpcsError):
        xmt.imm_index(imm[:-3])
    d = synth.imm_bytes_dense(counts, 10, 7)
    ad = xmt.imm_index(d)
    assert not ad["compressed"] and np.all(ad["dlen"] == 70)


def test_sparse_transpose_preserves_events(orc):
    c = cases.sparse_case(W=16, H=16, F=300, dpl=4, mu=0.3, tau_c=5.0, seed=2)
    qm = orc.build_qmap(c["dq"], c["sq"])
    fr = orc.imm_index(c["imm"])
    sel, of = orc.select_frames(fr.size, c["F"])
    csr = orc.sparse_transpose(c["imm"], fr, sel, of, qm)
    for p in range(qm["npix"]):
        a, b = csr["off"][p], csr["off"][p + 1]
        f = csr["frame"][a:b]
        assert np.all(np.diff(f) > 0)
        col = c["counts"][:, qm["order"][p]]
        assert np.array_equal(f, np.flatnonzero(col))
        assert np.array_equal(csr["cnt"][a:b], col[f])
    assert csr["off"][-1] == sum((c["counts"][:, qm["order"]] > 0).sum(0))


def test_frame_selection_known_answer(orc):
    sel, of = orc.select_frames(100, frames=3, darks=2, frame_from=5, stride=3, avg=2)
    # R = 6: blocks start at 5, 11, 17; frames j*stride within a block
    assert sel.tolist() == [5, 8, 11, 14, 17, 20]
    assert of.tolist() == [0, 0, 1, 1, 2, 2]
    sel, _ = orc.select_frames(100, frames=2, darks=4, frame_from=0)
    assert sel.tolist() == [4, 5]       # data follow the dark frames


def test_dense_transpose_dark_threshold(orc):
    W = H = 4
    dark = np.full((2, 16), 3, np.uint16)
    sig = np.array([[0, 1, 5, 3] * 4, [7, 3, 2, 4] * 4], np.uint16)
    imm = synth.imm_bytes_dense(np.concatenate([dark, sig]), W, H)
    dq = np.ones((4, 4), np.int32)
    sq = np.ones((4, 4), np.int32)
    qm = orc.build_qmap(dq, sq)
    fr = orc.imm_index(imm)
    sel, of = orc.select_frames(fr.size, 2, darks=2)
    csr = orc.dense_transpose(imm, fr, sel, of, qm, darks=2)
    # kept iff raw > dark (3); values scaled by n=2: 2*(raw-3)
    p0 = csr["cnt"][csr["off"][0]:csr["off"][1]]
    assert p0.tolist() == [8]                  # frame 0 raw 0 dropped, frame 1 raw 7 -> 2*4
    p2 = csr["cnt"][csr["off"][2]:csr["off"][3]]
    assert p2.tolist() == [4]                  # raw 5 kept (2*2), raw 2 dropped


def dark_straddle_case():
    """Pixel 5 has dark frames 32767 and 32769 (mean 32768, standard deviation
    1 as uint16 values; as int16 they would be 32767 and -32767, standard
    deviation 32767) and data 32768 + s for s = 0..3 in frames 0..3; every other
    pixel has dark 3 and data 3 + s."""
    W = H = 4
    dark = np.full((2, 16), 3, np.uint16)
    dark[0, 5], dark[1, 5] = 32767, 32769
    s = np.arange(4, dtype=np.uint16)[:, None]
    sig = np.repeat(3 + s, 16, axis=1).astype(np.uint16)
    sig[:, 5] = 32768 + s[:, 0]
    imm = synth.imm_bytes_dense(np.concatenate([dark, sig]), W, H)
    return dict(W=W, H=H, F=4, dq=np.ones((4, 4), np.int32), sq=np.ones((4, 4), np.int32), imm=imm)


def test_dense_dark_above_int16_known_answer(orc):
    """Dark mean, deviation, threshold and subtraction all use uint16 values:
    with lld 0 and sigma 1 pixel 5's threshold is 2 x 1 (n = 2 dark frames) and
    its values are 2 (raw - 32768), so s = 2, 3 are kept as 4, 6."""
    c = dark_straddle_case()
    qm = orc.build_qmap(c["dq"], c["sq"])
    fr = orc.imm_index(c["imm"])
    sel, of = orc.select_frames(fr.size, 4, darks=2)
    csr = orc.dense_transpose(c["imm"], fr, sel, of, qm, darks=2, lld=0.0, sigma=1.0)
    assert csr["dark_avg"][5] == 32768.0 and csr["dark_std"][5] == 1.0
    c5 = int(np.flatnonzero(qm["order"] == 5)[0])
    a, b = csr["off"][c5], csr["off"][c5 + 1]
    assert csr["frame"][a:b].tolist() == [2, 3] and csr["cnt"][a:b].tolist() == [4, 6]
    c0 = int(np.flatnonzero(qm["order"] == 0)[0])
    a, b = csr["off"][c0], csr["off"][c0 + 1]
    assert csr["frame"][a:b].tolist() == [1, 2, 3] and csr["cnt"][a:b].tolist() == [2, 4, 6]
