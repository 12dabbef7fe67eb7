"""The two-time oracle (Corr::twotime, src/xpcs/corr.cpp:562-924) pinned by
the reference's own pair loop and by closed forms: a geometric series pixel
(every C2T entry, g2full and g2partials window in closed form), constant
images under both smoothing methods, and single-event pixels (diagonal only)."""
import numpy as np
import pytest

import cases
import synth


def test_gram_equals_reference_pair_loop(orc):
    """BLAS Gram form == the reference's per-pixel j <= k event-pair loop
    (corr.cpp:816-835), bit-exact on integer counts."""
    c = cases.sparse_case(W=10, H=10, F=90, mu=0.4, tau_c=6.0, seed=21, S=4, Q=2)
    csr, qm, vs, _ = cases.twotime_inputs(orc, c)
    r = orc.twotime_f64(csr, qm, c["F"], [1, 2])
    assert r["qbins"] == [1, 2]
    for i, q in enumerate(r["qbins"]):
        want = orc.twotime_pairs(csr, qm, c["F"], q)
        assert np.array_equal(r["c2t"][i], want)


@pytest.mark.parametrize("wsize", [1, 3, 4])
def test_geometric_closed_form(orc, wsize):
    c = cases.geometric_case(16)
    csr, qm, vs, _ = cases.twotime_inputs(orc, c)
    r = orc.twotime_f64(csr, qm, c["F"], [1], wsize=wsize)
    C, full, part = cases.geometric_expect(16, wsize)
    assert np.array_equal(r["c2t"][0], C)
    np.testing.assert_allclose(r["g2full"][:, 0], full, rtol=1e-14)
    assert r["g2partials"].shape == (wsize, (16 - wsize) // wsize, 1)
    np.testing.assert_allclose(r["g2partials"][..., 0], part, rtol=1e-14)


def two_static_bins_case(a=3, b=5, F=24, n=4):
    """One dynamic q-bin split into two static bins of n pixels; every pixel of
    the first bin counts a in every frame, of the second b."""
    W, H = 2 * n, 1
    dq = np.ones((H, W), np.int32)
    sq = np.array([[1] * n + [2] * n], np.int32)
    counts = np.zeros((F, W), np.uint16)
    counts[:, :n] = a
    counts[:, n:] = b
    return dict(W=W, H=H, F=F, dpl=4, dq=dq, sq=sq, counts=counts, imm=synth.imm_bytes_sparse(counts, W, H))


@pytest.mark.parametrize("average", [False, True])
def test_smoothing_closed_forms(orc, average):
    a, b, F = 3, 5, 24
    c = two_static_bins_case(a, b, F)
    csr, qm, vs, _ = cases.twotime_inputs(orc, c)
    upper = np.triu(np.ones((F, F)))
    # no smoothing: C = (a^2 + b^2) / 2
    r0 = orc.twotime_f64(csr, qm, F, [1])
    np.testing.assert_allclose(r0["c2t"][0], upper * (a * a + b * b) / 2, rtol=1e-14)
    assert r0["sg"] is None
    # static map: each static bin divided by its own mean image -> all ones
    r1 = orc.twotime_f64(csr, qm, F, [1], smoothing=1, average=average)
    np.testing.assert_allclose(r1["c2t"][0], upper, rtol=1e-14)
    cols = 1 if average else F
    np.testing.assert_allclose(r1["sg"], np.array([[a] * cols, [b] * cols], float), rtol=1e-14)
    # symmetric: one mean image (a+b)/2 for the whole q-bin
    r2 = orc.twotime_f64(csr, qm, F, [1], smoothing=2, average=average)
    m = (a + b) / 2
    np.testing.assert_allclose(r2["c2t"][0], upper * ((a / m) ** 2 + (b / m) ** 2) / 2, rtol=1e-14)
    np.testing.assert_allclose(r2["sg"], np.full((1, cols), m), rtol=1e-14)
    np.testing.assert_allclose(r2["g2full"][:, 0], ((a / m) ** 2 + (b / m) ** 2) / 2, rtol=1e-14)


def test_single_event_pixels_are_diagonal(orc):
    """Pixels with one event each contribute only to C[f][f]."""
    F, W = 12, 6
    counts = np.zeros((F, W), np.uint16)
    frames = [0, 3, 3, 7, 11, 5]
    vals = [2, 3, 4, 1, 6, 9]
    for p, (f, v) in enumerate(zip(frames, vals)):
        counts[f, p] = v
    dq = np.ones((1, W), np.int32)
    sq = np.ones((1, W), np.int32)
    c = dict(W=W, H=1, F=F, dpl=4, dq=dq, sq=sq, imm=synth.imm_bytes_sparse(counts, W, 1))
    csr, qm, vs, _ = cases.twotime_inputs(orc, c)
    r = orc.twotime_f64(csr, qm, F, [1])
    want = np.zeros((F, F))
    for f, v in zip(frames, vals):
        want[f, f] += v * v
    want /= W
    np.testing.assert_allclose(r["c2t"][0], want, rtol=1e-15)
    assert np.all(r["c2t"][0][want == 0] == 0)
    np.testing.assert_allclose(r["g2full"][0, 0], np.trace(want) / F, rtol=1e-14)
    assert np.all(r["g2full"][1:, 0] == 0)


def test_unlisted_and_absent_qbins_are_skipped(orc):
    c = cases.sparse_case(W=10, H=10, F=40, mu=0.4, tau_c=6.0, seed=22, S=6, Q=3)
    csr, qm, vs, _ = cases.twotime_inputs(orc, c)
    r = orc.twotime_f64(csr, qm, c["F"], [3, 1, 99])
    assert r["qbins"] == [1, 3]
    assert r["g2full"].shape == (40, 2)


def test_c_pair_loop_matches_gram(orc):
    """The C restatement of the reference pair loop (CPU baseline of the
    two-time measurement) == the BLAS Gram form, with flat-field weights."""
    c = cases.sparse_case(W=12, H=12, F=80, mu=0.3, tau_c=6.0, seed=23, S=6, Q=3)
    ff = np.random.default_rng(24).uniform(0.5, 1.5, (12, 12))
    csr, qm, vs, w = cases.twotime_inputs(orc, c, flatfield=ff)
    r = orc.twotime_f64(csr, qm, c["F"], [1, 3], weight=w, value_scale=vs)
    qs, Cc = orc.twotime_pairs_c(csr, qm, c["F"], [1, 3], value_scale=vs, weight=w, nthreads=2)
    assert qs == r["qbins"]
    for i in range(len(qs)):
        np.testing.assert_allclose(Cc[i], r["c2t"][i], rtol=1e-13, atol=0)
