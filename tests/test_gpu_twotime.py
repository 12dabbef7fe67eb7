"""Two-time on the device vs the FP64 oracle, through the C ABI, for both ways
of forming the sums (1: FP64 matrix-core Gram product, 2: event pairs with
FP64 atomics).  Integer counts without smoothing give exact sums, so C2T
agrees to rounding of the final scale (1e-12); smoothed values agree to 1e-10
relative."""
import numpy as np
import pytest

import cases
import synth

pytestmark = pytest.mark.gpu

EXACT_RTOL = 1e-12
SMOOTH_RTOL = 1e-10


def gpu_twotime(xmt, c, qbins, smoothing=0, sg_average=False, window=1, method=0, chunk_bytes=0, **kw):
    """kw are plan arguments (average = frames averaged per output frame);
    sg_average is the smoothing filter "Average"."""
    print(f"twotime W={c['W']} H={c['H']} F={c['F']} q={list(qbins)} smoothing={smoothing} "
          f"sg_average={sg_average} window={window} method={method} chunk={chunk_bytes} kw={sorted(kw)}",
          flush=True)
    plan = xmt.Plan(c["W"], c["H"], c["F"], c.get("dpl", 4), c["dq"], c["sq"], **kw)
    plan.ingest_imm(c["imm"])
    r = plan.twotime(qbins, smoothing, sg_average, window, method, chunk_bytes)
    if method:
        assert all(m == method for m in r["methods"])
    return plan, r


def assert_close(got, want, rtol):
    assert got.shape == want.shape, (got.shape, want.shape)
    np.testing.assert_allclose(got, want, rtol=rtol, atol=0)


def compare(got, want, rtol):
    assert got["qbins"] == want["qbins"]
    assert len(got["c2t"]) == len(want["c2t"])
    for g, w in zip(got["c2t"], want["c2t"]):
        assert_close(g, w, rtol)
    assert_close(got["g2full"], want["g2full"], rtol)
    assert_close(got["g2partials"], want["g2partials"], rtol)
    if want["sg"] is None:
        assert got["sg"] is None
    else:
        assert_close(got["sg"], want["sg"], rtol)


def oracle_twotime(orc, c, qbins, smoothing=0, average=False, window=1, darks=0, avg=1, stride=1,
                   frame_from=0, flatfield=None):
    csr, qm, vs, w = cases.twotime_inputs(orc, c, darks=darks, avg=avg, stride=stride, frame_from=frame_from,
                                          flatfield=flatfield)
    return orc.twotime_f64(csr, qm, c["F"], qbins, smoothing, average, window, w, vs)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("wsize", [1, 3, 4])
def test_geometric_closed_form(xmt, wsize, method):
    c = cases.geometric_case(16)
    _, r = gpu_twotime(xmt, c, [1], window=wsize, method=method)
    C, full, part = cases.geometric_expect(16, wsize)
    assert np.array_equal(r["c2t"][0], C)
    assert_close(r["g2full"][:, 0], full, 1e-14)
    assert_close(r["g2partials"][..., 0], part, 1e-14)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("smoothing,average", [(0, False), (1, False), (1, True), (2, False), (2, True)])
def test_sparse_vs_oracle(xmt, orc, smoothing, average, method):
    c = cases.sparse_case(W=32, H=32, F=500, mu=0.08, tau_c=10.0, seed=71, S=6, Q=3)
    _, got = gpu_twotime(xmt, c, [3, 1, 2], smoothing, average, window=7, method=method)
    want = oracle_twotime(orc, c, [3, 1, 2], smoothing, average, 7)
    compare(got, want, EXACT_RTOL if smoothing == 0 else SMOOTH_RTOL)


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("F", [2, 17, 64, 65, 130])
def test_frame_counts_off_tile(xmt, orc, F, method):
    """Frame counts below, at and across the 64-frame output tile."""
    c = cases.sparse_case(W=12, H=12, F=F, mu=0.5, tau_c=5.0, seed=80 + F, S=4, Q=2)
    _, got = gpu_twotime(xmt, c, [1, 2], window=2 if F > 2 else 1, method=method)
    want = oracle_twotime(orc, c, [1, 2], window=2 if F > 2 else 1)
    compare(got, want, EXACT_RTOL)


@pytest.mark.parametrize("smoothing", [0, 1])
def test_pixel_chunks(xmt, orc, smoothing):
    """Values built in many small pixel chunks (sums accumulate across chunks)."""
    c = cases.sparse_case(W=24, H=24, F=200, mu=0.1, tau_c=8.0, seed=90, S=6, Q=3)
    _, got = gpu_twotime(xmt, c, [1, 2, 3], smoothing, window=5, chunk_bytes=37 * 200 * 8)
    want = oracle_twotime(orc, c, [1, 2, 3], smoothing, window=5)
    compare(got, want, EXACT_RTOL if smoothing == 0 else SMOOTH_RTOL)


@pytest.mark.parametrize("method", [1, 2])
def test_many_pixels_many_frames(xmt, orc, method):
    c = cases.sparse_case(W=64, H=64, F=1000, mu=0.03, tau_c=20.0, seed=91, S=6, Q=3)
    _, got = gpu_twotime(xmt, c, [2, 3], method=method)
    want = oracle_twotime(orc, c, [2, 3])
    compare(got, want, EXACT_RTOL)


@pytest.mark.parametrize("method", [1, 2])
def test_average_stride_frames(xmt, orc, method):
    """Averaged raw frames: two raw events of one output frame merge (pairs add
    their cross term twice)."""
    c = cases.sparse_case(W=20, H=20, F=700, mu=0.2, tau_c=8.0, seed=92, S=6, Q=3)
    c["F"] = 150
    _, got = gpu_twotime(xmt, c, [1, 3], 2, False, 4, method=method, stride=2, average=2, frame_from=3)
    want = oracle_twotime(orc, c, [1, 3], 2, False, 4, avg=2, stride=2, frame_from=3)
    compare(got, want, SMOOTH_RTOL)


def test_dense_with_darks(xmt, orc):
    W = H = 24
    F = 300
    dq, sq = synth.annular_qmaps(W, H, 9, 3)
    sig = synth.ou_counts(W, H, F, 5.0, 20.0, seed=93)
    dark = synth.dark_frames(W, H, 2, seed=94)
    frames = np.concatenate([dark, sig + dark[0]], axis=0)
    c = dict(W=W, H=H, F=F, dq=dq, sq=sq, imm=synth.imm_bytes_dense(frames, W, H))
    for smoothing in (0, 1):
        _, got = gpu_twotime(xmt, c, [1, 2, 3], smoothing, window=3, dark_frames=2)
        want = oracle_twotime(orc, c, [1, 2, 3], smoothing, window=3, darks=2)
        compare(got, want, EXACT_RTOL if smoothing == 0 else SMOOTH_RTOL)


@pytest.mark.parametrize("method", [1, 2])
def test_flatfield(xmt, orc, method):
    c = cases.sparse_case(W=32, H=32, F=300, mu=0.1, tau_c=10.0, seed=95, S=6, Q=3)
    ff = np.random.default_rng(96).uniform(0.5, 1.5, (32, 32))
    _, got = gpu_twotime(xmt, c, [1, 2, 3], 1, True, 2, method=method, flatfield=ff)
    want = oracle_twotime(orc, c, [1, 2, 3], 1, True, 2, flatfield=ff)
    compare(got, want, SMOOTH_RTOL)


def test_rows_ingest(xmt, orc):
    """Rows handed over in compact pixel order (no IMM) take the same path."""
    c = cases.sparse_case(W=16, H=16, F=120, mu=0.3, tau_c=6.0, seed=97, S=4, Q=2)
    csr, qm, vs, _ = cases.twotime_inputs(orc, c)
    # rows are handed over per detector pixel (raster order, W*H+1 offsets)
    npd = c["W"] * c["H"]
    lens = np.zeros(npd, np.int64)
    lens[qm["order"]] = np.diff(csr["off"])
    roff = np.concatenate([[0], np.cumsum(lens)])
    fr = np.zeros(roff[-1], np.int32)
    ct = np.zeros(roff[-1], np.uint32)
    for ci, j in enumerate(qm["order"]):
        b, e = csr["off"][ci], csr["off"][ci + 1]
        fr[roff[j]:roff[j + 1]] = csr["frame"][b:e]
        ct[roff[j]:roff[j + 1]] = csr["cnt"][b:e]
    print("twotime rows ingest", flush=True)
    plan = xmt.Plan(c["W"], c["H"], c["F"], 4, c["dq"], c["sq"])
    with pytest.raises(ValueError):
        plan.ingest_rows(csr["off"], csr["frame"], csr["cnt"])     # compact offsets: wrong length
    plan.ingest_rows(roff, fr, ct)
    got = plan.twotime([1, 2], 2, True, 3, method=2)
    want = orc.twotime_f64(csr, qm, c["F"], [1, 2], 2, True, 3)
    compare(got, want, SMOOTH_RTOL)


def test_multitau_unchanged_and_repeatable(xmt, orc):
    """Two-time after execute leaves the multi-tau result as it was; a second
    two-time call replaces the first."""
    c = cases.sparse_case(W=16, H=16, F=200, mu=0.2, tau_c=6.0, seed=98, S=4, Q=2)
    plan = xmt.Plan(c["W"], c["H"], c["F"], 4, c["dq"], c["sq"])
    plan.ingest_imm(c["imm"])
    print("twotime after execute", flush=True)
    plan.execute()
    g2a, _ = plan.g2()
    r1 = plan.twotime([1, 2], 1, False, 2)
    g2b, _ = plan.g2()
    assert np.array_equal(g2a, g2b)
    r2 = plan.twotime([2])
    want = oracle_twotime(orc, c, [2])
    compare(r2, want, EXACT_RTOL)
    assert r1["qbins"] == [1, 2]


def test_long_rows_pairs(xmt, orc):
    """Rows longer than one 64-event block (several blocks per pixel)."""
    c = cases.sparse_case(W=8, H=8, F=400, mu=0.6, tau_c=10.0, seed=101, S=2, Q=1)
    _, got = gpu_twotime(xmt, c, [1], 2, False, 5, method=2)
    want = oracle_twotime(orc, c, [1], 2, False, 5)
    compare(got, want, SMOOTH_RTOL)


def test_auto_method(xmt, orc):
    """Sparse photon counts take the event pairs, dense frames the Gram product."""
    c = cases.sparse_case(W=32, H=32, F=800, mu=0.01, tau_c=10.0, seed=102, S=6, Q=3)
    _, got = gpu_twotime(xmt, c, [1, 2, 3])
    assert got["methods"] == [2, 2, 2]
    compare(got, oracle_twotime(orc, c, [1, 2, 3]), EXACT_RTOL)
    geo = cases.geometric_case(16)
    _, g = gpu_twotime(xmt, geo, [1])
    assert g["methods"] == [1]


def test_pairs_refused_for_dense_frames(xmt):
    W = H = 8
    dq, sq = synth.annular_qmaps(W, H, 2, 1)
    frames = synth.ou_counts(W, H, 20, 3.0, 5.0, seed=103)
    plan = xmt.Plan(W, H, 20, 4, dq, sq)
    plan.ingest_imm(synth.imm_bytes_dense(frames, W, H))
    print("twotime dense method=2 refused", flush=True)
    with pytest.raises(xmt.XpcsError):
        plan.twotime([1], method=2)


def test_no_listed_bin(xmt):
    c = cases.sparse_case(W=8, H=8, F=30, mu=0.3, tau_c=5.0, seed=99, S=2, Q=1)
    _, r = gpu_twotime(xmt, c, [5, 7])
    assert r["qbins"] == [] and r["c2t"] == [] and r["g2full"].shape == (30, 0)


def test_bad_arguments(xmt):
    c = cases.sparse_case(W=8, H=8, F=30, mu=0.3, tau_c=5.0, seed=99, S=2, Q=1)
    plan = xmt.Plan(c["W"], c["H"], c["F"], 4, c["dq"], c["sq"])
    with pytest.raises(xmt.XpcsError):
        plan.twotime([1])            # nothing ingested
    plan.ingest_imm(c["imm"])
    with pytest.raises(xmt.XpcsError):
        plan.twotime([1], window=0)
    with pytest.raises(xmt.XpcsError):
        plan.twotime([1], smoothing=3)
    with pytest.raises(xmt.XpcsError):
        plan.twotime([1], method=5)
