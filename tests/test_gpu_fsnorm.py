"""normalize_by_framesum on the device (FP64 per-pixel series) vs the FP64
oracle: sums within 1e-12, g2 within 1e-9 relative (FP64 both sides, summed in
different orders)."""
import numpy as np
import pytest

import cases
import synth

pytestmark = pytest.mark.gpu

SUM_RTOL = 1e-12
G2_RTOL = 1e-9


def run_gpu(xmt, c, **kw):
    print(f"fsnorm W={c['W']} H={c['H']} F={c['F']} kw={sorted(kw)}", flush=True)
    plan = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"], normalize_by_framesum=True, **kw)
    plan.ingest_imm(c["imm"])
    plan.execute()
    return plan


def check(plan, r):
    G2, IP, IF = plan.sums_f64()
    for got, want in ((G2, r["G2f"]), (IP, r["IPf"]), (IF, r["IFf"])):
        np.testing.assert_allclose(got, want, rtol=SUM_RTOL, atol=1e-9)
    g2, se = plan.g2()
    ok = ~np.isnan(r["g2f"])
    assert np.array_equal(np.isnan(g2), ~ok)
    np.testing.assert_allclose(g2[ok], r["g2f"][ok], rtol=G2_RTOL)
    np.testing.assert_allclose(se, r["stderrf"], rtol=1e-7, atol=1e-12)


def test_sparse(xmt, orc):
    c = cases.sparse_case(W=32, H=32, F=500, mu=0.1, tau_c=10.0, seed=601)
    plan = run_gpu(xmt, c)
    check(plan, orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"]))
    with pytest.raises(xmt.XpcsError):
        plan.sums()                   # the int64 reader refuses FP64 sums


def mixed_rows_case(W, H, F, seed):
    """Pixels in three bands of event counts: <= 64 (one event per lane),
    65..128 (two per lane) and > 128 (the FP64 series kernel)."""
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    lo = synth.ou_counts(W, H, F, 0.05, 10.0, seed)
    mid = synth.ou_counts(W, H, F, 0.2, 10.0, seed + 1)
    hi = synth.ou_counts(W, H, F, 0.8, 10.0, seed + 2)
    band = np.arange(W * H) % 3
    counts = np.where(band == 0, lo, np.where(band == 1, mid, hi)).astype(np.uint16)
    ev = (counts > 0).sum(axis=0)
    assert (ev <= 64).any() and ((ev > 64) & (ev <= 128)).any() and (ev > 128).any()
    return dict(W=W, H=H, F=F, dpl=4, dq=dq, sq=sq, counts=counts, imm=synth.imm_bytes_sparse(counts, W, H))


def test_sparse_row_classes(xmt, orc):
    c = mixed_rows_case(24, 24, 400, seed=611)
    plan = run_gpu(xmt, c)
    check(plan, orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"]))


def test_sparse_row_classes_averaged(xmt, orc):
    # avg 3 merges three raw frames into one key: runs of equal keys at level 0
    c = mixed_rows_case(20, 20, 600, seed=612)
    c["F"] = 200
    plan = run_gpu(xmt, c, average=3)
    check(plan, orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"], avg=3))


def test_sparse_rows_beyond_lds(xmt, orc):
    """Rows of ~800 events (more than the 512-entry LDS kernel holds) beside
    sparse ones: the FP64 series kernel takes the long rows."""
    W = H = 16
    F = 900
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    lo = synth.ou_counts(W, H, F, 0.05, 10.0, 614)
    hi = synth.ou_counts(W, H, F, 3.0, 10.0, 615)
    counts = np.where(np.arange(W * H) % 4 == 0, hi, lo).astype(np.uint16)
    assert (counts > 0).sum(axis=0).max() > 512
    c = dict(W=W, H=H, F=F, dpl=4, dq=dq, sq=sq, imm=synth.imm_bytes_sparse(counts, W, H))
    plan = run_gpu(xmt, c)
    check(plan, orc.run_fsnorm(c["imm"], dq, sq, F, 4))


def test_sparse_many_taus(xmt, orc):
    # more than 128 taus: every row runs through the FP64 series kernel
    c = cases.sparse_case(W=16, H=16, F=1200, dpl=32, mu=0.1, tau_c=20.0, seed=613)
    plan = run_gpu(xmt, c)
    assert plan.T > 128
    check(plan, orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"]))


def test_sparse_flatfield_stride_average(xmt, orc):
    c = cases.sparse_case(W=24, H=24, F=900, mu=0.15, tau_c=8.0, seed=602)
    c["F"] = 200
    ff = np.random.default_rng(603).uniform(0.5, 1.5, (24, 24))
    plan = run_gpu(xmt, c, stride=2, average=2, frame_from=3, flatfield=ff)
    check(plan, orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"], stride=2, avg=2, frame_from=3,
                               flatfield=ff))


def test_dense_with_darks(xmt, orc):
    W = H = 16
    F = 300
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    sig = synth.ou_counts(W, H, F, 4.0, 15.0, seed=604)
    dark = synth.dark_frames(W, H, 2, seed=605)
    c = dict(W=W, H=H, F=F, dpl=4, dq=dq, sq=sq,
             imm=synth.imm_bytes_dense(np.concatenate([dark, sig + dark[0]], axis=0), W, H))
    plan = run_gpu(xmt, c, dark_frames=2)
    check(plan, orc.run_fsnorm(c["imm"], dq, sq, F, 4, darks=2))


def test_dense_tiles_dpl8(xmt, orc):
    # 64x64 uncompressed frames at a fixed step: the staged-tile raw pass, then
    # FP64 series passes for the higher levels
    W = H = 64
    F = 1100
    dq, sq = synth.annular_qmaps(W, H, 8, 4)
    c = dict(W=W, H=H, F=F, dpl=8, dq=dq, sq=sq,
             imm=synth.imm_bytes_dense(synth.ou_counts(W, H, F, 3.0, 25.0, seed=614), W, H))
    plan = run_gpu(xmt, c)
    check(plan, orc.run_fsnorm(c["imm"], dq, sq, F, 8))


def test_dense_stride_average_flatfield(xmt, orc):
    # 20x20 pixels (not a multiple of 64): the per-lane raw pass with averaging
    W = H = 20
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    raw = synth.ou_counts(W, H, 820, 2.0, 12.0, seed=615)
    ff = np.random.default_rng(616).uniform(0.5, 1.5, (H, W))
    c = dict(W=W, H=H, F=200, dpl=4, dq=dq, sq=sq, imm=synth.imm_bytes_dense(raw, W, H))
    plan = run_gpu(xmt, c, stride=2, average=2, frame_from=3, flatfield=ff)
    check(plan, orc.run_fsnorm(c["imm"], dq, sq, 200, 4, stride=2, avg=2, frame_from=3, flatfield=ff))


def test_dense_dpl16_series_rows(xmt, orc):
    # delays per level other than 4 and 8: dense frames go through the
    # transposed per-pixel series instead of the streaming passes
    W = H = 16
    F = 600
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    c = dict(W=W, H=H, F=F, dpl=16, dq=dq, sq=sq,
             imm=synth.imm_bytes_dense(synth.ou_counts(W, H, F, 3.0, 20.0, seed=617), W, H))
    plan = run_gpu(xmt, c)
    check(plan, orc.run_fsnorm(c["imm"], dq, sq, F, 16))


def test_side_stats_are_the_raw_counts(xmt):
    c = cases.sparse_case(W=24, H=24, F=300, mu=0.2, tau_c=8.0, seed=606)
    a = xmt.Plan(c["W"], c["H"], c["F"], 4, c["dq"], c["sq"], side_stats=True, normalize_by_framesum=True)
    b = xmt.Plan(c["W"], c["H"], c["F"], 4, c["dq"], c["sq"], side_stats=True)
    for p in (a, b):
        p.ingest_imm(c["imm"])
        p.execute()
    sa, sb = a.side(window=30, frame_sum=True, partial=True), b.side(window=30, frame_sum=True, partial=True)
    for k in sa:
        np.testing.assert_allclose(sa[k], sb[k], rtol=1e-6)


def test_twotime_uses_normalised_values(xmt, orc):
    c = cases.sparse_case(W=24, H=24, F=200, mu=0.2, tau_c=8.0, seed=607, S=6, Q=3)
    plan = xmt.Plan(c["W"], c["H"], c["F"], 4, c["dq"], c["sq"], normalize_by_framesum=True)
    plan.ingest_imm(c["imm"])
    r = orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], 4)
    for method in (1, 2):
        print("fsnorm twotime method", method, flush=True)
        got = plan.twotime([1, 3], 1, False, 4, method)
        csr = dict(r["csr"])
        csr["cnt"] = csr["cnt"].astype(np.float64) * r["fw"][csr["frame"]]
        want = orc.twotime_f64(csr, r["qm"], c["F"], [1, 3], 1, False, 4)
        for x, y in zip(got["c2t"], want["c2t"]):
            np.testing.assert_allclose(x, y, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(got["g2full"], want["g2full"], rtol=1e-10)


def test_rows_ingest_is_refused(xmt):
    c = cases.sparse_case(W=8, H=8, F=50, mu=0.3, tau_c=5.0, seed=608, S=2, Q=1)
    plan = xmt.Plan(8, 8, 50, 4, c["dq"], c["sq"], normalize_by_framesum=True)
    plan.ingest_rows(np.zeros(65, np.int64), np.zeros(0, np.int32), np.zeros(0, np.uint32))
    with pytest.raises(xmt.XpcsError) as e:
        plan.execute()
    assert e.value.status == 5
