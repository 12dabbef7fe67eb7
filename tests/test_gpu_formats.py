"""UFXC / Rigaku words and raw frames on the device vs the oracle: the oracle
restates each reader (src/xpcs/io/ufxc.cpp, rigaku.cpp, hdf5.cpp) as an
equivalent compressed IMM and runs its exact path; sums must agree bit for bit."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu

G2_RTOL = 1e-6


def maps(W, H, S=6, Q=3):
    return synth.annular_qmaps(W, H, S, Q)


def check_same(plan, r):
    G2, IP, IF = plan.sums()
    assert np.array_equal(IP, r["IP"]), "IP differs"
    assert np.array_equal(IF, r["IF"]), "IF differs"
    assert np.array_equal(G2, r["G2"]), "G2 differs"
    g2, _ = plan.g2()
    ok = ~np.isnan(r["g2"])
    assert np.array_equal(np.isnan(g2), ~ok)
    np.testing.assert_allclose(g2[ok], r["g2"][ok], rtol=G2_RTOL)


@pytest.mark.parametrize("stride,avg", [(1, 1), (2, 1), (1, 2), (2, 3)])
def test_ufxc(xmt, orc, stride, avg):
    W, H, F = 32, 24, 150
    R = stride * avg
    dq, sq = maps(W, H)
    c = np.minimum(synth.ou_counts(W, H, F * R, 0.15, 8.0, seed=200 + R), 3)
    w = synth.ufxc_words(c, W, H, counter0=1990, split=(7, 40))
    print(f"ufxc W={W} H={H} F={F} stride={stride} avg={avg} words={w.size}", flush=True)
    plan = xmt.Plan(W, H, F, 4, dq, sq, stride=stride, average=avg, frame_from=5)   # frame_from has no effect
    plan.ingest_packed(w, xmt.FMT_UFXC)
    plan.execute()
    imm = orc.ufxc_imm(w, W, H, F * R)
    check_same(plan, orc.run_exact(imm, dq, sq, F, 4, stride=stride, avg=avg))


@pytest.mark.parametrize("stride,avg,start", [(1, 1, 0), (2, 1, 0), (1, 3, 0), (1, 1, 4), (2, 2, 1)])
def test_rigaku(xmt, orc, stride, avg, start):
    W, H = 40, 30
    dq, sq = maps(W, H)
    mapped = (dq.reshape(-1) > 0) & (sq.reshape(-1) > 0)
    c = synth.ou_counts(W, H, 700, 0.05, 10.0, seed=300 + stride + avg + start)
    c[[3, 50, 51, 200]] = 0                              # file frames without words
    w = synth.rigaku_words(c, W, H, first_frame=1)
    F = 120
    print(f"rigaku W={W} H={H} F={F} stride={stride} avg={avg} start={start} words={w.size}", flush=True)
    plan = xmt.Plan(W, H, F, 4, dq, sq, stride=stride, average=avg, frame_from=start)
    plan.ingest_packed(w, xmt.FMT_RIGAKU)
    plan.execute()
    imm = orc.rigaku_imm(w, W, H, mapped, start, F, stride, avg)
    check_same(plan, orc.run_exact(imm, dq, sq, F, 4))


@pytest.mark.parametrize("dtype", [np.uint16, np.uint32])
def test_raw_frames(xmt, orc, dtype):
    """HDF5-reader frames: non-zero values are events, skip to frame_from, no darks."""
    W, H, F = 24, 20, 200
    dq, sq = maps(W, H)
    c = synth.ou_counts(W, H, 2 * F + 3, 0.4, 12.0, seed=400)
    frames = c.reshape(-1, H, W).astype(dtype)
    print(f"frames dtype={np.dtype(dtype).name} W={W} H={H} F={F}", flush=True)
    plan = xmt.Plan(W, H, F, 4, dq, sq, stride=2, frame_from=3, side_stats=True)
    plan.ingest_frames(frames)
    plan.execute()
    imm = synth.imm_bytes_sparse(c, W, H)
    check_same(plan, orc.run_exact(imm, dq, sq, F, 4, stride=2, frame_from=3))
    # side statistics as the sparse path computes them from the same counts
    ref = xmt.Plan(W, H, F, 4, dq, sq, stride=2, frame_from=3, side_stats=True)
    ref.ingest_imm(imm)
    ref.execute()
    a, b = plan.side(window=20, frame_sum=True, partial=True), ref.side(window=20, frame_sum=True, partial=True)
    for k in a:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-6)


def test_raw_frames_value_range(xmt):
    W, H = 8, 8
    dq, sq = maps(W, H, 2, 1)
    frames = np.zeros((10, H, W), np.uint32)
    frames[4, 2, 3] = 70000
    plan = xmt.Plan(W, H, 10, 4, dq, sq)
    with pytest.raises(xmt.XpcsError) as e:
        plan.ingest_frames(frames)
    assert e.value.status == 5   # XPCS_ERR_UNSUPPORTED


def test_packed_side_stats_and_twotime(xmt, orc):
    """Side statistics and two-time of UFXC words equal those of the same data as IMM."""
    W, H, F = 32, 24, 160
    dq, sq = maps(W, H)
    c = np.minimum(synth.ou_counts(W, H, F, 0.2, 8.0, seed=500), 3)
    w = synth.ufxc_words(c, W, H, counter0=100)
    print("ufxc side stats and two-time", flush=True)
    a = xmt.Plan(W, H, F, 4, dq, sq, side_stats=True)
    a.ingest_packed(w, xmt.FMT_UFXC)
    a.execute()
    b = xmt.Plan(W, H, F, 4, dq, sq, side_stats=True)
    b.ingest_imm(orc.ufxc_imm(w, W, H, F))
    b.execute()
    sa, sb = a.side(window=16, frame_sum=True, partial=True), b.side(window=16, frame_sum=True, partial=True)
    for k in sa:
        np.testing.assert_allclose(sa[k], sb[k], rtol=1e-6)
    ta, tb = a.twotime([1, 3], 1, False, 4), b.twotime([1, 3], 1, False, 4)
    for x, y in zip(ta["c2t"], tb["c2t"]):
        np.testing.assert_allclose(x, y, rtol=1e-12)


def test_ufxc_pixel_out_of_range_is_an_error(xmt):
    W, H = 32, 24
    dq, sq = maps(W, H)
    w = np.array([(5 << 21) | (1 << 15) | 10, (5 << 21) | (1 << 15) | 32767], np.uint32)
    plan = xmt.Plan(W, H, 4, 2, dq, sq)
    plan.ingest_packed(w, xmt.FMT_UFXC)
    plan.execute()
    with pytest.raises(xmt.XpcsError) as e:
        plan.g2()
    assert e.value.status == 4


def test_rigaku_pixel_out_of_range_is_an_error(xmt):
    W, H = 8, 8
    dq, sq = maps(W, H, 2, 1)
    w = np.array([(1 << 40) | (2000 << 16) | 1, (2 << 40) | (3 << 16) | 1, (3 << 40) | (4 << 16) | 1], np.uint64)
    plan = xmt.Plan(W, H, 5, 2, dq, sq)
    plan.ingest_packed(w, xmt.FMT_RIGAKU)
    plan.execute()
    with pytest.raises(xmt.XpcsError) as e:
        plan.g2()
    assert e.value.status == 4


def test_packed_needs_no_darks(xmt):
    W, H = 8, 8
    dq, sq = maps(W, H, 2, 1)
    plan = xmt.Plan(W, H, 10, 4, dq, sq, dark_frames=2)
    with pytest.raises(xmt.XpcsError):
        plan.ingest_packed(np.zeros(4, np.uint32), xmt.FMT_UFXC)


@pytest.mark.parametrize("stride,avg,frame_from", [(1, 1, 0), (2, 3, 5)])
def test_events(xmt, orc, stride, avg, frame_from):
    """Frame-major event arrays: the same results as the compressed IMM file
    holding those events (with empty frames and a repeated ingest)."""
    W, H, F = 40, 28, 150
    dq, sq = maps(W, H)
    raw = frame_from + F * stride * avg
    counts = synth.ou_counts(W, H, raw, 0.3, 10.0, seed=410)
    counts[7] = 0
    counts[frame_from + 1] = 0
    imm = synth.imm_bytes_sparse(counts, W, H)
    nz = [np.flatnonzero(c) for c in counts]
    pixel = np.concatenate(nz).astype(np.int32)
    count = np.concatenate([c[i] for c, i in zip(counts, nz)]).astype(np.uint16)
    frame_off = np.concatenate([[0], np.cumsum([len(i) for i in nz])]).astype(np.int64)
    print(f"events W={W} H={H} F={F} stride={stride} avg={avg} from={frame_from}", flush=True)
    plan = xmt.Plan(W, H, F, 4, dq, sq, stride=stride, average=avg, frame_from=frame_from)
    for _ in range(2):
        plan.ingest_events(pixel, count, frame_off)
        plan.execute()
    check_same(plan, orc.run_exact(imm, dq, sq, F, 4, stride=stride, avg=avg, frame_from=frame_from))


def test_events_pixel_out_of_range_is_an_error(xmt):
    W, H = 8, 8
    dq, sq = maps(W, H, 2, 1)
    plan = xmt.Plan(W, H, 4, 2, dq, sq)
    plan.ingest_events(np.array([3, 64, 5], np.int32), np.array([1, 1, 1], np.uint16),
                       np.array([0, 1, 2, 3, 3], np.int64))
    plan.execute()
    with pytest.raises(xmt.XpcsError) as e:
        plan.g2()
    assert e.value.status == 4


def test_events_bad_offsets_are_refused(xmt):
    W, H = 8, 8
    dq, sq = maps(W, H, 2, 1)
    plan = xmt.Plan(W, H, 4, 2, dq, sq)
    with pytest.raises(xmt.XpcsError) as e:
        plan.ingest_events(np.array([3, 4], np.int32), np.array([1, 1], np.uint16), np.array([0, 2, 1, 2], np.int64))
    assert e.value.status == 1
