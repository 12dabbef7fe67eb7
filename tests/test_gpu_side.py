"""Side statistics on the device: the per-(frame, static bin) sums of the
thresholded, flat-field weighted values (the partition sums behind frameSum and
partition-mean-partial, sparse_filter.cpp:134-191 and the dense filter's frame
loop) against sums over the oracle's transpose of the same frames, dense and
sparse. The values are integers (times a flat-field weight), so the sums agree
to float32 rounding of the reported means."""
import numpy as np
import pytest

import cases
import synth

pytestmark = pytest.mark.gpu


def oracle_sums(orc, imm, dq, sq, F, **kw):
    """[F][S] sums of the n-scaled values over the oracle's rows."""
    r = orc.run_exact(imm, dq, sq, F, 4, **kw)
    qm, csr = r["qm"], r["csr"]
    rows = np.repeat(np.arange(qm["npix"]), np.diff(csr["off"]))
    pix = qm["order"][rows]
    b = np.asarray(sq).reshape(-1)[pix] - 1
    v = csr["cnt"].astype(np.float64)
    if kw.get("flatfield") is not None:
        v = v * np.asarray(kw["flatfield"], np.float64).reshape(-1)[pix]
    fs = np.zeros((F, qm["ns"]))
    np.add.at(fs, (csr["frame"], b), v)
    return fs, qm


def check_side(xmt, orc, W, H, F, dq, sq, imm, darks=0, avg=1, stride=1, frame_from=0, flatfield=None):
    print(f"side W={W} H={H} F={F} darks={darks} avg={avg} S={int(np.max(sq))}", flush=True)
    kw = dict(darks=darks, avg=avg, stride=stride, frame_from=frame_from, flatfield=flatfield)
    fs, qm = oracle_sums(orc, imm, dq, sq, F, **kw)
    plan = xmt.Plan(W, H, F, 4, dq, sq, stride=stride, average=avg, frame_from=frame_from, dark_frames=darks,
                    flatfield=flatfield, side_stats=True)
    plan.ingest_imm(imm)
    plan.execute()
    side = plan.side(window=1, frame_sum=True, partial=True)
    vscale = 1.0 / (avg * (darks if darks else 1))
    ppb = qm["pixels_per_bin"].astype(np.float64)
    want = fs * vscale / np.where(ppb > 0, ppb, 1.0)[None, :]
    np.testing.assert_allclose(side["partition_partial"], want, rtol=2e-6, atol=1e-30)
    np.testing.assert_allclose(side["frame_sum"][1], fs.sum(1) * vscale / (W * H), rtol=2e-6, atol=1e-30)
    return plan


def test_dense_darks_vector_loads(xmt, orc):
    # 64x64 frames: 16-byte aligned payloads, pixel count a multiple of 8
    W = H = 64
    F = 300
    dq, sq = synth.annular_qmaps(W, H, 12, 4)
    sig = synth.ou_counts(W, H, F, 4.0, 15.0, seed=701)
    dark = synth.dark_frames(W, H, 3, seed=702)
    imm = synth.imm_bytes_dense(np.concatenate([dark, sig + dark[0]], axis=0), W, H)
    check_side(xmt, orc, W, H, F, dq, sq, imm, darks=3)


@pytest.mark.parametrize("W,H,darks", [(32, 32, 2), (15, 15, 0)])
def test_dense_integer_sums(xmt, orc, W, H, darks):
    # F >= 512 without a flat field: the kernel keeps its sums in u32 until
    # the LDS table (16-byte loads for 32x32, per-pixel loads for 15x15)
    F = 600
    dq, sq = synth.annular_qmaps(W, H, 8, 4)
    sig = synth.ou_counts(W, H, F, 6.0, 20.0, seed=714 + W)
    frames = sig if not darks else np.concatenate([synth.dark_frames(W, H, darks, seed=715), sig + 2], axis=0)
    imm = synth.imm_bytes_dense(frames, W, H)
    check_side(xmt, orc, W, H, F, dq, sq, imm, darks=darks)


def test_dense_odd_pixels_flatfield(xmt, orc):
    # 15x15 = 225 pixels: per-pixel 2-byte loads, last lane partly outside
    W = H = 15
    F = 260
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    imm = synth.imm_bytes_dense(synth.ou_counts(W, H, F, 3.0, 10.0, seed=703), W, H)
    ff = np.random.default_rng(704).uniform(0.5, 1.5, (H, W))
    check_side(xmt, orc, W, H, F, dq, sq, imm, flatfield=ff)


def test_dense_average_stride(xmt, orc):
    W = H = 32
    F = 150
    dq, sq = synth.annular_qmaps(W, H, 8, 4)
    imm = synth.imm_bytes_dense(synth.ou_counts(W, H, 3 + 4 * F, 2.0, 10.0, seed=705), W, H)
    check_side(xmt, orc, W, H, F, dq, sq, imm, avg=2, stride=2, frame_from=3)


def test_dense_many_bins_broken_runs(xmt, orc):
    # 700 static bins of 4 pixels each scattered over the detector (the other
    # 272 pixels unmapped): bins change inside a lane's 8 pixels and between
    # most lanes, and 700 bins leave 5 frames per LDS table, so the frames
    # run in chunks
    W, H = 64, 48
    F = 220
    rng = np.random.default_rng(706)
    bins = np.concatenate([np.repeat(np.arange(1, 701), 4), np.zeros(W * H - 2800, np.int64)])
    sq = rng.permutation(bins).reshape(H, W).astype(np.int32)
    dq = np.where(sq > 0, 1 + (sq - 1) % 5, 0).astype(np.int32)
    imm = synth.imm_bytes_dense(synth.ou_counts(W, H, F, 3.0, 10.0, seed=707), W, H)
    check_side(xmt, orc, W, H, F, dq, sq, imm)


def test_dense_uniform_bins_long_segments(xmt, orc):
    # two static bins in halves of the detector: segments span whole waves
    W, H = 128, 16
    F = 200
    sq = np.ones((H, W), np.int32)
    sq[:, W // 2:] = 2
    dq = np.ones((H, W), np.int32)
    imm = synth.imm_bytes_dense(synth.ou_counts(W, H, F, 5.0, 10.0, seed=708), W, H)
    check_side(xmt, orc, W, H, F, dq, sq, imm)


def test_sparse_events(xmt, orc):
    # compressed frames: events in raster order, one workgroup per frame
    c = cases.sparse_case(W=48, H=40, F=400, mu=0.3, tau_c=8.0, seed=709, S=12, Q=4)
    check_side(xmt, orc, 48, 40, 400, c["dq"], c["sq"], c["imm"])


def test_sparse_events_flatfield_average(xmt, orc):
    c = cases.sparse_case(W=32, H=32, F=800, mu=0.4, tau_c=8.0, seed=710, S=9, Q=3)
    ff = np.random.default_rng(711).uniform(0.5, 1.5, (32, 32))
    check_side(xmt, orc, 32, 32, 190, c["dq"], c["sq"], c["imm"], avg=2, stride=2, frame_from=5, flatfield=ff)


def test_sparse_events_many_bins(xmt, orc):
    # 300 static bins (a 2-byte per-pixel bin table), scattered over the
    # detector so that neighbouring events rarely share a bin
    W, H = 48, 32
    F = 300
    rng = np.random.default_rng(712)
    bins = np.concatenate([np.repeat(np.arange(1, 301), 4), np.zeros(W * H - 1200, np.int64)])
    sq = rng.permutation(bins).reshape(H, W).astype(np.int32)
    dq = np.where(sq > 0, 1 + (sq - 1) % 6, 0).astype(np.int32)
    imm = synth.imm_bytes_sparse(synth.ou_counts(W, H, F, 0.3, 8.0, seed=713), W, H)
    check_side(xmt, orc, W, H, F, dq, sq, imm)
