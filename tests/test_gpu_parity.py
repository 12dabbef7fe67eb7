"""HIP path vs the CPU oracle on the same seeded inputs (bit-exact integer sums,
g2/stderr within 1e-6 relative of the FP64 oracle).  Calls go through the C ABI."""
import numpy as np
import pytest

import cases
import synth

pytestmark = pytest.mark.gpu

G2_RTOL = 1e-6      # g2(q, tau) vs the FP64 oracle
REF_RTOL = 1e-4     # reference float units vs the float replica of the reference


def run_gpu(xmt, c, **kw):
    plan = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"], **kw)
    plan.ingest_imm(c["imm"])
    plan.execute()
    return plan


def assert_g2_close(got, want, rtol=G2_RTOL):
    assert got.shape == want.shape
    nan_w = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan_w), "NaN pattern differs"
    g, w = got[~nan_w], want[~nan_w]
    rel = np.abs(g - w) / np.maximum(np.abs(w), 1e-300)
    assert rel.max(initial=0.0) <= rtol, f"max rel err {rel.max()}"


def check_against_oracle(xmt, orc, c, **kw):
    print(f"case W={c['W']} H={c['H']} F={c['F']} dpl={c['dpl']} kw={kw}", flush=True)
    plan = run_gpu(xmt, c, **kw)
    okw = {k: v for k, v in kw.items() if k in ("stride", "avg", "frame_from", "darks", "flatfield",
                                               "lld", "sigma")}
    if "average" in kw:
        okw["avg"] = kw["average"]
    if "dark_frames" in kw:
        okw["darks"] = kw["dark_frames"]
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"], **okw)
    lev, tau = plan.schedule()
    assert np.array_equal(lev, r["level"]) and np.array_equal(tau, r["tau"])
    assert np.array_equal(plan.pixel_order(), r["qm"]["order"])
    G2, IP, IF = plan.sums()
    assert np.array_equal(IP, r["IP"]), "IP differs"
    assert np.array_equal(IF, r["IF"]), "IF differs"
    assert np.array_equal(G2, r["G2"]), "G2 differs"
    g2, se = plan.g2()
    assert_g2_close(g2, r["g2"])
    assert_g2_close(se, r["stderr"])
    return plan, r


def test_small_sparse(xmt, orc):
    check_against_oracle(xmt, orc, cases.sparse_case())


def test_counterexample_exact(xmt, orc):
    c = cases.counterexample_case()
    plan, r = check_against_oracle(xmt, orc, c)
    lev, tau = plan.schedule()
    i = int(np.flatnonzero((lev == 1) & (tau == 10))[0])
    G2, _, _ = plan.sums()
    assert G2[0, i] == 16          # 4 pairs of binned value 2*2 (summed-bin units)


def test_config1_plain(xmt, orc):
    """256x256, 1000 frames, mu=0.012, tau_c=20, dpl=4, 18 static / 6 dynamic rings."""
    c = cases.sparse_case(W=256, H=256, F=1000, dpl=4, mu=0.012, tau_c=20.0, seed=1001, S=18, Q=6)
    plan, r = check_against_oracle(xmt, orc, c)
    assert plan.T == 34


@pytest.mark.parametrize("F,dpl", [(3000, 4), (16000, 8)])
def test_register_rows_boundary_ranks(xmt, orc, F, dpl):
    """The register row kernel finds each event's boundary rank in per-64-frame
    words staged in LDS when they fit beside its accumulators (3000 frames) and
    in the per-frame table otherwise (16000 frames: 250 words, more than the
    LDS room that keeps its occupancy); both give the oracle's sums."""
    c = cases.sparse_case(W=24, H=24, F=F, dpl=dpl, mu=0.004, tau_c=40.0, seed=41 + F, S=8, Q=4)
    plan, r = check_against_oracle(xmt, orc, c)
    px, _ = plan.row_stats()
    assert px[1] > 0


def test_long_rows(xmt, orc):
    """High occupancy: rows longer than the on-chip capacity take the dense-row kernel."""
    c = cases.sparse_case(W=16, H=16, F=3000, dpl=4, mu=2.0, tau_c=30.0, seed=3)
    plan, _ = check_against_oracle(xmt, orc, c)
    assert plan.info().long_pixels > 0


@pytest.mark.parametrize("F,mu,cls", [(4000, 0.2, 4), (3000, 1.0, 4), (7000, 3.0, 5)])
def test_long_row_classes(xmt, orc, F, mu, cls):
    """Rows of 513..1024 and 1025..2048 events (the two larger LDS kernels,
    one list shared with the series kernel) and longer rows (series kernel),
    exact against the oracle; the row statistics see the class."""
    c = cases.sparse_case(W=12, H=12, F=F, dpl=4, mu=mu, tau_c=20.0, seed=31 + F)
    plan, r = check_against_oracle(xmt, orc, c)
    px, ev = plan.row_stats()
    n = np.diff(r["csr"]["off"])
    want = np.array([(n == 0).sum(), ((n >= 1) & (n <= 64)).sum(), ((n > 64) & (n <= 128)).sum(),
                     ((n > 128) & (n <= 512)).sum(), ((n > 512) & (n <= 2048)).sum(), (n > 2048).sum()])
    assert np.array_equal(px, want) and ev.sum() == n.sum()
    assert px[cls] > 0


@pytest.mark.parametrize("dpl,F,mu", [(4, 1000, 0.05), (8, 3000, 0.03), (8, 2500, 0.03)])
def test_deep_level_handover(xmt, orc, dpl, F, mu):
    """The deep levels of rows up to 128 events run one lane per pixel from the
    first level with at most deep_frames binned frames: every hand-over level
    (none, 16, 48, 96 frames) gives the oracle's sums bit for bit, rows of the
    one- and two-register kernels alike."""
    c = cases.sparse_case(W=24, H=24, F=F, dpl=dpl, mu=mu, tau_c=15.0, seed=61 + F + dpl)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], F, dpl)
    n = np.diff(r["csr"]["off"])
    assert (n > 64).any() and ((n >= 1) & (n <= 64)).any()
    levels = []
    nlev = len(set(orc.schedule(F, dpl)[0].tolist()))
    for frames in (-1, 16, 48, 96):
        plan = xmt.Plan(c["W"], c["H"], F, dpl, c["dq"], c["sq"])
        plan.set_tuning(deep_frames=frames)
        plan.ingest_imm(c["imm"])
        plan.execute()
        G2, IP, IF = plan.sums()
        assert np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"])
        assert np.array_equal(G2, r["G2"]), f"G2 differs with deep_frames={frames}"
        g2, _ = plan.g2()
        assert_g2_close(g2, r["g2"])
        info = plan.info()
        levels.append(info.deep_level)
        if info.deep_level < nlev:      # a hand-over: a level with at most `frames` binned frames
            assert frames > 0 and (F >> info.deep_level) <= frames and info.deep_stride >= (F >> info.deep_level)
        else:
            assert info.deep_stride == 0
        plan.close()
    assert levels[0] == nlev                       # none: the level count
    assert nlev >= levels[1] >= levels[2] >= levels[3] >= 1 and levels[3] < nlev


@pytest.mark.parametrize("dpl", [4, 8])
def test_deep_level_large_totals(xmt, orc, dpl):
    """Deep levels of rows with large counts, handed over at 96 frames: a
    wave holding pixels of total 90000 (30 events) and 3000000 (100 events,
    deep lag sums beyond 2^32) and a wave whose largest pixel totals 65400 in
    20 consecutive frames, exact against the oracle."""
    F = 2500
    c = cases.sparse_case(W=24, H=24, F=F, dpl=dpl, mu=0.03, tau_c=15.0, seed=67 + dpl)
    order = orc.run_exact(c["imm"], c["dq"], c["sq"], F, dpl)["qm"]["order"]
    counts = c["counts"]
    rng = np.random.default_rng(5 + dpl)
    for k, nev, val in ((70, 30, 3000), (71, 100, 30000)):      # the second wave
        counts[:, order[k]] = 0
        counts[np.sort(rng.choice(F, nev, replace=False)), order[k]] = val
    counts[:, order[200]] = 0                                     # the fourth wave
    counts[1000:1020, order[200]] = 3270
    c["imm"] = synth.imm_bytes_sparse(counts, 24, 24)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], F, dpl)
    assert np.array_equal(r["qm"]["order"], order)
    tot = counts.astype(np.int64).sum(axis=0)[order]
    assert tot[64:128].max() == 3000000 and tot[192:256].max() == 65400
    assert np.diff(r["csr"]["off"]).max() <= 128
    plan = xmt.Plan(c["W"], c["H"], F, dpl, c["dq"], c["sq"])
    plan.set_tuning(deep_frames=96)
    plan.ingest_imm(c["imm"])
    plan.execute()
    info = plan.info()
    assert info.deep_stride > 0
    G2, IP, IF = plan.sums()
    assert np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"])
    assert np.array_equal(G2, r["G2"])
    lev, _ = plan.schedule()
    assert int(G2[:, lev >= info.deep_level].max()) >= 2 ** 32     # 64-bit deep lag sums reached
    plan.close()


def test_mid_rows(xmt, orc):
    """Rows of 65..512 events take the LDS kernel."""
    c = cases.sparse_case(W=16, H=16, F=2000, dpl=4, mu=0.08, tau_c=12.0, seed=13)
    check_against_oracle(xmt, orc, c)


def test_transpose_staged_ranges(xmt, orc):
    """About 100 events per pixel: buckets of 256 pixels hold ~25 k events, so
    the row placement assembles each bucket in several LDS ranges."""
    c = cases.sparse_case(W=32, H=32, F=2000, dpl=4, mu=0.05, tau_c=10.0, seed=21)
    ev = (c["counts"] > 0).sum(axis=0)
    assert 8192 < ev[:256].sum() < 4 * 7000
    check_against_oracle(xmt, orc, c)


def test_transpose_hot_pixel_direct(xmt, orc):
    """A pixel with a count in every one of 5000 frames: its bucket's rows are
    placed straight into HBM (a row longer than half the staging buffer)."""
    c = cases.sparse_case(W=16, H=16, F=5000, dpl=4, mu=0.01, tau_c=20.0, seed=22)
    c["counts"][:, 37] = np.maximum(c["counts"][:, 37], 1)
    c["imm"] = synth.imm_bytes_sparse(c["counts"], 16, 16)
    check_against_oracle(xmt, orc, c)


@pytest.mark.parametrize("bshift,cf", [(8, 20), (11, 7), (12, 50)])
def test_transpose_bucket_sizes(xmt, orc, bshift, cf):
    """Row placement with 256-pixel buckets (one wave per slice) and with
    2048 / 4096-pixel buckets (wave groups sharing a slice, events beyond the
    register-held ones re-read per range), exact against the oracle."""
    c = cases.sparse_case(W=64, H=64, F=400, dpl=4, mu=0.02, tau_c=10.0, seed=23)
    print(f"bucket shift {bshift}, chunk frames {cf}", flush=True)
    plan = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"])
    plan.set_tuning(tr_bucket_shift=bshift, tr_chunk_frames=cf)
    plan.ingest_imm(c["imm"])
    plan.execute()
    assert plan.get_tuning()["tr_bucket_shift"] == bshift
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    G2, IP, IF = plan.sums()
    assert np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"]) and np.array_equal(G2, r["G2"])


def test_many_taus(xmt, orc):
    """More than 128 taus: every non-empty row takes the LDS kernel (dpl=16)."""
    c = cases.sparse_case(W=12, H=12, F=5000, dpl=16, mu=0.01, tau_c=30.0, seed=17)
    plan, _ = check_against_oracle(xmt, orc, c)
    assert plan.T > 128


def test_mixed_rows(xmt, orc):
    c = cases.sparse_case(W=24, H=24, F=2000, dpl=8, mu=0.3, tau_c=5.0, seed=11)
    check_against_oracle(xmt, orc, c)


def test_stride_average_skip(xmt, orc):
    c = cases.sparse_case(W=20, H=20, F=1300, dpl=4, mu=0.2, tau_c=8.0, seed=5)
    c["F"] = 200
    check_against_oracle(xmt, orc, c, stride=2, average=3, frame_from=4)


def test_dense_with_darks(xmt, orc):
    W = H = 24
    F = 400
    dq, sq = synth.annular_qmaps(W, H, 9, 3)
    sig = synth.ou_counts(W, H, F, 5.0, 20.0, seed=1003)
    dark = synth.dark_frames(W, H, 2, seed=2003)
    frames = np.concatenate([dark, sig + dark[0]], axis=0)
    imm = synth.imm_bytes_dense(frames, W, H)
    c = dict(W=W, H=H, F=F, dpl=4, dq=dq, sq=sq, imm=imm)
    plan, r = check_against_oracle(xmt, orc, c, dark_frames=2)
    assert plan.info().long_pixels == plan.P


def dense_case(W, H, F_raw, dpl, mu, tau_c, seed, darks=2, S=9, Q=3):
    dq, sq = synth.annular_qmaps(W, H, S, Q)
    sig = synth.ou_counts(W, H, F_raw, mu, tau_c, seed=seed)
    dark = synth.dark_frames(W, H, darks, seed=seed + 1000) if darks else np.zeros((0, W * H), np.uint16)
    base = dark[0] if darks else 0
    frames = np.concatenate([dark, sig + base], axis=0)
    return dict(W=W, H=H, F=F_raw, dpl=dpl, dq=dq, sq=sq, imm=synth.imm_bytes_dense(frames, W, H))


@pytest.mark.parametrize("dpl", [4, 2])
def test_dense_darks_above_int16(xmt, orc, dpl):
    """Dark frames straddling 32768 at some pixels (32767 and 32769: uint16
    standard deviation 1) with a sigma threshold: the streaming kernel (dpl 4)
    and the transpose path (dpl 2) keep and subtract exactly as the oracle."""
    W = H = 16
    F = 600
    dq, sq = synth.annular_qmaps(W, H, 9, 3)
    sig = synth.ou_counts(W, H, F, 3.0, 15.0, seed=57).astype(np.int64)
    dark = np.full((2, W * H), 3, np.int64)
    hot = np.arange(5, W * H, 23)
    dark[0, hot], dark[1, hot] = 32767, 32769
    base = np.where(np.isin(np.arange(W * H), hot), 32768, 3)
    frames = np.concatenate([dark, sig + base[None, :]]).astype(np.uint16)
    c = dict(W=W, H=H, F=F, dpl=dpl, dq=dq, sq=sq, imm=synth.imm_bytes_dense(frames, W, H))
    check_against_oracle(xmt, orc, c, dark_frames=2, sigma=1.0, lld=0.5)


def test_dense_dpl8_two_passes(xmt, orc):
    """Streaming dense kernel, dpl=8, 9 levels (two passes of the level cascade)."""
    c = dense_case(16, 16, 3000, 8, 3.0, 40.0, seed=41)
    check_against_oracle(xmt, orc, c, dark_frames=2)


@pytest.mark.parametrize("F,dpl", [(13, 4), (37, 4), (1007, 4), (4099, 4), (29, 8), (2061, 8)])
def test_dense_partial_groups(xmt, orc, F, dpl):
    """Frame counts that are not multiples of the kernel's 16-frame group, on
    the raw pass and on the series passes (including series shorter than one
    group)."""
    c = dense_case(12, 12, F, dpl, 2.0, 12.0, seed=53 + F)
    check_against_oracle(xmt, orc, c, dark_frames=2)


def test_dense_hot_pixels_exact_redo(xmt, orc):
    """Pixels whose total count exceeds the exact-FP64 bound (94906265) are
    recomputed with 64-bit integer lag sums; results stay bit-exact."""
    W, H, F = 16, 16, 2500
    dq, sq = synth.annular_qmaps(W, H, 9, 3)
    sig = synth.ou_counts(W, H, F, 3.0, 20.0, seed=61).astype(np.int64)
    rng = np.random.default_rng(62)
    hot = rng.choice(np.flatnonzero((dq.reshape(-1) > 0) & (sq.reshape(-1) > 0)), 7, replace=False)
    sig[:, hot] = rng.integers(45000, 65000, size=(F, hot.size))
    c = dict(W=W, H=H, F=F, dpl=4, dq=dq, sq=sq, imm=synth.imm_bytes_dense(sig.astype(np.uint16), W, H))
    plan, r = check_against_oracle(xmt, orc, c)
    assert plan.info().exact_pixels == hot.size


def test_dense_fallback_dpl2(xmt, orc):
    """dpl outside the streaming kernel's instantiations takes the dense-row kernel."""
    c = dense_case(16, 16, 500, 2, 4.0, 10.0, seed=43)
    check_against_oracle(xmt, orc, c, dark_frames=2)


def test_dense_no_darks_average(xmt, orc):
    c = dense_case(12, 20, 900, 4, 1.0, 15.0, seed=47, darks=0)
    c["F"] = 300
    check_against_oracle(xmt, orc, c, stride=1, average=3)


def test_flatfield(xmt, orc):
    c = cases.sparse_case(W=32, H=32, F=600, dpl=4, mu=0.1, tau_c=10.0, seed=9)
    ff = np.random.default_rng(0).uniform(0.5, 1.5, (32, 32))
    check_against_oracle(xmt, orc, c, flatfield=ff)


def test_ref_units_vs_float_replica(xmt, orc):
    """--g2out arrays vs the float replica of multiTau2 without the stale-tail search."""
    c = cases.sparse_case(W=32, H=32, F=500, dpl=4, mu=0.05, tau_c=10.0, seed=7)
    plan = run_gpu(xmt, c)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    G2r, IPr, IFr = plan.ref_f32()
    G2c, IPc, IFc = orc.multitau_compat(r["csr"], r["qm"]["order"], c["W"] * c["H"], c["F"],
                                        r["level"], r["tau"], stale_tail=False)
    for got, want in ((G2r, G2c), (IPr, IPc), (IFr, IFc)):
        rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
        assert rel.max() <= REF_RTOL


@pytest.mark.parametrize("ff", [False, True])
def test_normalize_ref_f32_arrays(xmt, orc, ff):
    """Corr::normalizeG2s(G2, IP, IF) on reference-unit float arrays, by a
    plan that ingested nothing: the arrays read back from an executed plan
    normalise to that plan's g2 (within float rounding of the arrays), to the
    FP64 normalisation of the same arrays, and to the reference's float
    normalisation of them."""
    c = cases.sparse_case(W=32, H=32, F=600, dpl=4, mu=0.1, tau_c=10.0, seed=9)
    flat = np.random.default_rng(0).uniform(0.5, 1.5, (32, 32)) if ff else None
    plan = run_gpu(xmt, c, flatfield=flat)
    G2r, IPr, IFr = plan.ref_f32()
    g2, se = plan.g2()
    fresh = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"], flatfield=flat)
    g2n, sen = fresh.normalize_ref_f32(G2r, IPr, IFr)
    assert np.array_equal(fresh.g2()[0], g2n, equal_nan=True)
    assert_g2_close(g2n, g2, 1e-5)
    # stderr of cells whose exact ratios are all equal is 0; from float arrays
    # with a flat field the ratios differ by float rounding
    np.testing.assert_allclose(sen, se, rtol=REF_RTOL, atol=1e-7)
    qm = orc.build_qmap(c["dq"], c["sq"])
    rows = [np.ascontiguousarray(a[:, qm["order"]].T, np.float64) for a in (G2r, IPr, IFr)]
    g2w, sew = orc.normalize_f64_sums(*rows, np.ones(plan.T), qm)
    assert_g2_close(g2n, g2w, 1e-9)
    np.testing.assert_allclose(sen, sew, rtol=1e-9, atol=1e-13)
    g2f, sef = orc.normalize_compat(G2r, IPr, IFr, qm)
    assert_g2_close(g2n, g2f.astype(np.float64), REF_RTOL)
    a = np.ascontiguousarray(G2r)
    out = np.zeros(plan.T * plan.Q)
    st = xmt.lib().xpcs_mt_normalize_ref_f32(fresh._h, xmt._p(a), xmt._p(a), xmt._p(a), a.shape[1], plan.T - 1,
                                             xmt._p(out), xmt._p(out))
    assert st == 1 and b"taus" in xmt.lib().xpcs_last_error()
    fresh.close()
    plan.close()


def test_sparse_detector_beyond_transpose_limit_is_refused(xmt):
    """4100 x 4100 raster pixels need transpose buckets of more than 4096 rows:
    execute refuses the job (nothing is launched)."""
    W = H = 4100
    dq = np.zeros((H, W), np.int32)
    sq = np.zeros((H, W), np.int32)
    dq[:2, :2] = 1
    sq[:2, :2] = 1
    ev = [(np.array([0, W * H - 1], np.int32), np.array([1, 1], np.uint16))] * 8
    plan = xmt.Plan(W, H, 8, 2, dq, sq)
    plan.ingest_imm(synth.imm_bytes_events(ev, W, H))
    with pytest.raises(xmt.XpcsError) as e:
        plan.execute()
    assert e.value.status == 5 and "4096 x 4096" in str(e.value)
    plan.close()


def test_device_ingest_checks_dark_frames(xmt):
    """Dark frames of a device-resident dense image are bounds- and size-checked
    before anything is read."""
    import torch
    W = H = 8
    frames = np.full((6, W * H), 4, np.uint16)
    imm = synth.imm_bytes_dense(frames, W, H)
    dev = torch.from_numpy(np.frombuffer(imm, np.uint8).copy()).cuda()
    idx = xmt.imm_index(imm)
    off, dl = idx["payload_off"].copy(), idx["dlen"].copy()
    dq, sq = synth.annular_qmaps(W, H, 2, 1)
    plan = xmt.Plan(W, H, 3, 2, dq, sq, dark_frames=2)
    bad = off.copy()
    bad[1] = len(imm) - 10
    with pytest.raises(xmt.XpcsError) as e:
        plan.ingest_imm_device(dev.data_ptr(), dev.numel(), bad, dl, compressed=False)
    assert e.value.status == 3 and "dark frame 1" in str(e.value)
    short = dl.copy()
    short[0] = 5
    with pytest.raises(xmt.XpcsError) as e:
        plan.ingest_imm_device(dev.data_ptr(), dev.numel(), off, short, compressed=False)
    assert e.value.status == 3 and "dark frame 0" in str(e.value)
    plan.ingest_imm_device(dev.data_ptr(), dev.numel(), off, dl, compressed=False)
    plan.execute()
    plan.g2()
    plan.close()


def test_pixel_out_of_range_is_an_error(xmt):
    W = H = 8
    ev = [(np.array([3, 70], np.int32), np.array([1, 1], np.uint16))] * 20
    imm = synth.imm_bytes_events(ev, W, H)
    dq, sq = synth.annular_qmaps(W, H, 2, 1)
    plan = xmt.Plan(W, H, 20, 2, dq, sq)
    plan.ingest_imm(imm)
    plan.execute()
    with pytest.raises(xmt.XpcsError) as e:
        plan.g2()
    assert e.value.status == 4


def test_repeated_execute_is_stable(xmt, orc):
    c = cases.sparse_case(W=32, H=32, F=400, dpl=4, mu=0.08, tau_c=6.0, seed=21)
    plan = run_gpu(xmt, c)
    a = plan.sums()
    plan.execute()
    b = plan.sums()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("dpl,deep,row_order", [(4, 96, "sorted"), (8, 96, "local"), (8, -1, "shuffled"),
                                                 (4, 48, "local")])
def test_lds_rows_dense_levels_and_large_totals(xmt, orc, dpl, deep, row_order):
    """Rows of 129..2048 events handed over pixel-major: levels that turn
    dense run as LDS series, the deep levels go to the one-lane-per-pixel
    kernel, and rows whose total passes 2^32 (counts of 3e6, 1500 events)
    stay in the LDS kernel to the last level; rows in frame order, permuted
    within stretches of 40 events (what the transpose leaves: the register
    block sorts), or fully shuffled (the whole-row network); exact against
    the oracle."""
    rng = np.random.default_rng(300 + dpl + deep + len(row_order))
    W = H = 10
    F = 3000
    dq, sq = synth.annular_qmaps(W, H, 4, 2)
    npd = W * H
    lens = rng.integers(129, 2049, npd)
    lens[:8] = rng.integers(1, 129, 8)                 # a few register rows beside them
    big = {20, 21, 55}
    frames, counts = [], []
    for j, n in enumerate(lens):
        n = min(int(n), F)
        lens[j] = n
        frames.append(np.sort(rng.choice(F, n, replace=False)).astype(np.int32))
        c = rng.integers(1, 40, n).astype(np.uint32)
        if j in big:
            lens[j] = n = 1500
            frames[-1] = np.sort(rng.choice(F, n, replace=False)).astype(np.int32)
            c = np.full(n, 3_000_000, np.uint32)        # total 4.5e9 > 2^32
        counts.append(c)
    for j in range(npd):
        perm = np.arange(lens[j])
        if row_order == "local":
            for a in range(0, lens[j], 40):
                rng.shuffle(perm[a:a + 40])
        elif row_order == "shuffled":
            rng.shuffle(perm)
        frames[j] = frames[j][perm]
        counts[j] = counts[j][perm]
    roff = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    print(f"LDS rows dpl={dpl} deep_frames={deep} order={row_order}", flush=True)
    plan = xmt.Plan(W, H, F, dpl, dq, sq)
    plan.set_tuning(deep_frames=deep)
    plan.ingest_rows(roff, np.concatenate(frames), np.concatenate(counts))
    plan.execute()
    qm = orc.build_qmap(dq, sq)
    order = qm["order"]
    off = np.concatenate([[0], np.cumsum(lens[order])]).astype(np.int64)
    csr = dict(off=off, frame=np.concatenate([frames[j] for j in order]),
               cnt=np.concatenate([counts[j] for j in order]).astype(np.int64))
    lev, tau = orc.schedule(F, dpl)
    G2, IP, IF = orc.multitau_exact(csr, F, lev, tau)
    g, a, b = plan.sums()
    assert np.array_equal(a, IP) and np.array_equal(b, IF)
    assert np.array_equal(g, G2)
    px, _ = plan.row_stats()
    assert px[3] > 0 and px[4] > 0
    if deep > 0:
        assert plan.info().deep_stride > 0
    plan.close()


@pytest.mark.parametrize("wide", [False, True])
def test_sum_layouts_and_wide_rows(xmt, orc, wide):
    """IP/IF kept as 32-bit words (high words only for pixels whose total
    passes 2^32 - 1) against int64 sums: rows of every class (register,
    LDS, beyond the LDS capacity) with totals on both sides of 2^32; with
    `wide`, rows of at most 128 events total more than 2^32 (u32 counts
    handed over pixel-major) and the register kernels step aside.  Sums
    exact against the oracle, g2/stderr identical between the layouts and
    close to the oracle's normalisation."""
    rng = np.random.default_rng(77 + wide)
    W = H = 12
    F = 4000
    dq, sq = synth.annular_qmaps(W, H, 4, 2)
    npd = W * H
    lens = rng.integers(1, 129, npd)
    lens[40:60] = rng.integers(129, 2049, 20)          # LDS rows
    lens[60:63] = [2600, 3100, 3900]                   # beyond the LDS capacity
    lens[[41, 45]] = 1500
    if wide:
        lens[[3, 7]] = 5
    # large rows total 4.5e9 .. 5e9: above 2^32, and every lag sum stays below
    # total^2 / 4 < 2^63
    big = {41: 3_000_000, 45: 3_000_000, 61: 1_500_000}
    if wide:
        big.update({3: 1_000_000_000, 7: 1_000_000_000})
    frames, counts = [], []
    for j, n in enumerate(lens):
        frames.append(np.sort(rng.choice(F, int(n), replace=False)).astype(np.int32))
        c = rng.integers(1, 50, int(n)).astype(np.uint32)
        if j in big:
            c[:] = big[j]
        counts.append(c)
    roff = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    qm = orc.build_qmap(dq, sq)
    order = qm["order"]
    off = np.concatenate([[0], np.cumsum(lens[order])]).astype(np.int64)
    csr = dict(off=off, frame=np.concatenate([frames[j] for j in order]),
               cnt=np.concatenate([counts[j] for j in order]).astype(np.int64))
    lev, tau = orc.schedule(F, 4)
    G2, IP, IF = orc.multitau_exact(csr, F, lev, tau)
    g2o, seo = orc.normalize_f64(G2, IP, IF, orc.schedule_N(F, lev, tau), qm)
    res = []
    for sums64 in (0, 1):
        print(f"sum layout sums64={sums64} wide={wide}", flush=True)
        plan = xmt.Plan(W, H, F, 4, dq, sq)
        plan.set_tuning(sums64=sums64)
        plan.ingest_rows(roff, np.concatenate(frames), np.concatenate(counts))
        if sums64:
            plan.autotune()             # launch shapes only: the layout stays
        plan.execute()
        g, a, b = plan.sums()
        assert np.array_equal(a, IP) and np.array_equal(b, IF)
        assert np.array_equal(g, G2)
        res.append(plan.g2())
        assert plan.get_tuning()["sums64"] == sums64
        # the readers follow the layout of the run that wrote the sums
        plan.set_tuning(sums64=1 - sums64)
        g, a, b = plan.sums()
        assert np.array_equal(a, IP) and np.array_equal(b, IF)
        plan.close()
    tot = np.array([int(c.astype(np.int64).sum()) for c in counts])
    assert tot.max() > 2 ** 32 and (tot[lens <= 128].max() > 2 ** 32) == wide
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)
    assert_g2_close(res[0][0], g2o)
    assert_g2_close(res[0][1], seo)


@pytest.mark.parametrize("seed", [1, 2])
def test_shuffled_rows_two_per_lane(xmt, orc, seed):
    """Rows of 65..128 events in random frame order (the two-events-per-lane
    kernel's full sort), some frames repeated (merged at level 0)."""
    rng = np.random.default_rng(900 + seed)
    W = H = 12
    F = 700
    dq, sq = synth.annular_qmaps(W, H, 4, 2)
    npd = W * H
    lens = rng.integers(60, 129, npd)
    frames, counts = [], []
    for n in lens:
        fr = np.sort(rng.choice(F, n, replace=False))
        if n > 70:
            fr[5] = fr[4]                       # a repeated frame in the row
        rng.shuffle(fr)
        frames.append(fr.astype(np.int32))
        counts.append(rng.integers(1, 6, n).astype(np.uint32))
    roff = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    fr_all = np.concatenate(frames)
    ct_all = np.concatenate(counts)
    print(f"shuffled rows seed={seed}", flush=True)
    plan = xmt.Plan(W, H, F, 4, dq, sq)
    plan.ingest_rows(roff, fr_all, ct_all)
    plan.execute()
    qm = orc.build_qmap(dq, sq)
    order = qm["order"]
    off = np.concatenate([[0], np.cumsum(lens[order])]).astype(np.int64)
    csr = dict(off=off, frame=np.concatenate([frames[j] for j in order]),
               cnt=np.concatenate([counts[j] for j in order]).astype(np.int64))
    lev, tau = orc.schedule(F, 4)
    G2, IP, IF = orc.multitau_exact(csr, F, lev, tau)
    g, a, b = plan.sums()
    assert np.array_equal(a, IP) and np.array_equal(b, IF) and np.array_equal(g, G2)
