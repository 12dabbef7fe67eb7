"""The drop-in `corr` tool: configuration errors (CPU) and whole runs on the GPU
checked against the oracle, with the reference's output datasets, names,
shapes and dtypes (main.cpp:339-452, corr.cpp:1089-1090)."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import synth
import xpcs_h5

ROOT = Path(__file__).resolve().parent.parent
CORR = ROOT / "xpcs-eigen_amd" / "corr"


def make_job(tmp, W=24, H=24, F=300, dpl=4, mu=0.15, tau_c=8.0, seed=5, S=6, Q=3, **cfg):
    dq, sq = synth.annular_qmaps(W, H, S, Q)
    counts = synth.ou_counts(W, H, F, mu, tau_c, seed)
    imm = tmp / "data.imm"
    imm.write_bytes(synth.imm_bytes_sparse(counts, W, H))
    conf = tmp / "cfg.h5"
    xpcs_h5.write_config(conf, width=W, height=H, dqmap=dq, sqmap=sq, imm_path=imm,
                         data_end_todo=F, dpl=dpl, **cfg)
    return dict(W=W, H=H, F=F, dpl=dpl, dq=dq, sq=sq, counts=counts, imm=imm, conf=conf)


def run(*args):
    return subprocess.run([str(CORR), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_missing_key_is_reported(tmp_path):
    job = make_job(tmp_path, F=50)
    with xpcs_h5.H5(job["conf"]) as h:
        pass
    # rewrite the configuration without delays_per_level
    conf2 = tmp_path / "bad.h5"
    with xpcs_h5.H5(job["conf"]) as src, xpcs_h5.H5(conf2, create=True) as dst:
        for k in ("/xpcs/compression", "/xpcs/output_data"):
            dst.write_string(k, src.read_string(k))
    r = run(conf2)
    assert r.returncode == 1 and "configuration key missing" in r.stderr


def test_twotime_unknown_smoothing_is_refused(tmp_path):
    """main.cpp:161-164: two-time with a smoothing method other than StaticMap
    or symmetric (any case) stops with status 1 before any work."""
    job = make_job(tmp_path, F=50, analysis_type="Twotime", smoothing_method="Gaussian")
    r = run(job["conf"])
    assert r.returncode == 1 and "Smoothing method is not valid" in r.stderr


def test_twotime_keys_are_required(tmp_path):
    job = make_job(tmp_path, F=50)
    with xpcs_h5.H5(job["conf"]) as h:
        h.write_string("/xpcs/analysis_type", "Twotime")
    r = run(job["conf"])
    assert r.returncode == 1 and "configuration key missing: /xpcs/smoothing_method" in r.stderr


def test_usage():
    """No configuration argument, or an unknown flag: status 1."""
    assert run().returncode == 1
    assert run("--bogus", "x").returncode == 1


def test_h5_dtype_query(tmp_path):
    """xpcs_h5_dtype reports the stored element type of every kind of dataset
    the writer makes, so the output contract's dtypes can be checked."""
    f = tmp_path / "t.h5"
    with xpcs_h5.H5(f, create=True) as h:
        for dt in (np.int32, np.int64, np.float32, np.float64, np.uint16, np.uint32):
            h.write(f"/d/{np.dtype(dt).name}", np.arange(6, dtype=dt).reshape(2, 3))
        h.write_string("/d/s", "ENABLED")
    with xpcs_h5.H5(f) as h:
        for dt in (np.int32, np.int64, np.float32, np.float64, np.uint16, np.uint32):
            assert h.dtype(f"/d/{np.dtype(dt).name}") == np.dtype(dt)
        assert h.dtype("/d/s") is str
        with pytest.raises(RuntimeError):
            h.dtype("/d/missing")


# dtypes of the reference's outputs (h5_result.cpp:56-347: float32, float64 for
# the double overload; main.cpp:339-478)
OUT_DTYPES = {"pixelSum": "f4", "frameSum": "f4", "partition-mean-total": "f4",
              "partition-mean-partial": "f4", "partition_norm_factor": "f4", "timestamp_clock": "f8",
              "timestamp_tick": "f8", "tau": "f4", "norm-0-g2": "f4", "norm-0-stderr": "f4", "G2": "f4",
              "IP": "f4", "IF": "f4"}


@pytest.mark.gpu
def test_cli_outputs_match_oracle(tmp_path, orc):
    job = make_job(tmp_path, static_window=40)
    print("corr job", job["W"], job["H"], job["F"], flush=True)
    r = run(job["conf"], "--g2out")
    assert r.returncode == 0, r.stderr
    ref = orc.run_exact(job["imm"].read_bytes(), job["dq"], job["sq"], job["F"], job["dpl"])
    W, H, F = job["W"], job["H"], job["F"]
    T, Q = ref["g2"].shape
    with xpcs_h5.H5(job["conf"]) as h:
        g2 = h.read("/exchange/norm-0-g2", np.float32)
        se = h.read("/exchange/norm-0-stderr", np.float32)
        assert g2.shape == (T, Q) and se.shape == (T, Q)
        ok = ~np.isnan(ref["g2"])
        assert np.allclose(g2[ok], ref["g2"][ok], rtol=1e-6)
        assert np.allclose(se, ref["stderr"], rtol=1e-6, atol=1e-12)
        tau = h.read("/exchange/tau", np.float32)
        assert tau.shape == (1, T) and np.array_equal(tau[0], ref["tau"].astype(np.float32))
        # side statistics from the counts themselves
        c = job["counts"].astype(np.float64)            # [F, W*H]
        mask = (job["dq"].reshape(-1) > 0) & (job["sq"].reshape(-1) > 0)
        cm = c * mask[None, :]
        pix = h.read("/exchange/pixelSum", np.float32)
        assert pix.shape == (H, W) and np.allclose(pix.reshape(-1), cm.sum(0) / F, rtol=1e-6)
        fsum = h.read("/exchange/frameSum", np.float32)
        assert fsum.shape == (2, F)
        assert np.array_equal(fsum[0], np.arange(1, F + 1, dtype=np.float32))
        assert np.allclose(fsum[1], cm.sum(1) / (W * H), rtol=1e-6)
        sqf = job["sq"].reshape(-1)
        S = int(sqf.max())
        ppb = np.array([np.sum(mask & (sqf == s)) for s in range(1, S + 1)])
        tot = np.array([cm[:, sqf == s].sum() for s in range(1, S + 1)])
        pmt = h.read("/exchange/partition-mean-total", np.float32)
        assert pmt.shape == (1, S) and np.allclose(pmt[0], tot / (ppb * F), rtol=1e-6)
        pmp = h.read("/exchange/partition-mean-partial", np.float32)
        assert pmp.shape == (F // 40, S)
        for w in range(F // 40):
            part = np.array([cm[w * 40:(w + 1) * 40, sqf == s].sum() for s in range(1, S + 1)])
            assert np.allclose(pmp[w], part / (ppb * 40), rtol=1e-6)
        clk = h.read("/exchange/timestamp_clock")
        assert clk.shape == (2, F) and np.allclose(clk[1], np.arange(F) * 1e-3)
        assert h.dims("/exchange/partition_norm_factor") == (1, 1)
        for name, dt in OUT_DTYPES.items():
            assert h.dtype("/exchange/" + name) == np.dtype(dt), name
        G2r = h.read("/exchange/G2", np.float32)
        assert G2r.shape == (T, W * H)
        G2c, _, _ = orc.multitau_compat(ref["csr"], ref["qm"]["order"], W * H, F, ref["level"], ref["tau"],
                                        stale_tail=False)
        assert np.allclose(G2r, G2c, rtol=1e-4, atol=0)


@pytest.mark.gpu
def test_cli_darkout(tmp_path, orc):
    """Dense IMM with two leading dark frames and --darkout: DarkAvg and
    DarkStd are float64 [H, W] (main.cpp:459-478) and equal the oracle's dark
    image; g2 matches the oracle with the darks subtracted."""
    W = H = 16
    F = 200
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    sig = synth.ou_counts(W, H, F, 5.0, 15.0, seed=71)
    dark = synth.dark_frames(W, H, 2, seed=72)
    raw = np.concatenate([dark, sig + dark[0]], axis=0)
    imm = tmp_path / "dense.imm"
    imm.write_bytes(synth.imm_bytes_dense(raw, W, H))
    conf = tmp_path / "cfg.h5"
    xpcs_h5.write_config(conf, width=W, height=H, dqmap=dq, sqmap=sq, imm_path=imm, compression=False,
                         dark_begin_todo=1, dark_end_todo=2, data_begin_todo=3, data_end_todo=F + 2, dpl=4)
    print("corr dense job with darks", W, H, F, flush=True)
    r = run(conf, "--darkout")
    assert r.returncode == 0, r.stderr
    buf = imm.read_bytes()
    davg, dstd, _ = orc.dark_image(buf, orc.imm_index(buf), 2, W * H)
    ref = orc.run_exact(buf, dq, sq, F, 4, darks=2, frame_from=2)
    with xpcs_h5.H5(conf) as h:
        for name, want in (("DarkAvg", davg), ("DarkStd", dstd)):
            assert h.dtype("/exchange/" + name) == np.dtype("f8"), name
            got = h.read("/exchange/" + name)
            assert got.shape == (H, W) and np.allclose(got.reshape(-1), want, rtol=1e-12, atol=1e-12), name
        g2 = h.read("/exchange/norm-0-g2", np.float32)
        ok = ~np.isnan(ref["g2"])
        assert np.allclose(g2[ok], ref["g2"][ok], rtol=1e-6)


@pytest.mark.gpu
def test_cli_exchange_and_positional_imm(tmp_path, orc):
    job = make_job(tmp_path, F=120, dpl=2)
    moved = tmp_path / "moved.imm"
    os.replace(job["imm"], moved)
    r = run(job["conf"], moved, "--exchange", "/results/run1", "--quiet")
    assert r.returncode == 0, r.stderr
    ref = orc.run_exact(moved.read_bytes(), job["dq"], job["sq"], job["F"], job["dpl"])
    with xpcs_h5.H5(job["conf"]) as h:
        g2 = h.read("/results/run1/norm-0-g2")
        ok = ~np.isnan(ref["g2"])
        assert np.allclose(g2[ok], ref["g2"][ok], rtol=1e-6)
    bad = run(job["conf"], tmp_path / "nope.imm")
    assert bad.returncode == 3


@pytest.mark.gpu
@pytest.mark.parametrize("method,filt", [("StaticMap", "Average"), ("SYMMETRIC", "None")])
def test_cli_twotime_matches_oracle(tmp_path, orc, method, filt):
    """analysis_type Twotime: C2T_all/g2_%05d, g2full, g2partials and sg
    (float, corr.cpp:433-560, :781-924) against the FP64 oracle; no tau/g2."""
    job = make_job(tmp_path, F=120, analysis_type="Twotime", smoothing_method=method, smoothing_filter=filt,
                   qphi_bins=[3, 1], twotime_window=4, static_window=30)
    print("corr twotime job", method, filt, flush=True)
    r = run(job["conf"])
    assert r.returncode == 0, r.stderr
    ex = orc.run_exact(job["imm"].read_bytes(), job["dq"], job["sq"], job["F"], 2)
    smoothing = 1 if method == "StaticMap" else 2
    want = orc.twotime_f64(ex["csr"], ex["qm"], job["F"], [3, 1], smoothing, filt == "Average", 4)
    F = job["F"]
    with xpcs_h5.H5(job["conf"]) as h:
        for i, q in enumerate(want["qbins"]):
            c2t = h.read(f"/exchange/C2T_all/g2_{q:05d}", np.float32)
            assert c2t.shape == (F, F)
            np.testing.assert_allclose(c2t, want["c2t"][i], rtol=1e-6, atol=0)
        full = h.read("/exchange/g2full", np.float32)
        assert full.shape == (F, 2)
        np.testing.assert_allclose(full, want["g2full"], rtol=1e-6, atol=0)
        part = h.read("/exchange/g2partials", np.float32)
        assert part.shape == (4, (F - 4) // 4, 2)
        np.testing.assert_allclose(part, want["g2partials"], rtol=1e-6, atol=0)
        sg = h.read("/exchange/sg", np.float32)
        assert sg.shape == want["sg"].shape
        np.testing.assert_allclose(sg, want["sg"], rtol=1e-6, atol=0)
        assert h.dims("/exchange/pixelSum") == (job["H"], job["W"])
        with pytest.raises(Exception):
            h.dims("/exchange/tau")


def _g2_matches(conf, ref):
    with xpcs_h5.H5(conf) as h:
        g2 = h.read("/exchange/norm-0-g2", np.float32)
    ok = ~np.isnan(ref["g2"])
    assert g2.shape == ref["g2"].shape
    assert np.allclose(g2[ok], ref["g2"][ok], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["ufxc", "rigaku", "hdf5"])
def test_cli_other_formats(tmp_path, orc, fmt):
    """--ufxc / --rigaku / --hdf5 inputs (main.cpp:204-220) against the oracle's
    restatement of each reader as an equivalent IMM."""
    W, H, F = 24, 20, 240
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    counts = synth.ou_counts(W, H, F, 0.3, 8.0, 9)
    data = tmp_path / f"data.{fmt}"
    if fmt == "ufxc":
        counts = np.minimum(counts, 3)
        w = synth.ufxc_words(counts, W, H, counter0=2000)
        data.write_bytes(w.tobytes())
        imm = orc.ufxc_imm(w, W, H, F)
    elif fmt == "rigaku":
        w = synth.rigaku_words(counts, W, H, first_frame=1)
        data.write_bytes(w.tobytes())
        mapped = (dq.reshape(-1) > 0) & (sq.reshape(-1) > 0)
        imm = orc.rigaku_imm(w, W, H, mapped, 0, F)
    else:
        with xpcs_h5.H5(data, create=True) as h:
            h.write("/entry/data/data", counts.reshape(F, H, W).astype(np.uint16))
        imm = synth.imm_bytes_sparse(counts, W, H)
    conf = tmp_path / "cfg.h5"
    xpcs_h5.write_config(conf, width=W, height=H, dqmap=dq, sqmap=sq, imm_path=data, data_end_todo=F, dpl=4,
                         static_window=40)
    print("corr", fmt, flush=True)
    r = run(conf, f"--{fmt}")
    assert r.returncode == 0, r.stderr
    _g2_matches(conf, orc.run_exact(imm, dq, sq, F, 4))
    with xpcs_h5.H5(conf) as h:
        ts = h.read("/exchange/timestamp_clock", np.float64)
        assert ts.shape == (2, F) and np.array_equal(ts[0], np.arange(1, F + 1))


@pytest.mark.gpu
def test_cli_normalize_by_framesum(tmp_path, orc):
    job = make_job(tmp_path, F=300, normalize_by_framesum=1)
    print("corr normalize_by_framesum", flush=True)
    r = run(job["conf"])
    assert r.returncode == 0, r.stderr
    ref = orc.run_fsnorm(job["imm"].read_bytes(), job["dq"], job["sq"], job["F"], job["dpl"])
    with xpcs_h5.H5(job["conf"]) as h:
        g2 = h.read("/exchange/norm-0-g2", np.float32)
    ok = ~np.isnan(ref["g2f"])
    assert np.allclose(g2[ok], ref["g2f"][ok], rtol=1e-6)
