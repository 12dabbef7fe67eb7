"""The xpcs::Corr C++ facade (include/xpcs_corr.hpp) with the reference's
names and argument meanings (corr.h:66-126, sparse_data.h:57-79, row.h:55-77),
driven by xpcs-eigen_amd/corr_api_check as the reference's main drives Corr:
rows of a SparseData filled by the caller, multiTau2 into zeroed [tau][W*H]
float arrays (accumulating), normalizeG2s. CPU: the header compiles and value
checks fail before any device work; GPU: the outputs against the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cases

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "xpcs-eigen_amd" / "corr_api_check"


def write_input(path, W, H, F, dpl, avg, dq, sq, ff, off, frame, val):
    with open(path, "wb") as f:
        f.write(np.array([W, H, F, dpl, avg, ff is not None], np.int32).tobytes())
        f.write(np.ascontiguousarray(dq, np.int32).tobytes())
        f.write(np.ascontiguousarray(sq, np.int32).tobytes())
        if ff is not None:
            f.write(np.ascontiguousarray(ff, np.float64).tobytes())
        f.write(np.array([frame.size], np.int64).tobytes())
        f.write(np.ascontiguousarray(off, np.int64).tobytes())
        f.write(np.ascontiguousarray(frame, np.int32).tobytes())
        f.write(np.ascontiguousarray(val, np.float32).tobytes())


def read_output(path, npix):
    b = Path(path).read_bytes()
    T, Q, maxl = np.frombuffer(b, np.int32, 3)
    o = 12
    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(b, dtype, n, o)
        o += a.nbytes
        return a
    level, tau = take(np.int32, T), take(np.int32, T)
    G2, IP, IF = (take(np.float32, T * npix).reshape(T, npix) for _ in range(3))
    g2, se = (take(np.float64, T * Q).reshape(T, Q) for _ in range(2))
    assert o == len(b)
    return dict(maxl=int(maxl), level=level, tau=tau, G2=G2, IP=IP, IF=IF, g2=g2, se=se)


def raster_rows(r, npix):
    """The oracle's compact rows as raster-pixel rows with the reference's values."""
    qm, csr = r["qm"], r["csr"]
    order = qm["order"]
    lens = np.zeros(npix, np.int64)
    lens[order] = np.diff(csr["off"])
    off = np.concatenate([[0], np.cumsum(lens)])
    cs = np.argsort(order)    # compact rows in raster order
    sl = [np.arange(csr["off"][c], csr["off"][c + 1]) for c in cs]
    idx = np.concatenate(sl) if sl else np.zeros(0, np.int64)
    return off, csr["frame"][idx].astype(np.int32), csr["val"][idx].astype(np.float32)


def test_header_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                        "-x", "c++", str(ROOT / "include" / "xpcs_corr.hpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_non_integer_value_is_refused(tmp_path):
    # a value that is not count * flatfield / average: refused while the rows
    # are converted, before a plan (or any device memory) exists
    W = H = 4
    dq = np.ones((H, W), np.int32)
    sq = np.ones((H, W), np.int32)
    off = np.zeros(W * H + 1, np.int64)
    off[6:] = 1
    write_input(tmp_path / "in.bin", W, H, 20, 4, 1, dq, sq, None, off, np.array([3], np.int32),
                np.array([0.5], np.float32))
    p = subprocess.run([str(BIN), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 2
    assert "pixel 5" in p.stderr and "not count * flatfield / average" in p.stderr


def check_against_oracle(tmp_path, orc, c, F, dpl, avg=1, ff=None, repeats=1, **kw):
    W, H = c["W"], c["H"]
    npix = W * H
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], F, dpl, avg=avg, flatfield=ff, **kw)
    off, frame, val = raster_rows(r, npix)
    write_input(tmp_path / "in.bin", W, H, F, dpl, avg, c["dq"], c["sq"], ff, off, frame, val)
    print(f"corr_api_check W={W} H={H} F={F} dpl={dpl} avg={avg} ff={ff is not None} x{repeats}", flush=True)
    p = subprocess.run([str(BIN), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(repeats)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    out = read_output(tmp_path / "out.bin", npix)
    assert out["maxl"] == orc.level_max(F, dpl)
    assert np.array_equal(out["level"], r["level"]) and np.array_equal(out["tau"], r["tau"])
    G2c, IPc, IFc = orc.multitau_compat(r["csr"], r["qm"]["order"], npix, F, r["level"], r["tau"],
                                        stale_tail=False)
    for got, want in ((out["G2"], G2c), (out["IP"], IPc), (out["IF"], IFc)):
        np.testing.assert_allclose(got, repeats * want, rtol=1e-4, atol=1e-6)
    # normalizeG2s of the accumulated arrays (corr.cpp:927-1091): repeats
    # times every sum divides g2 and stderr by repeats
    ok = ~np.isnan(r["g2"])
    assert np.array_equal(np.isnan(out["g2"]), ~ok)
    np.testing.assert_allclose(out["g2"][ok], r["g2"][ok] / repeats, rtol=1e-4)
    # (stderr of cells whose exact ratios are all equal is 0; from float arrays
    # with a flat field the ratios differ by float rounding, ~1e-8)
    np.testing.assert_allclose(out["se"], r["stderr"] / repeats, rtol=1e-4, atol=1e-7)
    # the same arrays through the FP64 normalisation of FP64 sums (pinned to the
    # exact one) and through the float replica of the reference's normalizeG2s
    order = r["qm"]["order"]
    rows = [np.ascontiguousarray(out[k][:, order].T, np.float64) for k in ("G2", "IP", "IF")]
    g2w, sew = orc.normalize_f64_sums(*rows, np.ones(out["level"].size), r["qm"])
    np.testing.assert_allclose(out["g2"][ok], g2w[ok], rtol=1e-9)
    np.testing.assert_allclose(out["se"], sew, rtol=1e-9, atol=1e-13)
    g2f, sef = orc.normalize_compat(out["G2"], out["IP"], out["IF"], r["qm"])
    np.testing.assert_allclose(out["g2"][ok], g2f[ok], rtol=1e-4)
    np.testing.assert_allclose(out["se"], sef, rtol=1e-4, atol=1e-6)   # float Welford noise near 0


@pytest.mark.gpu
def test_multitau2_accumulates_and_normalises(tmp_path, orc):
    c = cases.sparse_case(W=32, H=24, F=400, mu=0.2, tau_c=8.0, seed=720, S=6, Q=3)
    check_against_oracle(tmp_path, orc, c, 400, 4, repeats=2)


@pytest.mark.gpu
def test_flatfield_average_values(tmp_path, orc):
    c = cases.sparse_case(W=24, H=24, F=700, dpl=8, mu=0.3, tau_c=10.0, seed=721, S=8, Q=4)
    ff = np.random.default_rng(722).uniform(0.5, 1.5, (24, 24))
    check_against_oracle(tmp_path, orc, c, 330, 8, avg=2, ff=ff)
