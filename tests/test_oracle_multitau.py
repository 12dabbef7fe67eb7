"""Pins of the oracle's correlator (xo_multitau_exact / xo_multitau_compat):
pair-enumeration golden sums, an independent numpy binning, the float replica
without the stale-tail search, the stale-tail counterexample, and invariants."""
from pathlib import Path

import numpy as np
import pytest

import cases
import synth

GOLDEN = Path(__file__).resolve().parent / "golden"


def csr_of(frames, counts):
    return dict(off=np.array([0, len(frames)], np.int64), frame=np.asarray(frames, np.int32),
                cnt=np.asarray(counts, np.int64), val=np.asarray(counts, np.float32))


def test_pair_enumeration_golden(orc):
    g = np.load(GOLDEN / "multitau_bruteforce.npz")
    for k in range(int(g["ncases"])):
        F, dpl = int(g[f"c{k}_F"]), int(g[f"c{k}_dpl"])
        lev, tau = orc.schedule(F, dpl)
        assert np.array_equal(lev, g[f"c{k}_level"]) and np.array_equal(tau, g[f"c{k}_tau"])
        G2, IP, IF = orc.multitau_exact(csr_of(g[f"c{k}_frames"], g[f"c{k}_counts"]), F, lev, tau)
        assert np.array_equal(G2[0], g[f"c{k}_G2"]), k
        assert np.array_equal(IP[0], g[f"c{k}_IP"]), k
        assert np.array_equal(IF[0], g[f"c{k}_IF"]), k


def reshape_sums(x0, F, lev, tau):
    """Binning by reshape: x_L = x0[:lf*2^L].reshape(lf, 2^L).sum(1)."""
    T = lev.size
    out = np.zeros((3, T), np.int64)
    for i in range(T):
        L = int(lev[i])
        lf = F >> L
        tp = int(tau[i]) >> L
        N = lf - tp
        xL = x0[: lf << L].reshape(lf, 1 << L).sum(1)
        out[:, i] = (np.dot(xL[:N], xL[tp:tp + N]), xL[:N].sum(), xL[tp:lf].sum())
    return out


def test_exact_vs_numpy_binning(orc):
    c = cases.sparse_case(W=16, H=16, F=777, dpl=4, mu=0.2, tau_c=15.0, seed=3)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    order = r["qm"]["order"]
    for p in range(0, order.size, 7):
        x0 = c["counts"][:, order[p]].astype(np.int64)
        want = reshape_sums(x0, c["F"], r["level"], r["tau"])
        assert np.array_equal(r["G2"][p], want[0])
        assert np.array_equal(r["IP"][p], want[1])
        assert np.array_equal(r["IF"][p], want[2])


def test_float_replica_without_stale_tail_equals_exact(orc):
    """The statement-by-statement float replica of multiTau2 with the search
    limited to [0, lastIndex) reproduces the exact sums scaled by 2^-L/N."""
    c = cases.sparse_case(W=16, H=16, F=600, dpl=4, mu=0.1, tau_c=10.0, seed=5)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    lev, tau, N = r["level"], r["tau"], r["N"].astype(np.float64)
    G2c, IPc, IFc = orc.multitau_compat(r["csr"], r["qm"]["order"], 256, c["F"], lev, tau, stale_tail=False)
    o = r["qm"]["order"]
    sc = 2.0 ** lev
    for got, want in ((G2c[:, o], r["G2"].T / sc[:, None] ** 2 / N[:, None]),
                      (IPc[:, o], r["IP"].T / sc[:, None] / N[:, None]),
                      (IFc[:, o], r["IF"].T / sc[:, None] / N[:, None])):
        assert np.allclose(got, want, rtol=2e-7, atol=0)


def test_stale_tail_counterexample(orc):
    """Level-0 frames [0..15, 100, 101, 110, 111], counts 1, F=256, dpl=4: at
    level 1, tau=10 the exact pair count is 4; the reference's whole-vector
    lower_bound over the stale tail finds 3 (src/xpcs/corr.cpp:406-410)."""
    c = cases.counterexample_case()
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    i = int(np.flatnonzero((r["level"] == 1) & (r["tau"] == 10))[0])
    assert r["G2"][0, i] == 4 * 4          # 4 pairs of binned value 2 (summed units)
    N = float(r["N"][i])
    stale, _, _ = orc.multitau_compat(r["csr"], r["qm"]["order"], 1, c["F"], r["level"], r["tau"], True)
    clean, _, _ = orc.multitau_compat(r["csr"], r["qm"]["order"], 1, c["F"], r["level"], r["tau"], False)
    assert round(float(stale[i, 0]) * N) == 3
    assert round(float(clean[i, 0]) * N) == 4


def test_constant_series_closed_form(orc):
    """x_0[t] = c for all t: G2 = N (2^L c)^2, IP = IF = N 2^L c."""
    F, dpl, cnt = 300, 4, 3
    lev, tau = orc.schedule(F, dpl)
    G2, IP, IF = orc.multitau_exact(csr_of(np.arange(F), np.full(F, cnt)), F, lev, tau)
    N = orc.schedule_N(F, lev, tau).astype(np.int64)
    s = (1 << lev.astype(np.int64)) * cnt
    assert np.array_equal(G2[0], N * s * s)
    assert np.array_equal(IP[0], N * s)
    assert np.array_equal(IF[0], N * s)


def test_scale_and_single_event(orc):
    F, dpl = 200, 2
    lev, tau = orc.schedule(F, dpl)
    fr = [3, 4, 9, 40, 41, 42, 150]
    ct = [1, 2, 1, 3, 1, 1, 2]
    a = orc.multitau_exact(csr_of(fr, ct), F, lev, tau)
    b = orc.multitau_exact(csr_of(fr, [5 * x for x in ct]), F, lev, tau)
    assert np.array_equal(b[0], 25 * a[0]) and np.array_equal(b[1], 5 * a[1]) and np.array_equal(b[2], 5 * a[2])
    g, ip, iff = orc.multitau_exact(csr_of([17], [4]), F, lev, tau)
    assert not g.any()
    for i in range(lev.size):
        L = int(lev[i]); lf = F >> L; tp = int(tau[i]) >> L
        assert ip[0, i] == (4 if (17 >> L) < lf - tp else 0)
        assert iff[0, i] == (4 if tp <= (17 >> L) < lf else 0)
