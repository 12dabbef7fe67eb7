"""Tau schedule (Corr::calculateLevelMax / delaysPerLevel, src/xpcs/corr.cpp:1133-1164):
known answers for the benchmark configurations, structural invariants, and the
product's C ABI against the oracle."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"

# Known answers for the benchmark configurations (derived from corr.cpp:1133-1160).
KNOWN = {
    (1000, 4): dict(T=34, max_level=7, levels=list(range(8)), head=[1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16],
                    tail=[320, 384, 448, 512, 640, 768]),
    (10000, 8): dict(T=88, max_level=10, levels=list(range(10)), head=list(range(1, 17)) + [18],
                     tail=[4608, 5120, 5632, 6144, 6656, 7168, 7680, 8192]),
    (20000, 4): dict(T=52, max_level=11, levels=list(range(12)), tail=[10240, 12288, 14336, 16384]),
    (500, 4): dict(T=30, max_level=6),
    (20000, 8): dict(T=96, max_level=11),
}


@pytest.mark.parametrize("F,dpl", sorted(KNOWN))
def test_known_answers(orc, F, dpl):
    k = KNOWN[(F, dpl)]
    lev, tau = orc.schedule(F, dpl)
    assert orc.level_max(F, dpl) == k["max_level"]
    assert lev.size == k["T"]
    if "levels" in k:
        assert sorted(set(lev.tolist())) == k["levels"]
    if "head" in k:
        assert tau[: len(k["head"])].tolist() == k["head"]
    if "tail" in k:
        assert tau[-len(k["tail"]):].tolist() == k["tail"]


def test_golden_tables(orc):
    tables = json.loads((GOLDEN / "tau_tables.json").read_text())
    for key, t in tables.items():
        F, dpl = (int(x) for x in key.split(","))
        lev, tau = orc.schedule(F, dpl)
        assert lev.size == t["T"]
        assert sorted(set(lev.tolist())) == t["levels"]
        assert tau[:8].tolist() == t["head"]
        assert tau[-4:].tolist() == t["tail"]


@pytest.mark.parametrize("dpl", [1, 2, 3, 4, 5, 8, 16])
def test_invariants(orc, dpl):
    for F in [1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 100, 257, 1000, 4097, 10000, 65536]:
        lev, tau = orc.schedule(F, dpl)
        if lev.size == 0:
            continue
        assert np.all(np.diff(tau) > 0)                      # strictly increasing
        assert np.all(np.diff(lev) >= 0) and np.all(np.diff(lev) <= 1)
        assert np.all(tau % (1 << lev) == 0)                  # tau' = tau / 2^L exact
        assert np.all(tau + (1 << lev) <= F)                  # kept taus fit
        assert np.sum(lev == 0) <= 2 * dpl
        for L in range(1, lev.max() + 1):
            assert np.sum(lev == L) <= dpl
        # lags within a level are contiguous: tau' = lo..hi
        for L in np.unique(lev):
            tp = tau[lev == L] >> L
            assert np.array_equal(tp, np.arange(tp[0], tp[0] + tp.size))
        # N = lastframe_L - tau' >= 1
        N = (F >> lev) - (tau >> lev)
        assert np.all(N >= 1)


def test_product_matches_oracle(xmt, orc):
    for dpl in [1, 2, 3, 4, 6, 8, 16]:
        for F in [1, 5, 16, 100, 999, 1000, 4096, 10000, 20000, 123457]:
            assert xmt.level_max(F, dpl) == orc.level_max(F, dpl)
            a = xmt.delays_per_level(F, dpl)
            b = orc.schedule(F, dpl)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_level_max_small(orc):
    assert orc.level_max(7, 4) == 0          # frames < 2*dpl
    assert orc.level_max(8, 4) == 0
