"""normalize_by_framesum oracle (src/xpcs/main.cpp:313-336) pinned: the FP64
multi-tau at unit frame weights equals the exact int64 sums; the numpy
normalisation equals the C one on the same sums; frame weights have mean-sum
closed forms; scaling every frame by a constant leaves g2 unchanged."""
import numpy as np

import cases


def test_unit_weights_equal_exact(orc):
    c = cases.sparse_case(W=16, H=16, F=300, mu=0.2, tau_c=8.0, seed=31, S=4, Q=2)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], 4)
    G2, IP, IF = orc.multitau_f64(r["csr"], c["F"], r["level"], r["tau"], np.ones(c["F"]))
    assert np.array_equal(G2, r["G2"].astype(np.float64))
    assert np.array_equal(IP, r["IP"].astype(np.float64))
    assert np.array_equal(IF, r["IF"].astype(np.float64))


def test_numpy_normalisation_equals_c(orc):
    c = cases.sparse_case(W=16, H=16, F=300, mu=0.2, tau_c=8.0, seed=32, S=4, Q=2)
    ff = np.random.default_rng(33).uniform(0.5, 1.5, (16, 16))
    for flat in (None, ff):
        r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], 4, flatfield=flat)
        w = None if flat is None else flat.reshape(-1)[r["qm"]["order"]]
        g2, se = orc.normalize_f64_sums(r["G2"], r["IP"], r["IF"], r["N"], r["qm"], w)
        ok = ~np.isnan(r["g2"])
        np.testing.assert_allclose(g2[ok], r["g2"][ok], rtol=1e-12)
        np.testing.assert_allclose(se, r["stderr"], rtol=1e-9, atol=1e-15)


def test_frame_weights_closed_form(orc):
    """Frame sums 1, 2, 3 (and an empty frame): weights mean/S = 1.5, 0.75, 0.5, 0."""
    csr = dict(off=np.array([0, 3, 5], np.int64), frame=np.array([0, 1, 2, 1, 2], np.int32),
               cnt=np.array([1, 1, 1, 1, 2], np.int64))
    fw = orc.frame_weights(csr, 4)
    np.testing.assert_allclose(fw, [1.5, 0.75, 0.5, 0.0])


def test_uniform_frame_scaling_leaves_g2(orc):
    """Every frame's counts doubled: the weights halve them back, g2 equals the
    weight-free normalisation of the same pixel data."""
    c = cases.sparse_case(W=16, H=16, F=200, mu=0.3, tau_c=8.0, seed=34, S=4, Q=2)
    r = orc.run_fsnorm(c["imm"], c["dq"], c["sq"], c["F"], 4)
    fw = r["fw"]
    assert abs(np.mean(fw[fw > 0]) - 1.0) < 0.2      # weights near 1 for stationary data
    G2, IP, IF = orc.multitau_f64(r["csr"], c["F"], r["level"], r["tau"], fw * 2.0)
    g2, _ = orc.normalize_f64_sums(G2, IP, IF, r["N"], r["qm"])
    ok = ~np.isnan(g2)
    np.testing.assert_allclose(g2[ok], r["g2f"][ok], rtol=1e-12)
