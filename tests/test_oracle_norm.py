"""Pins of the normalisation (Corr::normalizeG2s, src/xpcs/corr.cpp:927-1091):
closed forms, NaN rules, the Welford stderr against numpy's two-pass variance,
and the float replica against the FP64 form."""
import numpy as np

import cases


def tiny_qm(orc, dq, sq):
    return orc.build_qmap(np.asarray(dq, np.int32), np.asarray(sq, np.int32))


def test_constant_rows_give_g2_one(orc):
    F, dpl = 256, 4
    counts = np.full((F, 4), 2, np.uint16)
    import synth
    imm = synth.imm_bytes_sparse(counts, 2, 2)
    r = orc.run_exact(imm, np.array([[1, 1], [2, 2]]), np.array([[1, 2], [3, 3]]), F, dpl)
    assert np.allclose(r["g2"], 1.0, rtol=0, atol=1e-15)
    assert np.allclose(r["stderr"], 0.0, atol=1e-15)


def test_welford_matches_two_pass_variance(orc):
    c = cases.sparse_case(W=24, H=24, F=400, dpl=4, mu=0.3, tau_c=8.0, seed=31)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    qm = r["qm"]
    N = r["N"].astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        x = N[None, :] * r["G2"] / (r["IP"].astype(np.float64) * r["IF"])
    x = np.nan_to_num(x, nan=0.0)
    for q in range(1, qm["nq"] + 1):
        sel = np.concatenate([np.arange(b, e) for g, (b, e) in
                              enumerate(zip(qm["group_begin"], qm["group_end"])) if qm["group_q"][g] == q])
        xs = x[sel]
        want = np.sqrt(xs.var(axis=0) / xs.shape[0])       # population variance, every pixel counted
        assert np.allclose(r["stderr"][:, q - 1], want, rtol=1e-9, atol=1e-15)


def test_g2_static_then_dynamic_mean(orc):
    c = cases.sparse_case(W=24, H=24, F=400, dpl=4, mu=0.3, tau_c=8.0, seed=32)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    qm = r["qm"]
    N = r["N"].astype(np.float64)
    g2s = []
    for g in range(qm["ngroups"]):
        b, e = qm["group_begin"][g], qm["group_end"][g]
        sg, sp, sf = r["G2"][b:e].sum(0), r["IP"][b:e].sum(0), r["IF"][b:e].sum(0)
        g2s.append((e - b) * N * sg / (sp.astype(np.float64) * sf))
    g2s = np.array(g2s)
    for q in range(1, qm["nq"] + 1):
        m = qm["group_q"] == q
        want = np.nanmean(g2s[m], axis=0)
        assert np.allclose(r["g2"][:, q - 1], want, rtol=1e-12)


def test_nan_rules(orc):
    """A static bin with no events gives NaN and is skipped; a dynamic bin whose
    static bins are all NaN is NaN; its stderr is 0 (all x_p = 0)."""
    import synth
    F, dpl = 64, 2
    counts = np.zeros((F, 4), np.uint16)
    counts[::3, 0] = 1          # pixel 0: events; pixels 1..3: none
    imm = synth.imm_bytes_sparse(counts, 2, 2)
    dq = np.array([[1, 1], [2, 2]], np.int32)
    sq = np.array([[1, 2], [3, 3]], np.int32)
    r = orc.run_exact(imm, dq, sq, F, dpl)
    assert np.all(np.isfinite(r["g2"][:, 0]))      # q=1 uses s=1 only (s=2 NaN skipped)
    assert np.all(np.isnan(r["g2"][:, 1]))         # q=2: its only static bin is NaN
    assert np.all(r["stderr"][:, 1] == 0.0)


def test_float_replica_close_to_f64(orc):
    c = cases.sparse_case(W=32, H=32, F=500, dpl=4, mu=0.05, tau_c=10.0, seed=7)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    G2c, IPc, IFc = orc.multitau_compat(r["csr"], r["qm"]["order"], 1024, c["F"], r["level"], r["tau"], False)
    g2c, sec = orc.normalize_compat(G2c, IPc, IFc, r["qm"])
    assert np.allclose(g2c, r["g2"], rtol=1e-5)
    assert np.allclose(sec, r["stderr"], rtol=1e-4, atol=1e-7)


def test_cpu_pipeline_equals_replica_chain(orc):
    c = cases.sparse_case(W=32, H=32, F=300, dpl=4, mu=0.05, tau_c=10.0, seed=8)
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"])
    G2s, IPs, IFs = orc.multitau_compat(r["csr"], r["qm"]["order"], 1024, c["F"], r["level"], r["tau"], True)
    want = orc.normalize_compat(G2s, IPs, IFs, r["qm"])
    g2, se, _ = orc.cpu_pipeline(c["imm"], 32, 32, c["dq"], c["sq"], c["F"], c["dpl"], 2)
    assert np.array_equal(g2, want[0], equal_nan=True) and np.array_equal(se, want[1], equal_nan=True)
