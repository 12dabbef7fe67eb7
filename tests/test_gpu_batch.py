"""Batched execution (one captured hipGraph over many plans), explicit launch
shapes and the occupancy-keyed autotuner: every job's sums and g2 equal the
plain per-job path bit for bit, and the oracle on the cases it checks."""
import numpy as np
import pytest

import cases
import synth

pytestmark = pytest.mark.gpu


def dense_case(W, H, F, dpl, mu, tau_c, seed, darks=2, S=9, Q=3):
    dq, sq = synth.annular_qmaps(W, H, S, Q)
    sig = synth.ou_counts(W, H, F, mu, tau_c, seed=seed)
    dark = synth.dark_frames(W, H, darks, seed=seed + 1000)
    frames = np.concatenate([dark, sig + dark[0]], axis=0)
    return dict(W=W, H=H, F=F, dpl=dpl, dq=dq, sq=sq, imm=synth.imm_bytes_dense(frames, W, H), darks=darks)


def mixed_jobs():
    """Uneven jobs: sparse of several sizes/occupancies and dpl, dense dpl 4 and 8."""
    return [
        cases.sparse_case(W=32, H=32, F=600, dpl=4, mu=0.05, tau_c=10.0, seed=71),
        cases.sparse_case(W=48, H=40, F=1500, dpl=8, mu=0.02, tau_c=30.0, seed=72, S=8, Q=4),
        cases.sparse_case(W=20, H=20, F=300, dpl=4, mu=0.8, tau_c=5.0, seed=73),
        cases.sparse_case(W=64, H=64, F=900, dpl=8, mu=0.005, tau_c=100.0, seed=74, S=12, Q=4),
        dense_case(16, 16, 700, 4, 3.0, 20.0, seed=75),
        dense_case(24, 16, 1100, 8, 2.0, 40.0, seed=76),
        cases.sparse_case(W=8, H=8, F=2000, dpl=4, mu=2.0, tau_c=50.0, seed=77, S=4, Q=2),
    ]


def make_plan(xmt, c):
    plan = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"], dark_frames=c.get("darks", 0))
    plan.ingest_imm(c["imm"])
    return plan


def results(plan):
    G2, IP, IF = plan.sums()
    g2, se = plan.g2()
    return G2, IP, IF, g2, se


def assert_same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=x.dtype.kind == "f")


def test_batch_equals_plain_and_oracle(xmt, orc):
    jobs = mixed_jobs()
    print("batch of", len(jobs), "jobs", flush=True)
    plain = []
    for c in jobs:
        p = make_plan(xmt, c)
        p.execute()
        plain.append(results(p))
        p.close()
    plans = [make_plan(xmt, c) for c in jobs]
    with xmt.Batch(branches=4) as b:
        for p in plans:
            b.add(p)
        b.execute()
        b.synchronize()
        for p, want in zip(plans, plain):
            assert_same(results(p), want)
    for k in (0, 4):
        c = jobs[k]
        r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"], darks=c.get("darks", 0))
        assert np.array_equal(plain[k][0], r["G2"]) and np.array_equal(plain[k][1], r["IP"])
        assert np.array_equal(plain[k][2], r["IF"])


def test_batch_graph_replays(xmt):
    """Several replays of the captured graph, with eager work of another plan
    and host synchronisations between them, each equal to the eager result."""
    jobs = mixed_jobs()[:4]
    plans = [make_plan(xmt, c) for c in jobs]
    eager = []
    for p in plans:
        p.execute()
        eager.append(results(p))
    other = make_plan(xmt, mixed_jobs()[4])
    with xmt.Batch(branches=3) as b:
        for p in plans:
            b.add(p)
        for rep in range(3):
            print("replay", rep, flush=True)
            b.execute()
            b.synchronize()
            for p, want in zip(plans, eager):
                assert_same(results(p), want)
            other.execute()
            other.synchronize()


def test_batch_recaptures_after_reingest(xmt):
    jobs = mixed_jobs()
    plans = [make_plan(xmt, c) for c in jobs[:3]]
    with xmt.Batch(branches=2) as b:
        for p in plans:
            b.add(p)
        b.execute()
        b.synchronize()
        # new data for plan 0 (same shape): the batch must capture again
        c = cases.sparse_case(W=32, H=32, F=600, dpl=4, mu=0.07, tau_c=4.0, seed=99)
        plans[0].ingest_imm(c["imm"])
        b.execute()
        b.synchronize()
        got = results(plans[0])
    ref = make_plan(xmt, c)
    ref.execute()
    assert_same(got, results(ref))


@pytest.mark.parametrize("shapes", [
    dict(tr_bucket_shift=10, tr_chunk_frames=3, reg_blocks=7, lds_blocks=5),
    dict(tr_bucket_shift=9, tr_chunk_frames=1000, reg_blocks=65535, lds_blocks=1),
])
def test_sparse_shapes_keep_results(xmt, shapes):
    c = cases.sparse_case(W=40, H=40, F=1300, dpl=8, mu=0.3, tau_c=20.0, seed=81, S=8, Q=4)
    base = make_plan(xmt, c)
    base.execute()
    want = results(base)
    p = make_plan(xmt, c)
    p.set_tuning(**shapes)
    t = p.get_tuning()
    for k, v in shapes.items():
        assert t[k] == v
    print("shapes", t, flush=True)
    p.execute()
    assert_same(results(p), want)


@pytest.mark.parametrize("dpl,levels", [(4, 3), (4, 5), (8, 2)])
def test_dense_levels_keep_results(xmt, orc, dpl, levels):
    c = dense_case(16, 12, 1500, dpl, 2.0, 15.0, seed=83 + levels)
    p = make_plan(xmt, c)
    p.set_tuning(dense_levels=levels)
    assert p.get_tuning()["dense_levels"] == levels
    p.execute()
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], c["F"], c["dpl"], darks=2)
    G2, IP, IF = p.sums()
    assert np.array_equal(G2, r["G2"]) and np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"])


def test_tuning_rejects_bad_values(xmt):
    c = dense_case(8, 8, 100, 4, 2.0, 5.0, seed=90)
    p = make_plan(xmt, c)
    for bad in (dict(dense_levels=2), dict(tr_bucket_shift=7), dict(reg_blocks=70000), dict(lds_blocks=-1)):
        with pytest.raises(xmt.XpcsError):
            p.set_tuning(**bad)


def test_autotune_keeps_results_and_caches(xmt):
    jobs = [cases.sparse_case(W=48, H=48, F=1000, dpl=4, mu=0.05, tau_c=10.0, seed=91),
            dense_case(16, 16, 900, 4, 2.0, 10.0, seed=92)]
    for c in jobs:
        base = make_plan(xmt, c)
        base.execute()
        want = results(base)
        p = make_plan(xmt, c)
        best = p.autotune()
        print("autotuned", best, flush=True)
        assert_same(results(p), want)
        p.execute()
        assert_same(results(p), want)
        # same occupancy key: the cached shapes are taken as they are
        q = make_plan(xmt, c)
        assert q.autotune() == best
    plans = [make_plan(xmt, c) for c in jobs]
    with xmt.Batch(branches=2) as b:
        for p in plans:
            b.add(p)
        b.execute(autotune=True)
        b.synchronize()


def test_two_plans_alternating_overlap_ingest(xmt):
    """A series of jobs of one geometry through two plans used in turn: each
    job is ingested while the other plan's job is still executing, and every
    job's sums and g2 equal a lone plan's."""
    jobs = [cases.sparse_case(W=64, H=48, F=1200, dpl=8, mu=mu, tau_c=tc, seed=300 + j, S=8, Q=4)
            for j, (mu, tc) in enumerate([(0.02, 20.0), (0.2, 5.0), (0.005, 80.0), (0.05, 40.0), (0.6, 10.0)])]
    want = []
    for c in jobs:
        p = make_plan(xmt, c)
        p.execute()
        want.append(results(p))
        p.close()
    c0 = jobs[0]
    plans = [xmt.Plan(c0["W"], c0["H"], c0["F"], c0["dpl"], c0["dq"], c0["sq"]) for _ in range(2)]
    try:
        prev = None
        for k, c in enumerate(jobs):
            print("job", k, flush=True)
            cur = plans[k % 2]
            cur.ingest_imm(c["imm"])
            if prev is not None:
                assert_same(results(prev), want[k - 1])
            cur.execute()
            prev = cur
        assert_same(results(prev), want[-1])
    finally:
        for p in plans:
            p.close()


def test_timing_filter(xmt):
    """With a timing filter only the named kernel timer records; without it
    every timer does, and the sums are the same either way."""
    c = cases.sparse_case(W=48, H=48, F=1200, dpl=8, mu=0.03, tau_c=20.0, seed=95, S=8, Q=4)
    p = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"], timing=True)
    try:
        p.ingest_imm(c["imm"])
        p.execute()
        want = results(p)
        every = set(p.kernel_times())
        assert "k_mt_reg" in every and len(every) > 3
        p.set_timing_filter("k_mt_reg")
        p.kernel_times_reset()
        p.execute()
        assert_same(results(p), want)
        kt = p.kernel_times()
        assert set(kt) == {"k_mt_reg"} and kt["k_mt_reg"][1] == 1
        p.set_timing_filter(None)
        p.kernel_times_reset()
        p.execute()
        p.synchronize()
        assert set(p.kernel_times()) == every
    finally:
        p.close()
