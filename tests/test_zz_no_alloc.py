"""xpcs_mt_normalize_ref_f32_noalloc through the C ABI: device arrays in,
device results out, nothing allocated; equal bit for bit to the allocating
xpcs_mt_normalize_ref_f32 on two calls in a row (the second smaller, so the
shared scratch holds the first call's contents), and sizes beyond the scratch
refused before anything runs."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def device_normalize(xmt, plan, arrays):
    import torch
    dev = [torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda() for a in arrays]
    g2 = torch.full((plan.T, plan.Q), -1.0, dtype=torch.float64, device="cuda")
    se = torch.full((plan.T, plan.Q), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = xmt.lib().xpcs_mt_normalize_ref_f32_noalloc(plan._h, dev[0].data_ptr(), dev[1].data_ptr(),
                                                     dev[2].data_ptr(), plan.width * plan.height, plan.T,
                                                     g2.data_ptr(), se.data_ptr(), None)
    assert st == 0, xmt.lib().xpcs_last_error()
    torch.cuda.synchronize()
    return g2.cpu().numpy(), se.cpu().numpy()


def test_no_alloc_normalize_equals_allocating_entry(xmt):
    for W, F, seed, ff in ((40, 700, 801, True), (20, 300, 802, False)):
        print(f"no-alloc normalize W={W} F={F} flatfield={ff}", flush=True)
        c = cases.sparse_case(W=W, H=W, F=F, dpl=4, mu=0.1, tau_c=10.0, seed=seed)
        flat = np.random.default_rng(seed).uniform(0.5, 1.5, (W, W)) if ff else None
        plan = xmt.Plan(W, W, F, 4, c["dq"], c["sq"], flatfield=flat)
        plan.ingest_imm(c["imm"])
        plan.execute()
        arrays = plan.ref_f32()
        g2_exec, _ = plan.g2()
        want_g2, want_se = plan.normalize_ref_f32(*arrays)
        got_g2, got_se = device_normalize(xmt, plan, arrays)
        assert np.array_equal(got_g2, want_g2, equal_nan=True)
        assert np.array_equal(got_se, want_se, equal_nan=True)
        assert np.isfinite(got_g2).any()
        # the plan's own results are the allocating call's, untouched by the no-alloc one
        assert np.array_equal(plan.g2()[0], want_g2, equal_nan=True)
        plan.close()


def test_no_alloc_refuses_beyond_scratch(xmt):
    import torch
    import synth
    W = 1024
    dq, sq = synth.annular_qmaps(W, W, 18, 6)
    plan = xmt.Plan(W, W, 1000, 4, dq, sq)
    small = torch.zeros(16, dtype=torch.float32, device="cuda")
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    st = xmt.lib().xpcs_mt_normalize_ref_f32_noalloc(plan._h, small.data_ptr(), small.data_ptr(),
                                                     small.data_ptr(), W * W, plan.T, out.data_ptr(),
                                                     out.data_ptr(), None)
    assert st == 5
    assert b"maximum" in xmt.lib().xpcs_last_error()
    plan.close()
