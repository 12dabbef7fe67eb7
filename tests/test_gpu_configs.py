"""The five benchmark workloads under test, at their own sizes.

* sparse megapixel: 1024x1024, 10 000 frames, dpl 8, 36 / 12 rings, mu 0.005,
  per-ring tau_c 2..2000 -- integer sums bit-exact and g2 / stderr within 1e-6
  of the oracle on the whole job;
* dense: 512x512, 20 000 frames, mu 5, tau_c 50, two identical dark frames,
  dpl 4 -- a 128x128 crop of the same frames against the oracle, the full
  detector's sums on the crop's pixels equal to the crop's, and size-
  independent properties of the full job (doubling every value, darks included,
  multiplies G2 by 4 and IP / IF by 2 exactly and leaves g2 / stderr bit-equal;
  g2 finite; no pixel beyond the exact-FP64 bound);
* batched: all 64 drawn jobs, the autotuned graph-captured batch equal to the
  plain per-job path bit for bit, three of them against the oracle;
* streaming: three detector streams of >= 2000 frames at 256^2..512^2 through
  the serving loop with admission by HBM budget, each stream's final g2 equal
  to a plan of the same length bit for bit;
* plain 256x256: g2 / stderr against the reference's float output (the
  oracle's replica of multiTau2 + normalizeG2s with the stale-tail search of
  src/xpcs/corr.cpp:406-410) within 1e-4 outside the committed list of
  stale-tail cells, and within 1e-4 of the replica without that search on
  every cell.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu

G2_RTOL = 1e-6     # g2 / stderr vs the FP64 oracle
REF_RTOL = 1e-4    # g2 / stderr vs the reference's float output


def say(*a):
    print(*a, flush=True)


def assert_close(got, want, rtol):
    nan_w = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan_w), "NaN pattern differs"
    rel = np.abs(got[~nan_w] - want[~nan_w]) / np.maximum(np.abs(want[~nan_w]), 1e-300)
    assert rel.max(initial=0.0) <= rtol, f"max rel err {rel.max()}"


def assert_same(got, want):
    for x, y in zip(got, want):
        assert x.shape == y.shape
        assert np.array_equal(x, y, equal_nan=x.dtype.kind == "f")


def results(plan):
    G2, IP, IF = plan.sums()
    g2, se = plan.g2()
    return G2, IP, IF, g2, se


@pytest.mark.timeout(900)
def test_config2_sparse_megapixel_full(xmt, orc):
    import torch
    W = H = 1024
    F, dpl = 10000, 8
    dq, sq = synth.annular_qmaps(W, H, 36, 12)
    tau_c = synth.ring_tau(dq, 2.0, 2000.0)
    say("sparse megapixel: generating 1024x1024 x 10000 frames")
    imm, off, dl = synth.imm_device_sparse(W, H, F, 0.005, tau_c, 1002, device="cuda")
    torch.cuda.synchronize()
    plan = xmt.Plan(W, H, F, dpl, dq, sq)
    plan.ingest_imm_device(imm.data_ptr(), imm.numel(), off, dl, compressed=True)
    plan.execute()
    G2, IP, IF, g2, se = results(plan)
    assert plan.T == 88
    host = imm.cpu().numpy()
    del imm
    say("sparse megapixel: oracle")
    r = orc.run_exact(host, dq, sq, F, dpl)
    assert np.array_equal(plan.pixel_order(), r["qm"]["order"])
    assert r["csr"]["off"][-1] > 3.0e7          # ~0.5 % occupancy of 8.2e9 pixel-frames
    assert np.array_equal(IP, r["IP"]), "IP differs"
    assert np.array_equal(IF, r["IF"]), "IF differs"
    assert np.array_equal(G2, r["G2"]), "G2 differs"
    assert_close(g2, r["g2"], G2_RTOL)
    assert_close(se, r["stderr"], G2_RTOL)
    plan.close()


def _dense_crop_imm(imm, total, W, H, r0, c0, w, h):
    """A w x h window (rows r0.., columns c0..) of every frame of a dense IMM
    image in device memory, as a dense IMM image of its own."""
    import torch
    rec = synth.HEADER_BYTES + 2 * W * H
    pay = imm.view(total, rec)[:, synth.HEADER_BYTES:].view(torch.int16).view(total, H, W)
    crop = pay[:, r0:r0 + h, c0:c0 + w].reshape(total, w * h)
    rec2 = synth.HEADER_BYTES + 2 * w * h
    out = torch.empty(total * rec2, dtype=torch.uint8, device=imm.device)
    v = out.view(total, rec2)
    v[:, :synth.HEADER_BYTES] = torch.from_numpy(
        synth._headers(0, h, w, np.full(total, w * h), np.arange(total))).to(imm.device)
    v[:, synth.HEADER_BYTES:] = crop.contiguous().view(torch.uint8)
    offs = np.arange(total, dtype=np.int64) * rec2 + synth.HEADER_BYTES
    return out, offs, np.full(total, w * h, np.uint32)


@pytest.mark.timeout(900)
def test_config3_dense_full_frames(xmt, orc):
    import torch
    W = H = 512
    F, dpl, darks = 20000, 4, 2
    dq, sq = synth.annular_qmaps(W, H, 27, 9)
    say("dense: generating 512x512 x 20000 frames + 2 dark frames")
    imm, off, dl = synth.imm_device_dense(W, H, F, 5.0, 50.0, 1003, darks=darks, device="cuda")
    torch.cuda.synchronize()
    total = F + darks
    plan = xmt.Plan(W, H, F, dpl, dq, sq, dark_frames=darks)
    plan.ingest_imm_device(imm.data_ptr(), imm.numel(), off, dl, compressed=False)
    plan.execute()
    G2, IP, IF, g2, se = results(plan)
    assert plan.T == 52
    assert plan.info().exact_pixels == 0
    assert np.isfinite(g2).all() and np.isfinite(se).all()

    # every value doubled, dark frames included: the dark-subtracted values
    # double, so the sums scale by exactly 4 / 2 / 2 and g2, stderr are bit-equal
    say("dense: doubled values")
    imm2 = imm.clone()
    rec = synth.HEADER_BYTES + 2 * W * H
    imm2.view(total, rec)[:, synth.HEADER_BYTES:].view(torch.int16).mul_(2)
    plan2 = xmt.Plan(W, H, F, dpl, dq, sq, dark_frames=darks)
    plan2.ingest_imm_device(imm2.data_ptr(), imm2.numel(), off, dl, compressed=False)
    plan2.execute()
    G2d, IPd, IFd, g2d, sed = results(plan2)
    plan2.close()
    del imm2
    assert np.array_equal(G2d, 4 * G2) and np.array_equal(IPd, 2 * IP) and np.array_equal(IFd, 2 * IF)
    assert np.array_equal(g2d, g2) and np.array_equal(sed, se)

    # a 128 x 128 window of the same frames against the oracle; the full job's
    # per-pixel sums of those pixels are the window's
    r0 = c0 = 192
    w = h = 128
    say("dense: 128x128 window vs the oracle")
    cimm, coff, cdl = _dense_crop_imm(imm, total, W, H, r0, c0, w, h)
    cdq = np.ascontiguousarray(dq[r0:r0 + h, c0:c0 + w])
    csq = np.ascontiguousarray(sq[r0:r0 + h, c0:c0 + w])
    cplan = xmt.Plan(w, h, F, dpl, cdq, csq, dark_frames=darks)
    cplan.ingest_imm_device(cimm.data_ptr(), cimm.numel(), coff, cdl, compressed=False)
    cplan.execute()
    cG2, cIP, cIF, cg2, cse = results(cplan)
    host = cimm.cpu().numpy()
    r = orc.run_exact(host, cdq, csq, F, dpl, darks=darks)
    assert np.array_equal(cplan.pixel_order(), r["qm"]["order"])
    assert np.array_equal(cIP, r["IP"]) and np.array_equal(cIF, r["IF"]) and np.array_equal(cG2, r["G2"])
    assert_close(cg2, r["g2"], G2_RTOL)
    assert_close(cse, r["stderr"], G2_RTOL)
    corder = cplan.pixel_order()
    full_flat = (corder // w + r0) * W + (corder % w + c0)
    row_of = {int(f): i for i, f in enumerate(plan.pixel_order())}
    rows = np.array([row_of[int(f)] for f in full_flat])
    assert np.array_equal(G2[rows], cG2) and np.array_equal(IP[rows], cIP) and np.array_equal(IF[rows], cIF)
    cplan.close()
    plan.close()


@pytest.mark.timeout(900)
def test_config4_batch_all_jobs(xmt, orc):
    import torch
    jobs = synth.batched_jobs(64)
    plans, imms = [], []
    say("batched: generating 64 jobs")
    for jb in jobs:
        imm, off, dl = synth.imm_device_sparse(jb["W"], jb["W"], jb["F"], jb["mu"], jb["tau"], jb["seed"],
                                               device="cuda")
        dq, sq = synth.annular_qmaps(jb["W"], jb["W"], jb["S"], jb["Q"])
        p = xmt.Plan(jb["W"], jb["W"], jb["F"], jb["dpl"], dq, sq)
        p.ingest_imm_device(imm.data_ptr(), imm.numel(), off, dl, compressed=True)
        plans.append(p)
        imms.append((imm, off, dl, dq, sq))
    torch.cuda.synchronize()
    say("batched: autotuned batch")
    with xmt.Batch(branches=8) as b:
        for p in plans:
            b.add(p)
        b.execute(autotune=True)
        b.synchronize()
        b.execute()
        b.synchronize()
        for k, p in enumerate(plans):
            got = results(p)
            p.execute()                 # the plain path on the same plan and shapes
            assert_same(got, results(p))
    say("batched: oracle on the three smallest jobs")
    ev = [int(x[2].sum()) for x in imms]
    for k in np.argsort(ev)[:3]:
        jb = jobs[k]
        imm, off, dl, dq, sq = imms[k]
        r = orc.run_exact(imm.cpu().numpy(), dq, sq, jb["F"], jb["dpl"])
        G2, IP, IF, g2, se = results(plans[k])
        assert np.array_equal(G2, r["G2"]) and np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"])
        assert_close(g2, r["g2"], G2_RTOL)
    for p in plans:
        p.close()


@pytest.mark.timeout(900)
def test_config5_streams_through_service(xmt, orc):
    import torch
    import serve
    specs = [(256, 2400, 4, 0.01, 20.0), (384, 2000, 8, 0.02, 200.0), (512, 2200, 4, 0.005, 3.0)]
    streams = []
    for k, (W, F, dpl, occ, tau) in enumerate(specs):
        imm, off, dl = synth.imm_device_sparse(W, W, F, occ / (1 - occ), tau, 5000 + k, device="cuda")
        dq, sq = synth.annular_qmaps(W, W, 18, 6)
        streams.append(dict(W=W, F=F, dpl=dpl, imm=imm, off=off, dl=dl, dq=dq, sq=sq))
    torch.cuda.synchronize()
    sizes = [xmt.stream_bytes(s["W"], s["W"], s["F"], s["dpl"], s["dq"], s["sq"]) for s in streams]
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]      # the third waits for a release
    svc = serve.StreamService(budget, max_active=3)
    for k, s in enumerate(streams):
        svc.submit(k, s["W"], s["W"], s["dpl"], s["dq"], s["sq"],
                   serve.DeviceFrames(s["imm"], s["off"], s["dl"], s["F"], start=0.5 * k, fps=200.0), now=0.0)
    assert len(svc.queue) == 1
    say("streaming: serving loop")
    now = 0.0
    for _ in range(200000):
        svc.step(now)
        if svc.idle():
            break
        now += 0.02
    assert svc.idle() and svc.peak_used <= budget
    assert max(svc.admit_wait) > 0.0
    for k, s in enumerate(streams):
        res = svc.results[k]
        assert res["frames"] == s["F"]
        p = xmt.Plan(s["W"], s["W"], s["F"], s["dpl"], s["dq"], s["sq"])
        p.ingest_imm_device(s["imm"].data_ptr(), s["imm"].numel(), s["off"], s["dl"], compressed=True)
        p.execute()
        g2, se = p.g2()
        assert np.array_equal(res["g2"], g2, equal_nan=True) and np.array_equal(res["stderr"], se, equal_nan=True)
        if k == 0:
            r = orc.run_exact(s["imm"].cpu().numpy(), s["dq"], s["sq"], s["F"], s["dpl"])
            G2, IP, IF = p.sums()
            assert np.array_equal(G2, r["G2"]) and np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"])
        p.close()
    assert svc.percentiles()["frames"] == sum(s["F"] for s in streams)


def test_config1_vs_reference_float_output(xmt, orc):
    import make_stale_tail_cells as mk
    listed_cells = json.loads((Path(mk.__file__).parent / "stale_tail_cells_plain256.json").read_text())["cells"]
    c = mk.plain_case()
    plan = xmt.Plan(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"])
    plan.ingest_imm(c["imm"])
    plan.execute()
    g2, se = plan.g2()
    with_tail, without_tail, _ = mk.reference_float_outputs(c, orc)
    listed = np.zeros(g2.shape, bool)
    for t, q in listed_cells:
        listed[t, q] = True
    say(f"plain 256x256: {listed.sum()} stale-tail cells of {listed.size} excluded")
    for got, ref_with, ref_without in ((g2, with_tail[0], without_tail[0]), (se, with_tail[1], without_tail[1])):
        ref_with = ref_with.astype(np.float64)
        ref_without = ref_without.astype(np.float64)
        assert_close(got[~listed], ref_with[~listed], REF_RTOL)
        assert_close(got, ref_without, REF_RTOL)
    plan.close()
