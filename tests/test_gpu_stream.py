"""Streams: online multi-tau with HBM-resident level state.  A snapshot after n
frames equals a plan of n frames bit for bit (sums and g2), whatever the push
pieces; the serving loop gives every stream the same result alone and in a mix
with queued admission."""
import numpy as np
import pytest

import cases
import synth

pytestmark = pytest.mark.gpu


class HostFrames:
    """Frames of a counts array arriving at `fps` from `start` (virtual clock)."""

    def __init__(self, counts, W, H, start=0.0, fps=100.0):
        self.counts, self.W, self.H, self.start, self.fps = counts, W, H, start, fps
        self.frames = counts.shape[0]

    def arrival(self, t):
        return self.start + t / self.fps

    def available(self, now):
        if now < self.start:
            return 0
        return int(min(self.frames, np.floor((now - self.start) * self.fps + 1e-9) + 1))

    def push(self, stream, a, b):
        stream.push_imm(synth.imm_bytes_sparse(self.counts[a:b], self.W, self.H))


def plan_results(xmt, c, n):
    p = xmt.Plan(c["W"], c["H"], n, c["dpl"], c["dq"], c["sq"])
    p.ingest_imm(synth.imm_bytes_sparse(c["counts"][:n], c["W"], c["H"]))
    p.execute()
    G2, IP, IF = p.sums()
    g2, se = p.g2()
    p.close()
    return G2, IP, IF, g2, se


def assert_same(got, want):
    for x, y in zip(got, want):
        assert x.shape == y.shape
        assert np.array_equal(x, y, equal_nan=x.dtype.kind == "f")


@pytest.mark.parametrize("dpl", [4, 8])
def test_snapshots_equal_plans_of_the_prefix(xmt, orc, dpl):
    c = cases.sparse_case(W=24, H=20, F=1400, dpl=dpl, mu=0.3, tau_c=12.0, seed=101 + dpl)
    F = c["F"]
    s = xmt.Stream(c["W"], c["H"], F, dpl, c["dq"], c["sq"])
    rng = np.random.default_rng(5)
    checks = {5, 37, 64, 129, 700, 1001, F}
    n = 0
    while n < F:
        k = int(min(F - n, rng.integers(1, 150)))
        # stop exactly at the check points
        nxt = min([x for x in checks if x > n] + [F])
        k = min(k, nxt - n)
        s.push_imm(synth.imm_bytes_sparse(c["counts"][n:n + k], c["W"], c["H"]))
        n += k
        if n in checks:
            print("snapshot at", n, flush=True)
            s.snapshot()
            G2, IP, IF = s.sums()
            g2, se = s.g2()
            assert_same((G2, IP, IF, g2, se), plan_results(xmt, c, n))
    r = orc.run_exact(c["imm"], c["dq"], c["sq"], F, dpl)
    assert np.array_equal(G2, r["G2"]) and np.array_equal(IP, r["IP"]) and np.array_equal(IF, r["IF"])
    s.close()


def test_frequent_snapshots(xmt):
    """A snapshot after every push of 1..3 frames recomputes only the levels
    that changed since the previous one; compared with plans of the prefix
    around power-of-two frame counts (where many levels change at once)."""
    c = cases.sparse_case(W=16, H=16, F=600, dpl=4, mu=0.3, tau_c=10.0, seed=131)
    F = c["F"]
    s = xmt.Stream(c["W"], c["H"], F, 4, c["dq"], c["sq"])
    rng = np.random.default_rng(7)
    checks = {8, 31, 32, 33, 63, 64, 127, 128, 129, 255, 256, 300, 511, 512, 600}
    n = 0
    while n < F:
        k = int(min(F - n, rng.integers(1, 4)))
        nxt = min([x for x in checks if x > n] + [F])
        k = min(k, nxt - n)
        s.push_imm(synth.imm_bytes_sparse(c["counts"][n:n + k], c["W"], c["H"]))
        n += k
        if s.frames()[1] > 0:
            s.snapshot()
        if n in checks:
            print("snapshot at", n, flush=True)
            G2, IP, IF = s.sums()
            g2, se = s.g2()
            assert_same((G2, IP, IF, g2, se), plan_results(xmt, c, n))
    s.close()


def test_device_push_equals_host_push(xmt):
    import torch
    W = H = 32
    F = 900
    dq, sq = synth.annular_qmaps(W, H, 6, 3)
    imm, off, dl = synth.imm_device_sparse(W, H, F, 0.2, 9.0, 11, device="cuda")
    torch.cuda.synchronize()
    host = imm.cpu().numpy().tobytes()
    a = xmt.Stream(W, H, F, 4, dq, sq)
    b = xmt.Stream(W, H, F, 4, dq, sq)
    for lo in range(0, F, 130):
        hi = min(F, lo + 130)
        a.push_imm_device(imm.data_ptr(), imm.numel(), off[lo:hi], dl[lo:hi])
    a_first = int(off[0]) - 1024
    b.push_imm(host[a_first:])
    a.snapshot()
    b.snapshot()
    assert_same(a.sums() + a.g2(), b.sums() + b.g2())
    p = xmt.Plan(W, H, F, 4, dq, sq)
    p.ingest_imm(host)
    p.execute()
    assert_same(a.sums() + a.g2(), p.sums() + p.g2())


def test_stream_errors(xmt):
    c = cases.sparse_case(W=16, H=16, F=100, dpl=4, mu=0.3, tau_c=5.0, seed=7)
    s = xmt.Stream(16, 16, 60, 4, c["dq"], c["sq"])
    s.push_imm(synth.imm_bytes_sparse(c["counts"][:50], 16, 16))
    with pytest.raises(xmt.XpcsError, match="RANGE"):
        s.push_imm(synth.imm_bytes_sparse(c["counts"][50:70], 16, 16))
    # a pixel index outside the detector is reported when results are read
    bad = bytearray(synth.imm_bytes_sparse(c["counts"][50:52], 16, 16))
    import struct
    dlen = struct.unpack_from("<I", bad, 152)[0]
    assert dlen > 0
    struct.pack_into("<i", bad, 1024, 16 * 16 + 5)
    s.push_imm(bytes(bad))
    s.snapshot()
    with pytest.raises(xmt.XpcsError, match="RANGE"):
        s.g2()


def test_service_mix_equals_alone(xmt):
    import serve
    streams = []
    for k, (W, F, dpl, mu, tau) in enumerate([(24, 700, 4, 0.3, 10.0), (32, 500, 8, 0.1, 30.0),
                                              (16, 900, 4, 1.0, 3.0), (28, 650, 8, 0.2, 50.0),
                                              (20, 400, 4, 0.5, 8.0)]):
        c = cases.sparse_case(W=W, H=W, F=F, dpl=dpl, mu=mu, tau_c=tau, seed=300 + k)
        streams.append(c)
    sizes = [xmt.stream_bytes(c["W"], c["H"], c["F"], c["dpl"], c["dq"], c["sq"]) for c in streams]
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]      # at most two of the largest at once
    svc = serve.StreamService(budget, max_active=3)
    for k, c in enumerate(streams):
        svc.submit(k, c["W"], c["H"], c["dpl"], c["dq"], c["sq"],
                   HostFrames(c["counts"], c["W"], c["H"], start=0.5 * k, fps=200.0), now=0.0)
    assert len(svc.queue) > 0, "the budget should hold some streams back"
    now = 0.0
    for _ in range(100000):
        svc.step(now)
        if svc.idle():
            break
        now += 0.01
    assert svc.idle()
    assert svc.peak_used <= budget
    for k, c in enumerate(streams):
        res = svc.results[k]
        print("stream", k, "frames", res["frames"], flush=True)
        G2, IP, IF, g2, se = plan_results(xmt, c, c["F"])
        assert np.array_equal(res["g2"], g2, equal_nan=True) and np.array_equal(res["stderr"], se, equal_nan=True)
    pc = svc.percentiles()
    assert pc["frames"] == sum(c["F"] for c in streams)
