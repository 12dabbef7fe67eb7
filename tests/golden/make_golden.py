"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

Everything here is computed by pure-Python definitions that do not share code
with the oracle or the HIP path:

* multitau_bruteforce.npz -- tiny seeded event rows and their exact int64 sums by
  pair enumeration over level-0 frames (no iterative binning).  With
  bin_L(t) = t >> L, lf_L = F >> L and tau' = tau >> L:
      G2 = sum over ordered pairs (a, b) of events with bin_L(b) - bin_L(a) = tau'
           and bin_L(b) < lf_L of c_a * c_b
      IP = sum of c_a with bin_L(a) < lf_L - tau'
      IF = sum of c_a with tau' <= bin_L(a) < lf_L
  (x_L[k] = sum of the level-0 counts whose frame bins to k, so the products
  of binned values expand into these pair sums.)
* tau_tables.json -- schedules quoted as known answers for the benchmark
  configurations (length, head and tail taus).
"""
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def schedule(F, dpl):
    """Direct statement of the tau schedule: level 0 has 2*dpl taus at step 1,
    level i has dpl taus at step 2^i, a tau being kept while tau + 2^i <= F."""
    import math
    if F < 2 * dpl:
        max_level = 0
    else:
        max_level = int(math.floor(math.log2(F) - math.log2(1.0 + 1.0 / dpl)) - math.log2(dpl))
    out = []
    prev = 0
    for i in range(max_level + 1):
        step = 2 ** i
        for _ in range(2 * dpl if i == 0 else dpl):
            if prev + 2 * step > F:
                break
            prev += step
            out.append((i, prev))
    return out


def pair_sums(frames, counts, F, sched):
    T = len(sched)
    G2 = np.zeros(T, np.int64)
    IP = np.zeros(T, np.int64)
    IF = np.zeros(T, np.int64)
    ev = list(zip(frames, counts))
    for i, (L, tau) in enumerate(sched):
        lf = F >> L
        tp = tau >> L
        g = ip = iff = 0
        for (ta, ca) in ev:
            ba = ta >> L
            if ba < lf - tp:
                ip += ca
            if tp <= ba < lf:
                iff += ca
            for (tb, cb) in ev:
                bb = tb >> L
                if bb - ba == tp and bb < lf:
                    g += ca * cb
        G2[i], IP[i], IF[i] = g, ip, iff
    return G2, IP, IF


def main():
    rng = np.random.default_rng(20261202)
    cases = []
    for (F, dpl, n_ev, dup) in [(64, 2, 10, False), (100, 4, 25, False), (257, 4, 60, True),
                                (500, 8, 40, False), (33, 1, 12, True), (1000, 4, 5, False)]:
        frames = np.sort(rng.choice(F, size=n_ev, replace=False)).astype(np.int64)
        if dup:   # repeated frames (as avg_frames > 1 or repeated IMM entries produce)
            frames = np.sort(np.concatenate([frames, frames[: n_ev // 4]]))
        counts = rng.integers(1, 7, size=frames.size).astype(np.int64)
        sched = schedule(F, dpl)
        G2, IP, IF = pair_sums(frames.tolist(), counts.tolist(), F, sched)
        cases.append(dict(F=F, dpl=dpl, frames=frames, counts=counts, G2=G2, IP=IP, IF=IF,
                          level=np.array([s[0] for s in sched], np.int32),
                          tau=np.array([s[1] for s in sched], np.int32)))
    arrays = {}
    for k, c in enumerate(cases):
        for key, val in c.items():
            arrays[f"c{k}_{key}"] = np.asarray(val)
    arrays["ncases"] = np.array(len(cases))
    np.savez(HERE / "multitau_bruteforce.npz", **arrays)

    tables = {}
    for (F, dpl) in [(1000, 4), (10000, 8), (20000, 4), (500, 4), (20000, 8), (7, 4), (16, 4)]:
        s = schedule(F, dpl)
        tables[f"{F},{dpl}"] = {"T": len(s), "levels": sorted({L for L, _ in s}),
                                "head": [t for _, t in s[:8]], "tail": [t for _, t in s[-4:]]}
    (HERE / "tau_tables.json").write_text(json.dumps(tables, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
