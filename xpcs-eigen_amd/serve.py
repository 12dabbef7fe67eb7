"""Streaming service over xpcs_mt.Stream: admission by HBM budget and continuous
batching of stream updates (no reference counterpart; the reference correlates
one finished file per process, src/xpcs/main.cpp:100-452).

Each admitted stream keeps its multi-tau level state resident in HBM and runs
on its own HIP stream, so the updates of different streams overlap on the
device.  `step(now)` is one turn of the request loop:

* newly arrived frames of every active stream are pushed (one update per
  stream per turn, whatever arrived since the last turn);
* a stream whose previous g2 snapshot has finished gets a new one covering all
  frames pushed so far -- frames that arrive while a snapshot is on the GPU
  join the next one instead of queueing behind it;
* finished snapshots are detected without blocking; every frame's latency is
  (turn that saw its first covering snapshot finished) - (its arrival time),
  and its service latency the same from the later of its arrival and its
  stream's admission (frames of a queued stream wait for a slot first);
* a stream whose last frame is covered is read out and closed, which frees its
  budget for streams waiting in the admission queue (first come, first served).

A source supplies frames: `frames` (total), `arrival(t)` (time frame t becomes
available), `available(now)` (frames available at `now`) and
`push(stream, a, b)` (push frames [a, b) into an xpcs_mt.Stream).
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass, field

import numpy as np

import xpcs_mt


@dataclass
class _Entry:
    sid: object
    width: int
    height: int
    dpl: int
    dq: np.ndarray
    sq: np.ndarray
    source: object
    nbytes: int
    stream: xpcs_mt.Stream | None = None
    pushed: int = 0
    snapped: int = 0
    reported: int = 0
    submitted: float = 0.0
    admitted: float | None = None
    lat: list = field(default_factory=list)


class StreamService:
    def __init__(self, budget_bytes: int, max_active: int = 16):
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.used = 0
        self.queue: deque[_Entry] = deque()
        self.active: dict = {}
        self.results: dict = {}
        self.latencies: list[float] = []
        self.service_latencies: list[float] = []
        self.admit_wait: list[float] = []
        self.peak_used = 0
        self.frames_done = 0

    def submit(self, sid, width, height, dpl, dqmap, sqmap, source, now=0.0):
        """Queue a stream; it is admitted as soon as a slot and its bytes are free."""
        nb = xpcs_mt.stream_bytes(width, height, source.frames, dpl, dqmap, sqmap)
        if nb > self.budget:
            raise ValueError(f"stream {sid} needs {nb} bytes, more than the whole budget {self.budget}")
        self.queue.append(_Entry(sid, width, height, dpl, np.asarray(dqmap), np.asarray(sqmap), source, nb,
                                 submitted=now))
        self._admit(now)

    def _admit(self, now):
        while self.queue and len(self.active) < self.max_active and self.used + self.queue[0].nbytes <= self.budget:
            e = self.queue.popleft()
            e.stream = xpcs_mt.Stream(e.width, e.height, e.source.frames, e.dpl, e.dq, e.sq)
            e.admitted = now
            self.admit_wait.append(now - e.submitted)
            self.used += e.nbytes
            self.peak_used = max(self.peak_used, self.used)
            self.active[e.sid] = e

    def step(self, now):
        """One turn of the request loop at time `now` (seconds, the sources' clock)."""
        closed = []
        for e in self.active.values():
            s = e.stream
            done = s.poll()
            if done > e.reported:
                arr = np.array([e.source.arrival(t) for t in range(e.reported, done)])
                lat = now - arr
                e.lat.extend(lat.tolist())
                self.latencies.extend(lat.tolist())
                self.service_latencies.extend((now - np.maximum(arr, e.admitted)).tolist())
                self.frames_done += done - e.reported
                e.reported = done
            avail = e.source.available(now)
            if avail > e.pushed:
                e.source.push(s, e.pushed, avail)
                e.pushed = avail
            in_flight = e.snapped > e.reported
            if e.pushed > e.snapped and not in_flight and s.frames()[1] > 0:
                s.snapshot()
                e.snapped = e.pushed
            if e.reported == e.source.frames:
                closed.append(e.sid)
        for sid in closed:
            e = self.active.pop(sid)
            g2, se = e.stream.g2()
            self.results[sid] = dict(g2=g2, stderr=se, frames=e.reported, latency=np.array(e.lat))
            e.stream.close()
            self.used -= e.nbytes
        if closed or self.queue:
            self._admit(now)

    def idle(self):
        return not self.active and not self.queue

    def percentiles(self, service=False):
        """p50 / p99 / max of the frames' latencies (service=True: from admission on)."""
        src = self.service_latencies if service else self.latencies
        if not src:
            return None
        a = np.array(src)
        return {"p50_ms": 1e3 * float(np.percentile(a, 50)), "p99_ms": 1e3 * float(np.percentile(a, 99)),
                "max_ms": 1e3 * float(a.max()), "frames": int(a.size)}


class DeviceFrames:
    """A source whose frames are already in device memory (a compressed IMM
    image, its payload offsets and event counts), arriving at `fps` from
    `start` on the serving loop's clock.  `total_frames` is the stream's full
    length (its admission size); only the frames of `off` exist."""

    def __init__(self, imm, off, dlen, total_frames, start, fps):
        self.imm, self.off, self.dlen = imm, off, dlen
        self.frames = total_frames
        self.generated = len(off)
        self.start, self.fps = start, fps

    def arrival(self, t):
        return self.start + t / self.fps

    def available(self, now):
        if now < self.start:
            return 0
        return int(min(self.generated, np.floor((now - self.start) * self.fps + 1e-9) + 1))

    def push(self, stream, a, b):
        stream.push_imm_device(self.imm.data_ptr(), self.imm.numel(), self.off[a:b], self.dlen[a:b])
