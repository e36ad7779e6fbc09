"""Host-side pieces of the job server (no GPU): footprint estimate and
latency statistics."""
from lumol_amd.serving import Job, device_bytes, latency_stats
from lumol_amd.workloads import argon, lj_mc_chain, spce


def test_device_bytes_grows_with_size_and_kspace():
    small, big = argon(ncell=3), argon(ncell=6)
    assert device_bytes(big, "md") > device_bytes(small, "md") > 0
    w = spce(nmol=216, L=18.64, cutoff=8.0, kmax=6)
    s, _, _ = lj_mc_chain(648)
    assert device_bytes(w, "md") > device_bytes(s, "md")
    assert device_bytes(s, "mc") > device_bytes(s, "md")


def test_latency_stats():
    jobs = []
    for k in range(100):
        j = Job("mc", None, 1000 + k)
        j.submitted, j.finished, j.natoms = 10.0 + 0.01 * k, 10.5 + 0.02 * k, 100
        jobs.append(j)
    st = latency_stats(jobs)
    assert st["jobs"] == 100
    assert abs(st["p50_s"] - (0.5 + 0.01 * 49.5)) < 1e-9
    assert st["p50_s"] < st["p99_s"] <= st["max_s"]
    assert st["atom_steps"] == sum(1000 + k for k in range(100))


# --- fault isolation of the lanes (host logic only: the device calls are
# replaced by stand-ins, so no GPU is needed) ---------------------------------------------
import threading

import numpy as np
import pytest

import lumol_amd.serving as serving
from lumol_amd.serving import Server


class FakeSystem:
    """Host stand-in with the System methods the lanes call."""
    closed = 0

    def __init__(self, n, bad=False):
        self.n, self.bad = n, bad
        self.positions = np.zeros((n, 3))
        self.velocities = np.zeros((n, 3))
        self.stream_slot = None
        self.pairs, self.coulomb = {}, None

    def size(self):
        return self.n

    def cache_init(self):
        return 0.0

    def cache_energy(self):
        return -1.0

    def close(self):
        FakeSystem.closed += 1


def _server():
    # no start(): the lane loops run on host threads of their own
    return Server(md_lanes=1, memory_budget=1 << 40)


def _bad_build():
    raise RuntimeError("cannot build this system")


def test_md_lane_survives_build_and_run_failures(monkeypatch):
    monkeypatch.setattr(serving, "device_bytes", lambda s, kind: 1000)

    def run_md(self, job, slot):
        s = job._system
        job.natoms = s.size()
        try:
            if s.bad:
                raise serving.nat.LumolGpuError(5, "positions or velocities became non-finite")
            job.result = {"positions": s.positions.copy()}
        finally:
            self._close_system(job)
        job.finished = 1.0
        job._done.set()

    monkeypatch.setattr(Server, "_run_md", run_md)
    srv = _server()
    FakeSystem.closed = 0
    jobs = [serving.Job("md", _bad_build, 10, name="bad-build"),
            serving.Job("md", lambda: FakeSystem(5), 10, name="ok-1"),
            serving.Job("md", lambda: FakeSystem(7, bad=True), 10, name="blows-up"),
            serving.Job("md", _bad_build, 10, name="bad-build-2"),
            serving.Job("md", lambda: FakeSystem(9), 10, name="ok-2")]
    for j in jobs:
        srv.submit(j)
    t = threading.Thread(target=srv._md_loop, args=(0,), daemon=True)
    t.start()
    for j in jobs:
        assert j._done.wait(10), j.name
    with srv._cv:
        srv._stop = True
        srv._cv.notify_all()
    t.join(10)
    assert not t.is_alive()
    assert [j.error is None for j in jobs] == [False, True, False, False, True]
    assert "cannot build" in str(jobs[0].error) and "non-finite" in str(jobs[2].error)
    assert jobs[1].result["positions"].shape == (5, 3)
    assert FakeSystem.closed == 3  # every built system is closed, failed or not
    assert srv._mem_used == 0


def test_mc_failure_fails_only_the_offending_chain(monkeypatch):
    monkeypatch.setattr(serving, "device_bytes", lambda s, kind: 1000)
    calls = []

    def fake_batch(systems, seeds, attempt0, nattempts, beta, delta):
        calls.append(len(systems))
        if any(s.bad for s in systems):
            raise serving.nat.LumolGpuError(1, "invalid temperature or amplitude")
        return np.array([na // 2 for na in nattempts], dtype=np.uint64)

    monkeypatch.setattr(serving, "mc_run_batch", fake_batch)
    srv = Server(md_lanes=1, memory_budget=1 << 40, mc_slice=100)
    FakeSystem.closed = 0
    jobs = [serving.Job("mc", (lambda k=k: FakeSystem(10 + k, bad=(k == 2))), 250,
                        mc={"beta": 1.0, "delta": 0.1, "seed": k}, name=f"chain-{k}") for k in range(5)]
    jobs.append(serving.Job("mc", _bad_build, 100, mc={"beta": 1.0, "delta": 0.1}, name="bad-build"))
    for j in jobs:
        srv.submit(j)
    t = threading.Thread(target=srv._mc_loop, daemon=True)
    t.start()
    for j in jobs:
        assert j._done.wait(10), j.name
    with srv._cv:
        srv._stop = True
        srv._cv.notify_all()
    t.join(10)
    assert not t.is_alive()
    assert [j.error is None for j in jobs] == [True, True, False, True, True, False]
    for j in (jobs[0], jobs[1], jobs[3], jobs[4]):
        assert j.result["accepted"] == 50 + 50 + 25  # three slices: 100, 100, 50 attempts
    assert FakeSystem.closed == 5
    assert srv._mem_used == 0
    # the failed batch was narrowed down chain by chain, then the rest batched again
    assert calls[0] == 5 and calls[1:6] == [1, 1, 1, 1, 1] and calls[6] == 4
