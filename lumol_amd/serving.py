"""Continuous-batching job server on one GPU.

A job is a system spec (a callable that builds a ``System`` on the host)
plus the work to run on it: MD steps (VelocityVerlet / LeapFrog, optional
Berendsen thermostat, FP64 or mixed forces) or Monte Carlo attempts
(Translate moves, Metropolis at beta).  Jobs of uneven size are submitted at
any time; the server

* admits a job when its estimated device footprint fits the memory budget
  (positions, velocities, neighbour list, cell list and rho(k) caches stay
  resident in HBM for the job's whole life);
* runs MD jobs on ``md_lanes`` worker threads, each with its own library
  context and HIP stream, so that the kernels of several jobs overlap on the
  device;
* runs every admitted Monte Carlo job as one chain of a batched launch
  (``lg_mc_run_batch``: one persistent workgroup per chain); a launch gives
  each chain up to ``mc_slice`` attempts, so chains that arrive while others
  run join at the next launch and finished chains leave it;
* records submit, start and finish times, so that job latency
  (submit -> result) percentiles and aggregate atom-steps/s can be reported.

Results do not depend on what else runs: MD kernels are deterministic and
use no cross-job state, and a chain's random numbers depend only on its seed
and attempt index (counter-based Philox), so a job gives the same bits alone
and in the mix.
"""
import collections
import ctypes
import math
import threading
import time

import numpy as np

from . import _native as nat
from .system import (BerendsenThermostat, LeapFrog, MolecularDynamics, Simulation, VelocityVerlet, mc_run_batch)

MC_SLOT = 0  # library context of the Monte Carlo lane; MD lanes use 1..md_lanes


class Job:
    """One request.  `build()` returns a host-side System (no device work);
    `steps` is MD steps for kind "md" and attempted moves for kind "mc".

    md: dict(integrator="vv"|"lf", dt=1.0, thermostat=None|(T, tau),
             precision=LG_PRECISION_FP64|LG_PRECISION_MIXED)
    mc: dict(beta=..., delta=..., seed=...)
    """

    def __init__(self, kind, build, steps, md=None, mc=None, name=""):
        if kind not in ("md", "mc"):
            raise ValueError("job kind must be 'md' or 'mc'")
        self.kind = kind
        self.build = build
        self.steps = int(steps)
        self.md = dict(md or {})
        self.mc = dict(mc or {})
        self.name = name
        self.natoms = None
        self.submitted = self.started = self.finished = None
        self.result = None
        self.error = None
        self._done = threading.Event()

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError(f"job {self.name} not finished")
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.submitted

    @property
    def atom_steps(self):
        return self.natoms * self.steps if self.kind == "md" else self.steps


def device_bytes(system, kind):
    """Estimated device memory of a job (bytes): per-atom state and
    neighbour list, k-space tables, Monte Carlo cell list."""
    n = system.size()
    vol = system.cell.volume() if not system.cell.is_infinite() else None
    rc = max([p.cutoff() for p in system.pairs.values()] + [getattr(system.coulomb, "cutoff", 0.0) or 0.0])
    per_atom = 8 * 32 + 64  # pos, vel, acc, force, ref, 3 snapshots; meta, cell list
    if vol and rc > 0:
        rl = 1.08 * rc
        cap = min(n, int(1.5 * 4.0 / 3.0 * math.pi * rl ** 3 * n / vol + 64))
    else:
        cap = n
    total = n * per_atom + 4 * n * cap
    coul = system.coulomb
    if coul is not None and hasattr(coul, "kmax"):
        kmax = coul.kmax
        nk = 2.1 * kmax ** 3  # half-space sphere
        total += n * (5 * kmax + 3) * 24 + int(nk * (n / 64 + 1) * 16) + n * 32 * 32
    if kind == "mc":
        total += n * 40 * 5
    return int(total * 1.25) + (8 << 20)


def device_free_bytes(device=0):
    """Free and total device memory (hipMemGetInfo).  The library is loaded
    first, so that its loader sets the hardware-queue count before this call
    initialises the HIP runtime."""
    nat.lib()
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    hip.hipSetDevice(device)
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return free.value, total.value


class Server:
    """See the module docstring.  Use as a context manager, or start() /
    close().  `memory_budget` defaults to 80% of the free device memory."""

    def __init__(self, device=0, md_lanes=4, mc_max_chains=1024, mc_slice=4096, memory_budget=None):
        self.device = device
        self.md_lanes = int(md_lanes)
        self.mc_max_chains = int(mc_max_chains)
        self.mc_slice = int(mc_slice)
        self.memory_budget = memory_budget
        self._cv = threading.Condition()
        self._md_queue = collections.deque()
        self._mc_queue = collections.deque()
        self._mem_used = 0
        self._stop = False
        self._threads = []
        self.jobs = []
        self.mc_launches = 0

    # -- lifecycle ------------------------------------------------------------------------
    def start(self):
        if self.memory_budget is None:
            free, _ = device_free_bytes(self.device)
            self.memory_budget = int(0.8 * free)
        # create the contexts up front (one stream each)
        for slot in range(self.md_lanes + 1):
            nat.context(self.device, slot)
        self._threads = [threading.Thread(target=self._mc_loop, name="mc-lane", daemon=True)]
        self._threads += [threading.Thread(target=self._md_loop, args=(k,), name=f"md-lane-{k}", daemon=True)
                          for k in range(self.md_lanes)]
        for t in self._threads:
            t.start()
        return self

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._threads = []

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.close()

    def submit(self, job):
        job.submitted = time.perf_counter()
        with self._cv:
            if self._stop:
                raise RuntimeError("server is closed")
            self.jobs.append(job)
            (self._md_queue if job.kind == "md" else self._mc_queue).append(job)
            self._cv.notify_all()
        return job

    def drain(self, timeout=None):
        """Waits for every submitted job."""
        for job in list(self.jobs):
            job._done.wait(timeout)

    # -- admission ------------------------------------------------------------------------
    def _reserve(self, nbytes):
        """Caller holds self._cv.  True when the job fits the budget now (a job
        larger than the whole budget runs alone)."""
        if self._mem_used == 0 or self._mem_used + nbytes <= self.memory_budget:
            self._mem_used += nbytes
            return True
        return False

    def _release(self, nbytes):
        with self._cv:
            self._mem_used -= nbytes
            self._cv.notify_all()

    @staticmethod
    def _fail(job, err):
        job.error = err
        job.finished = time.perf_counter()
        Server._close_system(job)
        job._done.set()

    @staticmethod
    def _close_system(job):
        s, job._system = job._system, None
        if s is not None:
            try:
                s.close()
            except Exception:  # noqa: BLE001 - the job already carries its error
                pass

    def _head(self, queue, kind):
        """Caller holds self._cv.  The job at the head of `queue` with its
        host-side System built and its footprint estimated; a job whose
        build() or estimate raises is removed and failed (returns None)."""
        cand = queue[0]
        if cand._system is None:
            try:
                cand._system = cand.build()
                cand._bytes = device_bytes(cand._system, kind)
            except Exception as err:  # noqa: BLE001 - reported on the job
                queue.popleft()
                cand._bytes = 0
                self._fail(cand, err)
                return None
        return cand

    # -- MD lanes -------------------------------------------------------------------------
    def _md_loop(self, lane):
        slot = 1 + lane
        while True:
            with self._cv:
                job = None
                while job is None:
                    if self._md_queue:
                        cand = self._head(self._md_queue, "md")
                        if cand is None:
                            continue
                        if self._reserve(cand._bytes):
                            job = self._md_queue.popleft()
                            break
                    elif self._stop:
                        return
                    self._cv.wait(0.05)
            try:
                self._run_md(job, slot)
            except Exception as err:  # noqa: BLE001 - reported on the job
                self._fail(job, err)
            finally:
                self._release(job._bytes)

    def _run_md(self, job, slot):
        s = job._system
        s.stream_slot = slot
        job.natoms = s.size()
        job.started = time.perf_counter()
        md = job.md
        if "precision" in md:
            s.set_precision(md["precision"])
        integ = LeapFrog(md.get("dt", 1.0)) if md.get("integrator", "vv") == "lf" else VelocityVerlet(
            md.get("dt", 1.0))
        prop = MolecularDynamics(integ)
        if md.get("thermostat"):
            prop.set_thermostat(BerendsenThermostat(*md["thermostat"]))
        try:
            Simulation(prop).run(s, job.steps)
            job.result = {"positions": s.positions.copy(), "velocities": s.velocities.copy()}
        finally:
            self._close_system(job)
        job.finished = time.perf_counter()
        job._done.set()

    # -- Monte Carlo lane -----------------------------------------------------------------
    def _mc_loop(self):
        active = []
        while True:
            with self._cv:
                while True:
                    while self._mc_queue and len(active) < self.mc_max_chains:
                        cand = self._head(self._mc_queue, "mc")
                        if cand is None:
                            continue
                        if not self._reserve(cand._bytes):
                            break
                        active.append(self._mc_queue.popleft())
                        active[-1]._fresh = True
                    if active:
                        break
                    if self._stop and not self._mc_queue:
                        return
                    self._cv.wait(0.05)
            # set-up of the chains admitted just now (outside the lock)
            ready = []
            for job in active:
                if getattr(job, "_fresh", False):
                    try:
                        s = job._system
                        s.stream_slot = MC_SLOT
                        job.natoms = s.size()
                        job.started = time.perf_counter()
                        s.cache_init()
                        job._done_attempts = 0
                        job._accepted = 0
                        job._fresh = False
                    except Exception as err:  # noqa: BLE001
                        self._fail(job, err)
                        self._release(job._bytes)
                        continue
                ready.append(job)
            active = ready
            if not active:
                continue
            n_att = [min(self.mc_slice, j.steps - j._done_attempts) for j in active]
            acc = self._mc_launch(active, n_att)
            ok = [a is not None for a in acc]
            active = [j for j, k in zip(active, ok) if k]
            n_att = [na for na, k in zip(n_att, ok) if k]
            acc = [a for a in acc if a is not None]
            still = []
            for j, a, na in zip(active, acc, n_att):
                j._accepted += int(a)
                j._done_attempts += na
                if j._done_attempts >= j.steps:
                    s = j._system
                    j.result = {"positions": s.positions.copy(), "accepted": j._accepted,
                                "energy": s.cache_energy()}
                    s.close()
                    j._system = None
                    j.finished = time.perf_counter()
                    j._done.set()
                    self._release(j._bytes)
                else:
                    still.append(j)
            active = still


    def _mc_launch(self, active, n_att):
        """One batched launch of the chains `active`; the accepted counts, or
        None for a chain that failed.  A failed batch does not fail every
        chain: the library validates every chain before it launches, so a
        failure leaves the chains where they were, and each chain is then
        run on its own so that only the offending ones fail."""
        def run(jobs, natt):
            return mc_run_batch([j._system for j in jobs], [j.mc.get("seed", 0) for j in jobs],
                                [j._done_attempts for j in jobs], natt, [j.mc["beta"] for j in jobs],
                                [j.mc["delta"] for j in jobs])
        try:
            out = [int(a) for a in run(active, n_att)]
            self.mc_launches += 1
            return out
        except Exception:  # noqa: BLE001 - narrowed down below
            pass
        out = []
        for j, na in zip(active, n_att):
            try:
                out.append(int(run([j], [na])[0]))
                self.mc_launches += 1
            except Exception as err:  # noqa: BLE001 - reported on the job
                self._fail(j, err)
                self._release(j._bytes)
                out.append(None)
        return out


Job._system = None
Job._bytes = 0


def latency_stats(jobs):
    """p50 / p99 / max job latency (s) and aggregate atom-steps/s over the
    interval from the first submission to the last result."""
    done = [j for j in jobs if j.finished is not None and j.error is None]
    if not done:
        return {}
    lat = np.array([j.latency for j in done])
    t0 = min(j.submitted for j in done)
    t1 = max(j.finished for j in done)
    work = sum(j.atom_steps for j in done)
    return {"jobs": len(done), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "max_s": float(lat.max()), "wall_s": t1 - t0, "atom_steps": int(work),
            "atom_steps_per_s": work / (t1 - t0)}
