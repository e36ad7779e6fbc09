"""The mixed serving stream at its full job count: 256 MD requests (LJ argon
of 64..8192 atoms and SPC/E water of 216..12000 molecules, in submission
order of the seeded stream) through the job server, with the steps of each
request scaled by 0.005 (U{1000..20000} -> 5..100 steps) so that the test
stays within a minute; every 16th request runs the FP64 form, the others the
mixed form.  Checks:
  * every job completes without error;
  * a sample of jobs gives bit-identical results when re-run alone (the
    stated tolerance of "same alone and in the mix" is bit-exact);
  * the three smallest FP64 jobs match the oracle's trajectory to 1e-8 A.
"""
import threading
import time

import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
from lumol_amd._native import LG_PRECISION_MIXED
from lumol_amd.serving import Job, Server, latency_stats
from lumol_amd.workloads import serving_stream

pytestmark = pytest.mark.gpu

SCALE = 0.005


def alone(name, s, steps, opts):
    if "precision" in opts:
        s.set_precision(opts["precision"])
    Simulation(MolecularDynamics(VelocityVerlet(opts["dt"]))).run(s, steps)
    return s.positions.copy(), s.velocities.copy()


def test_c5_stream_256_jobs():
    spec = serving_stream(256, SCALE, precision=LG_PRECISION_MIXED, fp64_every=16)
    # independent copies for the alone runs and the oracle (the server's
    # systems are consumed by the jobs)
    fresh = serving_stream(256, SCALE, precision=LG_PRECISION_MIXED, fp64_every=16)
    init = [(sys_.positions.copy(), sys_.velocities.copy()) for _, sys_, _, _ in spec]
    jobs = [Job("md", (lambda s=system: s), steps, md=opts, name=name) for name, system, steps, opts in spec]
    rng = np.random.default_rng(55)
    gaps = rng.exponential(2.0 / len(jobs), size=len(jobs))
    with Server(md_lanes=4) as server:
        def feed():
            for j, g in zip(jobs, gaps):
                time.sleep(g)
                server.submit(j)
        t = threading.Thread(target=feed)
        t.start()
        t.join()
        server.drain(timeout=600)
    errs = [(j.name, j.error) for j in jobs if j.error is not None]
    assert not errs, errs[:3]
    st = latency_stats(jobs)
    print(f"256 jobs: p50 {st['p50_s']:.3f} s, p99 {st['p99_s']:.3f} s, "
          f"{st['atom_steps_per_s'] / 1e6:.1f} M atom-steps/s")
    # bit-identical alone: every 32nd job, mixed and FP64 both represented
    sample = list(range(0, 256, 32)) + [16, 48]
    for k in sample:
        name, system, steps, opts = fresh[k]
        pos, vel = alone(name, system, steps, opts)
        assert np.array_equal(pos, jobs[k].result["positions"]), name
        assert np.array_equal(vel, jobs[k].result["velocities"]), name
        system.close()
    # the smallest FP64 jobs against the oracle
    fp64 = [k for k in range(0, 256, 16)]
    fp64.sort(key=lambda k: spec[k][1].size())
    for k in fp64[:3]:
        name, _, steps, opts = spec[k]
        o_sys = fresh[k][1]
        o_sys.positions, o_sys.velocities = init[k]
        o = lo.OracleSystem(o_sys, nthreads=16)
        o.md_reset()
        o.md_run(0, opts["dt"], steps)
        dp = np.abs(jobs[k].result["positions"] - o.positions()).max()
        assert dp < 1e-8, (name, dp)
