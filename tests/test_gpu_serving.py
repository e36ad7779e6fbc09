"""Serving: every job gives the same result alone and in the mix.

The server overlaps MD jobs on several HIP streams and runs Monte Carlo jobs
as chains of one batched launch that other chains join and leave between
slices.  The stated tolerance for "same" is bit-exact: the MD kernels are
deterministic and use no cross-job state, and a chain's random numbers depend
only on (seed, attempt).
"""
import threading
import time

import numpy as np
import pytest

from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
from lumol_amd._native import LG_PRECISION_MIXED
from lumol_amd.serving import Job, Server, latency_stats
from lumol_amd.workloads import argon, lj_mc_chain, nacl, spce

pytestmark = pytest.mark.gpu


def md_jobs():
    return [
        Job("md", lambda: argon(ncell=3), 300, md={"dt": 1.0}, name="argon-108"),
        Job("md", lambda: argon(ncell=5), 200, md={"dt": 1.0, "integrator": "lf"}, name="argon-500-lf"),
        Job("md", lambda: nacl(ncell=4, cutoff=10.0, kmax=7), 120,
            md={"dt": 1.0, "thermostat": (1200.0, 100.0)}, name="nacl-512-nvt"),
        Job("md", lambda: spce(nmol=216, L=18.64, cutoff=8.0, kmax=6), 100, md={"dt": 0.5}, name="water-648"),
        Job("md", lambda: spce(nmol=216, L=18.64, cutoff=8.0, kmax=6), 100,
            md={"dt": 0.5, "precision": LG_PRECISION_MIXED}, name="water-648-mixed"),
    ]


def mc_jobs():
    out = []
    for k, n in enumerate([108, 400, 1000, 256, 1500, 700]):
        def build(n=n, k=k):
            return lj_mc_chain(n, seed=300 + k)[0]
        _, beta, delta = lj_mc_chain(n, seed=300 + k)
        out.append(Job("mc", build, 1200 + 350 * k, mc={"beta": beta, "delta": delta, "seed": 9000 + k},
                       name=f"mc-{n}"))
    return out


def run_alone(job):
    s = job.build()
    if job.kind == "mc":
        s.cache_init()
        acc = s.mc_run(job.mc["seed"], 0, job.steps, job.mc["beta"], job.mc["delta"])
        return {"positions": s.positions.copy(), "accepted": acc, "energy": s.cache_energy()}
    md = job.md
    if "precision" in md:
        s.set_precision(md["precision"])
    from lumol_amd import BerendsenThermostat, LeapFrog
    integ = LeapFrog(md["dt"]) if md.get("integrator") == "lf" else VelocityVerlet(md["dt"])
    prop = MolecularDynamics(integ)
    if md.get("thermostat"):
        prop.set_thermostat(BerendsenThermostat(*md["thermostat"]))
    Simulation(prop).run(s, job.steps)
    return {"positions": s.positions.copy(), "velocities": s.velocities.copy()}


def same(a, b):
    for key, v in b.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(a[key], v), key
        else:
            assert a[key] == v, key


def test_results_same_alone_and_in_the_mix():
    jobs = md_jobs() + mc_jobs()
    late = jobs[1::3]
    early = [j for j in jobs if j not in late]
    with Server(md_lanes=3, mc_slice=400) as server:
        for j in early:
            server.submit(j)

        def later():
            time.sleep(0.05)  # arrive while the first jobs are on the device
            for j in late:
                server.submit(j)

        t = threading.Thread(target=later)
        t.start()
        t.join()
        server.drain(timeout=300)
        assert server.mc_launches > 1
    for j in jobs:
        assert j.error is None, (j.name, j.error)
        same(j.result, run_alone(j))
    st = latency_stats(jobs)
    assert st["jobs"] == len(jobs) and 0 < st["p50_s"] <= st["p99_s"] <= st["max_s"]


def test_memory_admission_serialises_but_completes():
    # a budget below any two jobs: they run one after the other, same results
    jobs = mc_jobs()[:3] + md_jobs()[:2]
    with Server(md_lanes=2, mc_slice=300, memory_budget=1) as server:
        for j in jobs:
            server.submit(j)
        server.drain(timeout=300)
    for j in jobs:
        assert j.error is None, (j.name, j.error)
        same(j.result, run_alone(j))
    # no two jobs overlapped in time
    spans = sorted((j.started, j.finished) for j in jobs)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1 - 1e-3


def _blowup():
    # two atoms on top of each other: non-finite forces on the first step
    from lumol_amd import LennardJones, PairInteraction, System, UnitCell, units
    s = System.from_arrays(["Ar"] * 3, [(5.0, 5.0, 5.0), (5.0, 5.0, 5.0), (9.0, 9.0, 9.0)], UnitCell.cubic(20.0))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(3.4, units.from_(1.0, "kJ/mol")), 8.0))
    return s


def test_failing_jobs_do_not_disturb_the_mix():
    from lumol_amd._native import LG_E_NONFINITE, LumolGpuError

    def bad_build():
        raise ValueError("the spec cannot be built")

    good = md_jobs()[:3] + mc_jobs()[:3]
    bad = [Job("md", _blowup, 100, md={"dt": 1.0}, name="blows-up"),
           Job("md", bad_build, 100, md={"dt": 1.0}, name="bad-spec"),
           Job("mc", lambda: lj_mc_chain(200, seed=77)[0], 500, mc={"beta": float("nan"), "delta": 0.3, "seed": 1},
               name="mc-bad-beta")]
    jobs = good[:2] + bad[:1] + good[2:4] + bad[1:] + good[4:]
    with Server(md_lanes=2, mc_slice=300) as server:
        for j in jobs:
            server.submit(j)
        server.drain(timeout=300)
    assert isinstance(bad[0].error, LumolGpuError) and bad[0].error.code == LG_E_NONFINITE
    assert isinstance(bad[1].error, ValueError)
    assert isinstance(bad[2].error, LumolGpuError)
    for j in good:
        assert j.error is None, (j.name, j.error)
        same(j.result, run_alone(j))


def test_large_water_system_sets_up_in_milliseconds():
    # A 36 000-atom water job's set-up (topology, tables, first list build
    # and forces) once took ~2 s, all in the angle/dihedral search of the
    # topology, which compared every pair of bonds; over the bonds incident to
    # each bond's atoms it takes tens of milliseconds.  The bound leaves a
    # wide margin on a loaded host.  Its angles, and their forces, are those
    # of the oracle.
    from lumol_amd.workloads import water_box
    import lumol_oracle as lo
    s = water_box(12000, seed=7)
    s.set_precision(LG_PRECISION_MIXED)
    t0 = time.perf_counter()
    s.forces()
    assert time.perf_counter() - t0 < 0.5
    s.close()
    small = water_box(300, seed=7)
    o = lo.OracleSystem(small, nthreads=8)
    f, fo = small.forces(), o.forces()
    assert np.abs(f - fo).max() / np.abs(fo).max() < 1e-10
