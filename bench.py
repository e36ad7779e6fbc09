#!/usr/bin/env python3
"""Throughput of Lumol's force-and-energy hot path on one MI355X.

Workloads (`--workload`):

* ``spce`` (default, the headline configuration): flexible SPC/E water NVE
  -- 3000 molecules (9000 atoms), box 44.8 A, Ewald rc 10 A kmax 10,
  inter-molecular LJ + Coulomb, harmonic bonds and angles, Velocity-Verlet
  dt = 0.5 fs, 2000 timed steps.  ``value`` is the mixed form's aggregate
  atom-steps/s over `--batch` systems (seeds 3, 4, ...) advancing together;
  the line carries the FP64 and mixed single-system rates, every form's
  error against the FP64 oracle (force RMS/max, energy, virial) and NVE
  drift in kJ/mol/atom/ns over 2000 steps, the ratio mixed-batched / FP64,
  the dominant kernel's roofline, the serving stream's p50/p99 latency at a
  stated step scale and the oracle's CPU throughput on this host.
* ``nacl``: NaCl Ewald NVT -- 4096 ions (2048 Na+, 2048 Cl-) on 8^3
  rocksalt cells, a = 5.64 A (box 45.12 A), seeded +-0.05 A jitter,
  Lorentz-Berthelot LJ + Ewald (rc 12 A, kmax 9, alpha = pi/rc), Berendsen
  1200 K (tau = 100 steps), Velocity-Verlet dt = 1 fs; FP64 and mixed forms.
* ``serve``: the job server under a Poisson stream of uneven jobs -- MD
  jobs of LJ argon (64..8192 atoms) and SPC/E Ewald water (216..12000
  molecules) in the mixed form, with 1k..20k steps scaled by
  `--steps-scale`, plus Monte Carlo chains as in ``mc`` with `--sweeps`
  sweeps; arrivals at a fixed rate sized to `--serve-seconds`.  Reports
  p50/p99 job latency (submit -> result) per class and aggregate
  atom-steps/s.  Host-side construction of each job's system happens before
  its submission and is not timed.
* ``spce-batched``: `--batch` copies of the SPC/E system (different seeds)
  run at once, each on its own library context (HIP stream), in the mixed
  form; the FP64 form of one system alone is timed in the same run, and the
  line reports the aggregate mixed-batched rate and its ratio to it.
* ``ka``: one large system -- the 1,000,000-atom Kob-Andersen 80:20 LJ
  mixture at rho* = 1.2, cutoff 2.5 sigma_AA, NVE VV dt = 1 fs, in the mixed
  form (FP32 forces, compensated, FP64 accumulation and integration); the
  FP64 form is timed beside it.  The force error is stated against the FP64
  GPU form at this size (which meets the oracle at 1e-10 on smaller KA
  systems); the CPU baseline times the oracle's O(N^2) force evaluation on a
  16384-atom KA system and scales it by N^2.
* ``mc``: batched Monte Carlo -- independent LJ argon chains (sizes uniform
  in 108..4000 atoms, rho* = 0.8, T* = 1.2, rc = 2.5 sigma, translation
  amplitude 0.3 sigma, Philox per chain), run together on the device, one
  workgroup per chain.  MC atom-steps = attempted single-particle moves
  (N per sweep); `--sweeps` sweeps per chain (a bounded sample of the 10k).

One step = one MD step on the device: integrator halves, neighbour-list
maintenance, pair kernel (LJ + Ewald real space), intra-molecular terms,
k-space rho(k) and forces, thermostat.  Both arithmetic forms are measured:
``fp64`` (all FP64) and ``mixed`` (FP32 pair/erfc/trig with FP64 positions,
phase reduction and accumulation).

Prints ONE JSON line.  Rates are over `--steps` timed steps after
`--warmup` untimed steps, with inputs resident in HBM.  The roofline names
the kernel class with the largest device time per step in a pass of the
same length right after the timed region (HIP events on the library's
stream around every launch), with algorithmic work per launch counted on the
oracle's pair set.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s, FP64 vs mixed at stated force error; p50/p99 job latency"
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}  # MI355X vector peaks (FP64, FP32)
HBM_PEAK_GBS = 8000.0

# Algorithmic FLOP per work unit (add/sub/mul = 1, FMA = 2, div/sqrt/exp/erfc = 1),
# counted on the reference's pair set (unordered pairs within the cutoff):
#   fused LJ + Ewald real-space pair, forces only (the MD step computes no
#   energy or virial): geometry 22 (d 3, min-image 12, r^2 5, sqrt 1, 1/r 1),
#   LJ 10, Ewald real 16, force vector and +/- accumulation 10  -> 58
#   (with energy and virial: 91, the figure of the survey's measurement table)
#   Coulomb-only pair: 58 - 10 = 48.  k-space per (k-vector, atom): rho 16, force 21.
FLOP_PAIR_LJ_EWALD_F = 58
FLOP_PAIR_EWALD_F = 48
FLOP_KSPACE_RHO = 16
FLOP_KSPACE_FORCE = 21

WORKLOADS = {
    "nacl": "NaCl Ewald NVT, 4096 ions, 8x8x8 rocksalt a=5.64 A, LJ(LB)+Ewald rc 12 A kmax 9, "
            "Berendsen 1200 K, VV dt 1 fs",
    "spce": "flexible SPC/E water NVE, 3000 molecules (9000 atoms), box 44.8 A, LJ O-O + Ewald rc 10 A "
            "kmax 10 (inter-molecular), harmonic O-H/H-O-H, VV dt 0.5 fs, 2000 steps",
    "mc": "batched MC: LJ argon NVT chains, sizes U{108..4000} (seed 4), rho*=0.8, T*=1.2, rc=2.5 sigma, "
          "translation 0.3 sigma, Philox4x32-10 per chain",
    "spce-batched": "flexible SPC/E water NVE, 3000 molecules (9000 atoms) per system, several systems at once, "
                    "Ewald rc 10 A kmax 10, VV dt 0.5 fs",
    "ka": "Kob-Andersen 80:20 LJ binary mixture, 1,000,000 atoms, rho*=1.2, rc=2.5 sigma_AA, NVE VV dt 1 fs",
    "serve": "job server: Poisson stream of MD jobs (LJ argon 64..8192 atoms, SPC/E Ewald 216..12000 molecules, "
             "mixed form) and LJ MC chains (108..4000 atoms)",
}

# LJ-only pair, forces only: geometry 22, LJ 10, force vector and +/- accumulation 10
FLOP_PAIR_LJ_F = 42
# Velocity-Verlet, algorithmic HBM bytes per atom-step (FP64): read x, v, F
# (72) and 1/m (8), write x and v (48)
BYTES_INTEGRATE = 128

# MC move: LJ energy at the old and the new position for every neighbour
# within rc: 2 x 29 FLOP per neighbour (geometry 22 + LJ energy 7)
FLOP_MC_PAIR = 29


def host_cores():
    """The host cores this job owns: the cores it may run on, limited to the
    host's CPU share per GPU when the environment states one
    (OMP_NUM_THREADS; 16 of a GPU box's cores belong to its one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    if share > 0:
        n = min(n, share)
    return max(1, n)


def nproc():
    return os.cpu_count() or 1


def build_system(workload):
    from lumol_amd.workloads import nacl, spce
    return nacl() if workload == "nacl" else spce()


def make_sim(workload):
    from lumol_amd import BerendsenThermostat, MolecularDynamics, Simulation, VelocityVerlet
    if workload == "nacl":
        md = MolecularDynamics(VelocityVerlet(1.0))
        md.set_thermostat(BerendsenThermostat(1200.0, 100.0))
    else:
        md = MolecularDynamics(VelocityVerlet(0.5))
    return Simulation(md)


def oracle_md(o, workload, nsteps):
    if workload == "nacl":
        o.md_run(0, 1.0, nsteps, 1200.0, 100.0)
    else:
        o.md_run(0, 0.5, nsteps)


def cpu_baseline(workload, target_seconds=12.0):
    """The oracle's threaded FP64 MD step (the reference's algorithm: O(N^2)
    pair loops with per-thread force buffers, Ewald recomputing e^{ik.r} per
    call) on a bounded sample of the same workload: 1 warm-up step, then 5
    repetitions of R steps (R sized to ~target/5 s each); median."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lumol_oracle as lo

    s = build_system(workload)
    cores = host_cores()
    o = lo.OracleSystem(s, nthreads=cores)
    o.md_reset()
    t0 = time.perf_counter()
    oracle_md(o, workload, 1)
    t1 = time.perf_counter() - t0
    reps = 5
    r = max(1, int(round(target_seconds / reps / max(t1, 1e-3))))
    rates = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle_md(o, workload, r)
        dt = time.perf_counter() - t0
        rates.append(s.size() * r / dt)
    return {"value": statistics.median(rates), "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "nproc": nproc(),
            "sample": f"oracle FP64 MD step on the same {s.size()}-atom system: 1 warm-up step, "
                      f"median of {reps} x {r} steps, {cores} threads (the host's CPU share for this GPU; "
                      f"nproc {nproc()})"}


def force_error(s, oracle_forces, oracle_energy, oracle_virial=None):
    """Force RMS / max relative error against the FP64 oracle (max over atoms
    with |F| > 1e-3 F_rms), energy relative error and, given the oracle's
    virial, the virial tensor's relative error (Frobenius)."""
    import numpy as np
    fg = s.forces()
    fo = oracle_forces
    d = fg - fo
    rms = math.sqrt((d * d).sum() / (fo * fo).sum())
    norms = np.linalg.norm(fo, axis=1)
    frms = math.sqrt((norms ** 2).mean())
    mask = norms > 1e-3 * frms
    mx = float((np.linalg.norm(d, axis=1)[mask] / norms[mask]).max())
    eg = s.potential_energy()
    out = {"force_rms_rel": rms, "force_max_rel": mx, "energy_rel": abs(eg - oracle_energy) / abs(oracle_energy)}
    if oracle_virial is not None:
        wg = s.virial()
        out["virial_rel"] = float(np.linalg.norm(wg - oracle_virial) / np.linalg.norm(oracle_virial))
    return out


def nve_drift(make, precision, nsteps=2000, every=100, dt=0.5):
    """NVE drift of a fresh system from make(): least-squares slope of the
    total energy per atom over `nsteps` Velocity-Verlet steps, in
    kJ/mol/atom/ns, with its standard error."""
    import numpy as np
    from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet, units
    s = make()
    s.set_precision(precision)
    sim = Simulation(MolecularDynamics(VelocityVerlet(dt)))
    t, e = [], []
    for k in range(nsteps // every + 1):
        t.append(k * every * dt * 1e-6)
        e.append(units.to(s.total_energy(), "kJ/mol") / s.size())
        if k < nsteps // every:
            sim.run(s, every)
    s.close()
    t, e = np.array(t), np.array(e)
    A = np.vstack([t, np.ones_like(t)]).T
    coef, *_ = np.linalg.lstsq(A, e, rcond=None)
    resid = e - A @ coef
    se = math.sqrt((resid ** 2).sum() / (len(t) - 2) / ((t - t.mean()) ** 2).sum())
    return {"kj_per_mol_atom_ns": float(coef[0]), "stderr": se, "steps": nsteps, "dt_fs": dt}


def calibrate_copy(mib=256):
    """One device-to-device hipMemcpy of `mib` MiB (the plain copy kernel the
    FETCH_SIZE / WRITE_SIZE counters are scaled by)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def load_traffic(workload, form, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/traffic_<workload>_<form>.json), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{form}.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


TUNING_CACHE = None  # --tuning-cache


ROOFLINE_CLASSES = ("pairs", "kspace_rho", "kspace_force", "integrate")


def class_work(cls, workload, form, natoms, n_pairs, nk, batch=1):
    """(bound, algorithmic work per launch, peak, unit, description) of a
    kernel class: VALU FLOP for the pair and k-space kernels (FP32 peak in
    the mixed form, FP64 otherwise), HBM bytes for the integrator."""
    peak = PEAK_TFLOPS[form]
    if cls == "pairs":
        # water: every O-O pair within rc also carries Coulomb; the pair set
        # counted is all inter-molecular pairs within rc (oracle's set)
        flop = FLOP_PAIR_LJ_EWALD_F if workload == "nacl" else FLOP_PAIR_EWALD_F
        return ("valu", batch * n_pairs * flop / 1e12, peak, "TFLOP/s",
                f"{flop} FLOP x {n_pairs} pairs within rc (forces only) x {batch} system(s) per launch")
    if cls == "kspace_force":
        return ("valu", batch * nk * natoms * FLOP_KSPACE_FORCE / 1e12, peak, "TFLOP/s",
                f"{FLOP_KSPACE_FORCE} FLOP x {nk} k-vectors x {natoms} atoms x {batch} system(s) per launch")
    if cls == "kspace_rho":
        return ("valu", batch * nk * natoms * FLOP_KSPACE_RHO / 1e12, peak, "TFLOP/s",
                f"{FLOP_KSPACE_RHO} FLOP x {nk} k-vectors x {natoms} atoms x {batch} system(s) per launch "
                f"(the launch also rebuilds the Verlet list when due)")
    return ("hbm", batch * natoms * BYTES_INTEGRATE / 1e9, HBM_PEAK_GBS, "GB/s",
            f"{BYTES_INTEGRATE} B x {natoms} atoms x {batch} system(s) per launch")


def roofline_of(prof, steps, workload, form, natoms, n_pairs, nk, batch=1, timing=""):
    """Dominant kernel class of a profiled pass (largest device time per
    step) and its roofline: algorithmic work per launch / average launch
    duration."""
    per_step = {c: prof[c][0] * 1e3 / steps for c in prof if prof[c][1]}
    dom = max((c for c in ROOFLINE_CLASSES if c in per_step), key=lambda c: per_step[c])
    ms, launches = prof[dom]
    avg_s = ms / 1e3 / launches
    bound, work, peak, unit, per = class_work(dom, workload, form, natoms, n_pairs, nk, batch)
    achieved = work / avg_s
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
            "traffic": load_traffic(workload, form, dom), "kernel": dom, "avg_launch_us": avg_s * 1e6,
            "launches": launches, "work": per, "timing": timing}, {c: round(v, 2) for c, v in per_step.items()}


def run_form(workload, form, steps, warmup, oracle_forces, oracle_energy, oracle_virial=None):
    from lumol_amd._native import KERNEL_CLASSES, LG_PRECISION_FP64, LG_PRECISION_MIXED
    s = build_system(workload)
    s.set_precision(LG_PRECISION_MIXED if form == "mixed" else LG_PRECISION_FP64)
    natoms = s.size()
    rc = 12.0 if workload == "nacl" else 10.0
    n_pairs = len(s.pairs_within(rc))
    nk = s.ewald_kvector_count()
    err = force_error(s, oracle_forces, oracle_energy, oracle_virial)
    # launch knobs for this size bucket (lumol_amd.tuning), chosen before
    # timing, each candidate timed by the wall time of 100 MD steps (the
    # launch mix of the timed region, list rebuilds included), best of 3
    from lumol_amd import tuning
    sim = make_sim(workload)
    knobs = tuning.autotune(s, rounds=3, step=lambda: sim.run(s, 100))
    if TUNING_CACHE:
        tuning.save_cache(TUNING_CACHE)
    sim.run(s, warmup)
    # timed region: no instrumentation (an event pair per launch costs ~10%
    # of the step rate of these small systems)
    t0 = time.perf_counter()
    sim.run(s, steps)  # returns after the stream is synchronised
    elapsed = time.perf_counter() - t0
    # the same number of steps again with HIP events around every launch:
    # per-step device time of each kernel class, the dominant class and its
    # average launch duration
    s.profile_begin(list(KERNEL_CLASSES))
    sim.run(s, steps)
    prof = s.profile_end()
    roofline, per_step = roofline_of(
        prof, steps, workload, form, natoms, n_pairs, nk,
        timing=f"HIP events on the library stream around every launch, over a pass of the same {steps} steps "
               f"right after the timed region")
    s.close()
    return {"value": natoms * steps / elapsed, "ms_per_step": elapsed * 1e3 / steps,
            "error_vs_fp64_oracle": err, "roofline": roofline, "kernel_us_per_step": per_step,
            "atoms": natoms, "pairs_within_rc": n_pairs, "kvectors": nk, "tuning": knobs}


def mc_chain_sizes(nchains, seed=4):
    import numpy as np
    return np.random.default_rng(seed).integers(108, 4001, size=nchains)


def mc_cpu_baseline(sizes, target_seconds):
    """The oracle's Monte Carlo chain (the reference's EnergyCache with its
    N x N pair matrix, one move at a time) on a stratified sample: one chain
    per size decile, each on its own host thread, for as many attempts as fit
    in ~target_seconds; attempts per second summed over threads."""
    import threading

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lumol_oracle as lo
    from lumol_amd.workloads import lj_mc_chain

    cores = host_cores()
    picks = [int(np.quantile(sizes, q)) for q in np.linspace(0.05, 0.95, min(cores, 10))]
    chains = []
    for k, n in enumerate(picks):
        sysm, beta, delta = lj_mc_chain(n, seed=1000 + k)
        o = lo.OracleSystem(sysm, nthreads=1)
        o.cache_init()
        chains.append((o, beta, delta, n))
    # calibrate on a short run of the first chain; a move costs O(N)
    o0, beta0, delta0, n0 = chains[0]
    t0 = time.perf_counter()
    o0.mc_run(1, 0, 400, beta0, delta0)
    rate0 = 400 / (time.perf_counter() - t0)
    done = [0.0] * len(chains)

    def work(k):
        o, beta, delta, n = chains[k]
        attempts = max(200, int(rate0 * target_seconds * n0 / n))
        t0 = time.perf_counter()
        o.mc_run(7 + k, 1000, attempts, beta, delta)
        done[k] = (attempts, time.perf_counter() - t0)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(chains))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    total = sum(a / dt for a, dt in done)
    return {"value": total, "unit": "atom-steps/s", "cores": len(chains), "kind": "port",
            "sample": f"oracle MC chains of {picks} atoms (one per thread, one chain per size decile), "
                      f"~{target_seconds:.0f} s each; attempts/s summed"}


def bench_mc(args):
    """Batched MC chains: every chain runs `sweeps` sweeps (N attempts each);
    launches give each active chain `slice` attempts, chains leave the batch
    when done."""
    from lumol_amd import mc_run_batch
    from lumol_amd.workloads import lj_mc_chain

    sizes = mc_chain_sizes(args.chains)
    chains = [lj_mc_chain(int(n), seed=10000 + k) for k, n in enumerate(sizes)]
    for s, _, _ in chains:
        s.cache_init()
    # warm-up: a short slice on every chain
    mc_run_batch([c[0] for c in chains], list(range(len(chains))), 0, 16, [c[1] for c in chains],
                 [c[2] for c in chains])
    todo = [int(n) * args.sweeps for n in sizes]
    done = [16] * len(chains)
    target = [16 + t for t in todo]
    accepted = [0] * len(chains)
    launches = 0
    t0 = time.perf_counter()
    while True:
        active = [k for k in range(len(chains)) if done[k] < target[k]]
        if not active:
            break
        n_att = [min(args.slice, target[k] - done[k]) for k in active]
        acc = mc_run_batch([chains[k][0] for k in active], list(active), [done[k] for k in active], n_att,
                           [chains[k][1] for k in active], [chains[k][2] for k in active])
        for k, a, na in zip(active, acc, n_att):
            accepted[k] += int(a)
            done[k] += na
        launches += 1
    elapsed = time.perf_counter() - t0
    attempts = sum(todo)
    # drift check of the running energies against a fresh evaluation
    worst = 0.0
    for s, _, _ in chains[:: max(1, len(chains) // 16)]:
        e_cache, e_new = s.cache_energy(), s.potential_energy()
        worst = max(worst, abs(e_cache - e_new) / abs(e_new))
    acc_ratio = sum(accepted) / attempts
    line = {
        "metric": METRIC, "value": attempts / elapsed, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.sweeps, "warmup": 16, "ms_per_step": elapsed * 1e3 / args.sweeps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded lattice + jitter)",
        "config": {"workload": WORKLOADS["mc"], "chains": len(chains), "sweeps_per_chain": args.sweeps,
                   "atoms_total": int(sizes.sum()), "attempts_per_launch_per_chain": args.slice,
                   "launches": launches},
        "acceptance_ratio": acc_ratio,
        "max_rel_energy_cache_drift": worst,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = mc_cpu_baseline(sizes, args.cpu_seconds)
    for s, _, _ in chains:
        s.close()
    print(json.dumps(line), flush=True)


def oracle_reference(workload):
    """The FP64 oracle's forces, energy and virial of the workload's initial
    state (the error reference of every form)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lumol_oracle as lo
    o = lo.OracleSystem(build_system(workload), nthreads=host_cores())
    return o.forces(), o.potential_energy(), o.virial()


def run_batched(batch, steps, warmup, ref, md_graph=0):
    """Aggregate mixed-form throughput of `batch` SPC/E systems (seeds 3,
    4, ...) advancing together, each on its own library context (HIP
    stream), with one set of launch knobs tuned on the batch itself."""
    import threading

    from lumol_amd import tuning
    from lumol_amd._native import LG_PRECISION_MIXED
    from lumol_amd.workloads import spce

    def make(seed, slot):
        s = spce(seed=seed)
        s.stream_slot = slot
        s.set_precision(LG_PRECISION_MIXED)
        s.set_tuning(md_graph=md_graph)
        return s

    systems = [make(3 + k, 1 + k) for k in range(batch)]
    sims = [make_sim("spce") for _ in systems]
    for s, sm in zip(systems, sims):
        sm.run(s, warmup)

    def concurrent(nsteps, warm=0):
        # every system advances `nsteps` MD steps on its own host thread and
        # stream; the wall time from the common start to the last finish
        barrier = threading.Barrier(len(systems) + 1)

        def work(s, sm):
            if warm:
                sm.run(s, warm)
            barrier.wait()
            sm.run(s, nsteps)

        threads = [threading.Thread(target=work, args=(s, sm)) for s, sm in zip(systems, sims)]
        for t in threads:
            t.start()
        barrier.wait()
        t0 = time.perf_counter()
        for t in threads:
            t.join()
        return time.perf_counter() - t0

    # 500 steps a candidate, best of two timings, two passes of coordinate
    # descent: with 200 steps and one pass the choice moved between runs by
    # the noise of the batch, and one pass stopped at pair lanes 16 / split 8
    # / rho width 128 (28.9 us per system-step) against 32 / 4 / 256 (28.0)
    knobs = tuning.autotune_concurrent(systems, lambda: concurrent(500, warm=50), rounds=2, passes=2)
    # the stated error: the first system's initial state with these knobs
    chk = make(3, 1)
    chk.set_tuning(knobs["pair_lanes"], knobs["kspace_split"], knobs["rho_width"], knobs["nlist_build"], md_graph,
                   rho_kind=knobs["rho_kind"])
    err = force_error(chk, *ref)
    chk.close()
    concurrent(warmup)
    elapsed = concurrent(steps)
    natoms = sum(s.size() for s in systems)
    for s in systems:
        s.close()
    return {"value": natoms * steps / elapsed, "ms_per_step": elapsed * 1e3 / steps, "batch": batch,
            "atoms": natoms, "error_vs_fp64_oracle": err, "tuning": knobs,
            "how": f"{batch} systems (seeds 3..{2 + batch}), one host thread and HIP stream each, "
                   f"wall time from a common start to the last finish"}


def bench_spce_batched(args):
    """Aggregate mixed-form throughput of `batch` SPC/E systems against one
    FP64 system alone."""
    ref = oracle_reference("spce")
    single = run_form("spce", "fp64", args.steps, args.warmup, *ref)
    bat = run_batched(args.batch, args.steps, args.warmup, ref, args.md_graph)
    line = {
        "metric": METRIC, "value": bat["value"], "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": bat["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded grid + jitter)",
        "config": {"workload": WORKLOADS["spce-batched"], "batch": args.batch, "form": "mixed",
                   "md_graph": args.md_graph},
        "error_vs_fp64_oracle": bat["error_vs_fp64_oracle"],
        "fp64_single_system": single["value"],
        "tuning": {"fp64_single_system": single["tuning"], "mixed_batched": bat["tuning"]},
        "mixed_batched_over_fp64": bat["value"] / single["value"],
    }
    print(json.dumps(line), flush=True)


def bench_c3(args):
    """The headline measurement: flexible SPC/E water, 3000 molecules (9000
    atoms), Ewald.  `value` is the mixed form's aggregate rate over
    `--batch` systems advancing together; FP64 and mixed single-system rates,
    each form's error against the FP64 oracle and NVE drift, the dominant
    kernel's roofline, the serving stream's p50/p99 and the CPU baseline sit
    beside it."""
    from lumol_amd._native import LG_PRECISION_FP64, LG_PRECISION_MIXED
    from lumol_amd.workloads import spce
    forms = set(args.forms.split(","))
    if "mixed" not in forms or not forms <= {"fp64", "mixed", "batched"}:
        raise SystemExit("--forms: a comma list of fp64, mixed, batched that includes mixed")
    ref = oracle_reference("spce")
    fp64 = run_form("spce", "fp64", args.steps, args.warmup, *ref) if "fp64" in forms else None
    mixed = run_form("spce", "mixed", args.steps, args.warmup, *ref)
    bat = run_batched(args.batch, args.steps, args.warmup, ref, args.md_graph) if "batched" in forms else None
    drift = {} if args.no_drift else {
        "fp64": nve_drift(spce, LG_PRECISION_FP64), "mixed": nve_drift(spce, LG_PRECISION_MIXED)}
    keep = ("value", "ms_per_step", "error_vs_fp64_oracle", "roofline", "kernel_us_per_step", "tuning")
    head = bat or mixed
    line = {
        "metric": METRIC, "value": head["value"], "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded jittered grid, random orientations, Maxwell-Boltzmann velocities)",
        "config": {"workload": WORKLOADS["spce"], "form": "mixed, batched" if bat else "mixed",
                   "batch": args.batch if bat else 1,
                   "atoms_per_system": mixed["atoms"], "pairs_within_rc": mixed["pairs_within_rc"],
                   "kvectors": mixed["kvectors"]},
        "error_vs_fp64_oracle": head["error_vs_fp64_oracle"],
        "roofline": mixed["roofline"],
        "forms": {"mixed_single": dict({k: mixed[k] for k in keep}, nve_drift=drift.get("mixed"))},
        "roofline_note": "dominant kernel of the mixed single-system run (the kernel measured alone on the device)",
    }
    if fp64:
        line["forms"]["fp64_single"] = dict({k: fp64[k] for k in keep}, nve_drift=drift.get("fp64"))
        line["forms"]["mixed_single_over_fp64"] = mixed["value"] / fp64["value"]
    if bat:
        line["forms"]["mixed_batched"] = bat
        if fp64:
            line["forms"]["mixed_batched_over_fp64"] = bat["value"] / fp64["value"]
    if not args.no_serve:
        line["serve"] = run_serve(args.serve_jobs, args.serve_scale, args.serve_seconds, 0, 0, args.md_lanes,
                                  args.slice, "mixed")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline("spce", args.cpu_seconds)
    print(json.dumps(line), flush=True)


def bench_ka(args):
    """1M-atom Kob-Andersen, mixed and FP64 forms."""
    import numpy as np

    from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
    from lumol_amd._native import LG_PRECISION_FP64, LG_PRECISION_MIXED
    from lumol_amd.workloads import kob_andersen

    forms = {}
    f64 = None
    for form in ("fp64", "mixed"):
        s = kob_andersen(args.atoms)
        s.set_precision(LG_PRECISION_MIXED if form == "mixed" else LG_PRECISION_FP64)
        f = s.forces()
        if form == "fp64":
            f64 = f
            err = None
        else:
            d = f - f64
            err = {"force_rms_rel_vs_fp64_gpu": float(np.sqrt((d * d).sum() / (f64 * f64).sum()))}
        sim = Simulation(MolecularDynamics(VelocityVerlet(1.0)))
        # lanes per atom (and the list build) for this size bucket, each
        # candidate timed by the wall time of 50 MD steps, best of 2: the
        # built-in size default (4 lanes at 1M atoms) measured 660 us per
        # mixed step against 512 us with 16 lanes
        from lumol_amd import tuning
        knobs = tuning.autotune(s, rounds=2, step=lambda: sim.run(s, 50))
        if TUNING_CACHE:
            tuning.save_cache(TUNING_CACHE)
        s.profile_begin(["nlist", "pairs", "integrate"])
        sim.run(s, args.warmup)
        warm = s.profile_end()
        if form == "fp64":
            # pairs of the configuration being simulated (after the lattice start)
            n_pairs = s.pair_count_within(2.5 * 3.405)
        t0 = time.perf_counter()
        sim.run(s, args.steps)
        elapsed = time.perf_counter() - t0
        # pair kernel launch duration over a second pass of the same steps
        s.profile_begin(["pairs"])
        sim.run(s, args.steps)
        ms, launches = s.profile_end()["pairs"]
        avg = ms / 1e3 / max(launches, 1)
        achieved = n_pairs * FLOP_PAIR_LJ_F / avg / 1e12
        peak = PEAK_TFLOPS[form]
        forms[form] = {"value": s.size() * args.steps / elapsed, "ms_per_step": elapsed * 1e3 / args.steps,
                       "error": err, "tuning": knobs,
                       "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                                    "frac": achieved / peak, "traffic": load_traffic("ka", form, "pairs"),
                                    "kernel": "pairs", "avg_launch_us": avg * 1e6,
                                    "work": f"{FLOP_PAIR_LJ_F} FLOP x {n_pairs} pairs (r < 2.5 sigma) per launch"},
                       "kernel_us_per_step": {c: round(warm[c][0] * 1e3 / max(args.warmup, 1), 2)
                                              for c in warm if warm[c][1]}}
        s.close()
    m = forms["mixed"]
    line = {
        "metric": METRIC, "value": m["value"], "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded lattice + jitter)",
        "config": {"workload": WORKLOADS["ka"], "atoms": args.atoms, "form": "mixed", "pairs_within_rc": n_pairs},
        "error_vs_fp64": m["error"], "roofline": m["roofline"],
        "forms": {"fp64": forms["fp64"], "mixed": {k: m[k] for k in ("value", "ms_per_step", "kernel_us_per_step", "tuning")},
                  "mixed_over_fp64": m["value"] / forms["fp64"]["value"]},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = ka_cpu_baseline(args.atoms)
    print(json.dumps(line), flush=True)


def ka_cpu_baseline(natoms, sample=16384):
    """The oracle's O(N^2) force evaluation of a `sample`-atom KA system
    (median of 3), scaled by (N / sample)^2 to the full size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lumol_oracle as lo
    from lumol_amd.workloads import kob_andersen
    s = kob_andersen(sample)
    cores = host_cores()
    o = lo.OracleSystem(s, nthreads=cores)
    o.forces()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.forces()
        times.append(time.perf_counter() - t0)
    t = statistics.median(times) * (natoms / sample) ** 2
    return {"value": natoms / t, "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"oracle force evaluation of a {sample}-atom KA system (median of 3, {cores} threads), "
                      f"scaled by (N/{sample})^2 to {natoms} atoms (extrapolated)"}


def run_serve(md_jobs, steps_scale, seconds, chains, sweeps, md_lanes, mc_slice, form):
    """Poisson stream of MD and MC jobs through lumol_amd.serving.Server:
    p50/p99 job latency (submit -> result) per class and aggregate
    atom-steps/s.  Host-side construction of each job's system happens
    before its submission and is not timed."""
    import threading

    import numpy as np

    from lumol_amd._native import LG_PRECISION_FP64, LG_PRECISION_MIXED
    from lumol_amd.serving import Job, Server, latency_stats
    from lumol_amd.workloads import lj_mc_chain, serving_stream

    rng = np.random.default_rng(5)
    precision = LG_PRECISION_MIXED if form == "mixed" else LG_PRECISION_FP64
    jobs = [Job("md", (lambda s=system: s), steps, md=opts, name=name)
            for name, system, steps, opts in serving_stream(md_jobs, steps_scale, precision=precision, rng=rng)]
    for k, n in enumerate(mc_chain_sizes(chains)):
        system, beta, delta = lj_mc_chain(int(n), seed=20000 + k)
        jobs.append(Job("mc", (lambda s=system: s), int(n) * sweeps,
                        mc={"beta": beta, "delta": delta, "seed": 7000 + k}, name=f"mc-{n}"))
    order = rng.permutation(len(jobs))
    jobs = [jobs[i] for i in order]
    gaps = rng.exponential(seconds / len(jobs), size=len(jobs))
    with Server(md_lanes=md_lanes, mc_slice=mc_slice) as server:
        t0 = time.perf_counter()

        def feed():
            for j, g in zip(jobs, gaps):
                time.sleep(g)
                server.submit(j)

        feeder = threading.Thread(target=feed)
        feeder.start()
        feeder.join()
        server.drain()
        elapsed = time.perf_counter() - t0
        launches = server.mc_launches
    errors = [(j.name, repr(j.error)) for j in jobs if j.error is not None]
    return {
        "md_jobs": md_jobs, "mc_chains": chains, "md_steps": "U{1000..20000} x %g" % steps_scale,
        "mc_sweeps": sweeps, "md_lanes": md_lanes, "mc_slice": mc_slice, "form": form,
        "mean_interarrival_s": seconds / len(jobs), "elapsed_s": elapsed,
        "latency": {"all": latency_stats(jobs), "md": latency_stats([j for j in jobs if j.kind == "md"]),
                    "mc": latency_stats([j for j in jobs if j.kind == "mc"])},
        "mc_launches": launches, "errors": errors[:5],
    }


def bench_serve(args):
    r = run_serve(args.md_jobs, args.steps_scale, args.serve_seconds, args.chains, args.sweeps, args.md_lanes,
                  args.slice, args.form)
    total = r["latency"]["all"]
    njobs = args.md_jobs + args.chains
    line = {
        "metric": METRIC, "value": total["atom_steps_per_s"], "unit": "atom-steps/s", "n_gpus": 1,
        "steps": njobs, "warmup": 0, "ms_per_step": r["elapsed_s"] * 1e3 / njobs,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.form == "mixed" else "f64", "data": "synthetic (seeded lattices + jitter)",
        "config": {"workload": WORKLOADS["serve"], **{k: r[k] for k in ("md_jobs", "mc_chains", "md_steps",
                                                                        "mc_sweeps", "md_lanes", "mc_slice",
                                                                        "mean_interarrival_s", "form")}},
        "latency": r["latency"], "mc_launches": r["mc_launches"], "errors": r["errors"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed MD steps (the SPC/E configuration's run)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="spce")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--chains", type=int, default=1024, help="mc: number of chains")
    ap.add_argument("--sweeps", type=int, default=20, help="mc: sweeps per chain")
    ap.add_argument("--slice", type=int, default=4096, help="mc: attempts per chain per launch")
    ap.add_argument("--atoms", type=int, default=1_000_000, help="ka: atoms")
    ap.add_argument("--batch", type=int, default=4, help="spce-batched: systems at once")
    ap.add_argument("--md-graph", type=int, default=0, choices=[0, 1, 2],
                    help="spce-batched: 2 replays MD chunks from HIP graphs, 1 launches every kernel")
    ap.add_argument("--md-jobs", type=int, default=256, help="serve: MD jobs")
    ap.add_argument("--md-lanes", type=int, default=4, help="serve: concurrent MD jobs (HIP streams)")
    ap.add_argument("--steps-scale", type=float, default=0.02, help="serve: factor on the 1k..20k MD steps")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serve: span of the arrivals")
    ap.add_argument("--form", choices=["fp64", "mixed"], default="mixed", help="serve: MD arithmetic form")
    ap.add_argument("--no-serve", action="store_true", help="spce: leave out the serving stream")
    ap.add_argument("--forms", default="fp64,mixed,batched",
                    help="spce: the forms measured (a comma list; mixed alone for a profiled run)")
    ap.add_argument("--no-drift", action="store_true", help="spce: leave out the NVE drift runs")
    ap.add_argument("--serve-jobs", type=int, default=256, help="spce: MD jobs of the serving stream")
    ap.add_argument("--serve-scale", type=float, default=0.05,
                    help="spce: factor on the serving stream's 1k..20k MD steps (0.05: 50..1000 steps, about "
                         "a third of one GPU's serving capacity over the 20 s of arrivals)")
    ap.add_argument("--tuning-cache", default=None,
                    help="JSON file of autotuner choices: read if it exists, written after tuning (so that a "
                         "profiled rerun launches the same kernel variants as the timed run)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 256 MiB device copy (scale factor of the HBM byte counters)")
    args = ap.parse_args()
    if args.tuning_cache:
        global TUNING_CACHE
        TUNING_CACHE = args.tuning_cache
        if os.path.exists(args.tuning_cache):
            from lumol_amd import tuning
            tuning.load_cache(args.tuning_cache)
    if args.calibrate_copy:
        calibrate_copy()
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the GPU boxes of this build have one MI355X, and independent "
              "systems/jobs shard across GPUs without communication. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "mc":
        bench_mc(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "ka":
        bench_ka(args)
        return
    if args.workload == "spce-batched":
        bench_spce_batched(args)
        return
    if args.workload == "spce":
        bench_c3(args)
        return

    ref = oracle_reference(args.workload)
    forms = {f: run_form(args.workload, f, args.steps, args.warmup, *ref) for f in ("fp64", "mixed")}
    m = forms["mixed"]
    line = {
        "metric": METRIC,
        "value": m["value"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded lattice/grid + jitter, Maxwell-Boltzmann velocities)",
        "config": {"workload": WORKLOADS[args.workload], "form": "mixed", "atoms": m["atoms"],
                   "pairs_within_rc": m["pairs_within_rc"], "kvectors": m["kvectors"]},
        "error_vs_fp64_oracle": m["error_vs_fp64_oracle"],
        "roofline": m["roofline"],
        "forms": {"fp64": {k: forms["fp64"][k] for k in ("value", "ms_per_step", "error_vs_fp64_oracle",
                                                         "roofline", "kernel_us_per_step", "tuning")},
                  "mixed": {k: m[k] for k in ("value", "ms_per_step", "kernel_us_per_step", "tuning")},
                  "mixed_over_fp64": m["value"] / forms["fp64"]["value"]},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
