"""Per-size-bucket autotuner of the device launch knobs.

The knobs (``lg_tuning``): lanes per atom of the pair and neighbour-list
kernels, row ranges of the k-space force kernel, the form of the rho(k)
kernel (rows, with its atoms per block, or the matrix cores).  They are
tuned one after the other: every candidate value is timed with the
library's HIP-event profiler over `reps` force evaluations of the system
itself, in `rounds` interleaved passes of which the fastest counts -- the
summed device time of the kernel classes the knob changes (the rho chunk
width, for example, changes both the rho launch and the pair launch that
reduces its partials) -- and the fastest candidate is kept.  Results are cached per bucket -- (Ewald
or not, precision, atoms rounded to a power of two, k-vectors rounded to a
power of two) -- so a server tunes once per bucket and reuses the choice for
every later job of that bucket.  Force evaluations leave the system's state
unchanged.
"""
import math
import threading
import time

from . import _native as nat

PAIR_LANES = (8, 16, 32, 64)
KSPACE_SPLITS = (2, 4, 8, 16, 32)
RHO_WIDTHS = (64, 128, 192, 256, 512)
NLIST_BUILDS = (0, 2)  # automatic, cell list (lg_tuning.nlist_build)
RHO_KINDS = (1, 2, 3)  # rows staged in LDS, matrix cores, rows from the table (lg_tuning.rho_kind)

_cache = {}
_lock = threading.Lock()


def bucket(system):
    n = system.size()
    ewald = system.coulomb is not None and hasattr(system.coulomb, "kmax")
    nk = system.ewald_kvector_count() if ewald else 0
    p2 = lambda x: 0 if x <= 0 else 1 << int(math.ceil(math.log2(x)))
    return (ewald, int(system.precision), p2(n), p2(nk))


def _time_all(system, reps, classes=None, step=None):
    """Device time (ms) of one force evaluation, summed over `classes` (the
    kernel classes a knob changes; all by default); or, given `step` (MD
    steps of the system, which return once the stream is synchronised), the
    wall time of one call of it: the MD step's own objective, list rebuilds
    and launch gaps included."""
    if step is not None:
        t0 = time.perf_counter()
        step()
        return (time.perf_counter() - t0) * 1e3
    system.profile_begin(list(classes or nat.KERNEL_CLASSES))
    for _ in range(reps):
        system.forces()
    prof = system.profile_end()
    return sum(ms for ms, _ in prof.values()) / reps


# the kernel classes each knob changes: timing only those keeps the noise of
# the other launches out of a choice between close candidates
_KNOB_CLASSES = {
    "pair_lanes": ("nlist", "pairs"),
    "kspace_split": ("kspace_force",),
    "rho_width": ("kspace_rho", "pairs"),  # the rho reduction runs in the pair launch
    "rho_kind": ("kspace_rho", "pairs"),
}



def autotune(system, reps=5, use_cache=True, rounds=3, step=None):
    """Picks and applies the knobs for `system`; returns them (dict).

    With `step` (a callable that advances the system by MD steps and
    returns once they are done, e.g. ``lambda: sim.run(system, 100)``) every
    candidate is timed by the wall time of those steps -- the launches of
    lg_md_run, with the list rebuilds, the fused first half-step and the
    force-only launch mix -- instead of by the device time of `reps`
    standalone force evaluations; the system then advances during tuning.
    Enough steps to hold several list rebuilds keep their count from
    deciding between close candidates."""
    key = bucket(system)
    keep = tuple(getattr(system, "tuning_", (0, 0, 0, 0, 0, 0))[3:5])  # nlist_build, md_graph: left as set

    def apply(b):
        system.set_tuning(b["pair_lanes"], b["kspace_split"], b["rho_width"], *keep, rho_kind=b["rho_kind"])

    with _lock:
        hit = _cache.get(key) if use_cache else None
    # after a change of knobs: the lists and tables are rebuilt before timing
    warm = system.forces if step is None else step
    classes = _KNOB_CLASSES
    if hit is None:
        warm()  # builds the lists and tables once
        best = {"pair_lanes": 0, "kspace_split": 0, "rho_width": 0, "rho_kind": 0}

        def pick(field, values, **extra):
            # each candidate timed in `rounds` interleaved passes, the
            # fastest pass kept: a single short timing flips the choice
            # between runs on small systems
            times = {v: float("inf") for v in values}
            for _ in range(rounds):
                for v in values:
                    trial = dict(best, **{field: v}, **extra)
                    apply(trial)
                    warm()
                    times[v] = min(times[v], _time_all(system, reps, classes[field], step))
            return min(times, key=times.get), times

        # pair kernel lanes
        best["pair_lanes"], _ = pick("pair_lanes", PAIR_LANES)
        if key[0]:
            best["kspace_split"], _ = pick("kspace_split", KSPACE_SPLITS)
            # rho(k): each width of the row form, then at the best width
            # the row form without its LDS staging, then the matrix-core form
            width, times = pick("rho_width", RHO_WIDTHS, rho_kind=1)
            kind, t_kind = 1, times[width]
            for k in (3, 2):
                trial = dict(best, rho_width=width if k != 2 else 0, rho_kind=k)
                apply(trial)
                warm()
                t = float("inf")
                if system.tuning()["rho_kind"] == k:
                    for _ in range(rounds):
                        t = min(t, _time_all(system, reps, classes["rho_kind"], step))
                if t < t_kind:
                    kind, t_kind = k, t
            best["rho_kind"] = kind
            best["rho_width"] = width if kind != 2 else 0
        hit = best
        with _lock:
            _cache[key] = hit
    apply(hit)
    return dict(hit)


def autotune_concurrent(systems, run, rounds=2, passes=2):
    """Picks and applies one set of knobs for systems that run at the same
    time, each on its own stream: the knobs fastest for one system alone are
    not those for several sharing the device (four SPC/E systems: wider rho
    blocks).  `run()` advances every system by the same fixed work,
    concurrently, and returns its wall time.  The knobs are chosen in turn
    (the list build, pair lanes, then the k-space split, rho width and rho
    form; `passes` rounds of this coordinate descent), every candidate timed
    in `rounds` interleaved passes and the fastest pass kept.  Returns the
    knobs (dict)."""
    graphs = [getattr(s, "tuning_", (0, 0, 0, 0, 0, 0))[4] for s in systems]
    best = {"pair_lanes": 0, "kspace_split": 0, "rho_width": 0, "nlist_build": 0, "rho_kind": 0}

    def apply(b):
        for s, g in zip(systems, graphs):
            s.set_tuning(b["pair_lanes"], b["kspace_split"], b["rho_width"], b["nlist_build"], g,
                         rho_kind=b["rho_kind"])

    # the list build first (it sets the cost of every rebuild step, which
    # the other knobs' timings include): the automatic choice against the
    # cell-list build, whose rebuilds cost the vector ALU less (the direct
    # build reads every atom's cell tag in every cell's workgroup) at one
    # more launch per step
    fields = [("nlist_build", NLIST_BUILDS), ("pair_lanes", PAIR_LANES)]
    if bucket(systems[0])[0]:
        # the matrix-core rho(k) leaves the vector ALU to the other systems'
        # kernels: it can win in a batch where it loses alone
        fields += [("kspace_split", KSPACE_SPLITS), ("rho_width", RHO_WIDTHS), ("rho_kind", RHO_KINDS)]
    # coordinate descent, `passes` times over the knobs: the first pass
    # tunes each knob with the later ones still at their built-in choice
    for _ in range(passes):
        for field, values in fields:
            times = {v: math.inf for v in values}
            for _ in range(rounds):
                for v in values:
                    apply(dict(best, **{field: v}))
                    times[v] = min(times[v], run())
            best[field] = min(times, key=times.get)
    apply(best)
    return dict(best)


def cached():
    """The choices made so far, by bucket."""
    with _lock:
        return dict(_cache)


def save_cache(path):
    """Writes the choices made so far to a JSON file (a list of
    [bucket, knobs] pairs)."""
    import json
    with _lock:
        items = [[list(k), v] for k, v in _cache.items()]
    with open(path, "w") as f:
        json.dump(items, f)


def load_cache(path):
    """Adds the choices of a file written by save_cache, so that another
    process (a profiled rerun, another server) uses the same knobs for the
    same buckets instead of timing them again."""
    import json
    with open(path) as f:
        items = json.load(f)
    with _lock:
        for k, v in items:
            _cache[tuple(k)] = dict(v)
