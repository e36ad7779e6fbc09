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
