"""GPU parity of the molecular virial (compute.rs:278-364, ewald.rs:533-570,
736-753, wolf.rs:285-324) and of what is built on the virials: pressure and
stress (compute.rs:383-478), with the degrees of freedom the propagator sets
(system.rs:249-255, monte_carlo.rs:247-280).

Bar: 1e-10 relative (largest tensor entry) against the oracle, which the CPU
tests pin by the trace identity and the rigid-molecule strain derivative.
"""
import numpy as np
import pytest

import lumol_oracle as lo
from helpers import other_potentials_system, random_molecular_system
from lumol_amd import K_BOLTZMANN, MonteCarlo, Simulation, Translate
from lumol_amd.workloads import nacl, spce
from test_gpu_parity import wrel

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("coulomb", ["ewald", "wolf", None])
def test_molecular_virial_water_like(coulomb):
    s = random_molecular_system(n_mol=40, L=16.0, coulomb=coulomb)
    o = lo.OracleSystem(s, nthreads=8)
    wg, wo = s.molecular_virial(), o.molecular_virial()
    assert wrel(wg, wo) < TOL, wrel(wg, wo)
    # and differs from the atomic definition (the molecules are not points)
    assert wrel(s.atomic_virial(), wo) > 1e-3


def test_molecular_virial_spce():
    s = spce(nmol=216, L=18.64, cutoff=8.0, kmax=6)
    o = lo.OracleSystem(s, nthreads=16)
    assert wrel(s.molecular_virial(), o.molecular_virial()) < TOL


def test_molecular_virial_other_potentials():
    s = other_potentials_system(variant=1, coulomb="ewald")
    o = lo.OracleSystem(s, nthreads=8)
    assert wrel(s.molecular_virial(), o.molecular_virial()) < TOL


def test_molecular_virial_of_atoms_is_atomic():
    s = nacl(ncell=4, cutoff=10.0, kmax=7)
    wa, wm = s.atomic_virial(), s.molecular_virial()
    assert wrel(wm, wa) < 1e-12


def test_pressure_and_stress_follow_the_degrees_of_freedom():
    s = random_molecular_system(n_mol=40, L=16.0, coulomb="ewald")
    rng = np.random.default_rng(5)
    s.velocities = rng.normal(scale=1e-3, size=(s.size(), 3))
    o = lo.OracleSystem(s, nthreads=8)
    V = s.volume()
    T = 300.0
    # particles (MD): atomic virial, 3N degrees of freedom
    p_ref = (3 * s.size() * K_BOLTZMANN * T + np.trace(o.virial())) / (3 * V)
    assert abs(s.pressure_at_temperature(T) - p_ref) <= TOL * abs(p_ref)
    st = s.stress_at_temperature(T)
    st_ref = (s.size() * K_BOLTZMANN * T * np.eye(3) + o.virial()) / V
    assert wrel(st, st_ref) < TOL
    kin = sum(m * np.outer(v, v) for m, v in zip(s.masses, s.velocities))
    assert wrel(s.stress(), (kin + o.virial()) / V) < TOL
    # molecules (Monte Carlo of whole molecules): molecular virial, 3 per molecule
    mc = MonteCarlo(T, seed=1)
    mc.add(Translate(0.1))
    Simulation(mc).setup(s)
    nmol = len(s.molecules())
    p_ref = (3 * nmol * K_BOLTZMANN * T + np.trace(o.molecular_virial())) / (3 * V)
    assert s.degrees_of_freedom() == 3 * nmol
    assert abs(s.pressure_at_temperature(T) - p_ref) <= TOL * abs(p_ref)
