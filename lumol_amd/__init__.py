"""lumol_amd -- MI355X (gfx950) backend of Lumol's force-and-energy evaluation.

The compute path is ``liblumol_gpu.so`` (hand-written HIP kernels behind the
C ABI of ``include/lumol_gpu.h``); this package is the host-side mirror of
lumol-core / lumol-sim names on top of it.
"""
from .system import (  # noqa: F401
    ATOMIC_MASSES, BerendsenThermostat, BornMayerHuggins, Buckingham, CosineHarmonic, Ewald, FOUR_PI_EPSILON_0,
    Gaussian, Harmonic, K_BOLTZMANN, LeapFrog, LennardJones, Mie, MolecularDynamics, MonteCarlo, Morse,
    NullPotential, PairInteraction, PairRestriction, SharedEwald, Simulation, System, TableComputation, Torsion,
    Translate, UnitCell, VelocityVerlet, Wolf, mc_run_batch, units,
)
from ._native import LumolGpuError  # noqa: F401

__version__ = "0.1.0"
