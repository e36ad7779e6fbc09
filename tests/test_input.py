"""The input layer (lumol_amd/input.py, lumol_amd/unitconv.py) against the
reference's own input data (tests/golden/scenarios, copied from the
reference's tests/data) and its known answers: the unit parser's values
(lumol-core/src/units.rs:414-500), the molecule structure its integration
tests assert after reading (tests/md-butane.rs:21-26, tests/md-methane.rs:21-26),
Ewald::with_accuracy (ewald.rs:996-1001, 1446-1462) and the velocity
controls (lumol-sim/src/md/controls.rs:96-125).

CPU only: systems are built and compared with the oracle, nothing runs on
a GPU.
"""
import math
import os

import numpy as np
import pytest

import lumol_oracle as lo
from helpers import GOLDEN, nist_spce_system, water_molecule
from lumol_amd import unitconv
from lumol_amd.consts import K_BOLTZMANN
from lumol_amd.input import (Input, InputError, InteractionsInput, boltzmann_velocities, ewald_with_accuracy,
                             read_pdb, read_structure, remove_rotation, remove_translation)

SCEN = os.path.join(GOLDEN, "scenarios")


def scen(*p):
    return os.path.join(SCEN, *p)


# --- units (units.rs tests) -----------------------------------------------------------------
def test_unit_values():
    assert unitconv.parse("A") == 1.0
    assert unitconv.parse("nm") == 10.0
    assert unitconv.parse("bohr/fs") == 0.52917720859
    assert unitconv.parse("(Ry / rad^-3   )") == 0.1312749878912494
    assert unitconv.parse("bar/(m * fs^2)") == 6.022141794216763e-19
    assert unitconv.parse("kJ/mol/deg^2") == 0.3282806352310398
    assert unitconv.parse("(kcal/mol/A)^2") == 1.7505856024515547e-7
    assert abs(unitconv.parse("kcal/mol/A^2") - 4.184e-4) <= 1e-9
    assert unitconv.from_str("10.0 A") == 10.0
    assert unitconv.from_str("10 A") == 10.0
    assert unitconv.from_str("1e1 A") == 10.0
    assert unitconv.from_str("10") == 10.0
    assert unitconv.to(25.0, "m") == 2.5e-9
    assert unitconv.to(25.0, "bar") == 4.1513469550000005e9
    assert unitconv.to(25.0, "kJ/mol") == 249999.99982494753


@pytest.mark.parametrize("bad", ["(", ")", "(bar/m", "m/K)", "m^4-8", "foo ^ bar", "m^z4", "HJK"])
def test_unit_errors(bad):
    with pytest.raises(unitconv.UnitError):
        unitconv.parse(bad)


@pytest.mark.parametrize("bad", ["10a.0 bar", "h10"])
def test_unit_from_str_errors(bad):
    with pytest.raises(unitconv.UnitError):
        unitconv.from_str(bad)


# --- structure files --------------------------------------------------------------------------
def test_pdb_cell_atoms_and_bonds():
    names, pos, cell, bonds = read_pdb(scen("md-water", "small.pdb"))
    assert cell == (28.0, 28.0, 28.0)
    assert len(names) == 81 and names[:3] == ["O", "H", "H"]
    assert tuple(pos[1]) == (0.635, 0.898, 0.0)
    assert len(bonds) == 54 and (0, 1) in bonds and (0, 2) in bonds


@pytest.mark.parametrize("name,nmol,bonds,angles,dihedrals", [
    ("md-butane", 50, 3, 2, 1),    # tests/md-butane.rs:21-26
    ("md-methane", 150, 4, 6, 0),  # tests/md-methane.rs:21-26
])
def test_guessed_bonds_give_the_reference_molecules(name, nmol, bonds, angles, dihedrals):
    config = Input(scen(name, "nve.toml"))
    s = config.read_system()
    mols = s.molecules()
    assert len(mols) == nmol
    o = lo.OracleSystem(s)
    assert o.n == s.size()
    per = s.size() // nmol
    assert all(b - a == per for a, b in mols)
    assert len(s.bonds) == nmol * bonds
    # the oracle derives angles and dihedrals from the bonds as the
    # reference's Bonding does; their energies are non-zero
    t = o.energy_terms()
    assert t["angles"] != 0.0
    assert (t["dihedrals"] != 0.0) == (dihedrals > 0)


def test_read_md_nacl_interactions():
    config = Input(scen("md-nacl", "nve-ewald-small.toml")).read()
    s = config.system
    assert s.size() == 64 and s.cell.lengths == (11.2804,) * 3
    assert sorted(set(s.charges)) == [-1.0, 1.0]
    assert s.coulomb.cutoff == 5.5 and s.coulomb.kmax == 10
    assert config.nsteps == 100
    p = s.pairs[("Na", "Cl")]
    assert p.cutoff() == 5.5 and p.potential.sigma == 3.5545
    assert p.potential.epsilon == unitconv.from_str("0.04425 kcal/mol")
    # Maxwell-Boltzmann velocities at exactly 300 K, without net momentum
    m = s.masses
    ke = 0.5 * (m[:, None] * s.velocities ** 2).sum()
    assert abs(2 * ke / (3 * s.size() * K_BOLTZMANN) - 300.0) < 1e-9
    assert np.abs((m[:, None] * s.velocities).sum(0)).max() < 1e-12


def test_read_md_water_restrictions():
    config = Input(scen("md-water", "nve-ewald.toml")).read()
    s = config.system
    assert len(s.molecules()) == 27
    from lumol_amd import PairRestriction
    assert s.pairs[("H", "H")].restriction() == PairRestriction.INTRA_MOLECULAR
    assert s.coulomb.restriction == PairRestriction.INTER_MOLECULAR
    assert s.angle_potentials[("H", "O", "H")].x0 == unitconv.from_str("109.5 deg")


def test_table_computation_input():
    config = Input(scen("md-helium", "nve-table.toml")).read()
    pot = config.system.pairs[("He", "He")].potential
    assert pot.table == (1000, 5.0)


@pytest.mark.parametrize("name", ["md-nacl/nve-ewald-small.toml", "md-nacl/nve-wolf-small.toml",
                                  "md-water/nve-wolf.toml", "md-helium/nve-shifted.toml",
                                  "md-helium/nve-table.toml", "md-butane/nve.toml"])
def test_scenarios_build_oracle_systems(name):
    # every scenario that reads also evaluates in the oracle, with F = sum 0
    s = Input(scen(*name.split("/"))).read().system
    o = lo.OracleSystem(s, nthreads=4)
    f = o.forces()
    assert np.isfinite(f).all()
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).sum(0).max() + 1e-15


# --- Ewald::with_accuracy ----------------------------------------------------------------------
def test_ewald_with_accuracy_water():
    # ewald.rs:996-1001
    s = water_molecule()
    ew = ewald_with_accuracy(8.5, 1e-6, s)
    assert abs(ew.alpha - 0.2998) < 1e-4 and ew.kmax == 5


@pytest.mark.parametrize("name,alpha,kmax", [("spce-1.xyz", 0.2826, 5), ("spce-2.xyz", 0.2900, 5),
                                             ("spce-3.xyz", 0.2943, 5), ("spce-4.xyz", 0.2912, 8)])
def test_ewald_with_accuracy_nist(name, alpha, kmax):
    # ewald.rs:1446-1462
    s = nist_spce_system(name, 9.0)
    ew = ewald_with_accuracy(9.0, 1e-5, s)
    assert abs(ew.alpha - alpha) < 1e-4 and ew.kmax == kmax


def test_ewald_accuracy_key_in_input():
    s = nist_spce_system("spce-1.xyz", 9.0)
    s.coulomb = None
    InteractionsInput({"coulomb": {"ewald": {"cutoff": "9 A", "accuracy": 1e-5}}}).read(s)
    assert abs(s.coulomb.alpha - 0.2826) < 1e-4 and s.coulomb.kmax == 5
    with pytest.raises(InputError):
        InteractionsInput({"coulomb": {"ewald": {"cutoff": "9 A", "accuracy": 1e-5, "kmax": 4}}}).read(s)


# --- velocities (controls.rs tests) ---------------------------------------------------------------
def test_remove_translation_known_answer():
    # controls.rs:96-108: two Ag atoms
    v = remove_translation([107.8682, 107.8682], np.array([[1.0, 2.0, 0.0], [1.0, 0.0, 0.0]]))
    assert v.tolist() == [[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]]


def test_remove_rotation_known_answer():
    # controls.rs:110-122
    v = remove_rotation([107.8682, 107.8682], np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]),
                        np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 2.0]]))
    assert v.tolist() == [[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]


def test_boltzmann_velocities_temperature_and_momentum():
    s = read_structure(scen("md-helium", "helium.xyz"), cell=(10.0, 10.0, 10.0))
    boltzmann_velocities(s, 273.0)
    m = s.masses
    v = s.velocities
    t = (m[:, None] * v * v).sum() / (3 * s.size() * K_BOLTZMANN)
    assert abs(t - 273.0) < 1e-9
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12 * np.abs(m[:, None] * v).max() * s.size()
    # the second removal leaves no translation either (sum of m (d x w) = 0)
    assert math.isfinite(t)


# --- errors the reference reports -----------------------------------------------------------------
def _write(tmp_path, text, name="in.toml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_input_errors(tmp_path):
    with pytest.raises(InputError, match="version"):
        Input(_write(tmp_path, "[input]\nversion = 2\n"))
    with pytest.raises(InputError, match="missing 'input'"):
        Input(_write(tmp_path, "[foo]\nbar = 1\n"))
    base = ('[input]\nversion = 1\n[[systems]]\nfile = "%s"\ncell = 10\n' %
            scen("md-helium", "helium.xyz").replace("\\", "/"))
    bad_pot = base + '[systems.potentials.pairs]\nHe-He = {type = "foo", cutoff = "4 A"}\n'
    with pytest.raises(InputError, match="unknown potential type"):
        Input(_write(tmp_path, bad_pot)).read_system()
    bad_restr = base + '[systems.potentials.pairs]\nHe-He = {type = "null", cutoff = "4 A", restriction = "scale14"}\n'
    with pytest.raises(InputError, match="scale14"):
        Input(_write(tmp_path, bad_restr)).read_system()
    bad_cut = base + '[systems.potentials.pairs]\nHe-He = {type = "null"}\n'
    with pytest.raises(InputError, match="cutoff"):
        Input(_write(tmp_path, bad_cut)).read_system()
    bad_unit = base + '[systems.potentials.pairs]\nHe-He = {type = "lj", sigma = "2 parsec", epsilon = "1 K", cutoff = "4 A"}\n'
    with pytest.raises(InputError, match="parsec"):
        Input(_write(tmp_path, bad_unit)).read_system()
