"""CPU checks of the drop-in boundary: the library loads, exports exactly
the entry points include/lumol_gpu.h declares, and the host mirror rejects
what the reference rejects.  No compute call is made (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lumol_gpu.h")
LIB = os.path.join(ROOT, "lumol_amd", "liblumol_gpu.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "lg_forces" in syms and "lg_md_run" in syms and len(syms) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(LIB)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_python_binding_covers_header():
    from lumol_amd import _native
    assert sorted(_native.EXPORTS) == declared_symbols()


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    assert " lo_" not in out
    deps = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in deps


def test_version_and_error_without_device():
    lib = ctypes.CDLL(LIB)
    lib.lg_version.restype = ctypes.c_char_p
    lib.lg_last_error.restype = ctypes.c_char_p
    assert b"gfx950" in lib.lg_version()
    assert lib.lg_last_error() is not None


def test_host_mirror_validation():
    from lumol_amd import Ewald, LennardJones, PairInteraction, System, UnitCell, Wolf, BerendsenThermostat
    s = System(UnitCell.cubic(10.0))
    s.add_particle("Ar", (0, 0, 0))
    with pytest.raises(ValueError):
        s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(3.4, 1.0), 6.0))
    with pytest.raises(ValueError):
        Ewald(-1.0, 5)
    with pytest.raises(ValueError):
        Ewald(8.0, 0)
    with pytest.raises(ValueError):
        Wolf(-2.0)
    with pytest.raises(ValueError):
        BerendsenThermostat(300.0, 0.5)
    with pytest.raises(ValueError):
        UnitCell.cubic(-1.0)
    # system.rs:157-167: the coulomb cutoff must fit half the cell too
    with pytest.raises(ValueError):
        s.set_coulomb_potential(Ewald(6.0, 5))
    with pytest.raises(ValueError):
        s.set_coulomb_potential(Wolf(5.5))
    s.set_coulomb_potential(Wolf(5.0))
    BerendsenThermostat(0.0, 10.0)  # thermostats.rs:103 allows T = 0
