"""Error statuses of the C entry points, called through ctypes exactly as a
foreign binding would (no Python mirror in between).  The reference panics
where these return LG_E_INVALID / LG_E_UNSUPPORTED / LG_E_STATE:
  pair cutoff above half the cell       system.rs:122-131
  coulomb cutoff above half the cell    system.rs:157-167
  Scale14 with Ewald                    ewald.rs:388
  Ewald in an infinite cell             ewald.rs:124 (UnitCell::infinite)
  virial of an infinite cell            compute.rs:198
  commit without a staged move          (cache.rs: update after cost)
"""
import ctypes

import numpy as np
import pytest

from lumol_amd import _native as nat

pytestmark = pytest.mark.gpu


def new_system(n=2, cell=(20.0, 20.0, 20.0)):
    L = nat.lib()
    h = ctypes.c_void_p()
    assert L.lg_system_create(nat.context(0, 0), n, ctypes.byref(h)) == nat.LG_OK
    assert L.lg_system_set_cell(h, *cell) == nat.LG_OK
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]][:n] + [[3.0 * k, 1.0, 1.0] for k in range(n - 2)])
    assert L.lg_system_set_positions(h, nat.dptr(np.ascontiguousarray(pos))) == nat.LG_OK
    q = np.array([1.0, -1.0] + [0.0] * (n - 2))
    assert L.lg_system_set_charges(h, nat.dptr(q)) == nat.LG_OK
    return h


def status_and_message(rc):
    return rc, nat.lib().lg_last_error().decode()


def test_cutoff_errors():
    L = nat.lib()
    h = new_system()
    lj = np.array([3.4, 1e-4])
    rc, msg = status_and_message(L.lg_pair_set(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 10.5, 0, 0, 0, 1.0))
    assert rc == nat.LG_E_INVALID and "half the cell" in msg
    assert L.lg_pair_set(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 9.0, 0, 0, 0, 1.0) == nat.LG_OK
    rc, msg = status_and_message(L.lg_coulomb_ewald(h, 10.5, 5, float("nan"), 0))
    assert rc == nat.LG_E_INVALID and "coulomb" in msg
    rc, msg = status_and_message(L.lg_coulomb_wolf(h, 11.0, 0, 1.0))
    assert rc == nat.LG_E_INVALID
    assert L.lg_coulomb_wolf(h, 9.0, 0, 1.0) == nat.LG_OK
    # shrinking the cell afterwards: the evaluation re-checks both cutoffs
    assert L.lg_system_set_cell(h, 17.0, 17.0, 17.0) == nat.LG_OK
    f = np.zeros(6)
    rc, msg = status_and_message(L.lg_forces(h, nat.dptr(f), 0))
    assert rc == nat.LG_E_INVALID and "half the cell" in msg
    assert L.lg_pair_set(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 8.0, 0, 0, 0, 1.0) == nat.LG_OK
    rc, msg = status_and_message(L.lg_forces(h, nat.dptr(f), 0))
    assert rc == nat.LG_E_INVALID and "coulomb" in msg
    assert L.lg_coulomb_wolf(h, 8.0, 0, 1.0) == nat.LG_OK
    assert L.lg_forces(h, nat.dptr(f), 0) == nat.LG_OK
    assert np.isfinite(f).all() and abs(f[0] + f[3]) < 1e-12 * abs(f[0])
    L.lg_system_destroy(h)


def test_ewald_restrictions_and_cells():
    L = nat.lib()
    h = new_system()
    rc, msg = status_and_message(L.lg_coulomb_ewald(h, 8.0, 5, float("nan"), 6))
    assert rc == nat.LG_E_UNSUPPORTED and "Scaling" in msg
    rc, _ = status_and_message(L.lg_coulomb_ewald(h, 8.0, 0, float("nan"), 0))
    assert rc == nat.LG_E_INVALID
    rc, _ = status_and_message(L.lg_coulomb_ewald(h, -1.0, 5, float("nan"), 0))
    assert rc == nat.LG_E_INVALID
    # Ewald needs a periodic cell
    assert L.lg_system_set_cell(h, 0.0, 0.0, 0.0) == nat.LG_OK
    assert L.lg_coulomb_ewald(h, 8.0, 5, float("nan"), 0) == nat.LG_OK
    f = np.zeros(6)
    rc, msg = status_and_message(L.lg_forces(h, nat.dptr(f), 0))
    assert rc == nat.LG_E_INVALID and "infinite" in msg
    assert L.lg_coulomb_clear(h) == nat.LG_OK
    w = np.zeros(9)
    rc, msg = status_and_message(L.lg_virial(h, nat.dptr(w)))
    assert rc == nat.LG_E_INVALID and "infinite" in msg
    L.lg_system_destroy(h)


def test_argument_and_state_errors():
    L = nat.lib()
    h = new_system()
    lj = np.array([3.4, 1e-4])
    assert L.lg_pair_set(h, 0, 0, 99, nat.dptr(lj), 5.0, 0, 0, 0, 1.0) == nat.LG_E_UNSUPPORTED
    assert L.lg_pair_set(h, 3, 0, nat.LG_POT_LJ, nat.dptr(lj), 5.0, 0, 0, 0, 1.0) == nat.LG_E_INVALID
    assert L.lg_pair_set(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 5.0, 0, 0, 9, 1.0) == nat.LG_E_INVALID
    assert L.lg_system_set_cell(h, -1.0, 2.0, 2.0) == nat.LG_E_INVALID
    bad = np.array([0.0, float("nan"), 0.0, 1.0, 1.0, 1.0])
    assert L.lg_system_set_positions(h, nat.dptr(bad)) == nat.LG_E_NONFINITE
    assert L.lg_move_commit(h, 1) == nat.LG_E_STATE
    assert L.lg_md_run(h, 10) == nat.LG_E_STATE  # no lg_md_setup yet
    assert L.lg_md_setup(h, 7, 1.0, -1.0, 1.0) == nat.LG_E_UNSUPPORTED
    assert L.lg_md_setup(h, 0, 1.0, 300.0, 0.5) == nat.LG_E_INVALID  # tau < 1
    assert L.lg_md_setup(h, 0, -1.0, -1.0, 1.0) == nat.LG_E_INVALID
    L.lg_system_destroy(h)
