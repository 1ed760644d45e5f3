"""The Ewald sum evaluated with the engine's own splitting parameter equals
the reference's split (alpha = 5/L, |m| <= 8 cube) once both series are
converged: forces and energies agree to round-off for every alpha' the
engine may pick when both series run to x_cut = 6, and the unconverged
reference parameters are left alone.  At the default force cutoff
x_cut = 5.25 the omitted force terms are exp(-x_cut^2) = 1.1e-12 of the
leading ones, so the forces are held to 5 exp(-x_cut^2) of the RMS force;
energies always sum to x_cut >= 6.
"""
import math
import ctypes as C

import numpy as np
import pytest

import pyxmd
from orc import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = pyxmd.Lib()
    yield lib
    lib.md_set_ewald_split_x(-1.0)
    lib.md_set_ewald_real_cutoff_x(DEFAULT_XCUT)


DEFAULT_XCUT = 5.25


def ueg(lib, N, P, nsum=8, seed=5):
    L = pyxmd.ueg_box(N, 2.0)
    T = pyxmd.ueg_temperature(N, L, 0.5)
    return pyxmd.System(lib, "ueg", N, P, T, 1.0, seed=seed, L=L, nsum=nsum)


def force_energy(lib, s):
    assert lib.md_reset_force_x(s.sim) == 0
    assert lib.md_pimd_calc_pair_force_x(s.pimd, lib.cb["md_coulomb_3d_ewald_force_x"]) == 0
    f = s.host_beads()["f"].reshape(-1).copy()
    st = lib.md_alloc_stats_x(1)
    assert lib.md_pimd_calc_pair_energy_x(s.pimd, lib.cb["md_coulomb_3d_ewald_energy_x"], st) == 0
    return f, st.contents.es[0]


def plan(lib, s):
    a, rc, nk = C.c_double(), C.c_double(), C.c_int()
    assert lib.md_pimd_get_ewald_plan_x(s.pimd, C.byref(a), C.byref(rc), C.byref(nk)) == 0
    return a.value, rc.value, nk.value


@pytest.mark.parametrize("xcut", [6.0, DEFAULT_XCUT])
@pytest.mark.parametrize("N,P", [(64, 4), (300, 2)])
def test_split_independence(lib, N, P, xcut):
    assert lib.md_set_ewald_real_cutoff_x(xcut) == 0
    ftol = max(1e-12, 5.0 * math.exp(-xcut * xcut))
    s = ueg(lib, N, P)
    lib.md_set_ewald_split_x(0.0)
    a0, _, nk0 = plan(lib, s)
    assert a0 == pytest.approx(5.0 / s.L, rel=1e-15) and nk0 == (17 * 17 * 17 - 1) // 2
    f0, e0 = force_energy(lib, s)
    # alpha' L = 4.5 .. 16 take k tables of one to four 16-slot tiles per
    # column (half-widths MZ = 7, 15, 23, 31) through the k kernels
    for aL in (4.5, 6.0, 8.0, 12.0, 16.0, -1.0):
        lib.md_set_ewald_split_x(aL)
        a, rc, nk = plan(lib, s)
        if aL > 0:
            assert a * s.L == pytest.approx(aL, rel=1e-14)
        f, e = force_energy(lib, s)
        err = rel_err(f, f0)
        print(f"N={N} P={P} x_cut={xcut} alpha L={a * s.L:.2f} r_c={rc:.2f} n_k={nk}: "
              f"force rel err {err:.2e}, energy {e:.15f} vs {e0:.15f}")
        assert err < ftol
        assert e == pytest.approx(e0, rel=1e-12, abs=1e-12)
    lib.md_set_ewald_split_x(-1.0)
    assert lib.md_set_ewald_real_cutoff_x(DEFAULT_XCUT) == 0


def test_unconverged_reference_split_is_kept(lib):
    """n_sum = 4 at alpha = 5/L leaves k terms of size exp(-pi^2) ~ 5e-5:
    the engine must reproduce that truncated sum, so it keeps the split."""
    s = ueg(lib, 32, 2, nsum=4)
    lib.md_set_ewald_split_x(-1.0)
    a, _, nk = plan(lib, s)
    assert a == pytest.approx(5.0 / s.L, rel=1e-15) and nk == (9 * 9 * 9 - 1) // 2
