"""Ewald forces in a rectangular, non-cubic box against the oracle's
restatement of the reference's direct sum (md_simulation_x.c:223-346).

The k kernels take the phase tables per dimension (x and z for m >= 0, y for
m >= 0 with negative m_y by conjugation), size the table from the longest
side, and the real-space image enumeration from the shortest; a cubic box
cannot tell a mixed-up dimension from a right one.  Two parameter sets at
alpha = 5 / L_min:
  * n = 8: the first omitted k shell is not below e^-30 along the longest
    side, so the engine evaluates the reference's own cube and image sum and
    the forces agree to round-off (1e-12 of the RMS force);
  * n = 12: converged, so the engine evaluates the same sum with its own
    split (pinned at alpha' L_min = 7.5: the ball of k, 17 along the longest
    side, so three 16-slot tiles per column; real space within 5.25 / alpha');
    the omitted terms are below exp(-5.25^2) = 1.1e-12 of the leading ones
    (bar 1e-11).
"""
import ctypes as C

import numpy as np
import pytest

import pyxmd
from orc import Oracle, beads_from, dptr, make_params, ptr, rel_err

pytestmark = pytest.mark.gpu

BOX = np.array([9.0, 10.8, 7.65])


@pytest.fixture(scope="module")
def lib():
    lib = pyxmd.Lib()
    yield lib
    lib.md_set_ewald_split_x(-1.0)


@pytest.mark.parametrize("nsum,tol", [(8, 1e-12), (12, 1e-11)])
def test_noncubic_box_forces_match_oracle(lib, nsum, tol):
    N, P = 24, 2
    lmin = float(BOX.min())
    alpha = 5.0 / lmin
    x = np.random.default_rng(31).uniform(0.0, 1.0, (N * P, 3)) * BOX
    s = pyxmd.System(lib, "ueg", N, P, 0.5, 1.0, seed=3, L=lmin, nsum=nsum, positions=x, alpha_L=5.0)
    lib.md_init_nd_rect_box_x(s.sim.contents.box, dptr(BOX.copy()))
    assert lib.md_set_ewald_split_x(7.5 if nsum == 12 else -1.0) == 0
    a, rc, nk = C.c_double(), C.c_double(), C.c_int()
    assert lib.md_pimd_get_ewald_plan_x(s.pimd, C.byref(a), C.byref(rc), C.byref(nk)) == 0
    assert lib.md_reset_force_x(s.sim) == 0
    assert lib.md_pimd_calc_pair_force_x(s.pimd, lib.cb["md_coulomb_3d_ewald_force_x"]) == 0
    f = s.host_beads()["f"].reshape(-1)
    o = Oracle()
    p = make_params(alpha=alpha, n=nsum)
    b = beads_from(x.reshape(-1), m=np.ones(N * P))
    o.pair_force(2, C.byref(p), N, P, dptr(BOX.copy()), ptr(b))
    err = rel_err(f, b["f"].reshape(-1))
    print(f"box {BOX.tolist()} n={nsum}: plan alpha L_min={a.value * lmin:.3f} r_c={rc.value:.3f} n_k={nk.value}, "
          f"force rel err vs oracle {err:.2e}", flush=True)
    if nsum == 8:
        assert a.value == pytest.approx(alpha, rel=1e-15)
    else:
        assert a.value * lmin == pytest.approx(7.5, rel=1e-12)
    assert err < tol
