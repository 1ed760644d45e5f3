"""The exchange recursion's two block solves against the CPU oracle.

Each 64-position block is normally solved in the frame of its
single-particle chain, with coefficients the set-up already exponentiated;
a block whose scaled values leave the frame's range, or a next-block row
whose look-ahead sum does, is solved or folded again exactly (per-row
shifts by a (max,+) scan, coefficients from the tables).  The range limit
is lowered here through md_set_recursion_frame_limit_x: below 1 every
block takes the exact path (every scaled value is >= 1), moderate limits
mix the two paths block by block and row by row.  V_B, the backward
potentials and the exchange forces must match the oracle on every mix,
to the bars of the multi-block parity test.
"""
import numpy as np
import pytest

import pyxmd
from orc import Oracle, beads_from, dptr, exchange_truth, ptr, rel_err, tri_size

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = pyxmd.Lib()
    yield lib
    lib.md_set_recursion_frame_limit_x(0.0)


CASES = [("trap", 300, 4, 1.0), ("trap", 200, 3, 0.0), ("ueg", 700, 2, 0.25)]
LIMITS = [0.5, 3.0, 40.0, 1e6]


def system(lib, kind, N, P, xi):
    if kind == "trap":
        return pyxmd.System(lib, "trap", N, P, 0.5, xi, seed=11), np.array([10.0] * 3)
    Lb = pyxmd.ueg_box(N, 2.0)
    s = pyxmd.System(lib, "ueg", N, P, pyxmd.ueg_temperature(N, Lb, 0.5), xi, seed=11, L=Lb, nsum=8)
    return s, np.array([Lb] * 3)


@pytest.mark.parametrize("kind,N,P,xi", CASES)
def test_frame_and_exact_paths(lib, kind, N, P, xi):
    s, L = system(lib, kind, N, P, xi)
    img = s.image
    hb = s.host_beads()
    o = Oracle()
    b = beads_from(hb["x"].reshape(-1), m=hb["m"])
    E = np.zeros(tri_size(N))
    Ei = np.zeros(N)
    o.fast_fill_ENk(N, P, s.omegaP, img, dptr(L), ptr(b), dptr(E), dptr(Ei))
    V = np.zeros(N + 1)
    o.fill_VB(N, s.beta, xi, dptr(E), dptr(V))
    Vt = np.zeros(N + 1)
    b["f"] = 0
    o.fast_force_VB(N, P, s.beta, xi, s.omegaP, img, dptr(L), dptr(E), dptr(V), ptr(b), dptr(Vt), None)
    fo = b["f"].reshape(-1).copy()
    ft, _ = exchange_truth(o, hb["x"].reshape(-1), hb["m"], N, P, s.beta, xi, s.omegaP, img, L[0])
    ref = rel_err(fo, ft)
    tol = max(1e-11, 4 * ref)
    try:
        for limit in LIMITS:
            print(f"{kind} N={N} P={P} xi={xi} frame limit {limit:g}", flush=True)
            assert lib.md_set_recursion_frame_limit_x(limit) == 0
            assert lib.md_pimd_fast_fill_ENk_x(s.pimd, img) == 0
            assert lib.md_pimd_fill_VB_x(s.pimd) == 0
            assert lib.md_reset_force_x(s.sim) == 0
            assert lib.md_pimd_fast_fill_force_VB_x(s.pimd, img) == 0
            assert lib.md_hip_synchronize_x() == 0
            f = s.host_beads()["f"].reshape(-1).copy()
            s.host_pimd()
            vbn = pyxmd.vbn(s.pimd).copy()
            Vt_g = np.zeros(N + 1)
            assert lib.md_pimd_get_backward_potentials_x(s.pimd, dptr(Vt_g)) == 0
            np.testing.assert_allclose(vbn, V, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(Vt_g, Vt, rtol=1e-12, atol=1e-12)
            err = rel_err(f, ft)
            print(f"  exchange force vs 80-bit truth: engine {err:.2e}, reference fast path {ref:.2e}", flush=True)
            assert err < tol
    finally:
        lib.md_set_recursion_frame_limit_x(0.0)
