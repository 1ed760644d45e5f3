"""md_pimd_run_steps_x keeps the executable step graph of its last call on an
object and updates it in place from the next call's fresh capture.  A run of
several graph-replayed calls with changing step parameters (time step, Ewald
split, pair interaction) must follow the eagerly launched trajectory bit for
bit, and a second object must not pick up the first one's graph."""
import ctypes as C

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu


def ueg_box(N, rs=2.0):
    return (4 * np.pi * N / 3.0) ** (1.0 / 3.0) * rs


def state(s):
    b = s.host_beads()
    ch = pyxmd.nhcs(s.sim)
    return b["x"].copy(), b["v"].copy(), b["f"].copy(), ch["theta"].copy(), ch["vtheta"].copy()


def make(lib, kind, N, P, seed):
    L = ueg_box(N) if kind == "ueg" else 10.0
    return pyxmd.System(lib, kind, N, P, 0.4 if kind == "ueg" else 1.0, 1.0, seed=seed, L=L, nsum=8)


# (time step, Ewald split alpha L, steps) of consecutive calls; calls of 32
# or more graph steps replay four steps per graph and run the rest eagerly
CALLS = [(0.01, -1.0, 4), (0.02, -1.0, 3), (0.02, 6.0, 5), (0.005, -1.0, 39), (0.005, -1.0, 2), (0.01, 0.0, 3),
         (0.01, 0.0, 35)]


@pytest.mark.parametrize("kind,N,P", [("ueg", 70, 8), ("trap", 40, 16)])
def test_graph_reuse_matches_eager(kind, N, P):
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    lib.md_set_coulomb_truc_x(5.0 / ueg_box(70))
    lib.md_set_coulomb_n_sum_x(8)
    graph, eager, other = make(lib, kind, N, P, 3), make(lib, kind, N, P, 3), make(lib, kind, 2 * N, P, 4)
    try:
        for dt, aL, n in CALLS:
            assert lib.md_set_ewald_split_x(aL) == 0
            cg = graph.step_config(dt, use_graph=1)
            ce = eager.step_config(dt, use_graph=0)
            co = other.step_config(dt, use_graph=1)
            assert lib.md_pimd_run_steps_x(graph.pimd, C.byref(cg), n) == 0
            assert lib.md_pimd_run_steps_x(other.pimd, C.byref(co), n) == 0
            assert lib.md_pimd_run_steps_x(eager.pimd, C.byref(ce), n) == 0
            assert lib.md_hip_synchronize_x() == 0
            for a, b in zip(state(graph), state(eager)):
                np.testing.assert_array_equal(a, b)
        assert np.all(np.isfinite(state(other)[0]))
    finally:
        lib.md_set_ewald_split_x(-1.0)


@pytest.mark.parametrize("kind,N,P", [("ueg", 70, 8)])
def test_concurrent_step_matches_serial_order(kind, N, P):
    """The three-stream step (real-space partial sums, reciprocal and exchange
    forces in side buffers, added inside NHC half 2) follows, bit for bit, the
    serial single-stream order that adds each part into F as it is made."""
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    lib.md_set_coulomb_truc_x(5.0 / ueg_box(70))
    lib.md_set_coulomb_n_sum_x(8)
    conc, ser = make(lib, kind, N, P, 5), make(lib, kind, N, P, 5)
    cc = conc.step_config(0.01, use_graph=1)
    cs = ser.step_config(0.01, use_graph=0)
    stages = np.zeros(11)
    assert lib.md_pimd_run_steps_x(conc.pimd, C.byref(cc), 37) == 0
    assert lib.md_pimd_profile_steps_x(ser.pimd, C.byref(cs), 37, 1, pyxmd.dptr(stages)) == 0
    assert lib.md_hip_synchronize_x() == 0
    for a, b in zip(state(conc), state(ser)):
        np.testing.assert_array_equal(a, b)
