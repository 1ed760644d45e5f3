"""Single-species periodic Gaussian pair interaction (md_periodic_gaussian_force_x,
md_simulation_x.c:174-180; energy md_statistics_x.c) on the GPU against the
compiled reference: 12 particles, P=8, xi=0.5, in a periodic box of edge 3
(minimum-image pair distances and springs, no trap), golden case pgauss12.
Pair force (tile kernel, periodic-Gaussian mode) and energy, the total force,
and 5-step trajectories through the per-stage calls and the fused,
graph-replayed step.  Bars as in the other golden cases: forces 1e-11 of the
RMS, energies 1e-12, trajectories 1e-10 (x) / 1e-9 (v, f, chains)."""
import ctypes as C

import numpy as np
import pytest

import pyxmd
from orc import nhc_from, read_golden, rel_err

pytestmark = pytest.mark.gpu
NAME = "pgauss12"


@pytest.fixture(scope="module")
def lib():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(8)
    return lib


def system(lib):
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith(NAME + "/")}
    N, P, T, vi, kind, image, L, nsum, fgroup, dt = g["meta"]
    assert int(kind) == 2 and int(image) == 1
    s = pyxmd.System(lib, "pgauss", int(N), int(P), T, vi, L=L, nsum=int(nsum), fgroup=int(fgroup),
                     positions=g["x0"])
    pyxmd.load_state(s, g["x0"], g["v0"], nhc_from(g["nhc0"]))
    return s, g, dt


def stat(lib, fn, *args, pos=1):
    st = lib.md_alloc_stats_x(1)
    a = list(args)
    a.insert(pos, st)
    assert fn(*a) == 0
    return st.contents.es[0]


def test_pair_force_energy_and_total(lib):
    s, g, _ = system(lib)
    assert lib.md_reset_force_x(s.sim) == 0
    assert lib.md_pimd_calc_pair_force_x(s.pimd, s.pair_cb()) == 0
    err = rel_err(s.host_beads()["f"].reshape(-1), g["f_pair"])
    print(f"periodic Gaussian pair force rel err {err:.2e}")
    assert err < 1e-11
    e = stat(lib, lib.md_pimd_calc_pair_energy_x, s.pimd, s.pair_energy_cb(), pos=2)
    assert e == pytest.approx(g["e_pair"][0], rel=1e-12, abs=1e-12)
    s.full_force()
    assert rel_err(s.host_beads()["f"].reshape(-1), g["f_total"]) < 1e-11


@pytest.mark.parametrize("mode", ["api", "graph"])
def test_trajectory(lib, mode):
    s, g, dt = system(lib)
    steps = len(g["e_vir_series"])
    if mode == "graph":
        cfg = s.step_config(dt, use_graph=1)
        assert lib.md_pimd_compute_forces_x(s.pimd, C.byref(cfg)) == 0
        assert lib.md_pimd_run_steps_x(s.pimd, C.byref(cfg), steps) == 0
        assert lib.md_hip_synchronize_x() == 0
    else:
        s.full_force()
        for _ in range(steps):
            s.step(dt)
    b = s.host_beads()
    assert rel_err(b["x"].reshape(-1), g["x_end"]) < 1e-10
    assert rel_err(b["v"].reshape(-1), g["v_end"]) < 1e-9
    assert rel_err(b["f"].reshape(-1), g["f_end"]) < 1e-9
    ch = pyxmd.nhcs(s.sim)
    ref = nhc_from(g["nhc_end"])
    assert rel_err(ch["vtheta"], ref["vtheta"]) < 1e-9
    assert rel_err(ch["theta"], ref["theta"]) < 1e-9
