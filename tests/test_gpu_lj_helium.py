"""Lennard-Jones and helium pair forces and energies on the GPU (one species
through md_pimd_calc_pair_force_x / _energy_x, two species through the _2
entry points) against the compiled reference's golden vectors: 27 particles
on a jittered 3 x 3 x 3 lattice, P = 4, periodic cubic box; two species
A = 15 particles (mass 1), B = 12 (mass 2)."""
import numpy as np
import pytest

import pyxmd
from orc import HE_AZIZ, dptr, read_golden, rel_err

pytestmark = pytest.mark.gpu

CB = {"lj": "md_LJ_periodic", "he": "md_periodic_helium"}


def system(lib, n, P, x, m, L):
    sim = lib.md_alloc_simulation_x(n * P, 1, 0, 0)
    lib.md_init_nd_rect_box_x(sim.contents.box, dptr(L))
    pimd = lib.md_alloc_pimd_x(sim, n, P, 1.0, 1.0)
    lib.md_init_particle_mass_x(sim, float(m[0]), 0, n * P)
    b = pyxmd.beads(sim)
    b["x"] = np.asarray(x).reshape(n * P, 3)
    b["f"] = 0
    pyxmd.set_beads(sim, b)
    return sim, pimd


@pytest.mark.parametrize("pot", list(CB))
def test_lj_helium(pot):
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith(pot + "/")}
    lib = pyxmd.Lib()
    lib.md_set_helium_parameters_x(*HE_AZIZ)
    fcb, ecb = lib.cb[CB[pot] + "_force_x"], lib.cb[CB[pot] + "_energy_x"]
    L = g["L"]
    sim, pimd = system(lib, 27, 4, g["x0"], g["m0"], L)
    assert lib.md_reset_force_x(sim) == 0
    assert lib.md_pimd_calc_pair_force_x(pimd, fcb) == 0
    assert lib.md_simulation_sync_host_x(sim, 1)
    err = rel_err(pyxmd.beads(sim)["f"].reshape(-1), g["f_pair"])
    print(f"{pot}: pair force rel err {err:.2e}")
    assert err < 1e-12
    st = lib.md_alloc_stats_x(1)
    st.contents.es[0] = 0.0
    assert lib.md_pimd_calc_pair_energy_x(pimd, ecb, st) == 0
    assert st.contents.es[0] == pytest.approx(g["e_pair"][0], rel=1e-12)
    sa, pa = system(lib, 15, 4, g["xA0"], g["mA0"], L)
    sb, pb = system(lib, 12, 4, g["xB0"], g["mB0"], L)
    assert lib.md_pimd_calc_pair_force_2_x(pa, pb, fcb) == 0
    assert lib.md_simulation_sync_host_x(sa, 1) and lib.md_simulation_sync_host_x(sb, 1)
    ea = rel_err(pyxmd.beads(sa)["f"].reshape(-1), g["fA"])
    eb = rel_err(pyxmd.beads(sb)["f"].reshape(-1), g["fB"])
    print(f"{pot}: two-species force rel err A {ea:.2e}, B {eb:.2e}")
    assert ea < 1e-12 and eb < 1e-12
    st.contents.es[0] = 0.0
    assert lib.md_pimd_calc_pair_energy_2_x(pa, pb, ecb, st) == 0
    assert st.contents.es[0] == pytest.approx(g["e2"][0], rel=1e-12)
