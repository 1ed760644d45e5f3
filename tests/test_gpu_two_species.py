"""Two-species pair forces and energy (md_pimd_calc_pair_force_2_x /
_energy_2_x) on the GPU against the compiled reference's golden vectors
(species A: 5 particles of mass 1, B: 3 of mass 2, P = 4; Gaussian and
periodic Gaussian, box 3 x 3.5 x 4)."""
import numpy as np
import pytest

import pyxmd
from orc import dptr, read_golden, rel_err

pytestmark = pytest.mark.gpu


def species(lib, n, P, x, m, L):
    sim = lib.md_alloc_simulation_x(n * P, 1, 0, 0)
    lib.md_init_nd_rect_box_x(sim.contents.box, dptr(L))
    pimd = lib.md_alloc_pimd_x(sim, n, P, 1.0, 1.0)
    lib.md_init_particle_mass_x(sim, float(m[0]), 0, n * P)
    b = pyxmd.beads(sim)
    b["x"] = np.asarray(x).reshape(n * P, 3)
    b["f"] = 0
    pyxmd.set_beads(sim, b)
    return sim, pimd


@pytest.mark.parametrize("tag,cb", [("g", "gaussian"), ("pg", "periodic_gaussian")])
def test_two_species(tag, cb):
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith("two/")}
    lib = pyxmd.Lib()
    lib.md_set_gaussian_force_strength_x(3.0)
    lib.md_set_guassian_force_range_x(0.5)
    L = np.array([3.0, 3.5, 4.0])
    sa, pa = species(lib, 5, 4, g["xA0"], g["mA0"], L)
    sb, pb = species(lib, 3, 4, g["xB0"], g["mB0"], L)
    assert lib.md_reset_force_x(sa) == 0 and lib.md_reset_force_x(sb) == 0
    assert lib.md_pimd_calc_pair_force_2_x(pa, pb, lib.cb[f"md_{cb}_force_x"]) == 0
    assert lib.md_simulation_sync_host_x(sa, 1) and lib.md_simulation_sync_host_x(sb, 1)
    fa = pyxmd.beads(sa)["f"].reshape(-1)
    fb = pyxmd.beads(sb)["f"].reshape(-1)
    ea, eb = rel_err(fa, g["fA_" + tag]), rel_err(fb, g["fB_" + tag])
    print(f"{tag}: force rel err A {ea:.2e}, B {eb:.2e}")
    assert ea < 1e-13 and eb < 1e-13
    st = lib.md_alloc_stats_x(1)
    st.contents.es[0] = 0.0
    assert lib.md_pimd_calc_pair_energy_2_x(pa, pb, lib.cb[f"md_{cb}_energy_x"], st) == 0
    assert st.contents.es[0] == pytest.approx(g["e_" + tag][0], rel=1e-13)
