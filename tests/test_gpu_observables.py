"""Per-sample observables on the GPU (md_pimd_calc_density_distribution_x,
_pair_correlation_x, _Sk_structure_x) against the compiled reference's
golden vectors: histograms bit for bit, S(q) to 1e-13."""
import numpy as np
import pytest

import pyxmd
from orc import dptr, read_golden

pytestmark = pytest.mark.gpu

SETUP = {"obs_pbc": dict(image=1, center=[2.5, 2.75, 3.0]), "obs_open": dict(image=0, center=[0.1, -0.2, 0.3])}


@pytest.mark.parametrize("name", list(SETUP))
def test_observables(name):
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith(name + "/")}
    lib = pyxmd.Lib()
    N, P = 20, 6
    L = np.array([5.0, 5.5, 6.0])
    sim = lib.md_alloc_simulation_x(N * P, 1, 0, 0)
    lib.md_init_nd_rect_box_x(sim.contents.box, dptr(L))
    pimd = lib.md_alloc_pimd_x(sim, N, P, 1.0, 1.0)
    lib.md_init_particle_mass_x(sim, 1.0, 0, N * P)
    b = pyxmd.beads(sim)
    b["x"] = g["x0"].reshape(N * P, 3)
    pyxmd.set_beads(sim, b)
    su = SETUP[name]

    def fxs(st, n):
        return np.ctypeslib.as_array(st.contents.fxs[0], shape=(n,)).copy()

    a = lib.md_alloc_stats_x(24)
    assert lib.md_pimd_calc_density_distribution_x(pimd, a, 3.0, su["image"], dptr(np.array(su["center"]))) == 0
    assert np.array_equal(fxs(a, 24), g["density"])
    c = lib.md_alloc_stats_x(30)
    assert lib.md_pimd_calc_pair_correlation_x(pimd, c, 3.5, 7.0, 4, su["image"]) == 0
    assert np.array_equal(fxs(c, 30), g["paircorr"])
    s = lib.md_alloc_stats_x(16)
    assert lib.md_pimd_calc_Sk_structure_x(pimd, s, 0.2, 0.35, 3) == 0
    np.testing.assert_allclose(fxs(s, 16), g["sk"], rtol=1e-13, atol=1e-13)
