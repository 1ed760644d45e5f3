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


def test_observables_two_systems():
    """ITCF, two-system and 2-D pair correlations and the S(q) variants
    (axes and spherical shells) on three systems in one box (A: 7 particles,
    B: 5, C: 4; P = 6) against the compiled reference's golden vectors."""
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith("obs2/")}
    lib = pyxmd.Lib()
    P = 6
    L = np.array([5.0, 5.5, 6.0])
    sys = []
    for tag, n in (("A0", 7), ("B0", 5), ("C0", 4)):
        sim = lib.md_alloc_simulation_x(n * P, 1, 0, 0)
        lib.md_init_nd_rect_box_x(sim.contents.box, dptr(L))
        pimd = lib.md_alloc_pimd_x(sim, n, P, 1.0, 1.0)
        lib.md_init_particle_mass_x(sim, 1.0, 0, n * P)
        b = pyxmd.beads(sim)
        b["x"] = g["x" + tag].reshape(n * P, 3)
        pyxmd.set_beads(sim, b)
        sys.append(pimd)
    a, b, c = sys

    def run(n, fn, *args):
        st = lib.md_alloc_stats_x(n)
        assert fn(*args[:args.index("ST")], st, *args[args.index("ST") + 1:]) == 0
        return np.ctypeslib.as_array(st.contents.fxs[0], shape=(n,)).copy()

    assert np.array_equal(run(36, lib.md_pimd_calc_ITCF_x, a, "ST", 0.5, 4.5, 3), g["itcf"])
    assert np.array_equal(run(36, lib.md_pimd_calc_ITCF_2_x, a, b, "ST", 0.5, 4.5, 3), g["itcf2"])
    assert np.array_equal(run(20, lib.md_pimd_calc_pair_correlation_2_x, a, b, "ST", 3.5, 7.0, 2, 1), g["paircorr2"])
    assert np.array_equal(run(20, lib.md_pimd_calc_pair_correlation_2_x, a, b, "ST", 6.0, 3.0, 5, 0),
                          g["paircorr2_open"])
    box = np.array([0.5, 4.5, 0.5, 5.0])
    assert np.array_equal(run(20, lib.md_pimd_calc_pair_correlation_2d_x, a, "ST", 3.0, 5.0, dptr(box)),
                          g["paircorr_2d"])
    assert np.array_equal(run(25, lib.md_pimd_calc_2d_pair_correlation_x, a, "ST", 3.0, 3.0, 4, 1), g["pc2d"])
    assert np.array_equal(run(25, lib.md_pimd_calc_2d_pair_correlation_2_x, a, b, "ST", 3.0, 3.0, 4, 0), g["pc2d_2"])
    np.testing.assert_allclose(run(16, lib.md_pimd_calc_Sk_structure_2_x, a, b, "ST", 0.2, 0.35, 3), g["sk2"],
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(run(16, lib.md_pimd_calc_Sk_structure_full_2_x, a, b, "ST", 0.2, 0.35, 3),
                               g["sk_full2"], rtol=1e-13, atol=1e-13)
    sk = lib.md_alloc_sk_sphere_3d_x(0.0, 0.6, 2.0)
    nq = sk.contents.N
    assert np.array_equal(np.ctypeslib.as_array(sk.contents.qs, shape=(nq,)), g["sphere_qs"])
    assert np.array_equal(np.ctypeslib.as_array(sk.contents.qis, shape=(3 * nq,)), g["sphere_qis"])
    np.testing.assert_allclose(run(nq + 3, lib.md_pimd_calc_sphere_Sk_structure_x, a, "ST", 3, sk), g["sphere_sk"],
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(run(nq, lib.md_pimd_calc_sphere_Sk_structure_full_2_x, a, b, "ST", 8, sk),
                               g["sphere_sk_full2"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(run(nq - 2, lib.md_pimd_calc_sphere_Sk_structure_3_x, a, b, c, "ST", 2, sk),
                               g["sphere_sk3"], rtol=1e-13, atol=1e-13)
