"""The CPU oracle against golden vectors written by the compiled reference.

Every per-step stage of the reference (nocl build) is reproduced by the oracle
on the same seeded inputs.  Golden vectors: tests/golden/golden.bin, written by
tests/golden/make_golden.c linked against the reference library.
"""
import ctypes as C

import numpy as np
import pytest

from orc import (PAIR_KIND, beads_from, dptr, make_params, nhc_from, ptr, rel_err, tri_size)

CASES = ["trap16", "trap16_g3", "trap8_xi0", "trap8_xi05", "trap24_xi025", "ueg8_n2",
         "ueg8_n4_xi0", "ueg16_n8", "pgauss12"]


def case(golden, name):
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(name + "/")}
    N, P, T, vi, kind, image, L, nsum, fgroup, dt = g["meta"]
    c = dict(N=int(N), P=int(P), T=T, vi=vi, kind=int(kind), image=int(image), L=L,
             nsum=int(nsum), fgroup=int(fgroup), dt=dt, beta=g["beta_omegaP"][0],
             omegaP=g["beta_omegaP"][1])
    c["Lv"] = np.array([L, L, L])
    c["params"] = make_params(alpha=5.0 / L, n=int(nsum))
    # fixture kinds: 0 trap + Gaussian, 1 UEG Ewald, 2 periodic box + periodic Gaussian
    c["pair"] = ("gaussian", "ewald", "periodic_gaussian")[c["kind"]]
    return c, g


def beads0(g):
    return beads_from(g["x0"], g["v0"], np.zeros_like(g["x0"]), g["m0"])


def forces(b):
    return b["f"].reshape(-1).copy()


@pytest.mark.parametrize("name", CASES)
def test_pair_and_trap_forces(golden, oracle, name):
    c, g = case(golden, name)
    b = beads0(g)
    oracle.pair_force(PAIR_KIND[c["pair"]], ptr_p(c), c["N"], c["P"], dptr(c["Lv"]), ptr(b))
    assert rel_err(forces(b), g["f_pair"]) < 1e-13
    if c["kind"] == 0:
        b = beads0(g)
        oracle.trap_force(ptr_p(c), c["N"], c["P"], ptr(b))
        assert rel_err(forces(b), g["f_trap"]) < 1e-15
    elif c["kind"] == 1:
        for split in ("R", "NI"):
            b = beads0(g)
            oracle.pair_force(PAIR_KIND["ewald_" + split], ptr_p(c), c["N"], c["P"], dptr(c["Lv"]), ptr(b))
            assert rel_err(forces(b), g["f_pair_" + split]) < 1e-13
        assert abs(oracle.madelung(ptr_p(c), dptr(c["Lv"])) - g["madelung"][0]) < 1e-14


def ptr_p(c):
    import ctypes
    return ctypes.byref(c["params"])


@pytest.mark.parametrize("name", CASES)
def test_exchange(golden, oracle, name):
    c, g = case(golden, name)
    N, P = c["N"], c["P"]
    b = beads0(g)
    ENk = np.zeros(tri_size(N))
    Eint = np.zeros(N)
    oracle.fast_fill_ENk(N, P, c["omegaP"], c["image"], dptr(c["Lv"]), ptr(b), dptr(ENk), dptr(Eint))
    np.testing.assert_allclose(ENk, g["ENk"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(Eint, g["Eint"], rtol=1e-14, atol=0)
    ENks = np.zeros(tri_size(N))
    oracle.fill_ENk_slow(N, P, c["omegaP"], c["image"], dptr(c["Lv"]), ptr(b), dptr(ENks))
    np.testing.assert_allclose(ENks, g["ENk_slow"], rtol=1e-14, atol=0)
    VBN = np.zeros(N + 1)
    oracle.fill_VB(N, c["beta"], c["vi"], dptr(g["ENk"]), dptr(VBN))
    np.testing.assert_allclose(VBN, g["VBN"], rtol=1e-14, atol=1e-14)
    Vt = np.zeros(N + 1)
    G = np.zeros(N * N)
    oracle.fast_force_VB(N, P, c["beta"], c["vi"], c["omegaP"], c["image"], dptr(c["Lv"]),
                         dptr(g["ENk"]), dptr(g["VBN"]), ptr(b), dptr(Vt), dptr(G))
    assert rel_err(forces(b), g["f_exch"]) < 1e-13
    if "f_exch_slow" in g:
        b = beads0(g)
        oracle.slow_force_VB(N, P, c["beta"], c["vi"], c["omegaP"], c["image"], dptr(c["Lv"]),
                             dptr(g["ENk"]), dptr(g["VBN"]), ptr(b))
        assert rel_err(forces(b), g["f_exch_slow"]) < 1e-13


@pytest.mark.parametrize("name", CASES)
def test_estimators(golden, oracle, name):
    c, g = case(golden, name)
    N, P = c["N"], c["P"]
    b = beads_from(g["x0"], g["v0"], g["f_total"], g["m0"])
    e = oracle.VBN_energy(N, P, c["beta"], c["vi"], dptr(g["ENk"]), dptr(g["VBN"]))
    assert e == pytest.approx(g["e_VBN"][0], rel=1e-13, abs=1e-13)
    E2 = np.zeros(tri_size(N))
    oracle.fast_fill_ENk2(N, P, c["omegaP"], c["image"], dptr(c["Lv"]), ptr(b), dptr(E2))
    np.testing.assert_allclose(E2, g["ENk2"], rtol=1e-13, atol=1e-12)
    e = oracle.vi_virial_energy(N, P, c["beta"], c["vi"], c["image"], dptr(c["Lv"]), dptr(g["ENk"]),
                                dptr(g["ENk2"]), dptr(g["VBN"]), ptr(b))
    assert e == pytest.approx(g["e_vir"][0], rel=1e-13, abs=1e-13)
    e = oracle.vi_virial_pressure(N, P, c["beta"], c["vi"], c["image"], dptr(c["Lv"]), dptr(g["ENk"]),
                                  dptr(g["ENk2"]), dptr(g["VBN"]), ptr(b))
    assert e == pytest.approx(g["e_press"][0], rel=1e-12, abs=1e-12)
    E2s = np.zeros(tri_size(N))
    oracle.fill_ENk2_slow(N, P, c["omegaP"], c["image"], dptr(c["Lv"]), ptr(b), dptr(E2s))
    np.testing.assert_allclose(E2s, g["ENk2_slow"], rtol=1e-13, atol=1e-12)
    e = oracle.vi_virial_energy(N, P, c["beta"], c["vi"], c["image"], dptr(c["Lv"]), dptr(g["ENk"]),
                                dptr(g["ENk2_slow"]), dptr(g["VBN"]), ptr(b))
    assert e == pytest.approx(g["e_vir_slow"][0], rel=1e-13, abs=1e-13)
    if c["kind"] == 0:
        e = oracle.trap_energy(ptr_p(c), N, P, ptr(b))
        assert e == pytest.approx(g["e_trap"][0], rel=1e-14)
    e = oracle.pair_energy(PAIR_KIND[c["pair"]], ptr_p(c), N, P, dptr(c["Lv"]), ptr(b))
    assert e == pytest.approx(g["e_pair"][0], rel=1e-13, abs=1e-13)
    assert oracle.kinetic(N * P, ptr(b)) == pytest.approx(g["e_kin"][0], rel=1e-14)
    ch = nhc_from(g["nhc0"])
    assert oracle.nhc_kinetic(len(ch), ptr(ch), c["T"], 8) == pytest.approx(g["e_nhc"][0], rel=1e-14)
    assert oracle.virial_energy(N, P, dptr(c["Lv"]), ptr(b)) == pytest.approx(g["e_virial_plain"][0], rel=1e-13)
    assert oracle.centroid_virial_energy(N, P, c["beta"], 3, ptr(b)) == pytest.approx(
        g["e_virial_centroid"][0], rel=1e-13)
    for a, s in zip(g["vi_sign_args"], g["vi_sign"]):
        got = oracle.vi_sign(N, c["beta"], a, dptr(g["ENk"]), dptr(g["VBN"]))
        assert got == pytest.approx(s, rel=1e-11, abs=1e-300)


@pytest.mark.parametrize("name", [n for n in CASES if n != "ueg16_n8"])
def test_trajectory(golden, oracle, name):
    """Full MD steps (NHC VV3 halves, PBC, all forces) vs the reference trajectory."""
    c, g = case(golden, name)
    N, P = c["N"], c["P"]
    steps = len(g.get("e_vir_series", []))
    if steps == 0:
        pytest.skip("no trajectory in fixture")
    b = beads0(g)
    ch = nhc_from(g["nhc0"])
    Lv = c["Lv"]
    p = ptr_p(c)

    def full_force():
        b["f"] = 0
        if c["kind"] == 0:
            oracle.trap_force(p, N, P, ptr(b))
        oracle.pair_force(PAIR_KIND[c["pair"]], p, N, P, dptr(Lv), ptr(b))
        ENk = np.zeros(tri_size(N))
        Eint = np.zeros(N)
        oracle.fast_fill_ENk(N, P, c["omegaP"], c["image"], dptr(Lv), ptr(b), dptr(ENk), dptr(Eint))
        VBN = np.zeros(N + 1)
        oracle.fill_VB(N, c["beta"], c["vi"], dptr(ENk), dptr(VBN))
        oracle.fast_force_VB(N, P, c["beta"], c["vi"], c["omegaP"], c["image"], dptr(Lv), dptr(ENk),
                             dptr(VBN), ptr(b), None, None)
        return ENk, VBN

    full_force()
    ev = []
    for s in range(steps):
        oracle.nhc_vv3_1(N * P, len(ch), ptr(ch), ptr(b), c["T"], c["dt"], 8)
        if c["kind"] != 0:
            oracle.periodic_boundary(N * P, dptr(Lv), ptr(b))
        ENk, VBN = full_force()
        oracle.nhc_vv3_2(N * P, len(ch), ptr(ch), ptr(b), c["T"], c["dt"], 8)
        E2 = np.zeros(tri_size(N))
        oracle.fast_fill_ENk2(N, P, c["omegaP"], c["image"], dptr(Lv), ptr(b), dptr(E2))
        ev.append(oracle.vi_virial_energy(N, P, c["beta"], c["vi"], c["image"], dptr(Lv), dptr(ENk),
                                          dptr(E2), dptr(VBN), ptr(b)))
    assert rel_err(b["x"].reshape(-1), g["x_end"]) < 1e-11
    assert rel_err(b["v"].reshape(-1), g["v_end"]) < 1e-10
    assert rel_err(b["f"].reshape(-1), g["f_end"]) < 1e-10
    ref = nhc_from(g["nhc_end"])
    assert rel_err(ch["vtheta"], ref["vtheta"]) < 1e-10
    assert rel_err(ch["theta"], ref["theta"]) < 1e-10
    np.testing.assert_allclose(ev, g["e_vir_series"], rtol=1e-9)


def test_rng(golden, oracle):
    oracle.set_seed(1234)
    u = [oracle.random_uniform(-1.0, 2.0) for _ in range(8)]
    gs = [oracle.random_gaussian() for _ in range(9)]
    np.testing.assert_array_equal(u, golden["rng/uniform"])
    np.testing.assert_array_equal(gs, golden["rng/gaussian"])
    oracle.random_gaussian()
    oracle.set_seed(77)
    m = np.r_[np.ones(6), np.full(6, 2.5)]
    b = beads_from(np.zeros(36), m=m)
    oracle.init_maxwell_vel(12, ptr(b), 0.7)
    np.testing.assert_array_equal(b["v"].reshape(-1), golden["rng/maxwell_v"])
    oracle.set_seed(5)
    b = beads_from(np.zeros(45))
    oracle.init_uniform_pos(5, 3, dptr(np.array([0.5, 1.0, -0.5])), dptr(np.array([2.0, 3.0, 4.0])), 0.1, ptr(b))
    np.testing.assert_array_equal(b["x"].reshape(-1), golden["rng/uniform_pos"])


def test_two_species_pair_force_and_energy(golden, oracle):
    """orc_pair_force_2 / orc_pair_energy_2 against the compiled reference's
    md_pimd_calc_pair_force_2_x / _energy_2_x, plus momentum balance."""
    o = oracle
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith("two/")}
    NA, NB, P = 5, 3, 4
    L = np.array([3.0, 3.5, 4.0])
    p = make_params(g=3.0, s=0.5)
    for tag, kind in (("g", PAIR_KIND["gaussian"]), ("pg", PAIR_KIND["periodic_gaussian"])):
        a = beads_from(g["xA0"], m=g["mA0"])
        b = beads_from(g["xB0"], m=g["mB0"])
        o.pair_force_2(kind, C.byref(p), NA, NB, P, dptr(L), ptr(a), ptr(b))
        assert rel_err(a["f"].reshape(-1), g["fA_" + tag]) < 1e-14
        assert rel_err(b["f"].reshape(-1), g["fB_" + tag]) < 1e-14
        mom = (a["m"][:, None] * a["f"]).sum(0) + (b["m"][:, None] * b["f"]).sum(0)
        assert np.abs(mom).max() < 1e-13 * np.abs(a["f"]).max()
        e = o.pair_energy_2(kind, C.byref(p), NA, NB, P, dptr(L), ptr(a), ptr(b))
        assert e == pytest.approx(g["e_" + tag][0], rel=1e-14)


OBS_SETUP = {"obs_pbc": dict(image=1, center=[2.5, 2.75, 3.0]), "obs_open": dict(image=0, center=[0.1, -0.2, 0.3])}


@pytest.mark.parametrize("name", list(OBS_SETUP))
def test_observables(golden, oracle, name):
    """orc_density_distribution / orc_pair_correlation / orc_sk_structure
    against the compiled reference (md_pimd_calc_density_distribution_x,
    _pair_correlation_x, _Sk_structure_x); histograms bit for bit."""
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(name + "/")}
    N, P = 20, 6
    L = np.array([5.0, 5.5, 6.0])
    b = beads_from(g["x0"], m=g["m0"])
    su = OBS_SETUP[name]
    fx = np.zeros(24)
    oracle.density_distribution(N, P, ptr(b), dptr(L), su["image"], dptr(np.array(su["center"])), 3.0, 24, dptr(fx))
    assert np.array_equal(fx, g["density"])
    fx = np.zeros(30)
    oracle.pair_correlation(N, P, ptr(b), dptr(L), su["image"], 3.5, 7.0, 4, 30, dptr(fx))
    assert np.array_equal(fx, g["paircorr"])
    fx = np.zeros(16)
    oracle.sk_structure(N, P, ptr(b), 0.2, 0.35, 3, 16, dptr(fx))
    np.testing.assert_allclose(fx, g["sk"], rtol=1e-14, atol=1e-14)
    # S(q -> 0) of bead 1 against bead 1 is N (all phases aligned)
    fx = np.zeros(1)
    oracle.sk_structure(N, P, ptr(b), 0.0, 1.0, 1, 1, dptr(fx))
    assert fx[0] == pytest.approx(N / 3.0 * 3, rel=1e-15)


@pytest.mark.parametrize("pot", ["lj", "he"])
def test_lj_helium_pair(golden, oracle, pot):
    """Lennard-Jones and helium pair forces and energies (one species and two)
    against the compiled reference; momentum balance and a finite-difference
    check of force against energy on one bead."""
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(pot + "/")}
    kind = PAIR_KIND["lj" if pot == "lj" else "helium"]
    N, P, NA, NB = 27, 4, 15, 12
    L = g["L"]
    p = make_params()
    b = beads_from(g["x0"], m=g["m0"])
    oracle.pair_force(kind, C.byref(p), N, P, dptr(L), ptr(b))
    assert rel_err(forces(b), g["f_pair"]) < 1e-13
    assert np.abs((b["m"][:, None] * b["f"]).sum(0)).max() < 1e-12 * np.abs(b["f"]).max()
    e = oracle.pair_energy(kind, C.byref(p), N, P, dptr(L), ptr(b))
    assert e == pytest.approx(g["e_pair"][0], rel=1e-13)
    # the energy is sum(e) / P and the acceleration F / (m P), so -dE/dx = m f for bead (1, 1)
    h = 1e-6 * (1.0 if pot == "lj" else 3.0)
    x = g["x0"].copy()
    fd = np.zeros(3)
    for d in range(3):
        ep = []
        for sgn in (1, -1):
            x2 = x.copy()
            x2[d] += sgn * h
            bb = beads_from(x2, m=g["m0"])
            ep.append(oracle.pair_energy(kind, C.byref(p), N, P, dptr(L), ptr(bb)))
        fd[d] = -(ep[0] - ep[1]) / (2 * h)
    np.testing.assert_allclose(b["f"][0] * b["m"][0], fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())
    a = beads_from(g["xA0"], m=g["mA0"])
    c = beads_from(g["xB0"], m=g["mB0"])
    oracle.pair_force_2(kind, C.byref(p), NA, NB, P, dptr(L), ptr(a), ptr(c))
    assert rel_err(a["f"].reshape(-1), g["fA"]) < 1e-13
    assert rel_err(c["f"].reshape(-1), g["fB"]) < 1e-13
    e = oracle.pair_energy_2(kind, C.byref(p), NA, NB, P, dptr(L), ptr(a), ptr(c))
    assert e == pytest.approx(g["e2"][0], rel=1e-13)


def obs2_inputs(golden):
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith("obs2/")}
    return g, [beads_from(g[t], m=g["m" + t[1:]]) for t in ("xA0", "xB0", "xC0")]


def test_observables_two_systems(golden, oracle):
    """ITCF, two-system and 2-D pair correlations and the S(q) variants
    against the compiled reference; histograms bit for bit."""
    o = oracle
    g, (a, b, c) = obs2_inputs(golden)
    NA, NB, NC, P = 7, 5, 4, 6
    L = np.array([5.0, 5.5, 6.0])

    def hist(n, fn, *args):
        fx = np.zeros(n)
        fn(*args, dptr(fx))
        return fx

    assert np.array_equal(hist(36, o.itcf, NA, NA, P, ptr(a), ptr(a), 0.5, 4.5, 3, 36, 0), g["itcf"])
    assert np.array_equal(hist(36, o.itcf, NA, NB, P, ptr(a), ptr(b), 0.5, 4.5, 3, 36, 1), g["itcf2"])
    assert np.array_equal(hist(20, o.pair_correlation_2, NA, NB, P, ptr(a), ptr(b), dptr(L), 1, 3.5, 7.0, 2, 20),
                          g["paircorr2"])
    assert np.array_equal(hist(20, o.pair_correlation_2, NA, NB, P, ptr(a), ptr(b), dptr(L), 0, 6.0, 3.0, 5, 20),
                          g["paircorr2_open"])
    box = np.array([0.5, 4.5, 0.5, 5.0])
    assert np.array_equal(hist(20, o.pair_correlation_2d, NA, P, ptr(a), 3.0, 5.0, dptr(box), 20), g["paircorr_2d"])
    assert np.array_equal(hist(25, o.lib.orc_2d_pair_correlation, NA, NA, P, ptr(a), ptr(a), dptr(L), 1, 3.0, 3.0, 4,
                               25, 1), g["pc2d"])
    assert np.array_equal(hist(25, o.lib.orc_2d_pair_correlation, NA, NB, P, ptr(a), ptr(b), dptr(L), 0, 3.0, 3.0, 4,
                               25, 0), g["pc2d_2"])
    np.testing.assert_allclose(hist(16, o.sk_structure_2, NA, NB, P, ptr(a), ptr(b), 0.2, 0.35, 3, 16, 0), g["sk2"],
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(hist(16, o.sk_structure_2, NA, NB, P, ptr(a), ptr(b), 0.2, 0.35, 3, 16, 1),
                               g["sk_full2"], rtol=1e-13, atol=1e-13)
    qis = g["sphere_qis"]
    nq = len(g["sphere_qs"])
    np.testing.assert_allclose(np.linalg.norm(qis.reshape(-1, 3), axis=1), g["sphere_qs"], rtol=1e-15)
    sk = hist(nq + 3, o.sphere_sk, 0, NA, 0, 0, P, ptr(a), None, None, 3, dptr(qis), nq)
    np.testing.assert_allclose(sk, g["sphere_sk"], rtol=1e-13, atol=1e-13)
    sk = hist(nq, o.sphere_sk, 1, NA, NB, 0, P, ptr(a), ptr(b), None, 8, dptr(qis), nq)
    np.testing.assert_allclose(sk, g["sphere_sk_full2"], rtol=1e-13, atol=1e-13)
    sk = hist(nq - 2, o.sphere_sk, 2, NA, NB, NC, P, ptr(a), ptr(b), ptr(c), 2, dptr(qis), nq - 2)
    np.testing.assert_allclose(sk, g["sphere_sk3"], rtol=1e-13, atol=1e-13)
    # S(q) of a bead set against itself at q -> 0 is the set size
    z = hist(1, o.sk_structure_2, NA, NB, P, ptr(a), ptr(b), 0.0, 1.0, 1, 1, 1)
    assert z[0] == pytest.approx(NA + NB, rel=1e-15)
    # every ITCF entry is counted once (N^2 entries of weight 1/P)
    assert hist(36, o.itcf, NA, NA, P, ptr(a), ptr(a), 0.5, 4.5, 3, 36, 0).sum() == pytest.approx(NA * NA / P)


@pytest.mark.parametrize("name", ["ueg8_n2", "ueg8_n4_xi0", "ueg16_n8"])
def test_ewald_energy_splits_identities(golden, name):
    """The reference's Ewald energy splits satisfy Ewald = R + NI and
    R2 = Ewald - sum over pairs of 1/r_min (per slice, divided by P), with
    r_min the minimum-image distance (md_statistics_x.c:398-495)."""
    c, g = case(golden, name)
    N, P, L = c["N"], c["P"], c["L"]
    eR, eR2, eNI = g["e_pair_R_R2_NI"]
    e = g["e_pair"][0]
    assert eR + eNI == pytest.approx(e, rel=1e-12, abs=1e-12)
    x = g["x0"].reshape(N, P, 3)
    coul = 0.0
    for j in range(P):
        d = x[:, None, j, :] - x[None, :, j, :]
        d -= L * np.rint(d / L)
        r = np.sqrt((d ** 2).sum(-1))
        iu = np.triu_indices(N, 1)
        coul += (1.0 / r[iu]).sum()
    assert eR2 == pytest.approx(e - coul / P, rel=1e-11, abs=1e-11)
