"""Full-size parity at the benchmark configurations against the compiled
reference (tests/golden/configs.bin, written by tests/golden/make_golden.c in
its `configs` mode from the reference's own nocl build):

* trapped bosons N=256, P=64, beta=2, Gaussian g=3 s=0.5, at the five xi of
  the fermion-extrapolation grid {0, 0.25, 0.5, 0.75, 1};
* spin-polarised UEG N=128, P=64, r_s=2, theta=0.5, Ewald alpha=5/L, |m|<=8
  (converged: evaluated with the engine's own split), xi in {0, 0.5, 1}.

Same initial state on both sides (positions and velocities stored whole,
chains from md_init_nhc_x).  Per xi: the total
force (trap + pair + exchange; every stride-th bead), V_B^(N) and the
thermodynamic, virial and pair estimators; at the last xi a short fused-step
trajectory.  Force bar: the total force (trap + pair from the reference,
plus the exchange and ring-spring part from the oracle's 80-bit exchange
path, which the reference's cycle-walk algorithm pins to ~1e-15 on the golden
cases) is the truth; the engine is held to max(1e-11, four times the reference
fast path's own distance from it).  At beta=2 with N(0,1) beads
beta V_B ~ 1.6e6, so the connection-matrix exponents carry ~1e-10 of absolute
round-off in any double-precision code.  Energies 1e-12 relative;
trajectories 1e-10 (x) / 1e-9 (v)."""
import os

import numpy as np
import pytest

import pyxmd
from orc import Oracle, exchange_truth, read_golden, rel_err

pytestmark = pytest.mark.gpu

CONFIGS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs.bin")


def stat(lib, fn, *args, pos=1):
    st = lib.md_alloc_stats_x(1)
    a = list(args)
    a.insert(pos, st)
    assert fn(*a) == 0
    return st.contents.es[0]


@pytest.mark.parametrize("name", ["cfg_trap256", "cfg_ueg128"])
def test_config_parity(name):
    gold = read_golden(CONFIGS)
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith(name + "/")}
    N, P, T, _, kind, image, L, nsum, stride, dt = g["meta"]
    N, P, image, nsum, stride = int(N), int(P), int(image), int(nsum), int(stride)
    kind = "trap" if kind == 0 else "ueg"
    xis = g["xis"]
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(8)
    s = pyxmd.System(lib, kind, N, P, T, float(xis[0]), L=L, nsum=nsum, positions=g["x0"])
    pyxmd.load_state(s, g["x0"], g["v0"])  # chains: md_init_nhc_x, as on the reference side
    sub = slice(0, None, stride)
    o = Oracle()
    for q, xi in enumerate(xis):
        s.pimd.contents.vi = float(xi)
        s.full_force()
        f = s.host_beads()["f"][sub].reshape(-1)
        err = rel_err(f, g[f"f_total_{q}"])
        s.host_pimd()
        e = g[f"e_{q}"]
        vbn = s.pimd.contents.VBN[N]
        e_vbn = stat(lib, lib.md_pimd_calc_VBN_energy_x, s.pimd)
        assert lib.md_pimd_fast_fill_ENk2_x(s.pimd, image) == 0
        e_vir = stat(lib, lib.md_pimd_calc_vi_virial_energy_x, s.pimd, image)
        e_pair = stat(lib, lib.md_pimd_calc_pair_energy_x, s.pimd, s.pair_energy_cb(), pos=2)
        fx, _ = exchange_truth(o, g["x0"], np.ones(N * P), N, P, 1.0 / T, float(xi), np.sqrt(P) * T, image, L)
        truth = g["f_ext"] + fx.reshape(-1, 3)[sub].reshape(-1)
        err, ref = rel_err(f, truth), rel_err(g[f"f_total_{q}"], truth)
        tol = max(1e-11, 4 * ref)
        print(f"{name} xi={xi}: total force vs 80-bit truth: engine {err:.2e}, reference fast path {ref:.2e} "
              f"(reference cycle walk {rel_err(g[f'f_total_slow_{q}'], truth):.2e}); V_B {vbn:.12g} vs {e[0]:.12g}")
        assert err < tol
        tol = tol + ref  # engine vs reference, for the virial bar below
        for got, ref in zip((vbn, e_vbn, e_pair), (e[0], e[1], e[3])):
            assert got == pytest.approx(ref, rel=1e-12, abs=1e-12)
        # the virial estimator adds sum_j>1 (x_j - x_1) . f_j m / 2, so the force bar
        # carries into it: tol * RMS(f) * sum |x_j - x_1| m / 2
        b = s.host_beads()
        x = b["x"].reshape(N, P, 3)
        m = b["m"].reshape(N, P)
        d = np.abs(x[:, 1:] - x[:, :1])
        if image:
            d = np.minimum(d, L - d)
        vtol = tol * np.sqrt(np.mean(g[f"f_total_{q}"] ** 2)) * float((d.sum(-1) * m[:, 1:]).sum()) / 2
        assert abs(e_vir - e[2]) <= vtol + 1e-12 * abs(e[2])
    # trajectory at the last xi, fused and graph-replayed, from the state above
    cfg = s.step_config(dt, use_graph=1)
    steps = len(g["e_vir_series"])
    assert lib.md_pimd_run_steps_x(s.pimd, pyxmd.C.byref(cfg), steps) == 0
    b = s.host_beads()
    ex, ev = rel_err(b["x"][sub].reshape(-1), g["x_end"]), rel_err(b["v"][sub].reshape(-1), g["v_end"])
    print(f"{name}: {steps}-step trajectory rel err x {ex:.2e}, v {ev:.2e}")
    assert ex < 1e-10 and ev < 1e-9
