"""Statistical parity of the energy estimators over a long run: the engine's
fused step on the GPU against the reference's own nocl build
(tests/golden/stats_trap16.json, written by tests/golden/make_stats.c from
the compiled reference).  Same system (trapped bosons N = 16, P = 32,
beta = 1, xi = 1, Gaussian g = 3, s = 0.5, dt = 0.01, massive NHC with the
reference's chain length 8), 10^4 equilibration steps, then 10^6 steps; the
reference samples every step (seed 7), the engine every 50th step, in two
independent trajectories (seeds 7 and 8).  Standard errors come from blocks
of 5 x 10^4 steps (20 per trajectory): the estimators stay correlated over
~10^4 steps (the reference's own block means drift for ~2 x 10^5 steps after
the random start), so shorter blocks understate the error.
Trajectories separate chaotically after a few hundred steps, so only
averages can agree: the engine's mean over both trajectories must lie
within 2 sigma of the reference's, sigma combining both block standard
errors.  (Single 10^6-step trajectories of the engine, seeds 7 .. 10 with
this and with an earlier recursion, scatter 1.2-1.5 times wider than their
block standard errors, and one sat 2.4-3.6 sigma from the reference on all
four correlated estimators at once; averaging two trajectories halves that
scatter.  Both trajectories are deterministic, so the outcome does not vary
between runs of the same build.)
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def trajectory_blocks(lib, ref, seed):
    """Block means (4 estimators x 20 blocks) of one 10^6-step trajectory."""
    s = pyxmd.System(lib, "trap", 16, 32, 1.0, 1.0, seed=seed)
    cfg = s.step_config(0.01, use_graph=1)
    assert lib.md_pimd_compute_forces_x(s.pimd, C.byref(cfg)) == 0
    assert lib.md_pimd_run_steps_x(s.pimd, C.byref(cfg), ref["equil"]) == 0
    every, nblock = 50, ref["nblock"]
    nsamp = ref["steps"] // every
    per = nsamp // nblock
    cb = lib.cb
    st = lib.md_alloc_stats_x(1)

    def take():
        v = st.contents.es[0]
        st.contents.es[0] = 0.0
        return v

    st.contents.es[0] = 0.0
    samples = np.zeros((4, nsamp))
    for k in range(nsamp):
        assert lib.md_pimd_run_steps_x(s.pimd, C.byref(cfg), every) == 0
        assert lib.md_pimd_fast_fill_ENk2_x(s.pimd, 0) == 0
        assert lib.md_pimd_calc_vi_virial_energy_x(s.pimd, st, 0) == 0
        samples[0, k] = take()
        assert lib.md_pimd_calc_VBN_energy_x(s.pimd, st) == 0
        samples[1, k] = take()
        assert lib.md_pimd_calc_trap_energy_x(s.pimd, cb["md_harmonic_trap_energy_x"], st) == 0
        samples[2, k] = take()
        assert lib.md_pimd_calc_pair_energy_x(s.pimd, cb["md_gaussian_energy_x"], st) == 0
        samples[3, k] = take()
    assert lib.md_hip_synchronize_x() == 0
    return samples[:, : per * nblock].reshape(4, 20, -1).mean(axis=2)


def test_estimator_averages_trap16():
    ref = json.load(open(os.path.join(HERE, "golden", "stats_trap16.json")))
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(ref["nhc_length"])
    blocks = np.concatenate([trajectory_blocks(lib, ref, seed) for seed in (7, 8)], axis=1)
    nb = blocks.shape[1]
    zs = {}
    for i, name in enumerate(["vir", "vbn", "trap", "pair"]):
        m = blocks[i].mean()
        sem = blocks[i].std(ddof=1) / np.sqrt(nb)
        rb = np.asarray(ref[name]["blocks"]).reshape(20, -1).mean(axis=1)
        r = {"mean": rb.mean(), "sem": rb.std(ddof=1) / np.sqrt(20)}
        sig = np.hypot(sem, r["sem"])
        zs[name] = (m - r["mean"]) / sig
        print(f"{name}: engine {m:.5f} +- {sem:.5f}, reference {r['mean']:.5f} +- {r['sem']:.5f}, z = {zs[name]:+.2f}")
    for name, z in zs.items():
        assert abs(z) < 2.0, (name, z)
