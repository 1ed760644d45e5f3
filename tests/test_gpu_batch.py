"""Batched replicas (md_pimd_run_batch_x): nine independent systems of uneven
N, P, xi and interaction (trapped Gaussian bosons and Ewald UEG) advanced
together in one captured graph per batch step must follow exactly the
trajectories each system follows alone through md_pimd_run_steps_x (every
kernel of the step is deterministic and the replicas share no buffers)."""
import ctypes as C

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu


def ueg_box(N, rs=2.0):
    return (4 * np.pi * N / 3.0) ** (1.0 / 3.0) * rs


#          kind    N    P   T     xi    seed  nsum
REPLICAS = [("trap", 16, 32, 1.0, 1.0, 1, 2),
            ("trap", 40, 16, 0.5, 0.5, 2, 2),
            ("trap", 130, 8, 0.5, 0.0, 3, 2),
            ("ueg", 8, 4, 0.4, 1.0, 5, 4),
            ("ueg", 33, 6, 0.4, 0.25, 6, 8),
            ("trap", 70, 4, 1.0, 0.75, 7, 2),
            ("trap", 12, 3, 1.0, 1.0, 8, 2),
            ("ueg", 24, 10, 0.4, 0.5, 9, 8),
            ("trap", 5, 2, 1.0, 1.0, 10, 2)]


def build(lib):
    out = []
    for kind, N, P, T, xi, seed, nsum in REPLICAS:
        L = ueg_box(N) if kind == "ueg" else 10.0
        out.append(pyxmd.System(lib, kind, N, P, T, xi, seed=seed, L=L, nsum=nsum))
    return out


def state(s):
    b = s.host_beads()
    ch = pyxmd.nhcs(s.sim)
    return b["x"].copy(), b["v"].copy(), b["f"].copy(), ch["theta"].copy(), ch["vtheta"].copy()


def test_batch_matches_individual_runs():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    dt, steps = 0.01, 6
    alone = build(lib)
    batch = build(lib)
    x0 = [pyxmd.beads(s.sim)["x"].copy() for s in alone]
    # the same process-global interaction parameters for both runs
    lib.md_set_coulomb_truc_x(5.0 / ueg_box(33))
    lib.md_set_coulomb_n_sum_x(8)
    for s in alone:
        cfg = s.step_config(dt, use_graph=1)
        assert lib.md_pimd_run_steps_x(s.pimd, C.byref(cfg), steps) == 0
    assert lib.md_hip_synchronize_x() == 0
    cfgs = (pyxmd.StepConfig * len(batch))(*[s.step_config(dt, use_graph=1) for s in batch])
    ptrs = (pyxmd.PIMD_P * len(batch))(*[s.pimd for s in batch])
    assert lib.md_pimd_run_batch_x(ptrs, cfgs, len(batch), steps) == 0
    for (kind, N, P, *_), a, b, x in zip(REPLICAS, alone, batch, x0):
        sa, sb = state(a), state(b)
        same = all(np.array_equal(u, w) for u, w in zip(sa, sb))
        moved = not np.array_equal(sa[0], x)
        print(f"{kind} N={N} P={P}: identical={same}")
        assert same and moved
    # the batch API refuses a system listed twice
    dup = (pyxmd.PIMD_P * 2)(batch[0].pimd, batch[0].pimd)
    cf2 = (pyxmd.StepConfig * 2)(cfgs[0], cfgs[0])
    assert lib.md_pimd_run_batch_x(dup, cf2, 2, 1) == -1
    lib.md_set_nhc_length_x(8)


def test_batch_with_unbatched_interaction():
    """A replica whose interaction has no batched kernel (Lennard-Jones) is
    launched on its own inside the batch step; its trajectory and the others'
    still match their solo runs bit for bit."""
    from orc import read_golden
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    g = read_golden()
    x_lj, L_lj = g["lj/x0"], float(g["lj/L"][0])

    def build():
        lj = pyxmd.System(lib, "trap", 27, 4, 1.0, 1.0, seed=3, L=L_lj, positions=x_lj)
        ga = pyxmd.System(lib, "trap", 16, 8, 1.0, 0.5, seed=4)
        cl = pyxmd.StepConfig(None, lib.cb["md_LJ_periodic_force_x"], 1e-4, 1, 1, 1)
        return [lj, ga], [cl, ga.step_config(0.01, use_graph=1)]

    alone, cfg_a = build()
    batch, cfg_b = build()
    for s, c in zip(alone, cfg_a):
        assert lib.md_pimd_run_steps_x(s.pimd, C.byref(c), 4) == 0
    cfgs = (pyxmd.StepConfig * 2)(*cfg_b)
    ptrs = (pyxmd.PIMD_P * 2)(*[s.pimd for s in batch])
    assert lib.md_pimd_run_batch_x(ptrs, cfgs, 2, 4) == 0
    for a, b in zip(alone, batch):
        assert all(np.array_equal(u, w) for u, w in zip(state(a), state(b)))
    lib.md_set_nhc_length_x(8)
