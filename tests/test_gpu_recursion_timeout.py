"""Error path of the persistent exchange-recursion solve.  Every hand-off wait
inside the solve is bounded; a wait that gives up raises an error word that
lets every other wait fall through (so every kernel exits) and that the host
reports as -1.  The word must survive the later steps of the same call (the
per-step reset clears only the hand-off flags), be reported by the call or by
md_hip_synchronize_x for whichever object it belongs to, and be cleared once
reported.  The timeout is forced from the host through the test hook
md_pimd_inject_recursion_timeouts_x (the solve's own wait gives up at once),
on trapped Gaussian bosons with N = 200 (four 64-row blocks, so the solve
hands blocks between workgroups)."""
import ctypes as C

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu

N, P = 200, 8


def make(lib, seed):
    return pyxmd.System(lib, "trap", N, P, 1.0, 1.0, seed=seed)


def test_timeout_in_first_of_40_graph_steps_is_reported():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    bad = make(lib, 11)
    cfg = bad.step_config(0.01, use_graph=1)
    print(f"inject: N={N} P={P} one timeout, then a 40-step graph call", flush=True)
    assert lib.md_pimd_inject_recursion_timeouts_x(bad.pimd, 1) == 0
    assert lib.md_pimd_run_steps_x(bad.pimd, C.byref(cfg), 40) == -1
    # reported once: the word is clear afterwards
    assert lib.md_hip_synchronize_x() == 0
    # a normal object runs clean, with finite state
    good = make(lib, 12)
    cg = good.step_config(0.01, use_graph=1)
    assert lib.md_pimd_run_steps_x(good.pimd, C.byref(cg), 40) == 0
    assert lib.md_hip_synchronize_x() == 0
    b = good.host_beads()
    assert np.all(np.isfinite(b["x"])) and np.all(np.isfinite(b["f"]))


def test_timeout_in_eager_steps_is_reported_by_synchronize():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    bad = make(lib, 13)
    cfg = bad.step_config(0.01, use_graph=0)
    # injected into the second of 5 eager steps' forward solves
    assert lib.md_pimd_run_steps_x(bad.pimd, C.byref(cfg), 1) == 0
    assert lib.md_hip_synchronize_x() == 0
    assert lib.md_pimd_inject_recursion_timeouts_x(bad.pimd, 1) == 0
    assert lib.md_pimd_run_steps_x(bad.pimd, C.byref(cfg), 5) == 0  # eager: queued only
    assert lib.md_hip_synchronize_x() == -1
    assert lib.md_hip_synchronize_x() == 0


def test_timeout_of_an_earlier_object_is_not_masked_by_a_later_one():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    a, b = make(lib, 14), make(lib, 15)
    ca, cb = a.step_config(0.01, use_graph=0), b.step_config(0.01, use_graph=0)
    assert lib.md_pimd_inject_recursion_timeouts_x(a.pimd, 1) == 0
    assert lib.md_pimd_run_steps_x(a.pimd, C.byref(ca), 2) == 0
    assert lib.md_pimd_run_steps_x(b.pimd, C.byref(cb), 2) == 0  # last launched, clean
    assert lib.md_hip_synchronize_x() == -1
    assert lib.md_hip_synchronize_x() == 0
