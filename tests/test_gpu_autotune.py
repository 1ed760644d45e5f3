"""The autotuner of the Ewald real-space kernel's launch choices (image
enumeration x waves per workgroup): every valid choice gives bit-identical
forces, the tuned choice is within 5 % of the best candidate when all are
timed again, the kept choices survive a save / load round trip, and a batch
with tuning on advances bit-identically to one with the fixed rule.  UEG
N=256, P=8 and N=1024, P=32 (the benchmark's shape), r_s = 2."""
import ctypes as C

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu


def ueg(lib, N, P, seed):
    L = pyxmd.ueg_box(N, 2.0)
    return pyxmd.System(lib, "ueg", N, P, pyxmd.ueg_temperature(N, L, 0.5), 1.0, seed=seed, L=L, nsum=8)


def pair_forces(lib, s):
    assert lib.md_reset_force_x(s.sim) == 0
    assert lib.md_pimd_calc_pair_force_x(s.pimd, s.pair_cb()) == 0
    return s.host_beads()["f"].copy()


@pytest.mark.parametrize("N,P", [(256, 8), (1024, 32)])
def test_choices_bit_identical_and_tuned_choice_fastest(N, P, tmp_path):
    lib = pyxmd.Lib()
    s = ueg(lib, N, P, 3)
    ms = np.zeros(12)
    ch = (C.c_int * 2)()
    assert lib.md_pimd_autotune_x(s.pimd, 10, pyxmd.dptr(ms), ch) == 0
    valid = [k for k in range(12) if ms[k] > 0]
    print(f"N={N} P={P}: times (ms) {dict((k, round(ms[k], 4)) for k in valid)}; choice {ch[0]}, {ch[1]}")
    assert len(valid) >= 4 and 4 * ch[0] + ch[1] in valid
    ref = pair_forces(lib, s)
    try:
        for k in valid:
            assert lib.md_set_ewald_real_variant_x(k // 4, k % 4) == 0
            np.testing.assert_array_equal(pair_forces(lib, s), ref)
    finally:
        lib.md_set_ewald_real_variant_x(-1, -1)
    ms2 = np.zeros(12)
    ch2 = (C.c_int * 2)()
    assert lib.md_pimd_autotune_x(s.pimd, 10, pyxmd.dptr(ms2), ch2) == 0
    best = min(ms2[k] for k in valid)
    print(f"re-timed: tuned choice {ms2[4 * ch[0] + ch[1]]:.4f} ms, best {best:.4f} ms")
    assert ms2[4 * ch[0] + ch[1]] <= 1.05 * best + 0.002
    path = str(tmp_path / "tune.txt")
    assert lib.md_autotune_save_x(path.encode()) == 0
    assert lib.md_autotune_load_x(path.encode()) >= 1


def test_batch_with_tuning_matches_rule():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    out = []
    for tune in (0, 1):
        lib.md_set_autotune_x(tune)
        try:
            systems = [ueg(lib, 128, 8, 40), ueg(lib, 256, 16, 41)]
            cfgs = (pyxmd.StepConfig * 2)(*[s.step_config(0.01, use_graph=1) for s in systems])
            ptrs = (pyxmd.PIMD_P * 2)(*[s.pimd for s in systems])
            assert lib.md_pimd_run_batch_x(ptrs, cfgs, 2, 6) == 0
            out.append([s.host_beads()["x"].copy() for s in systems])
        finally:
            lib.md_set_autotune_x(0)
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
