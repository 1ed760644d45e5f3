"""Bit-exact restart: a run checkpointed after k steps (md_checkpoint_write_x)
and resumed in a fresh object (md_checkpoint_read_x) reproduces the
uninterrupted run's positions, velocities, forces and chains bit for bit
(every kernel of the step reduces in a fixed order)."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu

libc = C.CDLL(ctypes.util.find_library("c"))
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]


def make(lib, kind):
    if kind == "trap":
        return pyxmd.System(lib, "trap", 16, 8, 0.5, 1.0, seed=5)
    L = pyxmd.ueg_box(64, 2.0)
    return pyxmd.System(lib, "ueg", 64, 4, pyxmd.ueg_temperature(64, L, 0.5), 1.0, seed=5, L=L, nsum=8)


@pytest.mark.parametrize("kind,graph", [("trap", 0), ("ueg", 1)])
def test_resume_bit_exact(lib_fixture, tmp_path, kind, graph):
    lib = lib_fixture
    path = str(tmp_path / "ck.txt").encode()
    a = make(lib, kind)
    cfg = a.step_config(0.01 if kind == "trap" else 0.05, use_graph=graph)
    assert lib.md_pimd_compute_forces_x(a.pimd, C.byref(cfg)) == 0
    assert lib.md_pimd_run_steps_x(a.pimd, C.byref(cfg), 15) == 0
    f = libc.fopen(path, b"w")
    assert lib.md_checkpoint_write_x(f, a.sim) == 0
    libc.fclose(f)
    assert lib.md_pimd_run_steps_x(a.pimd, C.byref(cfg), 15) == 0
    ba, ca = a.host_beads().copy(), pyxmd.nhcs(a.sim).copy()

    b = make(lib, kind)
    f = libc.fopen(path, b"r")
    assert lib.md_checkpoint_read_x(f, b.sim) == 0
    libc.fclose(f)
    assert lib.md_pimd_run_steps_x(b.pimd, C.byref(cfg), 15) == 0
    bb, cb = b.host_beads(), pyxmd.nhcs(b.sim)
    for fld in ("x", "v", "f"):
        assert np.array_equal(ba[fld], bb[fld]), fld
    for fld in ("theta", "vtheta"):
        assert np.array_equal(ca[fld], cb[fld]), fld


@pytest.fixture(scope="module")
def lib_fixture():
    lib = pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    yield lib
    lib.md_set_nhc_length_x(8)
