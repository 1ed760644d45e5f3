"""Lossless checkpoint (md_checkpoint_write_x / _read_x): host-side
round trip bit for bit, and the file parses with the reference's own
md_fread_particle_pos_x / _vel_x / md_fread_nhcs_x (compiled reference in
oracle/_ref, when present).  No GPU needed: objects that never touched the
device keep their state in the host mirror."""
import ctypes as C
import ctypes.util
import os

import numpy as np
import pytest

import pyxmd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libxmd_ref.so")

libc = C.CDLL(ctypes.util.find_library("c"))
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]


def filled_sim(lib, nb, seed):
    sim = lib.md_alloc_simulation_x(nb, 1, 0, 0)
    rng = np.random.default_rng(seed)
    b = pyxmd.beads(sim)
    for f in ("x", "v", "f"):
        b[f] = rng.standard_normal((nb, 3)) * 10.0 ** rng.integers(-5, 5, size=(nb, 3))
    pyxmd.set_beads(sim, b)
    ch = pyxmd.nhcs(sim)
    for f in ("theta", "vtheta", "Q"):
        ch[f] = rng.standard_normal(ch[f].shape) / 3.0
    pyxmd.set_nhcs(sim, ch)
    return sim


def test_checkpoint_round_trip_bit_exact(tmp_path):
    lib = pyxmd.Lib()
    path = str(tmp_path / "ckpt.txt").encode()
    a = filled_sim(lib, 37, 1)
    f = libc.fopen(path, b"w")
    assert lib.md_checkpoint_write_x(f, a) == 0
    libc.fclose(f)
    b = lib.md_alloc_simulation_x(37, 1, 0, 0)
    f = libc.fopen(path, b"r")
    assert lib.md_checkpoint_read_x(f, b) == 0
    libc.fclose(f)
    ba, bb = pyxmd.beads(a), pyxmd.beads(b)
    for fld in ("x", "v", "f"):
        assert np.array_equal(ba[fld], bb[fld])
    ca, cb = pyxmd.nhcs(a), pyxmd.nhcs(b)
    for fld in ("f", "theta", "vtheta", "Q"):
        assert np.array_equal(ca[fld], cb[fld])


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="compiled reference (oracle/_ref) not built")
def test_checkpoint_read_by_reference(tmp_path):
    lib = pyxmd.Lib()
    ref = C.CDLL(REF_LIB)
    ref.md_alloc_simulation_x.restype = C.POINTER(pyxmd.Simulation)
    ref.md_alloc_simulation_x.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    for fn in ("md_fread_particle_pos_x", "md_fread_particle_vel_x", "md_fread_nhcs_x"):
        getattr(ref, fn).argtypes = [C.c_void_p, C.POINTER(pyxmd.Simulation)]
        getattr(ref, fn).restype = C.c_int
    path = str(tmp_path / "ckpt.txt").encode()
    a = filled_sim(lib, 21, 2)
    f = libc.fopen(path, b"w")
    assert lib.md_checkpoint_write_x(f, a) == 0
    libc.fclose(f)
    r = ref.md_alloc_simulation_x(21, 1, 0, 0)
    f = libc.fopen(path, b"r")
    assert ref.md_fread_particle_pos_x(f, r) == 0
    assert ref.md_fread_particle_vel_x(f, r) == 0
    # skip the force block (the reference has no reader for it)
    text = open(path, "rb").read().split(b"\n")
    libc.fclose(f)
    f = libc.fopen(path, b"r")
    libc.fseek.argtypes = [C.c_void_p, C.c_long, C.c_int]
    nhc_off = sum(len(line) + 1 for line in text[: 3 * (21 + 1)])
    libc.fseek(f, nhc_off, 0)
    assert ref.md_fread_nhcs_x(f, r) == 0
    libc.fclose(f)
    ba, br = pyxmd.beads(a), pyxmd.beads(r)
    assert np.array_equal(ba["x"], br["x"]) and np.array_equal(ba["v"], br["v"])
    ca, cr = pyxmd.nhcs(a), pyxmd.nhcs(r)
    for fld in ("theta", "vtheta", "Q"):
        assert np.array_equal(ca[fld], cr[fld])


def test_checkpoint_restores_random_state(tmp_path):
    """The checkpoint's rng line restores md_random_* exactly: after a read,
    the next uniform and Gaussian draws (including a pending Box-Muller
    partner) are the ones that followed the write."""
    lib = pyxmd.Lib()
    lib.md_random_uniform_x.restype = C.c_double
    lib.md_random_uniform_x.argtypes = [C.c_double, C.c_double]
    lib.md_random_gaussian_x.restype = C.c_double
    lib.md_set_seed_x.argtypes = [C.c_uint]
    path = str(tmp_path / "ckpt.txt").encode()
    sim = filled_sim(lib, 5, 3)
    lib.md_set_seed_x(91)
    for _ in range(7):
        lib.md_random_uniform_x(0.0, 1.0)
    lib.md_random_gaussian_x()  # leaves the partner value cached
    f = libc.fopen(path, b"w")
    assert lib.md_checkpoint_write_x(f, sim) == 0
    libc.fclose(f)
    after = [lib.md_random_gaussian_x(), lib.md_random_uniform_x(-1.0, 1.0), lib.md_random_gaussian_x(),
             lib.md_random_gaussian_x()]
    lib.md_set_seed_x(5)  # disturb the generator
    lib.md_random_gaussian_x()
    f = libc.fopen(path, b"r")
    assert lib.md_checkpoint_read_x(f, lib.md_alloc_simulation_x(5, 1, 0, 0)) == 0
    libc.fclose(f)
    again = [lib.md_random_gaussian_x(), lib.md_random_uniform_x(-1.0, 1.0), lib.md_random_gaussian_x(),
             lib.md_random_gaussian_x()]
    assert after == again
