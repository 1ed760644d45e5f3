"""The allocation-free V_B entry point md_pimd_fill_VB_noalloc_x against the
allocating one, md_pimd_fill_VB_x, on the same cycle energies: V_B^(0..N) and
the backward potentials Vt^(0..N) must agree bit for bit (the same set-up and
solve kernels, with the working memory in the library's static scratch
array).  Two calls in a row through the C ABI, the second with a smaller N
(the scratch keeps the first call's contents, including hand-off flags at
other positions), then a call with N above the maximum (rejected), then a
forward-only call and the first size again.  The caller's device buffers are
allocated by the test through the HIP runtime the library itself uses (plain
device pointers are passed)."""
import ctypes as C

import numpy as np
import pytest

import pyxmd

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libxmd_x.so is linked to
    (the same loaded instance, by its soname)."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), 8 * n) == 0
        return p

    def put(self, dst, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.rt.hipMemcpy(dst, a.ctypes.data, a.nbytes, self.H2D) == 0

    def get(self, src, n):
        a = np.empty(n)
        assert self.rt.hipMemcpy(a.ctypes.data, src, a.nbytes, self.D2H) == 0
        return a

    def free(self, *ps):
        for p in ps:
            assert self.rt.hipFree(p) == 0


def reference_vb(lib, kind, N, P, xi, seed):
    if kind == "trap":
        s = pyxmd.System(lib, "trap", N, P, 0.5, xi, seed=seed)
    else:
        L = pyxmd.ueg_box(N, 2.0)
        s = pyxmd.System(lib, "ueg", N, P, pyxmd.ueg_temperature(N, L, 0.5), xi, seed=seed, L=L, nsum=8)
    assert lib.md_pimd_fast_fill_ENk_x(s.pimd, s.image) == 0
    assert lib.md_pimd_fill_VB_x(s.pimd) == 0
    s.host_pimd()
    vt = np.zeros(N + 1)
    assert lib.md_pimd_get_backward_potentials_x(s.pimd, pyxmd.dptr(vt)) == 0
    return s, pyxmd.tri(s.pimd), pyxmd.vbn(s.pimd), vt


def noalloc_vb(lib, hip, N, beta, xi, enk, with_vt=True):
    d_enk, d_vbn, d_vt = hip.alloc(len(enk)), hip.alloc(N + 1), hip.alloc(N + 1)
    hip.put(d_enk, enk)
    hip.put(d_vbn, np.full(N + 1, np.nan))
    hip.put(d_vt, np.full(N + 1, np.nan))
    assert lib.md_pimd_fill_VB_noalloc_x(N, beta, xi, d_enk, d_vbn, d_vt if with_vt else None) == 0
    assert lib.md_hip_synchronize_x() == 0
    out = hip.get(d_vbn, N + 1), hip.get(d_vt, N + 1)
    hip.free(d_enk, d_vbn, d_vt)
    return out


def test_noalloc_matches_allocating_entry_point():
    lib = pyxmd.Lib()
    hip = Hip()
    cases = [("ueg", 1024, 4, 1.0, 21), ("trap", 300, 4, 0.5, 22)]
    for kind, N, P, xi, seed in cases:
        print(f"noalloc: {kind} N={N} P={P} xi={xi}", flush=True)
        s, enk, vbn, vt = reference_vb(lib, kind, N, P, xi, seed)
        got_vbn, got_vt = noalloc_vb(lib, hip, N, s.beta, xi, enk)
        np.testing.assert_array_equal(got_vbn, vbn)
        np.testing.assert_array_equal(got_vt, vt)
    # beyond the maximum size: rejected before anything is queued
    big = hip.alloc(1)
    assert lib.md_pimd_fill_VB_noalloc_x(1025, 1.0, 1.0, big, big, None) == -1
    hip.free(big)
    # forward only, then the first size again after the smaller call
    kind, N, P, xi, seed = cases[1]
    s, enk, vbn, _ = reference_vb(lib, kind, N, P, 0.0, seed + 10)
    got_vbn, got_vt = noalloc_vb(lib, hip, N, s.beta, 0.0, enk, with_vt=False)
    np.testing.assert_array_equal(got_vbn, vbn)
    assert np.all(np.isnan(got_vt))
    kind, N, P, xi, seed = cases[0]
    s, enk, vbn, vt = reference_vb(lib, kind, N, P, xi, seed + 10)
    got_vbn, got_vt = noalloc_vb(lib, hip, N, s.beta, xi, enk)
    np.testing.assert_array_equal(got_vbn, vbn)
    np.testing.assert_array_equal(got_vt, vt)
