"""Test helpers: golden-fixture reader and ctypes bindings of the CPU oracle.

The oracle (oracle/orc_xmd.c) is test infrastructure: a plain-C restatement of
the reference's nocl per-step path.  Only tests, the smoke check and the CPU
baseline leg of the benchmark load it.
"""
import ctypes as C
import os
import struct

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.bin")
ORC_LIB = os.path.join(ROOT, "oracle", "build", "liborc_xmd.so")

DP = C.POINTER(C.c_double)


def read_golden(path=GOLDEN):
    """Parse the fixture file: records of (64-byte name, int64 count, doubles)."""
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        name = data[off:off + 64].split(b"\0", 1)[0].decode()
        off += 64
        (n,) = struct.unpack_from("<q", data, off)
        off += 8
        out[name] = np.frombuffer(data, dtype="<f8", count=n, offset=off).copy()
        off += 8 * n
    return out


class Particle(C.Structure):
    _fields_ = [("m", C.c_double), ("x", C.c_double * 3), ("v", C.c_double * 3), ("f", C.c_double * 3)]


class Nhc(C.Structure):
    _fields_ = [("f", C.c_int), ("theta", C.c_double * 8), ("vtheta", C.c_double * 8), ("Q", C.c_double * 8)]


class Params(C.Structure):
    _fields_ = [("trap_omega", C.c_double), ("trap_center", C.c_double * 3),
                ("gauss_g", C.c_double), ("gauss_s", C.c_double),
                ("coul_q", C.c_double), ("ewald_alpha", C.c_double), ("ewald_n", C.c_int),
                ("he", C.c_double * 8)]


PAIR_KIND = {"gaussian": 0, "periodic_gaussian": 1, "ewald": 2, "ewald_R": 3, "ewald_NI": 4, "lj": 5, "helium": 6}

PARTICLE_DT = np.dtype([("m", "<f8"), ("x", "<f8", 3), ("v", "<f8", 3), ("f", "<f8", 3)])
NHC_DT = np.dtype([("f", "<i4"), ("theta", "<f8", 8), ("vtheta", "<f8", 8), ("Q", "<f8", 8)], align=True)
assert PARTICLE_DT.itemsize == 80 and NHC_DT.itemsize == C.sizeof(Nhc) == 200


def beads_from(x, v=None, f=None, m=None):
    """Structured particle array (reference bead order) from flat arrays."""
    n = len(x) // 3
    b = np.zeros(n, dtype=PARTICLE_DT)
    b["x"] = np.asarray(x).reshape(n, 3)
    if v is not None:
        b["v"] = np.asarray(v).reshape(n, 3)
    if f is not None:
        b["f"] = np.asarray(f).reshape(n, 3)
    b["m"] = 1.0 if m is None else m
    return b


def nhc_from(flat):
    """Structured chain array from the fixture layout (f, theta[8], vtheta[8], Q[8])."""
    r = np.asarray(flat).reshape(-1, 25)
    c = np.zeros(len(r), dtype=NHC_DT)
    c["f"] = r[:, 0].astype(np.int32)
    c["theta"] = r[:, 1:9]
    c["vtheta"] = r[:, 9:17]
    c["Q"] = r[:, 17:25]
    return c


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def dptr(a):
    return a.ctypes.data_as(DP)


class Oracle:
    """Thin wrapper over liborc_xmd.so."""

    def __init__(self, path=ORC_LIB):
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        d, i, vp = C.c_double, C.c_int, C.c_void_p
        sig = {
            "orc_minimum_image": (d, [d, d]),
            "orc_trap_force": (None, [vp, i, i, vp]),
            "orc_pair_force": (None, [i, vp, i, i, DP, vp]),
            "orc_trap_energy": (d, [vp, i, i, vp]),
            "orc_pair_energy": (d, [i, vp, i, i, DP, vp]),
            "orc_pair_force_2": (None, [i, vp, i, i, i, DP, vp, vp]),
            "orc_pair_energy_2": (d, [i, vp, i, i, i, DP, vp, vp]),
            "orc_density_distribution": (None, [i, i, vp, DP, i, DP, d, i, DP]),
            "orc_pair_correlation": (None, [i, i, vp, DP, i, d, d, i, i, DP]),
            "orc_sk_structure": (None, [i, i, vp, d, d, i, i, DP]),
            "orc_itcf": (None, [i, i, i, vp, vp, d, d, i, i, i, DP]),
            "orc_pair_correlation_2": (None, [i, i, i, vp, vp, DP, i, d, d, i, i, DP]),
            "orc_pair_correlation_2d": (None, [i, i, vp, d, d, DP, i, DP]),
            "orc_2d_pair_correlation": (None, [i, i, i, vp, vp, DP, i, d, d, i, i, i, DP]),
            "orc_sk_structure_2": (None, [i, i, i, vp, vp, d, d, i, i, i, DP]),
            "orc_sphere_sk": (None, [i, i, i, i, i, vp, vp, vp, i, DP, i, DP]),
            "orc_madelung": (d, [vp, DP]),
            "orc_fast_fill_ENk": (None, [i, i, d, i, DP, vp, DP, DP]),
            "orc_fill_ENk_slow": (None, [i, i, d, i, DP, vp, DP]),
            "orc_fill_VB": (None, [i, d, d, DP, DP]),
            "orc_fast_force_VB": (None, [i, i, d, d, d, i, DP, DP, DP, vp, DP, DP]),
            "orc_slow_force_VB": (None, [i, i, d, d, d, i, DP, DP, DP, vp]),
            "orc_VBN_energy": (d, [i, i, d, d, DP, DP]),
            "orc_exchange_forces_ld": (None, [i, i, d, d, d, i, DP, vp, DP, DP]),
            "orc_fast_fill_ENk2": (None, [i, i, d, i, DP, vp, DP]),
            "orc_fill_ENk2_slow": (None, [i, i, d, i, DP, vp, DP]),
            "orc_vi_virial_energy": (d, [i, i, d, d, i, DP, DP, DP, DP, vp]),
            "orc_vi_virial_pressure": (d, [i, i, d, d, i, DP, DP, DP, DP, vp]),
            "orc_virial_energy": (d, [i, i, DP, vp]),
            "orc_centroid_virial_energy": (d, [i, i, d, i, vp]),
            "orc_vi_sign": (d, [i, d, d, DP, DP]),
            "orc_nhc_vv3_1": (None, [i, i, vp, vp, d, d, i]),
            "orc_nhc_vv3_2": (None, [i, i, vp, vp, d, d, i]),
            "orc_periodic_boundary": (None, [i, DP, vp]),
            "orc_kinetic": (d, [i, vp]),
            "orc_nhc_kinetic": (d, [i, vp, d, i]),
            "orc_set_seed": (None, [C.c_uint]),
            "orc_random_uniform": (d, [d, d]),
            "orc_random_gaussian": (d, []),
            "orc_init_maxwell_vel": (None, [i, vp, d]),
            "orc_init_uniform_pos": (None, [i, i, DP, DP, d, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args

    def __getattr__(self, name):
        return getattr(self.lib, "orc_" + name)


# Aziz HFDHE2 helium parameters in K and Angstrom (eps, A, alpha, C6, C8, C10, D, r_m)
HE_AZIZ = (10.8, 544850.4, 13.353384, 1.3732412, 0.4253785, 0.178100, 1.241314, 2.9673)


def make_params(trap_omega=1.0, center=(0, 0, 0), g=3.0, s=0.5, q=1.0, alpha=0.0, n=2, he=HE_AZIZ):
    p = Params()
    p.trap_omega = trap_omega
    for k in range(3):
        p.trap_center[k] = center[k]
    p.gauss_g, p.gauss_s, p.coul_q, p.ewald_alpha, p.ewald_n = g, s, q, alpha, n
    for k in range(8):
        p.he[k] = he[k]
    return p


def tri_size(n):
    return n * (n + 1) // 2


def rel_err(a, b):
    """max |a-b| / RMS(b): the force-parity norm used throughout the tests."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    rms = np.sqrt(np.mean(b * b))
    if rms == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / rms)


def exchange_truth(oracle, x, m, N, P, beta, xi, omegaP, image, L):
    """Exchange + ring-spring forces (accelerations, reference bead order,
    3 N P) and V_B^(0..N) from the oracle's 80-bit exchange path: the truth
    double-precision exchange forces are measured against where beta V_B is
    large and the connection matrix's exponents lose digits in double."""
    b = beads_from(np.asarray(x).reshape(-1), m=m)
    f = np.zeros(3 * N * P)
    V = np.zeros(N + 1)
    oracle.exchange_forces_ld(N, P, beta, xi, omegaP, image, dptr(np.array([L, L, L], dtype=float)), ptr(b), dptr(f),
                              dptr(V))
    return f, V
