"""The drop-in boundary from C: a plain C99 driver (tests/driver/
pimd_step_driver.c) that uses only the reference's API is compiled with gcc
and the flags of the engine's xmd_config (--xmd-flags / --xmd-libs, as the
reference's xmd_config.c:9-23 gives them to its drivers) and linked against
libxmd_x.so.  The CPU test builds it (the headers are valid C, every symbol
it uses resolves); the GPU test runs the reference driver loop for the golden
case trap16 (N=16, P=32, 20 steps, NHC M=8) from the golden initial state and
compares the final state and the per-step virial estimator with the
compiled reference's.  Also: the library exports the C API only."""
import os
import subprocess

import numpy as np
import pytest

from orc import ROOT, read_golden, rel_err

XMD_CONFIG = os.path.join(ROOT, "md-lab_amd", "bin", "xmd_config")
LIB = os.path.join(ROOT, "md-lab_amd", "lib", "libxmd_x.so")
SRC = os.path.join(ROOT, "tests", "driver", "pimd_step_driver.c")


def build_driver(tmp_path):
    flags = subprocess.run([XMD_CONFIG, "--xmd-flags"], capture_output=True, text=True, check=True).stdout.split()
    libs = subprocess.run([XMD_CONFIG, "--xmd-libs"], capture_output=True, text=True, check=True).stdout.split()
    exe = str(tmp_path / "pimd_step_driver")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-O2", SRC, "-o", exe] + flags + libs, check=True)
    return exe


def test_driver_compiles_and_links(tmp_path):
    exe = build_driver(tmp_path)
    assert os.path.exists(exe)


def test_library_exports_only_the_c_api():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    names = [l.split()[-1] for l in out.splitlines() if l.strip()]
    other = [n for n in names if not n.startswith("md_")]
    assert not other, other[:20]


@pytest.mark.gpu
def test_driver_runs_reference_loop(tmp_path):
    exe = build_driver(tmp_path)
    gold = read_golden()
    g = {k.split("/", 1)[1]: v for k, v in gold.items() if k.startswith("trap16/")}
    N, P, T, vi, _, _, _, _, _, dt = g["meta"]
    N, P = int(N), int(P)
    steps = len(g["e_vir_series"])
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([g["x0"], g["v0"]]).astype("<f8").tofile(inp)
    print(f"C driver: trap16 N={N} P={P} {steps} steps", flush=True)
    r = subprocess.run([exe, str(N), str(P), repr(float(T)), repr(float(vi)), str(steps), repr(float(dt)), str(inp),
                        str(outp)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(outp, dtype="<f8")
    nb3 = 3 * N * P
    x, v, f, ev = out[:nb3], out[nb3:2 * nb3], out[2 * nb3:3 * nb3], out[3 * nb3:]
    assert rel_err(x, g["x_end"]) < 1e-10
    assert rel_err(v, g["v_end"]) < 1e-9
    assert rel_err(f, g["f_end"]) < 1e-9
    np.testing.assert_allclose(ev, g["e_vir_series"], rtol=1e-9)
