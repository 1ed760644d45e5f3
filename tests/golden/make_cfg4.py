"""Writes tests/golden/cfg4.bin: reference records for a seeded slice of the
batched-replica table of bench.py (replica_table(), seed 2024), one
make_golden process per replica (oracle/_ref/make_golden, the reference's
own nocl build; run in the build container, where the reference tree is).

    make -C oracle ref && python tests/golden/make_cfg4.py
"""
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "md-lab_amd"))

import bench  # noqa: E402

# trap N = 8 .. 512 and P = 16 .. 64; UEG r_s = 1, 2, 4 and P = 16 .. 64; xi 0 .. 1
PICK = [0, 1, 6, 22, 27, 28, 38, 44, 46, 51, 56, 62]
STRIDE = 7


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "make_golden")
    table = bench.replica_table()
    with tempfile.TemporaryDirectory() as tmp:
        procs, parts = [], []
        for i in PICK:
            r = table[i]
            part = os.path.join(tmp, f"r{i}.bin")
            parts.append(part)
            kind = 0 if r["kind"] == "trap" else 1
            procs.append(subprocess.Popen([exe, part, "replica", f"rep{i}", str(kind), str(r["N"]), str(r["P"]),
                                           repr(r["xi"]), repr(r["T"]), repr(r["L"]), str(100 + i), str(STRIDE)]))
        for p in procs:
            assert p.wait() == 0
        with open(os.path.join(ROOT, "tests", "golden", "cfg4.bin"), "wb") as f:
            for part in parts:
                with open(part, "rb") as g:
                    f.write(g.read())


if __name__ == "__main__":
    main()
