"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle.c, oracle_chem.c and oracle_fr.c into oracle/liboracle.so with gcc.  FP contraction is off
so the oracle evaluates every product and sum exactly as written, like the
reference's separate multiplies and adds.

The reference solver itself is built separately, where its sources and the
MPI/HDF5 installation are present, by oracle/build_ref.py.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build() -> str:
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-fPIC", "-shared",
           os.path.join(HERE, "oracle.c"), os.path.join(HERE, "oracle_chem.c"),
           os.path.join(HERE, "oracle_fr.c"), "-o", out + ".tmp", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
