"""Case directories for the reference harness (oracle/ref_harness.cpp,
built by oracle/build_ref.py into oracle/_ref/): a partition file of the
mesh written by this repository's HDF5 writer, plus the reference's
.param and .bc text files (ucs/param.tcc keywords, ucs/bc.tcc surface
lines).  Test infrastructure: used by tests/golden/make_ref_vectors.py and
by bench.py's CPU-baseline leg to run the reference solver itself.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

# BC type ids (ucs/bc_defines.h) -> the reference's .bc names
BC_NAMES = {1: "dirichlet", 2: "neumann", 3: "impermeableWall", 4: "noSlip", 6: "farField", 7: "sonicInflow",
            8: "sonicOutflow", 9: "symmetry"}


def param_text(name, keys, nsteps=1):
    lines = [f"{k} = {v}" for k, v in keys.items()]
    return (f"<<<BEGIN TEMPORAL CONTROL>>>\nnumTimeSteps = {nsteps}\nnewtonIterations = 1\n"
            f"<<<END TEMPORAL CONTROL>>>\n\n<<<BEGIN SOLUTION ORDERING>>>\nIterate {name}\n"
            f"<<<END SOLUTION ORDERING>>>\n\n<<<BEGIN SPACE {name}>>>\n" + "\n".join(lines) +
            "\n<<<END SPACE>>>\n")


def write_case(mesh, keys, bc, nsteps=1, name="c", extra=None, directory=None):
    """Writes <dir>/<name>.0.h5, <name>.param and <name>.bc; returns dir
    (with a trailing slash)."""
    d = (directory or tempfile.mkdtemp(prefix="pcfd_ref_")).rstrip("/") + "/"
    os.makedirs(d, exist_ok=True)
    mesh.write_h5(d + name, 1)
    with open(d + name + ".param", "w") as f:
        f.write(param_text(name, keys, nsteps))
    with open(d + name + ".bc", "w") as f:
        for tag, typ in bc.items():
            f.write(f"surface #{tag} = {BC_NAMES[typ]}{(extra or {}).get(tag, '')}\n")
    return d


def harness_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "ref_harness")


def run(directory, name, mode, *args):
    """Runs the harness; returns the meta values it wrote."""
    r = subprocess.run([harness_path(), directory, name, mode, *map(str, args)], capture_output=True, text=True,
                       cwd=directory)
    if r.returncode != 0:
        raise RuntimeError(f"ref_harness {mode} failed ({r.returncode}): " + (r.stdout[-2000:] + r.stderr[-2000:]))
    meta = {}
    for line in open(directory + "ref_meta.txt"):
        k, v = line.split()
        meta[k] = float(v)
    return meta


def read(directory, name, dtype=np.float64):
    return np.fromfile(directory + "ref_" + name + ".bin", dtype=dtype)
