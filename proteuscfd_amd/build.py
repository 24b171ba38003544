"""Build the native libraries in-tree.

* ``libpcfd_mesh.so`` -- host C++ mesh generator / median-dual metrics (g++).
* ``libpcfd.so``      -- HIP kernels for gfx950 + the C ABI (hipcc).
* ``libpcfd_h5.so``   -- partitioned HDF5 mesh files (g++, where HDF5 is installed).

All land next to this file so they travel with the source tree.  Building
never needs a GPU (hipcc cross-compiles for gfx950).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

MESH_SRC = ["mesh.cpp", "mesh_io.cpp"]
HIP_SRC = ["ctx.hip", "kernels_flux.hip", "kernels_visc.hip", "kernels_jac.hip", "kernels_sgs.hip",
           "kernels_misc.hip", "solver.hip", "kernels_chem.hip", "kernels_mlp.hip", "fr.hip"]

# gfx950 only.  Packed-FP32 instructions are disabled for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:6]) + " ...")
    return r


def build_mesh() -> str:
    out = os.path.join(HERE, "libpcfd_mesh.so")
    srcs = [os.path.join(CSRC, s) for s in MESH_SRC]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, *srcs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip() -> str:
    out = os.path.join(HERE, "libpcfd.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SRC]
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    _run([hipcc, *HIPCC_FLAGS, "-I", INCLUDE, *srcs,
          *[os.path.join(CSRC, s) for s in MESH_SRC], "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


HDF5_DIR = os.environ.get("PCFD_HDF5_DIR", "/opt/conda")
HDF5_SONAME = "libhdf5.so.103"


def hdf5_library() -> str | None:
    """The HDF5 shared library the partition-file reader links, or None where HDF5 is absent."""
    lib = os.path.join(HDF5_DIR, "lib", HDF5_SONAME)
    inc = os.path.join(HDF5_DIR, "include", "hdf5.h")
    return lib if os.path.exists(lib) and os.path.exists(inc) else None


def build_h5() -> str | None:
    """libpcfd_h5.so: partitioned HDF5 mesh files (mesh_h5.cpp) plus the median-dual build.
    Linked against HDF5 by path with no rpath: loaders preload HDF5 first
    (proteuscfd_amd.mesh.h5lib), so the image's own libstdc++ stays the one in use."""
    lib = hdf5_library()
    if lib is None:
        return None
    out = os.path.join(HERE, "libpcfd_h5.so")
    srcs = [os.path.join(CSRC, s) for s in ["mesh_h5.cpp", *MESH_SRC]]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, "-isystem",
          os.path.join(HDF5_DIR, "include"), *srcs, lib, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_examples() -> str:
    """C++ host program on include/pcfd_ucs.hpp, linked against libpcfd.so."""
    src = os.path.join(HERE, "examples", "ucs_newton.cpp")
    out = os.path.join(HERE, "examples", "ucs_newton")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, src, "-o", out + ".tmp", "-L", HERE, "-lpcfd",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_mesh()
    build_h5()
    build_hip()
    build_examples()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "mesh" in what:
        print(build_mesh())
    elif "h5" in what:
        print(build_h5())
    elif "hip" in what:
        print(build_hip())
    else:
        build_all()
