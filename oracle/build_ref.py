"""Build recipe for oracle/_ref/ref_harness -- the reference solver itself,
compiled from its own sources where they lie under /root/reference, driven by
oracle/ref_harness.cpp, which also compiles the reference-side binding
(proteuscfd_amd/examples/ucs_binding.hpp) against the reference's headers.
TEST INFRASTRUCTURE ONLY: used in the build container
to write golden vectors (tests/golden/make_ref_vectors.py); it never travels
to or runs on the GPU box, and the product never links it.

What it needs, all present in the build container and nothing substituted:
  * the reference's ucs/ and structuralDynamics/ sources (the solver core:
    mesh, parallel, eqnset, elements, h5layer, parameter parser, ...; the
    structural-dynamics sources only for the fluid-structure classes the
    solution-space factory names; none of the LAPACK-using files);
  * MPICH and HDF5 as installed under /opt/conda (headers + shared libraries);
  * the system g++.  The reference's mesh.tcc uses std::numeric_limits
    without including <limits>; the standard header is force-included
    (-include limits) instead of editing the reference.
The conda MPI/HDF5 libraries are reached through symlinks in oracle/_ref/lib
(the executable's rpath), so the system libstdc++ stays the one loaded.

Outputs: oracle/_ref/obj/*.o, oracle/_ref/lib/*, oracle/_ref/ref_harness.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
REF = os.environ.get("PCFD_REFERENCE", "/root/reference")
CONDA = os.environ.get("PCFD_CONDA", "/opt/conda")

UCS_SRC = ["exceptions.cpp", "etypes.cpp", "parallel.cpp", "mesh.cpp", "eqnset.cpp", "elements.cpp", "timer.cpp",
           "endian_util.cpp", "strings_util.cpp", "h5layer.cpp", "parameterParser.cpp", "base64.cpp",
           "oddsNends.cpp"]
SD_SRC = ["fluid_structure.cpp", "element_lib.cpp", "structparam.cpp", "utility.cpp", "explicit.cpp", "implicit.cpp",
          "forces.cpp", "bc.cpp", "io.cpp"]
CONDA_LIBS = ["libmpi.so.12", "libhdf5.so.103", "libgfortran.so.4", "libgomp.so.1", "libquadmath.so.0"]


def available() -> bool:
    return (os.path.isdir(os.path.join(REF, "ucs")) and os.path.exists(os.path.join(CONDA, "include", "mpi.h"))
            and os.path.exists(os.path.join(CONDA, "include", "hdf5.h"))
            and all(os.path.exists(os.path.join(CONDA, "lib", x)) for x in CONDA_LIBS))


def harness_path() -> str:
    return os.path.join(OUT, "ref_harness")


def _flags():
    return ["-std=c++11", "-O2", "-include", "limits", "-w", "-isystem", os.path.join(CONDA, "include"),
            "-I", os.path.join(os.path.dirname(HERE), "include"),
            "-I", os.path.join(REF, "ucs"), "-I", os.path.join(REF, "structuralDynamics"),
            "-I", os.path.join(REF, "TPLs", "tinyxml", "tinyxml-2.6.2")]


def _compile(src, obj, deps=()):
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(d) for d in deps])
    if os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return
    r = subprocess.run(["g++", *_flags(), "-c", src, "-o", obj + ".tmp"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("reference build failed: " + src)
    os.replace(obj + ".tmp", obj)


def build(jobs: int = 8) -> str | None:
    """Compile the reference core and the harness; None where the reference
    or the conda MPI/HDF5 installation is absent (e.g. on the GPU box)."""
    if not available():
        return None
    objdir = os.path.join(OUT, "obj")
    libdir = os.path.join(OUT, "lib")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(libdir, exist_ok=True)
    for lib in CONDA_LIBS:
        link = os.path.join(libdir, lib)
        if not os.path.islink(link):
            os.symlink(os.path.join(CONDA, "lib", lib), link)
    jobs_list = [(os.path.join(REF, "ucs", s), os.path.join(objdir, "ucs_" + s[:-4] + ".o")) for s in UCS_SRC]
    jobs_list += [(os.path.join(REF, "structuralDynamics", s), os.path.join(objdir, "sd_" + s[:-4] + ".o"))
                  for s in SD_SRC]
    jobs_list.append((os.path.join(HERE, "ref_harness.cpp"), os.path.join(objdir, "ref_harness.o")))
    binding = [os.path.join(os.path.dirname(HERE), "proteuscfd_amd", "examples", "ucs_binding.hpp"),
               os.path.join(os.path.dirname(HERE), "include", "pcfd.h")]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda a: _compile(*a, deps=binding if a[0].endswith("ref_harness.cpp") else ()),
                    jobs_list))
    out = harness_path()
    cmd = ["g++", *[o for _, o in jobs_list], os.path.join(libdir, "libmpi.so.12"),
           os.path.join(libdir, "libhdf5.so.103"), "-lz", "-Wl,-rpath," + libdir, "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("reference harness link failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
