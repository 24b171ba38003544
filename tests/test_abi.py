"""The C-ABI libraries load on a machine without a GPU and export every
function their headers declare (no compute calls are made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(pcfd_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("lib,headers", [("libpcfd.so", ["pcfd.h", "pcfd_mesh.h", "pcfd_chem.h", "pcfd_fr.h"]),
                                         ("libpcfd_mesh.so", ["pcfd_mesh.h"]),
                                         ("libpcfd_h5.so", ["pcfd_mesh.h", "pcfd_mesh_h5.h"])])
def test_library_exports_declared_symbols(lib, headers):
    path = os.path.join(ROOT, "proteuscfd_amd", lib)
    if lib == "libpcfd_h5.so":
        from proteuscfd_amd import build, mesh
        if build.hdf5_library() is None:
            pytest.skip("HDF5 is not installed: libpcfd_h5.so is not built")
        L = mesh.h5lib()
    else:
        assert os.path.exists(path), f"{lib} not built"
        L = ctypes.CDLL(path)
    names = [n for h in headers for n in declared(h)]
    assert len(names) > 5
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_binding_covers_the_abi():
    from proteuscfd_amd import gpu
    assert sorted(gpu.SYMBOLS) == declared("pcfd.h")
    assert sorted(gpu.CHEM_SYMBOLS) == declared("pcfd_chem.h")
    assert sorted(gpu.FR_SYMBOLS) == declared("pcfd_fr.h")


def test_version_string_without_gpu():
    L = ctypes.CDLL(os.path.join(ROOT, "proteuscfd_amd", "libpcfd.so"))
    L.pcfd_version.restype = ctypes.c_char_p
    assert L.pcfd_version().startswith(b"pcfd-gfx950")


def test_create_rejects_bad_mesh_before_touching_the_device():
    """Argument validation happens on the host: a bad descriptor is refused
    with an error code and message, no HIP call is made."""
    from proteuscfd_amd import gpu
    L = gpu.lib()
    desc = gpu.MeshDesc()
    desc.nnode = 0
    h = ctypes.c_void_p()
    assert L.pcfd_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert b"invalid" in L.pcfd_last_error(None)


def test_fr_create_rejects_bad_arguments_on_the_host():
    """pcfd_fr_create refuses null arguments before any HIP call."""
    from proteuscfd_amd import gpu
    L = gpu._fr_lib()
    h = ctypes.c_void_p()
    assert L.pcfd_fr_create(None, None, None, ctypes.byref(h)) == -1
    assert b"null" in L.pcfd_fr_last_error(None)
