"""ctypes binding of the host mesh library (include/pcfd_mesh.h).

Mirrors the reference's mesh set-up calls (ucs/mesh.tcc BuildEdges,
CalcAreasVolumes, CheckForClosure, ReorderMeshCuthillMcKee) for the synthetic
benchmark meshes.  Arrays are copied into numpy on construction.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

WEDGE, BUMP, BOX, PLATE = 1, 2, 3, 4
TET, PRISM = 2, 4
NFACTAG = 7
# pcfd_mesh_read formats
FMT_AUTO, FMT_GMSH, FMT_UGRID, FMT_FIELDVIEW, FMT_SU2 = 0, 1, 2, 3, 4


class MeshSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("seed", C.c_ulonglong), ("jitter", C.c_double), ("rcm", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("angle_deg", C.c_double), ("x0", C.c_double), ("bump_h", C.c_double),
        ("nlayers", C.c_int), ("first_h", C.c_double), ("growth", C.c_double),
    ]


class MeshData(C.Structure):
    _fields_ = [
        ("nnode", C.c_int), ("nelem", C.c_int), ("nedge", C.c_int), ("nbface", C.c_int),
        ("nbedge", C.c_int), ("nfactag", C.c_int),
        ("xyz", C.POINTER(C.c_double)), ("vol", C.POINTER(C.c_double)),
        ("edges", C.POINTER(C.c_int)), ("edge_a", C.POINTER(C.c_double)),
        ("bedges", C.POINTER(C.c_int)), ("bedge_a", C.POINTER(C.c_double)),
        ("bfactag", C.POINTER(C.c_int)),
        ("elem_type", C.POINTER(C.c_int)), ("elem_ptr", C.POINTER(C.c_int)),
        ("elem_nodes", C.POINTER(C.c_int)),
        ("bface_type", C.POINTER(C.c_int)), ("bface_ptr", C.POINTER(C.c_int)),
        ("bface_nodes", C.POINTER(C.c_int)), ("bface_tag", C.POINTER(C.c_int)),
        ("perm", C.POINTER(C.c_int)),
        ("closure_max", C.c_double), ("total_volume", C.c_double), ("elem_volume", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libpcfd_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError("libpcfd_mesh.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.pcfd_mesh_default_spec.argtypes = [C.c_int, C.POINTER(MeshSpec)]
        L.pcfd_mesh_build.argtypes = [C.POINTER(MeshSpec), C.POINTER(C.c_void_p)]
        L.pcfd_mesh_build.restype = C.c_int
        L.pcfd_mesh_get.argtypes = [C.c_void_p]
        L.pcfd_mesh_get.restype = C.POINTER(MeshData)
        L.pcfd_mesh_free.argtypes = [C.c_void_p]
        L.pcfd_mesh_error.restype = C.c_char_p
        L.pcfd_color_nodes.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pcfd_color_nodes.restype = C.c_int
        ipt, dpt = C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.pcfd_mesh_build_from.argtypes = [C.c_int, dpt, C.c_int, ipt, ipt, ipt, C.c_int, ipt, ipt, ipt, ipt,
                                           C.c_int, C.POINTER(C.c_void_p)]
        L.pcfd_mesh_build_from.restype = C.c_int
        L.pcfd_mesh_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.pcfd_mesh_read.restype = C.c_int
        _lib = L
    return _lib


_h5lib = None


def h5lib():
    """libpcfd_h5.so (partitioned HDF5 mesh files).  HDF5 itself is loaded first by
    absolute path, so the library's dependency on it resolves without an rpath."""
    global _h5lib
    if _h5lib is None:
        from . import build as B
        path = os.path.join(HERE, "libpcfd_h5.so")
        hdf5 = B.hdf5_library()
        if hdf5 is None or not os.path.exists(path):
            raise RuntimeError("libpcfd_h5.so is not built (HDF5 not found); run __graft_entry__.build()")
        C.CDLL(hdf5, mode=C.RTLD_GLOBAL)
        L = C.CDLL(path)
        ipt, dpt = C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.pcfd_mesh_write_h5.argtypes = [C.c_char_p, C.c_int, ipt, C.c_int, dpt, C.c_int, ipt, ipt, ipt, C.c_int,
                                         ipt, ipt, ipt, ipt, C.c_int]
        L.pcfd_mesh_write_h5.restype = C.c_int
        L.pcfd_mesh_read_h5.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        L.pcfd_mesh_read_h5.restype = C.c_int
        L.pcfd_mesh_get.argtypes = [C.c_void_p]
        L.pcfd_mesh_get.restype = C.POINTER(MeshData)
        L.pcfd_mesh_free.argtypes = [C.c_void_p]
        L.pcfd_mesh_error.restype = C.c_char_p
        _h5lib = L
    return _h5lib


def default_spec(kind: int, **kw) -> MeshSpec:
    s = MeshSpec()
    lib().pcfd_mesh_default_spec(kind, C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """Numpy view of a generated mesh (all arrays copied, C object freed)."""

    def __init__(self, spec: MeshSpec | None, _handle=None, _lib=None):
        L = _lib or lib()
        if _handle is None:
            h = C.c_void_p()
            if L.pcfd_mesh_build(C.byref(spec), C.byref(h)) != 0:
                raise RuntimeError(L.pcfd_mesh_error().decode())
        else:
            h = _handle
        try:
            d = L.pcfd_mesh_get(h).contents
            self.spec = spec
            self.nnode, self.nelem, self.nedge = d.nnode, d.nelem, d.nedge
            self.nbface, self.nbedge, self.nfactag = d.nbface, d.nbedge, d.nfactag
            self.xyz = _arr(d.xyz, 3 * d.nnode, np.float64).reshape(-1, 3)
            self.vol = _arr(d.vol, d.nnode, np.float64)
            self.edges = _arr(d.edges, 2 * d.nedge, np.int32).reshape(-1, 2)
            self.edge_a = _arr(d.edge_a, 4 * d.nedge, np.float64).reshape(-1, 4)
            self.bedges = _arr(d.bedges, 2 * d.nbedge, np.int32).reshape(-1, 2)
            self.bedge_a = _arr(d.bedge_a, 4 * d.nbedge, np.float64).reshape(-1, 4)
            self.bfactag = _arr(d.bfactag, d.nbedge, np.int32)
            self.elem_type = _arr(d.elem_type, d.nelem, np.int32)
            self.elem_ptr = _arr(d.elem_ptr, d.nelem + 1, np.int32)
            self.elem_nodes = _arr(d.elem_nodes, int(self.elem_ptr[-1]), np.int32)
            self.bface_type = _arr(d.bface_type, d.nbface, np.int32)
            self.bface_ptr = _arr(d.bface_ptr, d.nbface + 1, np.int32)
            self.bface_nodes = _arr(d.bface_nodes, int(self.bface_ptr[-1]), np.int32)
            self.bface_tag = _arr(d.bface_tag, d.nbface, np.int32)
            self.perm = _arr(d.perm, d.nnode, np.int32)
            self.closure_max = d.closure_max
            self.total_volume = d.total_volume
            self.elem_volume = d.elem_volume
        finally:
            L.pcfd_mesh_free(h)

    @classmethod
    def build(cls, kind: int, **kw) -> "Mesh":
        return cls(default_spec(kind, **kw))

    @classmethod
    def from_elements(cls, xyz, elems, bfaces, rcm=False) -> "Mesh":
        """elems: list of (type, nodes); bfaces: list of (type, nodes, tag); 0-based ids."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        et = np.array([t for t, _ in elems], dtype=np.int32)
        ep = np.concatenate([[0], np.cumsum([len(n) for _, n in elems])]).astype(np.int32)
        en = np.concatenate([np.asarray(n) for _, n in elems]).astype(np.int32)
        bt = np.array([t for t, _, _ in bfaces], dtype=np.int32)
        bp = np.concatenate([[0], np.cumsum([len(n) for _, n, _ in bfaces])]).astype(np.int32)
        bn = np.concatenate([np.asarray(n) for _, n, _ in bfaces]).astype(np.int32)
        bg = np.array([g for _, _, g in bfaces], dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        h = C.c_void_p()
        rc = lib().pcfd_mesh_build_from(len(xyz), xyz.ctypes.data_as(C.POINTER(C.c_double)), len(et), ip(et),
                                        ip(ep), ip(en), len(bt), ip(bt), ip(bp), ip(bn), ip(bg), int(rcm),
                                        C.byref(h))
        if rc != 0:
            raise RuntimeError(lib().pcfd_mesh_error().decode())
        return cls(None, _handle=h)

    @classmethod
    def read(cls, path: str, fmt: int = 0, rcm=False) -> "Mesh":
        """Read a GMSH 2.2/4.0/4.1, UGRID, FieldView or SU2 ASCII mesh file
        (format from the extension when fmt is FMT_AUTO)."""
        h = C.c_void_p()
        rc = lib().pcfd_mesh_read(os.fsencode(path), int(fmt), int(rcm), C.byref(h))
        if rc != 0:
            raise RuntimeError(lib().pcfd_mesh_error().decode())
        return cls(None, _handle=h)

    @classmethod
    def read_h5(cls, casename: str, rcm=False) -> "Mesh":
        """Read the partition files <casename>.<rank>.h5 (ucs.x's format) and join
        the partitions into one mesh."""
        L = h5lib()
        h = C.c_void_p()
        if L.pcfd_mesh_read_h5(os.fsencode(casename), int(rcm), C.byref(h)) != 0:
            raise RuntimeError(L.pcfd_mesh_error().decode())
        return cls(None, _handle=h, _lib=L)

    def write_h5(self, casename: str, np_parts: int = 1, part=None, reordered=False) -> None:
        """Write the mesh as np_parts partition files <casename>.<rank>.h5;
        part[nnode] gives each node's rank (required when np_parts > 1)."""
        L = h5lib()
        ip = lambda a: np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int))
        keep = [np.ascontiguousarray(a, dtype=np.int32) for a in
                (self.elem_type, self.elem_ptr, self.elem_nodes, self.bface_type, self.bface_ptr, self.bface_nodes,
                 self.bface_tag)]
        pp = None
        if part is not None:
            part = np.ascontiguousarray(part, dtype=np.int32)
            if part.shape != (self.nnode,):
                raise ValueError(f"part has shape {part.shape}; expected ({self.nnode},)")
            pp = ip(part)
        xyz = np.ascontiguousarray(self.xyz, dtype=np.float64)
        rc = L.pcfd_mesh_write_h5(os.fsencode(casename), int(np_parts), pp, self.nnode,
                                  xyz.ctypes.data_as(C.POINTER(C.c_double)), self.nelem, *[ip(a) for a in keep[:3]],
                                  self.nbface, *[ip(a) for a in keep[3:]], int(bool(reordered)))
        if rc != 0:
            raise RuntimeError(L.pcfd_mesh_error().decode())

    def colour(self) -> tuple[np.ndarray, int]:
        col = np.empty(self.nnode, dtype=np.int32)
        e = np.ascontiguousarray(self.edges, dtype=np.int32)
        nc = lib().pcfd_color_nodes(self.nnode, self.nedge, e.ctypes.data_as(C.POINTER(C.c_int)),
                                    col.ctypes.data_as(C.POINTER(C.c_int)))
        if nc < 0:
            raise RuntimeError(lib().pcfd_mesh_error().decode())
        return col, nc
