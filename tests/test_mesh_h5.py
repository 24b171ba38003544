"""Partitioned HDF5 mesh files (pcfd_mesh_write_h5 / pcfd_mesh_read_h5) on the CPU.

Pinned by the reference's own partition files: udecomp.x's output for its
structured cube series (unitTest/meshResources/cubeStructuredSeries/
cube_{LowFi,MedFi,HiFi}.0.h5, kept in tests/golden) must read back to the unit
volume the reference's gradient test expects (unitTest/gradientTest.h:44-46),
and cube_HiFi.0.h5, which udecomp.x wrote from cube.msh, must give exactly the
nodes and dual volumes our GMSH reader gives for that cube.msh.  Our own files
are checked by round trips (one partition: identical mesh; several
partitions: the joined mesh has the same edges, areas and volumes once the
node numbering is matched), and by the reference's ReadPartedMesh itself
where the reference harness is built (tests/golden/make_ref_vectors.py reads
every case through it).
"""
import gzip
import os
import shutil

import numpy as np
import pytest

from proteuscfd_amd import build as B
from proteuscfd_amd import mesh as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.skipif(B.hdf5_library() is None or
                                not os.path.exists(os.path.join(os.path.dirname(M.__file__), "libpcfd_h5.so")),
                                reason="HDF5 is not installed here")


def gunzip(name, tmp_path):
    out = tmp_path / name[:-3]
    with gzip.open(os.path.join(GOLD, name)) as a, open(out, "wb") as b:
        shutil.copyfileobj(a, b)
    return str(out)


def by_coords(m):
    return np.lexsort(m.xyz.T[::-1])


@pytest.mark.parametrize("fid,nnode,nprism,nbface", [("LowFi", 216, 250, 200), ("MedFi", 1331, 2000, 800)])
def test_reference_partition_files(fid, nnode, nprism, nbface):
    m = M.Mesh.read_h5(os.path.join(GOLD, f"cube_{fid}"))
    assert (m.nnode, m.nelem, m.nbface) == (nnode, nprism, nbface)
    assert np.all(m.elem_type == M.PRISM)
    assert abs(m.total_volume - 1.0) < 1e-14          # gradientTest.h:44-46
    assert abs(m.elem_volume - 1.0) < 1e-14
    assert m.closure_max < 1e-14
    assert sorted(set(m.bface_tag.tolist())) == [1, 13, 17, 21, 25, 26]


def test_reference_partition_file_matches_its_gmsh_source(tmp_path):
    h = M.Mesh.read_h5(gunzip("cube_HiFi.0.h5.gz", tmp_path)[:-5])
    g = M.Mesh.read(gunzip("cube_structured_v40.msh.gz", tmp_path))
    assert (h.nnode, h.nelem, h.nbface, h.nedge, h.nbedge) == (g.nnode, g.nelem, g.nbface, g.nedge, g.nbedge)
    a, b = by_coords(g), by_coords(h)
    np.testing.assert_array_equal(g.xyz[a], h.xyz[b])
    np.testing.assert_array_equal(g.vol[a], h.vol[b])


@pytest.mark.parametrize("kind", [M.BUMP, M.PLATE])
def test_round_trip_one_partition(kind, tmp_path):
    kw = dict(nx=5, ny=4, nz=6) if kind == M.BUMP else dict(nx=6, ny=3, nz=5, nlayers=3)
    m = M.Mesh.build(kind, **kw)
    m.write_h5(str(tmp_path / "c"))
    r = M.Mesh.read_h5(str(tmp_path / "c"))
    assert (m.nnode, m.nelem, m.nedge, m.nbface, m.nbedge) == (r.nnode, r.nelem, r.nedge, r.nbface, r.nbedge)
    np.testing.assert_array_equal(m.xyz, r.xyz)
    np.testing.assert_array_equal(m.edges, r.edges)
    np.testing.assert_array_equal(m.edge_a, r.edge_a)
    np.testing.assert_array_equal(m.vol, r.vol)
    np.testing.assert_array_equal(m.bedges, r.bedges)
    np.testing.assert_array_equal(m.bfactag, r.bfactag)
    np.testing.assert_array_equal(m.bedge_a, r.bedge_a)


def test_partitioned_files_join_back(tmp_path):
    m = M.Mesh.build(M.BUMP, nx=6, ny=5, nz=4)
    part = ((m.xyz[:, 0] > 0.5).astype(np.int32) + 2 * (m.xyz[:, 1] > 0.4)).astype(np.int32)
    m.write_h5(str(tmp_path / "p"), 4, part)
    assert all(os.path.exists(tmp_path / f"p.{r}.h5") for r in range(4))
    r = M.Mesh.read_h5(str(tmp_path / "p"))
    assert (m.nnode, m.nelem, m.nedge, m.nbface, m.nbedge) == (r.nnode, r.nelem, r.nedge, r.nbface, r.nbedge)
    # rank r's nodes come back as one block, in ascending global id
    new_of_old = np.empty(m.nnode, dtype=np.int64)
    order = np.concatenate([np.flatnonzero(part == k) for k in range(4)])
    new_of_old[order] = np.arange(m.nnode)
    np.testing.assert_array_equal(r.xyz[new_of_old], m.xyz)
    np.testing.assert_allclose(r.vol[new_of_old], m.vol, rtol=1e-14, atol=0)
    # the same edges with the same areas (orientation follows the numbering)
    e = new_of_old[m.edges]
    sign = np.where(e[:, 0] < e[:, 1], 1.0, -1.0)
    e = np.sort(e, axis=1)
    idx = {tuple(x): i for i, x in enumerate(r.edges.tolist())}
    k = np.array([idx[tuple(x)] for x in e.tolist()])
    np.testing.assert_allclose(r.edge_a[k, :3] * sign[:, None], m.edge_a[:, :3], rtol=0, atol=1e-14)
    np.testing.assert_allclose(r.edge_a[k, 3], m.edge_a[:, 3], rtol=1e-14, atol=0)
    assert abs(r.total_volume - m.total_volume) < 1e-14


def test_bad_inputs(tmp_path):
    m = M.Mesh.build(M.BUMP, nx=3, ny=3, nz=3)
    with pytest.raises(RuntimeError, match="partition vector"):
        m.write_h5(str(tmp_path / "x"), 2)
    with pytest.raises(ValueError):
        m.write_h5(str(tmp_path / "x"), 2, part=np.zeros(3, dtype=np.int32))
    with pytest.raises(RuntimeError, match="out of range"):
        m.write_h5(str(tmp_path / "x"), 2, part=np.full(m.nnode, 5, dtype=np.int32))
    with pytest.raises(RuntimeError, match="cannot open"):
        M.Mesh.read_h5(str(tmp_path / "missing"))
