"""The C++ host layer (include/pcfd_ucs.hpp) end to end on the GPU.

proteuscfd_amd/examples/ucs_newton is a ucs.x-style C++ program: it builds the
bump mesh, evaluates one residual phase by phase through SolutionSpace's
mirrored calls and runs Newton iterations; its output is checked here against
the oracle on the same mesh and state.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from proteuscfd_amd import cases
from proteuscfd_amd.cases import Params

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "proteuscfd_amd", "examples", "ucs_newton")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def initial_state(case, alpha=0.0, beta=0.0):
    """The wave the C++ program puts on the freestream (same formula)."""
    m = case.mesh
    g, M = 1.4, 0.5
    p = 1.0 / g
    a, b = np.deg2rad(alpha), np.deg2rad(beta)
    d = (np.cos(a) * np.cos(b), np.sin(b), np.sin(a) * np.cos(b))
    n = m.nnode + m.nbedge
    w = np.zeros(n)
    w[:m.nnode] = 0.02 * np.sin(3.0 * m.xyz[:, 0]) * np.cos(2.0 * m.xyz[:, 1])
    r, u, pp = 1.0 + w, M * (1.0 - w), p * (1.0 + 0.5 * w)
    vel = np.stack([u * d[0], u * d[1] + 0.1 * w, u * d[2]], axis=1)
    q = np.zeros((n, 10))
    q[:, 0] = r
    q[:, 1:4] = r[:, None] * vel
    q[:, 4] = pp / (g - 1.0) + 0.5 * r * (vel * vel).sum(axis=1)
    return cases.aux(q, g)


def run_and_check(tmp_path, arg, case, iters=3, resid_rtol=1e-8, alpha=0.0, beta=0.0):
    assert os.path.exists(EXE), "C++ example not built (run __graft_entry__.build())"
    out = tmp_path / "run.bin"
    print(f"ucs_newton {arg} iters={iters}", flush=True)
    r = subprocess.run([EXE, str(arg), str(iters), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    nnode, nbnode, nit = np.frombuffer(raw[:12], dtype=np.int32)
    off = 12

    def take(count, dt):
        nonlocal off
        a = np.frombuffer(raw[off:off + count * np.dtype(dt).itemsize], dtype=dt)
        off += count * np.dtype(dt).itemsize
        return a

    b = take(nnode * 5, np.float64).reshape(nnode, 5)
    qgrad = take(nnode * 27, np.float64).reshape(nnode, 27)
    resid = take(nit, np.float64)
    qout = take((nnode + nbnode) * 5, np.float64).reshape(-1, 5)
    order = take(nnode, np.int32)

    assert (nnode, nbnode) == (case.mesh.nnode, case.mesh.nbedge)
    q0 = initial_state(case, alpha, beta)
    pb = O.Problem.from_case(case)
    qo = pb.update_bcs(q0)
    go = pb.gradient(qo)
    lo = pb.limiter(qo, go)
    bo, _ = pb.residual(qo, go, lo, it=2)
    assert rel(qgrad, go) < 1e-12
    assert rel(b, bo) < 1e-12

    so = O.Solver(pb, q0, order=order)
    ro = [so.step()[0] for _ in range(nit)]
    assert rel(qout[:nnode], so.q[:nnode, :5]) < 1e-9
    np.testing.assert_allclose(resid, ro, rtol=resid_rtol)
    so.close()


PARAMS = Params(sorder=2, limiter=3, venkat_K=1.0, cfl=10.0, cfl_start=1.0, cfl_ramp=5, nsgs=4)


def test_cpp_host_program_matches_oracle(tmp_path):
    run_and_check(tmp_path, 10, cases.bump_case(n=10, params=PARAMS))


def test_cpp_host_program_on_a_mesh_file(tmp_path):
    """The same program on the reference's GMSH 4.0 structured cube, read by
    pcfd::ucs::Mesh::Read (RCM numbering, farfield on every face tag)."""
    import gzip
    import shutil
    path = tmp_path / "cube.msh"
    with gzip.open(os.path.join(HERE, "golden", "cube_structured_v40.msh.gz")) as a, open(path, "wb") as f:
        shutil.copyfileobj(a, f)
    case = cases.file_case(str(path), {t: cases.BC_FARFIELD for t in (1, 13, 17, 21, 25, 26)}, mach=0.5,
                           alpha_deg=5.0, beta_deg=3.0, params=PARAMS)
    run_and_check(tmp_path, path, case, alpha=5.0, beta=3.0)
