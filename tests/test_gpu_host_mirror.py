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


def initial_state(case):
    """The wave the C++ program puts on the freestream (same formula)."""
    m = case.mesh
    g, M = 1.4, 0.5
    p = 1.0 / g
    n = m.nnode + m.nbedge
    w = np.zeros(n)
    w[:m.nnode] = 0.02 * np.sin(3.0 * m.xyz[:, 0]) * np.cos(2.0 * m.xyz[:, 1])
    r, u, pp = 1.0 + w, M * (1.0 - w), p * (1.0 + 0.5 * w)
    q = np.zeros((n, 10))
    q[:, 0] = r
    q[:, 1] = r * u
    q[:, 2] = r * 0.1 * w
    q[:, 4] = pp / (g - 1.0) + 0.5 * r * (u * u + 0.01 * w * w)
    return cases.aux(q, g)


def test_cpp_host_program_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "C++ example not built (run __graft_entry__.build())"
    n, iters = 10, 3
    out = tmp_path / "run.bin"
    print(f"ucs_newton n={n} iters={iters}", flush=True)
    r = subprocess.run([EXE, str(n), str(iters), str(out)], capture_output=True, text=True, timeout=120)
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

    params = Params(sorder=2, limiter=3, venkat_K=1.0, cfl=10.0, cfl_start=1.0, cfl_ramp=5, nsgs=4)
    case = cases.bump_case(n=n, params=params)
    assert (nnode, nbnode) == (case.mesh.nnode, case.mesh.nbedge)
    q0 = initial_state(case)
    pb = O.Problem.from_case(case, clip_jacobi=True)
    qo = pb.update_bcs(q0)
    go = pb.gradient(qo)
    lo = pb.limiter(qo, go)
    bo, _ = pb.residual(qo, go, lo, it=2)
    assert rel(qgrad, go) < 1e-12
    assert rel(b, bo) < 1e-12

    so = O.Solver(pb, q0, order=order)
    ro = [so.step()[0] for _ in range(nit)]
    np.testing.assert_allclose(resid, ro, rtol=1e-8)
    assert rel(qout[:nnode], so.q[:nnode, :5]) < 1e-9
    so.close()
