"""Parity of the HIP path (through the C ABI) against the CPU oracle, phase by phase.

Bar: FP64 phases agree normwise to 1e-12 relative (the oracle restates the
reference's per-edge arithmetic; the GPU sums in the same order, differences
come from FMA contraction only).  The finite-difference Jacobian divides flux
differences by h = 1e-8, so flux rounding differences are amplified by ~1e8:
its bar is 1e-6 relative.
"""
import numpy as np
import pytest

import oracle as O
from proteuscfd_amd import cases
from proteuscfd_amd.cases import Params

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _cases():
    w = cases.wedge_case(params=Params(sorder=2, limiter=2, cfl=0.9, cfl_start=0.9, nsgs=0))
    b = cases.bump_case(n=12, params=Params(sorder=2, limiter=3, cfl=20.0, cfl_start=1.0, cfl_ramp=10, nsgs=4))
    return {"wedge": w, "bump": b}


CASES = _cases()


@pytest.fixture(scope="module", params=list(CASES))
def setup(request):
    from proteuscfd_amd import gpu
    case = CASES[request.param]
    pb = O.Problem.from_case(case)
    ctx = gpu.Context(case)
    q0 = cases.perturbed_state(case)
    print(f"[parity] case={request.param} nnode={case.mesh.nnode} nbedge={case.mesh.nbedge}", flush=True)
    yield request.param, case, pb, ctx, q0
    ctx.close()


def test_update_bcs(setup):
    name, case, pb, ctx, q0 = setup
    qo = pb.update_bcs(q0)
    ctx.upload_q(q0)
    ctx.update_bcs()
    qg = ctx.download_q()
    assert rel(qg[:, :5], qo[:, :5]) < 1e-13
    nn = case.mesh.nnode
    assert rel(qg[nn:, :5], qo[nn:, :5]) < 1e-13


def test_gradient(setup):
    name, case, pb, ctx, q0 = setup
    qo = pb.update_bcs(q0)
    go = pb.gradient(qo)
    ctx.upload_q(qo)
    gg = ctx.gradient()
    for t in range(9):
        assert rel(gg[:, 3 * t:3 * t + 3], go[:, 3 * t:3 * t + 3]) < 1e-12, f"term {t}"


@pytest.mark.parametrize("limiter", [1, 2, 3])
def test_limiter(setup, limiter):
    name, case, pb, ctx, q0 = setup
    p = Params(**{**case.params.__dict__, "limiter": limiter})
    pbl = O.Problem(case.mesh, case.bc_type, case.qref, p)
    qo = pbl.update_bcs(q0)
    go = pbl.gradient(qo)
    lo = pbl.limiter(qo, go)
    ctx.set_params(p)
    ctx.upload_q(qo)
    ctx.gradient()
    lg = ctx.limiter()
    ctx.set_params(case.params)
    assert rel(lg, lo) < 1e-10


@pytest.mark.parametrize("order", [1, 2])
def test_residual(setup, order):
    name, case, pb, ctx, q0 = setup
    qo = pb.update_bcs(q0)
    go = pb.gradient(qo)
    lo = pb.limiter(qo, go) if order == 2 else np.ones((case.mesh.nnode, 5))
    p = Params(**{**case.params.__dict__, "sorder": order})
    pbo = O.Problem(case.mesh, case.bc_type, case.qref, p)
    bo, no = pbo.residual(qo, go, lo, it=1)
    ctx.set_params(p)
    ctx.upload_q(qo)
    ctx.gradient()
    ctx.set_limiter(lo)
    bg, ng = ctx.residual(it=1)
    ctx.set_params(case.params)
    assert rel(bg, bo) < 1e-12
    assert abs(ng[0] - no[0]) <= 1e-12 * no[0]
    np.testing.assert_allclose(ng[1:], no[1:], rtol=1e-12)


def test_timestep(setup):
    name, case, pb, ctx, q0 = setup
    qo = pb.update_bcs(q0)
    dto, mino = pb.timestep(qo, 3.0)
    ctx.upload_q(qo)
    dtg, ming = ctx.timestep(3.0)
    assert rel(dtg, dto) < 1e-13
    assert abs(ming - mino) <= 1e-13 * mino


def test_jacobian(setup):
    name, case, pb, ctx, q0 = setup
    qo = pb.update_bcs(q0)
    dto, _ = pb.timestep(qo, 5.0)
    ia, ja, iau, Mo, qafter = pb.jacobian(qo, dto)
    ctx.upload_q(qo)
    ctx.set_timestep(dto)
    ctx.jacobian()
    Mg = ctx.download_jacobian(ia, ja)
    assert rel(Mg, Mo) < 1e-6
    diag = Mo[iau]
    assert rel(Mg[iau], diag) < 1e-6
    qg = ctx.download_q()
    assert rel(qg[:, :5], qafter[:, :5]) < 1e-13


def test_jacobian_raw_state(setup):
    """Jacobian straight from a state whose phantoms and Dirichlet nodes are
    stale, so the boundary re-evaluation inside the Jacobian does real work."""
    name, case, pb, ctx, q0 = setup
    dto, _ = pb.timestep(q0, 5.0)
    ia, ja, iau, Mo, qafter = pb.jacobian(q0, dto)
    ctx.upload_q(q0)
    ctx.set_timestep(dto)
    ctx.jacobian()
    Mg = ctx.download_jacobian(ia, ja)
    assert rel(Mg[iau], Mo[iau]) < 1e-6
    assert rel(Mg, Mo) < 1e-6
    qg = ctx.download_q()
    assert rel(qg[:, :5], qafter[:, :5]) < 1e-13


@pytest.mark.parametrize("precision", [0, 1])
def test_sgs(setup, precision):
    name, case, pb, ctx, q0 = setup
    qo = pb.update_bcs(q0)
    dto, _ = pb.timestep(qo, 5.0)
    ia, ja, iau, _, _ = pb.jacobian(qo, dto)
    p = Params(**{**case.params.__dict__, "precision": precision, "nsgs": 4})
    ctx.set_params(p)
    ctx.upload_q(qo)
    ctx.set_timestep(dto)
    ctx.jacobian()
    Mg = ctx.download_jacobian(ia, ja)
    bg, _ = ctx.residual(it=1)
    ctx.prepare_sgs()
    xg, _ = ctx.sgs(4)
    order = ctx.sgs_order()
    ctx.set_params(case.params)
    xo = O.sgs_on_matrix(ia, ja, iau, Mg, bg, 4, order)
    tol = 1e-11 if precision == 0 else 2e-4
    assert rel(xg, xo) < tol


def test_iterations(setup):
    name, case, pb, ctx, q0 = setup
    nsteps = 6
    ctx.set_params(case.params)
    q = case.initial_state()
    ctx.upload_q(q)
    ctx.reset_iteration()
    order = ctx.sgs_order() if case.params.nsgs > 0 else None
    so = O.Solver(pb, q, order=order)
    for _ in range(nsteps):
        rg = ctx.iterate()
        ro, _, _ = so.step()
        assert abs(rg[0] - ro) <= 1e-8 * ro
    qg = ctx.download_q()
    assert rel(qg[:, :5], so.q[:, :5]) < 1e-9
    so.close()
