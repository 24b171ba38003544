"""Pins for the CPU oracle (plain C restatement of the reference path).

Each check is independent of the oracle's own algebra: a textbook
Roe solver written as a wave decomposition (not the reference's T*|L|*Tinv
matrices), rotational invariance, freestream preservation and discrete
conservation on a closed mesh, LSQ exactness on linear fields, limiter
closed forms, dense linear algebra against numpy, a dense block
Gauss-Seidel, and a matrix-free finite difference of the residual.
"""
import numpy as np
import pytest

import oracle as O
from proteuscfd_amd import cases
from proteuscfd_amd import mesh as M
from proteuscfd_amd.cases import Params

G = 1.4


def prim_to_cons(rho, u, p, g=G):
    u = np.asarray(u, dtype=float)
    return np.array([rho, *(rho * u), p / (g - 1) + 0.5 * rho * u @ u])


def phys_flux_textbook(Q, n, g=G):
    rho, u = Q[0], Q[1:4] / Q[0]
    p = (g - 1) * (Q[4] - 0.5 * rho * u @ u)
    un = u @ n
    return np.array([rho * un, *(rho * un * u + p * n), (Q[4] + p) * un])


def roe_textbook(QL, QR, n, area, g=G, fix=True):
    """Roe flux from Roe averages and the characteristic wave decomposition
    (Roe 1981; Toro, Riemann Solvers, ch. 11), Harten-Hyman fix optional."""
    def prim(Q):
        rho = Q[0]
        u = Q[1:4] / rho
        p = (g - 1) * (Q[4] - 0.5 * rho * u @ u)
        return rho, u, p, (Q[4] + p) / rho
    rL, uL, pL, hL = prim(QL)
    rR, uR, pR, hR = prim(QR)
    sL, sR = np.sqrt(rL), np.sqrt(rR)
    r = sL * sR
    u = (sL * uL + sR * uR) / (sL + sR)
    h = (sL * hL + sR * hR) / (sL + sR)
    c = np.sqrt((g - 1) * (h - 0.5 * u @ u))
    un = u @ n
    dr, du, dp = rR - rL, uR - uL, pR - pL
    dun = du @ n
    a1 = (dp - r * c * dun) / (2 * c * c)          # u - c
    a5 = (dp + r * c * dun) / (2 * c * c)          # u + c
    a2 = dr - dp / (c * c)                         # entropy
    dut = du - dun * n                             # shear
    lam = np.array([un - c, un, un + c])
    if fix:
        cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
        lamL = np.array([uL @ n - cL, uL @ n, uL @ n + cL])
        lamR = np.array([uR @ n - cR, uR @ n, uR @ n + cR])
        eps = np.maximum(0, np.maximum(lam - lamL, lamR - lam))
        alam = np.where(np.abs(lam) < eps, 0.5 * (lam ** 2 / np.where(eps > 0, eps, 1) + eps), np.abs(lam))
    else:
        alam = np.abs(lam)
    d = alam[0] * a1 * np.array([1, *(u - c * n), h - un * c])
    d += alam[2] * a5 * np.array([1, *(u + c * n), h + un * c])
    d += alam[1] * (a2 * np.array([1, *u, 0.5 * u @ u]) + r * np.array([0, *dut, u @ dut]))
    return 0.5 * area * (phys_flux_textbook(QL, n) + phys_flux_textbook(QR, n) - d)


def rand_state(rng, mach=0.5):
    rho = rng.uniform(0.5, 2.0)
    u = rng.normal(size=3)
    u *= mach * rng.uniform(0.5, 1.5) / np.linalg.norm(u)
    p = rng.uniform(0.4, 1.2)
    return prim_to_cons(rho, u, p)


def rand_normal(rng):
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


@pytest.mark.parametrize("mach", [0.3, 0.9, 2.5])
def test_roe_matches_textbook_wave_decomposition(mach):
    rng = np.random.default_rng(11)
    for _ in range(200):
        QL = rand_state(rng, mach)
        rho, u = QL[0], QL[1:4] / QL[0]
        p = (G - 1) * (QL[4] - 0.5 * rho * u @ u)
        QR = prim_to_cons(rho * rng.uniform(0.8, 1.2), u * rng.uniform(0.8, 1.2, 3), p * rng.uniform(0.8, 1.2))
        n = rand_normal(rng)
        area = rng.uniform(0.1, 2.0)
        f = O.roe_flux(QL, QR, np.array([*n, area]))
        ft = roe_textbook(QL, QR, n, area)
        np.testing.assert_allclose(f, ft, rtol=1e-11, atol=1e-12 * np.abs(ft).max())


def test_roe_consistency_and_supersonic_upwinding():
    rng = np.random.default_rng(3)
    for _ in range(50):
        Q = rand_state(rng)
        n = rand_normal(rng)
        f = O.roe_flux(Q, Q, np.array([*n, 0.7]))
        np.testing.assert_allclose(f, 0.7 * phys_flux_textbook(Q, n), rtol=1e-13, atol=1e-14)
    # supersonic along +n on both sides: the Roe flux is the left flux
    for _ in range(50):
        n = rand_normal(rng)
        QL = prim_to_cons(rng.uniform(0.8, 1.2), 3.0 * n + 0.1 * rng.normal(size=3), rng.uniform(0.6, 0.8))
        QR = QL * (1 + 0.01 * rng.uniform(-1, 1, 5))
        f = O.roe_flux(QL, QR, np.array([*n, 1.0]))
        np.testing.assert_allclose(f, phys_flux_textbook(QL, n), rtol=1e-12, atol=1e-13)


def test_roe_rotational_invariance():
    rng = np.random.default_rng(5)
    for _ in range(50):
        QL, QR, n = rand_state(rng), rand_state(rng), rand_normal(rng)
        A, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        rot = lambda Q: np.array([Q[0], *(A @ Q[1:4]), Q[4]])
        f = O.roe_flux(QL, QR, np.array([*n, 1.3]))
        fr = O.roe_flux(rot(QL), rot(QR), np.array([*(A @ n), 1.3]))
        np.testing.assert_allclose(fr[[0, 4]], f[[0, 4]], rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(fr[1:4], A @ f[1:4], rtol=1e-12, atol=1e-13)


def box_case(bc, mach=0.5, n=4, direction=(1, 0.3, -0.2)):
    m = M.Mesh.build(M.BOX, nx=n, ny=n + 1, nz=n, jitter=0.1, seed=9)
    q0 = cases.qinf(G, mach, direction=direction)
    bct = np.array([0] + [bc] * 6, dtype=np.int32)
    return cases.Case("box", m, bct, np.tile(q0, (7, 1)), q0, Params(sorder=1, nsgs=2))


def test_freestream_preservation_and_conservation():
    case = box_case(cases.BC_FARFIELD)
    pb = O.Problem.from_case(case)
    q = pb.update_bcs(case.initial_state())
    b, norms = pb.residual(q)
    flux_scale = np.abs(O.roe_flux(case.qinf, case.qinf, np.array([1, 0, 0, 1.0]))).max()
    assert np.abs(b).max() < 1e-13 * flux_scale * case.mesh.edge_a[:, 3].max()
    # perturbed state: interior fluxes cancel, sum of b = -sum of boundary fluxes
    qp = pb.update_bcs(cases.perturbed_state(case, amp=0.1))
    b, _ = pb.residual(qp)
    m = case.mesh
    fb = sum(O.roe_flux(qp[m.bedges[h, 0]], qp[m.bedges[h, 1]], m.bedge_a[h]) for h in range(m.nbedge))
    np.testing.assert_allclose(b.sum(axis=0), -fb, atol=1e-13 * np.abs(fb).max())


def test_farfield_phantom_reproduces_freestream():
    """With the interior at Q_inf every characteristic branch returns Q_inf."""
    rng = np.random.default_rng(1)
    for mach in (0.3, 0.8, 1.5, 2.5):
        qinf = cases.qinf(G, mach, direction=rng.normal(size=3))
        for _ in range(10):
            n = rand_normal(rng)
            QL, QR = qinf.copy(), cases.perturbed_state(box_case(6), 0.0)[0] * 0 + qinf
            QR = QR.copy()
            QR[:5] *= 1 + 0.05 * rng.uniform(-1, 1, 5)
            O.lib().o_bc_state(6, O._d(QL), O._d(QR), O._d(qinf), O._d(np.array([*n, 1.0])), G)
            np.testing.assert_allclose(QR[:5], qinf[:5], rtol=1e-12)


def test_wall_phantom_is_tangential():
    """Slip-wall CVBC (Whitfield): p_b = p_i + rho c theta_i and
    u_b = u_i - n (p_b - p_i)/(rho c), so the phantom velocity is the interior
    velocity with its normal component removed, whatever the fixed point."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        n = rand_normal(rng)
        QL = cases.aux(np.r_[rand_state(rng, 0.3), np.zeros(5)], G)
        QR = QL.copy()
        QR[:5] *= 1 + 0.05 * rng.uniform(-1, 1, 5)
        O.lib().o_bc_state(3, O._d(QL), O._d(QR), O._d(QL), O._d(np.array([*n, 1.0])), G)
        ui, ub = QL[1:4] / QL[0], QR[1:4] / QR[0]
        assert abs(ub @ n) < 1e-14 * np.linalg.norm(ui)
        np.testing.assert_allclose(ub, ui - (ui @ n) * n, atol=1e-14)


def test_lsq_gradient_exact_for_linear_fields():
    case = cases.bump_case(n=6)
    pb = O.Problem.from_case(case)
    m = case.mesh
    rng = np.random.default_rng(4)
    A = rng.normal(size=(5, 3)) * 0.1
    c0 = np.array([1.0, 0.5, 0.2, -0.1, 3.0])
    q = case.initial_state()
    q[: m.nnode, :5] = c0 + m.xyz @ A.T
    g = pb.gradient(q)
    for t in range(5):
        np.testing.assert_allclose(g[:, 3 * t:3 * t + 3], np.tile(A[t], (m.nnode, 1)), atol=1e-12)


def test_limiter_closed_forms():
    case = cases.bump_case(n=5)
    m = case.mesh
    q = case.initial_state()
    for lim, expect in [(1, 1.0), (2, 0.75), (3, 1.0)]:
        p = Params(sorder=2, limiter=lim)
        pb = O.Problem(m, case.bc_type, case.qref, p)
        g = pb.gradient(q)
        l = pb.limiter(q, g)
        # uniform field: extrapolation = node value; Venkat gives phi(1)=3/4
        np.testing.assert_array_equal(l, expect)
    # positive linear field, exact gradients: Barth never limits
    A = np.array([0.02, -0.01, 0.03])
    q2 = q.copy()
    q2[: m.nnode, 0] = 1.0 + m.xyz @ A
    q2[: m.nnode, 4] = 2.5 + m.xyz @ (2 * A)
    pb = O.Problem(m, case.bc_type, case.qref, Params(sorder=2, limiter=1))
    l = pb.limiter(q2, pb.gradient(q2))
    np.testing.assert_allclose(l, 1.0, atol=1e-12)


def test_timestep_quiescent_closed_form():
    case = box_case(cases.BC_FARFIELD, mach=0.0, direction=(1, 0, 0))
    pb = O.Problem.from_case(case)
    m = case.mesh
    q = case.initial_state()
    dt, dtmin = pb.timestep(q, 2.0)
    area = np.zeros(m.nnode)
    np.add.at(area, m.edges[:, 0], m.edge_a[:, 3])
    np.add.at(area, m.edges[:, 1], m.edge_a[:, 3])
    np.add.at(area, m.bedges[:, 0], m.bedge_a[:, 3])
    c = np.sqrt(G * q[0, 6] / q[0, 0])
    np.testing.assert_allclose(dt, 2.0 * m.vol / (c * area), rtol=1e-13)
    assert dtmin == dt.min()


def test_dense_lu_against_numpy():
    rng = np.random.default_rng(8)
    for _ in range(100):
        A = rng.normal(size=(5, 5))
        A[rng.integers(5)] *= 1e-3  # force pivoting
        b = rng.normal(size=5)
        lu, p, bad = O.lu(A)
        x = O.lu_solve(lu, p, b)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


def _dense_from_crs(ia, ja, M_):
    n = len(ia) - 1
    A = np.zeros((5 * n, 5 * n))
    for i in range(n):
        for k in range(ia[i], ia[i + 1]):
            j = ja[k]
            A[5 * i:5 * i + 5, 5 * j:5 * j + 5] = M_[k]
    return A


def test_sgs_matches_dense_block_gauss_seidel_and_converges():
    case = cases.bump_case(n=3, params=Params(sorder=1, nsgs=3))
    pb = O.Problem.from_case(case)
    q = pb.update_bcs(cases.perturbed_state(case))
    dt, _ = pb.timestep(q, 2.0)
    ia, ja, iau, M_, _ = pb.jacobian(q, dt)
    n = case.mesh.nnode
    b = np.random.default_rng(0).normal(size=(n, 5))
    A = _dense_from_crs(ia, ja, M_)
    order = np.random.default_rng(1).permutation(n)
    x = np.zeros(5 * n)
    for _ in range(3):
        for sweep in (order, order[::-1]):
            for i in sweep:
                r = b[i] - A[5 * i:5 * i + 5] @ x + A[5 * i:5 * i + 5, 5 * i:5 * i + 5] @ x[5 * i:5 * i + 5]
                x[5 * i:5 * i + 5] = np.linalg.solve(A[5 * i:5 * i + 5, 5 * i:5 * i + 5], r)
    xo = O.sgs_on_matrix(ia, ja, iau, M_, b, 3, order)
    np.testing.assert_allclose(xo.reshape(-1), x, rtol=1e-10, atol=1e-12 * np.abs(x).max())
    xs = O.sgs_on_matrix(ia, ja, iau, M_, b, 200)
    np.testing.assert_allclose(xs.reshape(-1), np.linalg.solve(A, b.reshape(-1)), rtol=1e-8, atol=1e-10)


def test_jacobian_matches_matrix_free_residual_derivative():
    """(A - V/dt) v == -(b(q + eps v) - b(q))/eps, boundary states re-evaluated."""
    case = cases.bump_case(n=4, params=Params(sorder=1, nsgs=2))
    pb = O.Problem.from_case(case)
    m = case.mesh
    q = pb.update_bcs(cases.perturbed_state(case, amp=0.1))
    dt, _ = pb.timestep(q, 1.0)
    ia, ja, iau, M_, qj = pb.jacobian(q, dt)
    A = _dense_from_crs(ia, ja, M_)
    A -= np.kron(np.diag(m.vol / dt), np.eye(5))
    rng = np.random.default_rng(6)
    v = rng.normal(size=(m.nnode, 5)) * np.abs(qj[: m.nnode, :5])
    eps = 1e-6
    b0, _ = pb.residual(pb.update_bcs(qj))
    qp = qj.copy()
    qp[: m.nnode, :5] += eps * v
    qp = cases.aux(qp, G)
    bp, _ = pb.residual(pb.update_bcs(qp))
    jv = -(bp - b0) / eps
    np.testing.assert_allclose(A @ v.reshape(-1), jv.reshape(-1), rtol=2e-4, atol=2e-4 * np.abs(jv).max())


def test_explicit_update_applies_the_correction_twice():
    """ExplicitSolve applies x = b dt/vol and NewtonIterate applies it again
    (ucs/solve.tcc:93-98, ucs/solutionSpace.tcc:802-804)."""
    case = cases.wedge_case(nx=12, ny=5, nz=2)
    pb = O.Problem.from_case(case)
    q0 = cases.perturbed_state(case, amp=0.01)
    dt, _ = pb.timestep(q0, case.params.cfl_start)
    qb = pb.update_bcs(q0)
    b, _ = pb.residual(qb)
    s = O.Solver(pb, q0)
    s.step()
    nn = case.mesh.nnode
    expect = qb[:nn, :5] + 2 * b * (dt / case.mesh.vol)[:, None]
    np.testing.assert_allclose(s.q[:nn, :5], expect, rtol=1e-12, atol=1e-14)
    s.close()
