"""The HIP path against golden vectors written by the reference solver itself
(tests/golden/ref_*.npz, made by tests/golden/make_ref_vectors.py): every
phase of an implicit step on the reference's own arrays, the pressure clip
with chains in the reference's edge order, and the converged steady field of
both precision forms against the reference's natural-order run."""
import os

import numpy as np
import pytest

from proteuscfd_amd import cases
from proteuscfd_amd.cases import Params

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, f"ref_{name}.npz")))


def rel(a, b):
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300))


def with_ref_gas(case, d, **kw):
    """The case with the reference run's gamma and freestream."""
    case.params = Params(gamma=float(d["gamma"]), venkat_K=1.0, **kw)
    case.qref = np.tile(d["qinf"], (len(case.bc_type), 1))
    case.qinf = d["qinf"].copy()
    return case


PHASE_CASES = {
    "bump5_phases": (lambda: cases.bump_case(n=5, rcm=False),
                     dict(sorder=2, limiter=3, cfl=50.0, cfl_start=1.0, cfl_ramp=50, nsgs=8)),
    "wedge_phases": (lambda: cases.wedge_case(nx=8, ny=4, nz=2, rcm=False),
                     dict(sorder=2, limiter=2, cfl=2.0, cfl_start=2.0, nsgs=4)),
}


@pytest.mark.parametrize("name", list(PHASE_CASES))
def test_phases_match_reference(name):
    from proteuscfd_amd import gpu
    d = load(name)
    build, kw = PHASE_CASES[name]
    case = with_ref_gas(build(), d, **kw)
    nn = case.mesh.nnode
    print(f"[ref] {name} nnode={nn}", flush=True)
    ctx = gpu.Context(case)
    ctx.upload_q(d["q_in"])
    ctx.update_bcs()
    assert rel(ctx.download_q(), d["q_bc"]) <= 1e-13
    ctx.upload_q(d["q_bc"])
    assert rel(ctx.gradient(), d["grad"]) <= 1e-12
    assert rel(ctx.limiter(), d["limiter"]) <= 1e-12
    ctx.set_limiter(d["limiter"])
    b, norms = ctx.residual(it=int(d["iter"]))
    assert np.linalg.norm(b - d["residual"]) <= 1e-12 * np.linalg.norm(d["residual"])
    assert rel(norms, d["residual_norms"]) <= 1e-12
    dt, dtmin = ctx.timestep(float(d["cfl"]))
    assert rel(dt, d["dt"]) <= 1e-13
    ctx.set_timestep(d["dt"])
    ctx.jacobian()
    M = ctx.download_jacobian(d["ia"], d["ja"])
    ia = d["ia"]
    for i in range(nn):
        scale = np.abs(d["jacobian"][ia[i]:ia[i + 1]]).max()
        assert np.abs(M[ia[i]:ia[i + 1]] - d["jacobian"][ia[i]:ia[i + 1]]).max() <= 1e-6 * scale, i
    ctx.close()


@pytest.mark.parametrize("lim", [1, 2, 3])
def test_pressure_clip_chains_match_reference(lim):
    """The limiter field after the reference's sequential pressure clip
    (ucs/limiters.tcc:736-815) on a state where the clip zeroes nodes in
    chains: equal node for node, zeroed set included."""
    from proteuscfd_amd import gpu
    d = load("bump6_clip")
    case = with_ref_gas(cases.bump_case(n=6, rcm=False), d, sorder=2, limiter=lim)
    print(f"[ref] clip limiter={lim}", flush=True)
    ctx = gpu.Context(case)
    ctx.upload_q(d["q_bc"])
    g = ctx.gradient()
    assert rel(g, d["grad"]) <= 1e-12
    lg = ctx.limiter()
    ref = d[f"limiter_{lim}"]
    np.testing.assert_array_equal((lg == 0).all(axis=1), (ref == 0).all(axis=1))
    assert (ref == 0).all(axis=1).sum() >= 5
    assert rel(lg, ref) <= 1e-12
    ctx.set_limiter(ref)
    b, _ = ctx.residual(it=5)
    assert np.linalg.norm(b - d[f"residual_{lim}"]) <= 1e-12 * np.linalg.norm(d[f"residual_{lim}"])
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_converged_field_matches_reference(precision):
    """Both precision forms from freestream to an 8-order drop on the 6^3
    bump (multicolour SGS), against the reference's own converged field
    (natural-order SGS): per-variable error <= 1e-6 relative, and the
    number of Newton iterations within a few of the reference's."""
    from proteuscfd_amd import gpu
    d = load("bump6_solve")
    case = with_ref_gas(cases.bump_case(n=6, rcm=False), d, sorder=2, limiter=3, cfl=20.0, cfl_start=1.0,
                        cfl_ramp=20, nsgs=8, precision=precision)
    print(f"[ref] converged bump 6^3 precision={precision}", flush=True)
    ctx = gpu.Context(case)
    ctx.upload_q(case.initial_state())
    ctx.reset_iteration()
    hist, nit, _ = ctx.solve(400, 1e-8)
    q = ctx.download_q()[:case.mesh.nnode, :5]
    qr = d["q_final"][:case.mesh.nnode, :5]
    err = [float(np.abs(q[:, k] - qr[:, k]).max() / np.abs(qr[:, k]).max()) for k in range(5)]
    print(f"[ref] iterations gpu {nit} reference {len(d['history'])}, field error per variable {err}", flush=True)
    assert hist[-1] <= 1e-8 * hist[0]
    assert abs(nit - len(d["history"])) <= 5
    assert max(err) <= 1e-6
    ctx.close()


def plate(twall, nx=12, ny=3, nz=10, nlayers=4):
    return cases.plate_case(nx=nx, ny=ny, nz=nz, nlayers=nlayers, first_h=0.01, growth=1.3, twall=twall,
                            incidence_deg=1.0, outflow_bc=cases.BC_SONIC_OUTFLOW)


def with_ref_ns(case, d, pre="", **kw):
    case.params = Params(gamma=float(d[pre + "gamma"]), viscous=1, reynolds=float(d[pre + "Re"]), mach_ref=0.3,
                         sutherland_s=110.4 / float(d[pre + "ref_temperature"]), **kw)
    case.qref = np.tile(d[pre + "qinf"], (len(case.bc_type), 1))
    case.qinf = d[pre + "qinf"].copy()
    return case


@pytest.mark.parametrize("tag,twall", [("iso", 1.0), ("adi", -1.0)])
def test_viscous_phases_match_reference(tag, twall):
    from proteuscfd_amd import gpu
    d = load("plate_phases")
    g = lambda k: d[f"{tag}_{k}"]
    case = with_ref_ns(plate(twall, nx=6, ny=2, nz=6, nlayers=3), d, pre=f"{tag}_", sorder=2, limiter=0, cfl=10.0,
                       cfl_start=10.0, nsgs=4)
    nn = case.mesh.nnode
    print(f"[ref] plate {tag} nnode={nn}", flush=True)
    ctx = gpu.Context(case)
    ctx.upload_q(g("q_in"))
    ctx.update_bcs()
    assert rel(ctx.download_q(), g("q_bc")) <= 1e-13
    ctx.upload_q(g("q_bc"))
    assert rel(ctx.gradient(), g("grad")) <= 1e-12
    ctx.set_limiter(np.ones((nn, 5)))
    b, _ = ctx.residual(it=int(g("iter")))
    assert np.linalg.norm(b - g("residual")) <= 1e-12 * np.linalg.norm(g("residual"))
    dt, _ = ctx.timestep(float(g("cfl")))
    assert rel(dt, g("dt")) <= 1e-13
    ctx.set_timestep(g("dt"))
    ctx.jacobian()
    M = ctx.download_jacobian(g("ia"), g("ja"))
    ia, Mr = g("ia"), g("jacobian")
    for i in range(nn):
        scale = np.abs(Mr[ia[i]:ia[i + 1]]).max()
        assert np.abs(M[ia[i]:ia[i + 1]] - Mr[ia[i]:ia[i + 1]]).max() <= 1e-6 * scale, i
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_viscous_converged_field_matches_reference(precision):
    """Laminar NS plate channel, both precision forms, to an 8-order drop,
    against the reference's converged field: <= 1e-6 per variable."""
    from proteuscfd_amd import gpu
    d = load("plate_solve")
    case = with_ref_ns(plate(1.0), d, sorder=2, limiter=0, cfl=50.0, cfl_start=1.0, cfl_ramp=30, nsgs=8,
                       precision=precision)
    print(f"[ref] converged plate precision={precision}", flush=True)
    ctx = gpu.Context(case)
    ctx.upload_q(case.initial_state())
    ctx.reset_iteration()
    hist, nit, _ = ctx.solve(2000, 1e-8)
    q = ctx.download_q()[:case.mesh.nnode, :5]
    qr = d["q_final"][:case.mesh.nnode, :5]
    err = [float(np.abs(q[:, k] - qr[:, k]).max() / np.abs(qr[:, k]).max()) for k in range(5)]
    print(f"[ref] plate iterations gpu {nit} reference {len(d['history'])}, field error per variable {err}",
          flush=True)
    assert hist[-1] <= 1e-8 * hist[0]
    assert max(err) <= 1e-6
    ctx.close()


@pytest.mark.parametrize("precision,jac_freq", [(0, 1), (1, 1), (1, 4)])
def test_bench_plate_boundary_set_converges_to_a_reference_steady_state(precision, jac_freq):
    """The bench plate's boundary set on a small channel (no incidence,
    farfield in- and outflow, symmetry top and sides, isothermal floor), in
    the plain form, the chosen form and the chosen form with the bench's
    Jacobian refresh every fourth iteration, to an 8-order drop.  The
    farfield state at the floor's outflow corner is degenerate there (any
    phantom with zero normal velocity is kept, ucs/compressible.tcc:
    1318-1321), so the steady state is not unique and is not compared with
    the reference's field; instead the device's converged state (with its
    phantoms) must be a steady state of the oracle's restatement of the
    reference's equations: after the oracle's boundary update, residGlobal
    <= 2e-8 of the reference's first residual (the drop criterion, with a
    factor 2 for the two evaluations' rounding)."""
    import oracle as O
    from proteuscfd_amd import gpu
    d = load("plate_bench_solve")
    case = with_ref_ns(cases.plate_case(nx=12, ny=3, nz=10, nlayers=4, first_h=0.01, growth=1.3, twall=1.0,
                                        incidence_deg=0.0, outflow_bc=cases.BC_FARFIELD,
                                        top_bc=cases.BC_SYMMETRY), d, sorder=2, limiter=0, cfl=50.0, cfl_start=1.0,
                       cfl_ramp=30, nsgs=8, precision=precision, jac_freq=jac_freq)
    m = case.mesh
    print(f"[ref] bench plate set precision={precision} jac_freq={jac_freq}", flush=True)
    ctx = gpu.Context(case)
    ctx.upload_q(case.initial_state())
    ctx.reset_iteration()
    hist, nit, _ = ctx.solve(2000, 1e-8)
    q = ctx.download_q()
    ctx.close()
    assert hist[-1] <= 1e-8 * hist[0]
    assert nit <= 1.2 * len(d["history"])
    pb = O.Problem(m, case.bc_type, case.qref, case.params, twall=case.twall)
    qb = pb.update_bcs(q)
    _, nrm = pb.residual(qb, pb.gradient(qb), np.ones((m.nnode, 5)), it=nit)
    print(f"[ref] iterations gpu {nit} reference {len(d['history'])}; oracle residual at the device's state "
          f"{nrm[0]:.3e} (reference first residual {d['history'][0]:.3e})", flush=True)
    assert nrm[0] <= 2e-8 * d["history"][0]


def test_failure_paths_match_reference():
    """Negative pressure / negative density nodes (NaN Roe fluxes zeroed
    component-wise, ApplyDQ's density clip) and a NaN node (every correction
    row NaN, zeroed and counted): the GPU against the reference's arrays
    where they do not depend on the SGS row order, and against the oracle run
    in the GPU's colour order where they do."""
    import oracle as O
    from proteuscfd_amd import gpu
    d = load("bump5_failure")
    for tag in ("neg", "nan"):
        g = lambda k: d[f"{tag}_{k}"]
        case = with_ref_gas(cases.bump_case(n=5, rcm=False), {"gamma": g("gamma"), "qinf": g("qinf")}, sorder=1,
                            limiter=0, cfl=5.0, cfl_start=5.0, nsgs=4)
        nn = case.mesh.nnode
        print(f"[ref] failure {tag}", flush=True)
        ctx = gpu.Context(case)
        ctx.upload_q(g("q_in"))
        ctx.update_bcs()
        qbc = ctx.download_q()
        np.testing.assert_array_equal(np.isnan(qbc), np.isnan(g("q_bc")))
        assert rel(np.nan_to_num(qbc), np.nan_to_num(g("q_bc"))) <= 1e-13
        b, _ = ctx.residual(it=1)
        assert np.isfinite(b).all()
        assert np.linalg.norm(b - g("residual")) <= 1e-12 * np.linalg.norm(g("residual"))
        ctx.set_timestep(g("dt"))
        ctx.jacobian()
        if tag == "neg":
            ia, ja = g("ia"), g("ja")
            M = ctx.download_jacobian(ia, ja)
            assert np.isfinite(M).all()
            for i in range(nn):
                scale = np.abs(g("jacobian")[ia[i]:ia[i + 1]]).max()
                assert np.abs(M[ia[i]:ia[i + 1]] - g("jacobian")[ia[i]:ia[i + 1]]).max() <= 1e-6 * scale, i
        ctx.prepare_sgs()
        x, _ = ctx.sgs(int(g("nsgs")))
        nz = ctx.apply_dq()
        q = ctx.download_q()[:nn]
        if tag == "neg":
            xo = O.sgs_on_matrix(ia, ja, ia[:-1].copy(), M, b, int(g("nsgs")), order=ctx.sgs_order())
            assert rel(x, xo) <= 1e-11
            assert nz == 0
            assert q[62, 0] == 1e-10 and q[100, 0] == 1e-10
        else:
            assert nz == int(g("zerocount")) == nn
            np.testing.assert_array_equal(np.isnan(q[:, :5]), np.isnan(g("q_new")[:nn, :5]))
        ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_rerun_is_bitwise_identical(precision):
    """Two runs of 20 Newton iterations from the same state, with the
    Jacobian rows on their own stream beside the edge passes: the states and
    residual histories agree bit for bit (node sums are gathers in a fixed
    order, the scans have a fixed order, there are no floating-point
    atomics)."""
    from proteuscfd_amd import gpu
    case = cases.bump_case(n=24, params=Params(sorder=2, limiter=3, cfl=20.0, cfl_start=1.0, cfl_ramp=10, nsgs=8,
                                              precision=precision))
    print(f"[ref] determinism precision={precision} nnode={case.mesh.nnode}", flush=True)
    runs = []
    for _ in range(2):
        ctx = gpu.Context(case)
        ctx.upload_q(case.initial_state())
        ctx.reset_iteration()
        hist = np.array([ctx.iterate()[0] for _ in range(20)])
        runs.append((hist, ctx.download_q()))
        ctx.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][0][-1] < runs[0][0][0]
