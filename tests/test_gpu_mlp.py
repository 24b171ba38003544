"""The chemistry surrogate on the GPU: the MFMA forward and backward
kernels against the CPU restatement (oracle/mlp.py), training by the
library's own kernels, and what the trained network gains over exact
chemistry on the same cells (time, with its error after fallback)."""
import numpy as np
import pytest

from oracle import mlp as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b))


def test_forward_matches_oracle():
    from proteuscfd_amd import gpu
    net = gpu.MLP(seed=0)
    p = R.init(7).astype(np.float32)
    p[R.unpack(np.arange(R.NPARAM))["b1"].astype(int)] = 0.05
    net.params = p
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (1000, 7)).astype(np.float32)  # not a multiple of 16: ragged last tile
    print("[mlp] forward n=1000", flush=True)
    yg = net.forward(x)
    yo = R.forward(p, x)
    assert rel(yg, yo) < 2e-3
    assert np.abs(yg - yo).max() < 1e-2 * np.abs(yo).max()
    # and against the unrounded mathematics, at BF16 accuracy
    assert rel(yg, R.forward(p, x, exact=True)) < 3e-2
    net.close()


def test_gradient_matches_oracle():
    from proteuscfd_amd import gpu
    net = gpu.MLP(seed=0)
    p = R.init(9).astype(np.float32)
    net.params = p
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (777, 7)).astype(np.float32)
    t = rng.normal(size=(777, 5)).astype(np.float32)
    print("[mlp] loss_grad n=777", flush=True)
    lg, gg = net.loss_grad(x, t)
    lo, go = R.loss_grad(p, x, t)
    assert abs(lg - lo) <= 1e-4 * lo
    for name in ["W1", "b1", "W2", "b2", "W3", "b3"]:
        a, b = R.unpack(gg)[name], R.unpack(go)[name]
        assert rel(a, b) < 1e-2, name
    assert np.all(R.unpack(gg)["W1"][:, R.NIN:] == 0) and np.all(R.unpack(gg)["W3"][R.NOUT:] == 0)
    net.close()


def test_training_fits_a_smooth_map():
    from proteuscfd_amd import gpu
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (50000, 7)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], np.tanh(x[:, 2] - x[:, 3]), x[:, 4] ** 2 - 0.3,
                  np.cos(x[:, 5] + x[:, 6]), 0.5 * x[:, 0] * x[:, 6]], axis=1).astype(np.float32)
    net = gpu.MLP(seed=4)
    l0, _ = net.loss_grad(x, t)
    print("[mlp] train 1500 steps", flush=True)
    hist = net.train(x, t, steps=1500, batch=1024, lr=3e-3, seed=5)
    l1, _ = net.loss_grad(x, t)
    assert np.isfinite(hist).all()
    assert l1 < 0.02 * l0, (l0, l1)
    net.close()


def test_surrogate_gain_over_exact_chemistry():
    """The rate network (60000 steps on 2M exact-rate samples: T
    log-uniform in [300, 12000] K, rho log-uniform in [1e-5, 1] kg/m^3,
    mass fractions half Dirichlet, half partly dissociated air) and its error network (trained on the rate
    network's own errors) form the surrogate.  Two error scales: max_i
    |wdot_i| + rho * 1/s, and + rho / t_ref with t_ref = 1/350 s (the flow
    time of the Mach 8 wedge, below which a rate cannot move the mass
    fractions within a flow time).  For each, err_tol is calibrated on 1M
    held-out states so that the p99 error over all cells after fallback is
    at most 5% (those sent back are exact); on another 1M states the device
    path must then fall back on the cells the host restatement of its rule
    names (to 1% of the count) and meet the stated 5%.  The estimate must
    rank the errors (cells kept at err_tol 0.03: at most half the p90 error
    of those kept at 1).  Reported: the error / fallback curves and the
    device time beside the exact rates -- at the 5% bound most cells go back
    to the exact rates (the rate network's median error is ~10%), so there
    the surrogate costs more than exact chemistry."""
    from proteuscfd_amd import chem, gpu, surrogate as S
    model = chem.load_air5()
    ctx = gpu.ChemContext(model)
    print("[mlp] rate network (60000 steps) and error network (20000 steps) on 2M samples", flush=True)
    net, enet, tr, gm, gs, (h1, h2) = S.train_with_error_net(ctx, n=2_000_000)
    assert h1[-500:].mean() < 0.05 * h1[:50].mean()
    assert h2[-500:].mean() < h2[:50].mean()
    atoms = S.atom_counts(model.symbols)
    n = 1_000_000
    rhoi, T = S.sample_states(n, seed=42)
    rho = rhoi.sum(axis=1)
    runs_e = [S.exact_wdot(ctx, rhoi, T) for _ in range(3)]
    we, ms_e = runs_e[-1][0], min(r[1] for r in runs_e)
    x = tr.features(rhoi, T)
    ya, yb = net.forward(x), enet.forward(x)
    for floor in (1.0, 350.0):
        d = S.desc(tr, tol=0.1, floor=max(10.0, floor / 0.0289), rate_floor=floor, err_mid=gm, err_scale=gs)
        best, curve = S.calibrate(ctx, net, enet, tr, d, model.mw, atoms, n=1_000_000, seed=41, target_p99=0.05)
        for row in curve:
            print("[mlp] floor {fl:g}/s err_tol {err_tol:<6g} fallback {fallback:6.1%}  kept p90 {kept_p90:.2e} "
                  "p99 {kept_p99:.2e}  all p90 {p90:.2e} p99 {p99:.2e}".format(fl=floor, **row), flush=True)
        assert best is not None
        # the estimate ranks the errors: the cells kept at err_tol 0.03 have
        # at most half the p90 error of the cells kept without it
        loose = next(r for r in curve if r["err_tol"] == 1.0)
        tight = next(r for r in curve if r["err_tol"] == 0.03)
        assert tight["kept_p90"] <= 0.5 * loose["kept_p90"]
        d.err_tol = best
        runs_s = [S.surrogate_wdot(ctx, net, d, rhoi, T) for _ in range(3)]
        ws, nf, ms_s = runs_s[-1][0], runs_s[-1][1], min(r[2] for r in runs_s)
        _, fb_host, _ = S.fallback_rule(tr, d, model.mw, atoms, rhoi, T, ya, yb)
        err = S.rate_error(ws, we, rho, floor)
        print(f"[mlp] floor {floor:g}/s, err_tol {best:g}: 1M cells exact {ms_e:.3f} ms, surrogate+fallback "
              f"{ms_s:.3f} ms (gain {ms_e / ms_s:.2f}x), fallback {nf / n:.1%} (host rule {fb_host.mean():.1%}), "
              f"error p90 {np.quantile(err, 0.9):.2e} p99 {np.quantile(err, 0.99):.2e}", flush=True)
        assert abs(nf - int(fb_host.sum())) <= 0.01 * max(nf, 1000)
        assert np.quantile(err, 0.99) <= 0.05
    net.set_partner(None)
    d1 = S.desc(tr, tol=0.1, floor=10.0)
    w1, nf1, ms_1 = S.surrogate_wdot(ctx, net, d1, rhoi, T)
    err1 = S.rate_error(w1, we, rho)
    print(f"[mlp] rate network with the element-balance rule only: {ms_1:.3f} ms, fallback {nf1 / n:.1%}, "
          f"error p90 {np.quantile(err1, 0.9):.2e} p99 {np.quantile(err1, 0.99):.2e}", flush=True)
    enet.close()
    net.close()
    ctx.close()
