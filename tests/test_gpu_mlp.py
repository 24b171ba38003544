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
    """Train on exact rates from the GPU chemistry, then evaluate new cells
    with and without the network.  Reported: device time of the exact
    rates, of the surrogate with its exact fallback, the fallback share and
    the error of the rates against the exact ones.  Checked: training
    converges, the fallback rule catches the surrogate's large errors (the
    cells it keeps are within the element-balance bound's reach), and the
    surrogate path never returns rates further than that from exact."""
    from proteuscfd_amd import chem, gpu, surrogate as S
    ctx = gpu.ChemContext(chem.load_air5())
    print("[mlp] surrogate training", flush=True)
    net, tr, hist = S.train_surrogate(ctx, n=1_000_000, steps=20000, batch=8192, lr=3e-3, seed=40)
    assert hist[-200:].mean() < 0.05 * hist[:50].mean()
    n = 1_000_000
    rhoi, T = S.sample_states(n, seed=41)
    d = S.desc(tr, tol=0.1, floor=10.0)
    runs_e = [S.exact_wdot(ctx, rhoi, T) for _ in range(3)]
    runs_s = [S.surrogate_wdot(ctx, net, d, rhoi, T) for _ in range(3)]
    we_h, ms_e = runs_e[-1][0], min(r[1] for r in runs_e)
    ws_h, nf, ms_s = runs_s[-1][0], runs_s[-1][1], min(r[2] for r in runs_s)
    rho = rhoi.sum(axis=1)
    err = np.abs(ws_h - we_h).max(axis=1) / (np.abs(we_h).max(axis=1) + rho * 1.0)
    raw = S.Transforms()
    print(f"[mlp] 1M cells: exact {ms_e:.3f} ms, surrogate+fallback {ms_s:.3f} ms (gain {ms_e / ms_s:.2f}x), "
          f"fallback {nf / n:.1%}, rate error median {np.median(err):.2e} p90 {np.quantile(err, 0.9):.2e} "
          f"p99 {np.quantile(err, 0.99):.2e}; final training loss {hist[-200:].mean():.2e}", flush=True)
    assert 0 < nf < 0.6 * n
    assert np.quantile(err, 0.9) < 0.5
    net.close()
    ctx.close()
