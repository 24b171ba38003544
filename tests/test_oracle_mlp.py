"""Pins for the surrogate network's CPU restatement (oracle/mlp.py): its
gradient against central differences of its own loss in FP64, and the
BF16-rounded form close to the exact mathematics."""
import numpy as np

from oracle import mlp as R


def data(n=64, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, R.NIN))
    t = rng.normal(size=(n, R.NOUT))
    return x, t


def test_gradient_matches_central_differences():
    p = R.init(1).astype(np.float64)
    x, t = data()
    loss, g = R.loss_grad(p, x, t, exact=True)
    rng = np.random.default_rng(5)
    W = R.unpack(np.arange(R.NPARAM, dtype=np.float64))
    idx = np.concatenate([rng.choice(W["W1"][:, :R.NIN].ravel().astype(int), 10, replace=False),
                          rng.choice(W["W2"].ravel().astype(int), 10, replace=False),
                          rng.choice(W["W3"][:R.NOUT].ravel().astype(int), 10, replace=False),
                          W["b1"][:5].astype(int), W["b3"][:R.NOUT].astype(int)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_grad(pp, x, t, exact=True)[0] - R.loss_grad(pm, x, t, exact=True)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert np.all(g[W["W1"][:, R.NIN:].ravel().astype(int)] == 0)


def test_bf16_form_close_to_exact():
    p = R.init(2)
    x, t = data(256, 4)
    ye = R.forward(p, x, exact=True)
    yb = R.forward(p, x)
    assert np.abs(yb - ye).max() < 3e-2 * np.abs(ye).max()
    le, ge = R.loss_grad(p, x, t, exact=True)
    lb, gb = R.loss_grad(p, x, t)
    assert abs(lb - le) < 1e-2 * le
    assert np.linalg.norm(gb - ge) < 5e-2 * np.linalg.norm(ge)


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39], dtype=np.float32)
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
