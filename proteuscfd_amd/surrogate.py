"""Production-rate surrogate for 5-species air: data, transforms, training.

The training set is made by the repository's own exact chemistry on the
GPU (pcfd_chem_wdot): T log-uniform in [300, 12000] K, rho log-uniform in
[1e-5, 1] kg/m^3, mass fractions from a symmetric Dirichlet(1) for half of
the states and partly dissociated air for the other half.  Inputs
are (ln T, ln rho, Y_1..Y_5) scaled to about [-1, 1]; targets are
asinh(wdot_i / (rho s0)) / sigma_i, which compresses the ~90 decades the
rates span across that box.  The network is trained on the GPU by the
library's own MFMA forward/backward kernels (pcfd_mlp_train).  At run
time pcfd_chem_surrogate_wdot_device evaluates it and sends every cell
outside the box, or whose predicted rates break N/O element conservation
by more than the bound, to the exact rates.  A second network of the same
shape, trained on the first one's own errors, predicts each cell's rate
error; cells whose predicted error exceeds err_tol of their rate scale go
to the exact rates as well (train_with_error_net, calibrate).
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field

import numpy as np

T_MIN, T_MAX = 300.0, 12000.0
RHO_MIN, RHO_MAX = 1e-5, 1.0


# air's element mass fractions and the molar masses (kg/mol) of O, N, NO
Y_O_AIR, Y_N_AIR = 0.233, 0.767
M_O, M_N = 0.0159994, 0.0140067


def air_compositions(rng, n, symbols=("O2", "O", "N", "N2", "NO")):
    """Mass fractions of partly dissociated air with air's N/O element
    ratio: NO takes a mass fraction log-uniform in [1e-10, 0.1], the rest of
    the O atoms a share a of them as O and the rest of the N atoms a share c
    as N (a, c log-uniform in [1e-10, 1])."""
    x_no = np.exp(rng.uniform(np.log(1e-10), np.log(0.1), n))
    a = np.exp(rng.uniform(np.log(1e-10), 0.0, n))
    c = np.exp(rng.uniform(np.log(1e-10), 0.0, n))
    o_left = Y_O_AIR - x_no * M_O / (M_O + M_N)
    n_left = Y_N_AIR - x_no * M_N / (M_O + M_N)
    y = {"O2": o_left * (1 - a), "O": o_left * a, "N2": n_left * (1 - c), "N": n_left * c, "NO": x_no}
    return np.stack([y[s] for s in symbols], axis=1)


def sample_states(n, seed, air_share=0.5):
    """T log-uniform in [T_MIN, T_MAX], rho log-uniform in [RHO_MIN,
    RHO_MAX]; mass fractions from a symmetric Dirichlet(1) for a share
    1 - air_share of the states and partly dissociated air (the compositions
    an air flow visits, which the Dirichlet draws leave at the simplex's
    corner) for the rest."""
    rng = np.random.default_rng(seed)
    T = np.exp(rng.uniform(np.log(T_MIN), np.log(T_MAX), n))
    rho = np.exp(rng.uniform(np.log(RHO_MIN), np.log(RHO_MAX), n))
    Y = rng.dirichlet(np.ones(5), n)
    na = int(round(air_share * n))
    if na:
        Y[:na] = air_compositions(rng, na)
        rng.shuffle(Y, axis=0)
    return rho[:, None] * Y, T


@dataclass
class Transforms:
    lnT_mid: float = 0.5 * (np.log(T_MIN) + np.log(T_MAX))
    lnT_half: float = 0.5 * (np.log(T_MAX) - np.log(T_MIN))
    lnrho_mid: float = 0.5 * (np.log(RHO_MIN) + np.log(RHO_MAX))
    lnrho_half: float = 0.5 * (np.log(RHO_MAX) - np.log(RHO_MIN))
    # mass fractions enter through ln(Y + 1e-4), mapped to about [-1, 1]:
    # the rates are products of concentration powers
    y_eps: float = 1e-4
    y_mid: float = 0.5 * np.log(1e-4)
    y_scale: float = -0.5 * np.log(1e-4)
    rate_scale: float = 1.0
    sigma: np.ndarray = field(default_factory=lambda: np.ones(5))

    def features(self, rhoi, T):
        rho = rhoi.sum(axis=1)
        x = np.empty((len(T), 7), dtype=np.float32)
        x[:, 0] = (np.log(T) - self.lnT_mid) / self.lnT_half
        x[:, 1] = (np.log(rho) - self.lnrho_mid) / self.lnrho_half
        Y = rhoi / rho[:, None]
        g = np.log(Y + self.y_eps) if self.y_eps > 0 else Y
        x[:, 2:] = (g - self.y_mid) / self.y_scale
        return x

    def targets(self, wdot, rhoi):
        rho = rhoi.sum(axis=1)
        return (np.arcsinh(wdot / (rho[:, None] * self.rate_scale)) / self.sigma).astype(np.float32)

    def rates(self, y, rhoi):
        rho = rhoi.sum(axis=1)
        return rho[:, None] * self.rate_scale * np.sinh(y * self.sigma)

    def fit_sigma(self, wdot, rhoi):
        rho = rhoi.sum(axis=1)
        self.sigma = np.maximum(np.arcsinh(wdot / (rho[:, None] * self.rate_scale)).std(axis=0), 1e-3)


def atom_counts(symbols, elements=("N", "O")):
    """Atoms of each element in each species symbol (e.g. "N2" -> N: 2)."""
    out = np.zeros((len(elements), len(symbols)))
    for k, sym in enumerate(symbols):
        for el, n in re.findall(r"([A-Z][a-z]?)(\d*)", sym):
            if el in elements:
                out[elements.index(el), k] += int(n) if n else 1
    return out


def desc(tr: Transforms, tol=0.05, floor=10.0, symbols=None, err_tol=0.0, rate_floor=1.0, err_mid=None,
         err_scale=None):
    from . import chem, gpu
    d = gpu.SurrogateDesc()
    atoms = atom_counts(symbols if symbols is not None else chem.load_air5().symbols)
    for e in range(2):
        for k in range(5):
            d.atoms[e][k] = float(atoms[e, k])
    d.lnT_mid, d.lnT_half, d.lnrho_mid, d.lnrho_half = tr.lnT_mid, tr.lnT_half, tr.lnrho_mid, tr.lnrho_half
    d.y_mid, d.y_scale, d.y_eps, d.rate_scale = tr.y_mid, tr.y_scale, tr.y_eps, tr.rate_scale
    for k in range(5):
        d.sigma[k] = float(tr.sigma[k])
    d.T_min, d.T_max, d.rho_min, d.rho_max = T_MIN, T_MAX, RHO_MIN, RHO_MAX
    d.tol, d.floor = tol, floor
    d.err_tol, d.rate_floor = err_tol, rate_floor
    for k in range(5):
        d.err_mid[k] = float(err_mid[k]) if err_mid is not None else 0.0
        d.err_scale[k] = float(err_scale[k]) if err_scale is not None else 1.0
    return d


def train_surrogate(chem_ctx, n=2_000_000, steps=4000, batch=4096, lr=3e-3, seed=40):
    """Exact rates on the GPU -> transforms -> MLP trained on the GPU."""
    from . import gpu
    rhoi, T = sample_states(n, seed)
    wdot = chem_ctx.wdot(rhoi, T)
    tr = Transforms()
    tr.fit_sigma(wdot, rhoi)
    x, t = tr.features(rhoi, T), tr.targets(wdot, rhoi)
    net = gpu.MLP(seed=seed)
    hist = net.train(x, t, steps, batch, lr, seed=seed + 1)
    return net, tr, hist


def train_with_error_net(chem_ctx, n=2_000_000, steps=60000, batch=8192, lr=3e-3, err_steps=20000, seed=40):
    """The rate network, then its error network: trained on the same exact-
    rate samples to predict g_i = ln(|sigma_i (y_i - target_i)| + 1e-3), the
    rate network's own error per output in asinh units (normalised per
    output).  The rate network carries the error network as its partner.
    Returns (rate net, error net, transforms, err_mid, err_scale, loss
    histories)."""
    from . import gpu
    rhoi, T = sample_states(n, seed)
    wdot = chem_ctx.wdot(rhoi, T)
    tr = Transforms()
    tr.fit_sigma(wdot, rhoi)
    x, t = tr.features(rhoi, T), tr.targets(wdot, rhoi)
    net = gpu.MLP(seed=seed)
    h1 = net.train(x, t, steps, batch, lr, seed=seed + 1)
    g = np.log(np.abs(tr.sigma[None, :] * (net.forward(x).astype(np.float64) - t)) + 1e-3)
    gm, gs = g.mean(axis=0), g.std(axis=0)
    err = gpu.MLP(seed=seed + 37)
    h2 = err.train(x, ((g - gm) / gs).astype(np.float32), err_steps, batch, lr, seed=seed + 38)
    net.set_partner(err)
    return net, err, tr, gm, gs, (h1, h2)


def fallback_rule(tr: Transforms, d, mw, atoms, rhoi, T, ya, yb=None):
    """The kernel's fallback rule restated on the host from the rate
    network's outputs ya [n, 5] and the error network's yb: returns (rates,
    fallback mask, predicted rate error)."""
    rho = rhoi.sum(axis=1)
    w = tr.rates(ya.astype(np.float64), rhoi)
    m = w / mw
    imb = np.abs(m @ atoms.T).sum(axis=1)
    out = ~((T >= d.T_min) & (T <= d.T_max) & (rho >= d.rho_min) & (rho <= d.rho_max))
    bad = imb > d.tol * (np.abs(m).sum(axis=1) + rho * d.floor)
    werr = np.zeros(len(T))
    if yb is not None:
        e = np.exp(yb.astype(np.float64) * np.array(d.err_scale[:]) + np.array(d.err_mid[:]))
        z = np.abs(tr.sigma[None, :] * ya.astype(np.float64))
        with np.errstate(over="ignore"):
            werr = (rho[:, None] * tr.rate_scale * (np.sinh(z + e) - np.sinh(z))).max(axis=1)
        bad |= ~(werr <= d.err_tol * (np.abs(w).max(axis=1) + rho * d.rate_floor))
    return w, out | bad, werr


def rate_error(w, w_exact, rho, rate_floor=1.0):
    """Per cell: max_i |w_i - exact_i| / (max_i |exact_i| + rho * rate_floor)."""
    return np.abs(w - w_exact).max(axis=1) / (np.abs(w_exact).max(axis=1) + rho * rate_floor)


def calibrate(chem_ctx, net, err_net, tr, d, mw, atoms, n=1_000_000, seed=41, target_p99=0.05,
              grid=(0.001, 0.003, 0.01, 0.02, 0.03, 0.05, 0.1, 0.2, 0.3, 1.0)):
    """The error / fallback curve on held-out states: for each err_tol of the
    grid, the fallback share and the rate error (exact rates judging,
    d.rate_floor in its floor) over the cells the surrogate keeps and over all
    cells after fallback (those sent back are exact).  Returns (the largest
    err_tol whose p99 over all cells is <= target_p99 with some cells kept, or
    None; the curve)."""
    rhoi, T = sample_states(n, seed)
    we = chem_ctx.wdot(rhoi, T)
    x = tr.features(rhoi, T)
    ya, yb = net.forward(x), err_net.forward(x)
    rho = rhoi.sum(axis=1)
    curve, best = [], None
    for et in grid:
        d.err_tol = et
        w, fb, _ = fallback_rule(tr, d, mw, atoms, rhoi, T, ya, yb)
        e = np.where(fb, 0.0, rate_error(w, we, rho, d.rate_floor))
        ek = e[~fb] if (~fb).any() else np.zeros(1)
        row = dict(err_tol=et, fallback=float(fb.mean()), kept_p90=float(np.quantile(ek, 0.9)),
                   kept_p99=float(np.quantile(ek, 0.99)), p90=float(np.quantile(e, 0.9)),
                   p99=float(np.quantile(e, 0.99)), max=float(e.max()))
        curve.append(row)
        if row["p99"] <= target_p99 and row["fallback"] < 1.0:
            best = et
    return best, curve


def surrogate_wdot_device(chem_ctx, net, d, rhoi_ptr, T_ptr, wdot_ptr, n):
    """Device pointers in and out; returns (fallback count, device ms)."""
    from . import gpu
    nf = C.c_int()
    ms = C.c_double()
    L = gpu._mlp_lib()
    rc = L.pcfd_chem_surrogate_wdot_device(chem_ctx.h, net.h, C.byref(d), n, rhoi_ptr, T_ptr, wdot_ptr,
                                           C.byref(nf), C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"pcfd_chem_surrogate_wdot_device failed ({rc}): {L.pcfd_mlp_last_error(net.h).decode()}")
    return nf.value, ms.value


def surrogate_wdot(chem_ctx, net, d, rhoi, T):
    """Host arrays; returns (wdot, fallback count, device ms of the kernels)."""
    from . import gpu
    rhoi = np.ascontiguousarray(rhoi, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    w = np.zeros_like(rhoi)
    nf = C.c_int()
    ms = C.c_double()
    L = gpu._mlp_lib()
    dp = C.POINTER(C.c_double)
    rc = L.pcfd_chem_surrogate_wdot(chem_ctx.h, net.h, C.byref(d), len(T), rhoi.ctypes.data_as(dp),
                                    T.ctypes.data_as(dp), w.ctypes.data_as(dp), C.byref(nf), C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"pcfd_chem_surrogate_wdot failed ({rc}): {L.pcfd_mlp_last_error(net.h).decode()}")
    return w, nf.value, ms.value


def exact_wdot(chem_ctx, rhoi, T):
    """Host arrays; returns (wdot, device ms of the exact kernel)."""
    from . import gpu
    w = chem_ctx.wdot(rhoi, T)
    return w, gpu._mlp_lib().pcfd_chem_last_kernel_ms(chem_ctx.h)


def exact_wdot_device(chem_ctx, rhoi_ptr, T_ptr, wdot_ptr, n):
    from . import gpu
    L = gpu._mlp_lib()
    rc = L.pcfd_chem_wdot_device(chem_ctx.h, n, rhoi_ptr, T_ptr, wdot_ptr)
    if rc != 0:
        raise RuntimeError("pcfd_chem_wdot_device failed")
    return L.pcfd_chem_last_kernel_ms(chem_ctx.h)
