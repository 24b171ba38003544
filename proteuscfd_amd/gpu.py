"""ctypes binding of libpcfd.so (include/pcfd.h) -- the HIP path.

There is no CPU fallback: if the library (or a GPU) is missing every call
raises.  Arrays cross the boundary in the reference layouts (AoS q with
nvars per node, qgrad[node*27+term*3+dir], l/b/x[node*5+k]).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .cases import Case, Params

HERE = os.path.dirname(os.path.abspath(__file__))
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

PLAIN_FP64, CHOSEN_FP32 = 0, 1

# pcfd_phase_times slots (PCFD_NPHASE); the last three are kernel-only times
PHASES = ["bcs", "gradient", "limiter", "residual", "timestep", "jacobian", "prepare_sgs", "sgs", "update",
          "k_sgs_sweeps", "k_residual", "k_jacobian"]


class MeshDesc(C.Structure):
    _fields_ = [("nnode", C.c_int), ("nbnode", C.c_int), ("nedge", C.c_int), ("edges", ip), ("edge_a", dp),
                ("bedges", ip), ("bedge_a", dp), ("bfactag", ip), ("xyz", dp), ("vol", dp), ("nfactag", C.c_int),
                ("bc_type", ip), ("qref", dp), ("twall", dp)]


class CParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("sorder", C.c_int), ("nfirst", C.c_int), ("chi", C.c_double),
                ("limiter", C.c_int), ("venkat_K", C.c_double), ("limiter_refresh", C.c_int),
                ("limiter_freeze", C.c_int), ("cfl", C.c_double), ("cfl_start", C.c_double), ("cfl_ramp", C.c_int),
                ("nsgs", C.c_int), ("symmetry2d", C.c_int), ("jac_freq", C.c_int), ("precision", C.c_int),
                ("gradient_terms", C.c_int), ("viscous", C.c_int), ("re_tilde", C.c_double),
                ("prandtl", C.c_double), ("prandtl_t", C.c_double), ("sutherland_s", C.c_double),
                ("gradient_type", C.c_int), ("global_dt", C.c_int), ("fixed_dt", C.c_double),
                ("enable_vnn", C.c_int), ("vnn", C.c_double), ("dynamic_cfl", C.c_int),
                ("jac_field_type", C.c_int), ("flux_type", C.c_int), ("jac_boundary_type", C.c_int),
                ("rk_stages", C.c_int)]


SYMBOLS = [
    "pcfd_last_error", "pcfd_version", "pcfd_create", "pcfd_destroy", "pcfd_set_params", "pcfd_info",
    "pcfd_upload_q", "pcfd_download_q", "pcfd_update_bcs", "pcfd_gradient", "pcfd_limiter", "pcfd_set_limiter",
    "pcfd_residual", "pcfd_timestep", "pcfd_set_timestep", "pcfd_jacobian", "pcfd_download_jacobian",
    "pcfd_prepare_sgs", "pcfd_sgs", "pcfd_sgs_order", "pcfd_apply_dq", "pcfd_explicit_update", "pcfd_iterate",
    "pcfd_reset_iteration", "pcfd_get_iteration", "pcfd_solve", "pcfd_phase_times", "pcfd_enable_phase_timing",
    "pcfd_synchronize",
]

_lib = None


def lib_path() -> str:
    return os.path.join(HERE, "libpcfd.so")


def lib():
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: the HIP path is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.pcfd_last_error.argtypes = [vp]
        L.pcfd_last_error.restype = C.c_char_p
        L.pcfd_version.restype = C.c_char_p
        L.pcfd_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
        L.pcfd_destroy.argtypes = [vp]
        L.pcfd_set_params.argtypes = [vp, C.POINTER(CParams)]
        L.pcfd_info.argtypes = [vp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.pcfd_upload_q.argtypes = [vp, dp, C.c_int]
        L.pcfd_download_q.argtypes = [vp, dp, C.c_int]
        L.pcfd_update_bcs.argtypes = [vp]
        L.pcfd_gradient.argtypes = [vp, dp]
        L.pcfd_limiter.argtypes = [vp, dp]
        L.pcfd_set_limiter.argtypes = [vp, dp]
        L.pcfd_residual.argtypes = [vp, C.c_int, dp, dp]
        L.pcfd_timestep.argtypes = [vp, C.c_double, dp, dp]
        L.pcfd_set_timestep.argtypes = [vp, dp]
        L.pcfd_jacobian.argtypes = [vp]
        L.pcfd_download_jacobian.argtypes = [vp, ip, ip, dp]
        L.pcfd_prepare_sgs.argtypes = [vp]
        L.pcfd_sgs.argtypes = [vp, C.c_int, dp, dp]
        L.pcfd_sgs_order.argtypes = [vp, ip]
        L.pcfd_apply_dq.argtypes = [vp, ip]
        L.pcfd_explicit_update.argtypes = [vp, ip]
        L.pcfd_iterate.argtypes = [vp, dp]
        L.pcfd_reset_iteration.argtypes = [vp]
        L.pcfd_get_iteration.argtypes = [vp]
        L.pcfd_solve.argtypes = [vp, C.c_int, C.c_double, dp, ip, dp]
        L.pcfd_phase_times.argtypes = [vp, dp, C.c_int]
        L.pcfd_enable_phase_timing.argtypes = [vp, C.c_int]
        L.pcfd_synchronize.argtypes = [vp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def to_cparams(p: Params, gradient_terms: int = 9) -> CParams:
    return CParams(p.gamma, p.sorder, p.nfirst, p.chi, p.limiter, p.venkat_K, p.limiter_refresh, p.limiter_freeze,
                   p.cfl, p.cfl_start, p.cfl_ramp, p.nsgs, p.symmetry2d, p.jac_freq, p.precision,
                   9 if p.viscous else gradient_terms, int(p.viscous), p.re_tilde, p.prandtl, p.prandtl_t,
                   p.sutherland_s, p.gradient_type, p.global_dt, p.fixed_dt, p.enable_vnn, p.vnn, p.dynamic_cfl,
                   p.jac_field_type, p.flux_type, p.jac_boundary_type, p.rk_stages)


class Context:
    """One pcfd context on the current HIP device."""

    def __init__(self, case: Case, params: Params | None = None, gradient_terms: int = 9):
        m = case.mesh
        self.case = case
        self.nnode, self.nbnode = m.nnode, m.nbedge
        self._keep = dict(edges=_i32(m.edges), edge_a=_f64(m.edge_a), bedges=_i32(m.bedges),
                          bedge_a=_f64(m.bedge_a), bfactag=_i32(m.bfactag), xyz=_f64(m.xyz), vol=_f64(m.vol),
                          bc_type=_i32(case.bc_type), qref=_f64(case.qref[:, :5]),
                          twall=_f64(case.twall if case.twall is not None else np.full(len(case.bc_type), -1.0)))
        k = self._keep
        desc = MeshDesc(m.nnode, m.nbedge, m.nedge, _i(k["edges"]), _d(k["edge_a"]), _i(k["bedges"]),
                        _d(k["bedge_a"]), _i(k["bfactag"]), _d(k["xyz"]), _d(k["vol"]), len(k["bc_type"]),
                        _i(k["bc_type"]), _d(k["qref"]), _d(k["twall"]))
        h = C.c_void_p()
        rc = lib().pcfd_create(C.byref(desc), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"pcfd_create failed ({rc}): {lib().pcfd_last_error(None).decode()}")
        self.h = h
        self.set_params(params or case.params, gradient_terms)

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {lib().pcfd_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            lib().pcfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: Params, gradient_terms: int = 9):
        self.params = p
        cp = to_cparams(p, gradient_terms)
        self._chk(lib().pcfd_set_params(self.h, C.byref(cp)), "pcfd_set_params")

    def info(self):
        nc = C.c_int()
        nb = C.c_longlong()
        by = C.c_longlong()
        self._chk(lib().pcfd_info(self.h, C.byref(nc), C.byref(nb), C.byref(by)), "pcfd_info")
        return dict(ncolor=nc.value, nblocks=nb.value, device_bytes=by.value)

    def upload_q(self, q):
        q = _f64(q)
        self._chk(lib().pcfd_upload_q(self.h, _d(q), q.shape[1]), "pcfd_upload_q")

    def download_q(self, nvars=10):
        q = np.zeros((self.nnode + self.nbnode, nvars))
        self._chk(lib().pcfd_download_q(self.h, _d(q), nvars), "pcfd_download_q")
        return q

    def update_bcs(self):
        self._chk(lib().pcfd_update_bcs(self.h), "pcfd_update_bcs")

    def gradient(self, download=True):
        g = np.zeros((self.nnode, 27)) if download else None
        self._chk(lib().pcfd_gradient(self.h, _d(g) if download else None), "pcfd_gradient")
        return g

    def limiter(self, download=True):
        l = np.zeros((self.nnode, 5)) if download else None
        self._chk(lib().pcfd_limiter(self.h, _d(l) if download else None), "pcfd_limiter")
        return l

    def set_limiter(self, l):
        self._chk(lib().pcfd_set_limiter(self.h, _d(_f64(l))), "pcfd_set_limiter")

    def residual(self, it=1, download=True):
        b = np.zeros((self.nnode, 5)) if download else None
        norms = np.zeros(6)
        self._chk(lib().pcfd_residual(self.h, it, _d(b) if download else None, _d(norms)), "pcfd_residual")
        return b, norms

    def timestep(self, cfl, download=True):
        dt = np.zeros(self.nnode) if download else None
        dtmin = C.c_double()
        self._chk(lib().pcfd_timestep(self.h, cfl, _d(dt) if download else None, C.byref(dtmin)), "pcfd_timestep")
        return dt, dtmin.value

    def set_timestep(self, dt):
        self._chk(lib().pcfd_set_timestep(self.h, _d(_f64(dt))), "pcfd_set_timestep")

    def jacobian(self):
        self._chk(lib().pcfd_jacobian(self.h), "pcfd_jacobian")

    def download_jacobian(self, ia, ja):
        ia, ja = _i32(ia), _i32(ja)
        M = np.zeros((len(ja), 5, 5))
        self._chk(lib().pcfd_download_jacobian(self.h, _i(ia), _i(ja), _d(M)), "pcfd_download_jacobian")
        return M

    def prepare_sgs(self):
        self._chk(lib().pcfd_prepare_sgs(self.h), "pcfd_prepare_sgs")

    def sgs(self, nsgs, download=True):
        x = np.zeros((self.nnode, 5)) if download else None
        delta = C.c_double()
        self._chk(lib().pcfd_sgs(self.h, nsgs, _d(x) if download else None, C.byref(delta)), "pcfd_sgs")
        return x, delta.value

    def sgs_order(self):
        o = np.zeros(self.nnode, dtype=np.int32)
        self._chk(lib().pcfd_sgs_order(self.h, _i(o)), "pcfd_sgs_order")
        return o

    def apply_dq(self):
        n = C.c_int()
        self._chk(lib().pcfd_apply_dq(self.h, C.byref(n)), "pcfd_apply_dq")
        return n.value

    def explicit_update(self):
        n = C.c_int()
        self._chk(lib().pcfd_explicit_update(self.h, C.byref(n)), "pcfd_explicit_update")
        return n.value

    def reset_iteration(self):
        self._chk(lib().pcfd_reset_iteration(self.h), "pcfd_reset_iteration")

    @property
    def iteration(self):
        return lib().pcfd_get_iteration(self.h)

    def iterate(self):
        r = np.zeros(6)
        self._chk(lib().pcfd_iterate(self.h, _d(r)), "pcfd_iterate")
        return r

    def solve(self, max_iter, drop=1e-8):
        hist = np.zeros(max_iter)
        n = C.c_int()
        secs = C.c_double()
        rc = lib().pcfd_solve(self.h, max_iter, drop, _d(hist), C.byref(n), C.byref(secs))
        self._chk(rc, "pcfd_solve")
        return hist[: n.value], n.value, secs.value

    def phase_times(self, reset=False):
        ms = np.zeros(len(PHASES))
        self._chk(lib().pcfd_phase_times(self.h, _d(ms), int(reset)), "pcfd_phase_times")
        return dict(zip(PHASES, ms))

    def enable_phase_timing(self, on=True):
        self._chk(lib().pcfd_enable_phase_timing(self.h, int(on)), "pcfd_enable_phase_timing")

    def synchronize(self):
        self._chk(lib().pcfd_synchronize(self.h), "pcfd_synchronize")


# ---------------------------------------------------------------------------
# device buffers for the *_device entry points, through the HIP runtime that
# libpcfd.so itself links (no second runtime in the process)

_hip = None


def _hip_rt():
    global _hip
    if _hip is None:
        lib()  # loads libamdhip64.so.7 as libpcfd.so's dependency
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


class DeviceArray:
    """A device copy of a numpy array (hipMalloc'd; freed by close())."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.shape, self.dtype, self.nbytes = a.shape, a.dtype, a.nbytes
        H = _hip_rt()
        p = C.c_void_p()
        if H.hipMalloc(C.byref(p), max(self.nbytes, 1)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr = p.value
        if self.nbytes and H.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) != 0:  # host -> device
            raise RuntimeError("hipMemcpy failed")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        H = _hip_rt()
        if H.hipDeviceSynchronize() != 0 or (self.nbytes and H.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)):
            raise RuntimeError("hipMemcpy failed")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            _hip_rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# finite-rate chemistry (include/pcfd_chem.h)

class ChemDesc(C.Structure):
    _fields_ = [("nspecies", C.c_int), ("nreactions", C.c_int), ("mw", dp), ("nasa_lo", dp), ("nasa_hi", dp),
                ("rxn_type", ip), ("third_body", ip), ("A", dp), ("EA", dp), ("n", dp), ("nloc", ip), ("loc", ip),
                ("nup", dp), ("nupp", dp), ("tb_eff", dp), ("ref_density", C.c_double), ("ref_time", C.c_double),
                ("ref_temperature", C.c_double)]


CHEM_SYMBOLS = ["pcfd_chem_create", "pcfd_chem_destroy", "pcfd_chem_last_error", "pcfd_chem_wdot",
                "pcfd_chem_source", "pcfd_chem_source_device", "pcfd_chem_last_kernel_ms", "pcfd_chem_fr_aux",
                "pcfd_chem_t_from_p"]


def _chem_lib():
    L = lib()
    if not getattr(L, "_chem_ready", False):
        vp = C.c_void_p
        L.pcfd_chem_create.argtypes = [C.POINTER(ChemDesc), C.POINTER(vp)]
        L.pcfd_chem_destroy.argtypes = [vp]
        L.pcfd_chem_last_error.argtypes = [vp]
        L.pcfd_chem_last_error.restype = C.c_char_p
        L.pcfd_chem_wdot.argtypes = [vp, C.c_int, dp, dp, dp]
        L.pcfd_chem_source.argtypes = [vp, C.c_int, dp, dp, dp, dp]
        L.pcfd_chem_source_device.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.pcfd_chem_last_kernel_ms.argtypes = [vp]
        L.pcfd_chem_fr_aux.argtypes = [vp, C.c_int, dp, C.c_double, C.c_double, dp]
        L.pcfd_chem_t_from_p.argtypes = [vp, C.c_int, dp, dp, dp, C.c_double, dp, C.POINTER(C.c_int)]
        L.pcfd_chem_last_kernel_ms.restype = C.c_double
        L._chem_ready = True
    return L


class ChemContext:
    """A chemistry model on the current HIP device (proteuscfd_amd.chem.ChemModel)."""

    def __init__(self, model):
        L = _chem_lib()
        self.model = model
        MAXLOC = 8
        k = dict(mw=_f64(model.mw), lo=_f64(model.lo.reshape(-1)), hi=_f64(model.hi.reshape(-1)),
                 rt=_i32(model.rtype), tb=_i32(model.tb), A=_f64(model.A), EA=_f64(model.EA), n=_f64(model.n),
                 nloc=_i32(model.nloc), loc=_i32(model.loc[:, :MAXLOC].reshape(-1)),
                 nup=_f64(model.nup[:, :MAXLOC].reshape(-1)), nupp=_f64(model.nupp[:, :MAXLOC].reshape(-1)),
                 eff=_f64(model.eff[:, :MAXLOC].reshape(-1)))
        self._keep = k
        d = ChemDesc(model.ns, model.nr, _d(k["mw"]), _d(k["lo"]), _d(k["hi"]), _i(k["rt"]), _i(k["tb"]),
                     _d(k["A"]), _d(k["EA"]), _d(k["n"]), _i(k["nloc"]), _i(k["loc"]), _d(k["nup"]), _d(k["nupp"]),
                     _d(k["eff"]), model.ref_density, model.ref_time, model.ref_temperature)
        h = C.c_void_p()
        rc = L.pcfd_chem_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"pcfd_chem_create failed ({rc}): {L.pcfd_chem_last_error(None).decode()}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {_chem_lib().pcfd_chem_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            _chem_lib().pcfd_chem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_states(self, name, a, ncols, n=None):
        """The C side strides host arrays by the model's species count: a mismatched
        shape would read or write past the buffers, so it is refused here."""
        if a.ndim != 2 or a.shape[1] != ncols or (n is not None and a.shape[0] != n):
            want = f"({n if n is not None else 'n'}, {ncols})"
            raise ValueError(f"{name} has shape {a.shape}; this model ({self.model.ns} species) needs {want}")

    def _check_vector(self, name, a, n):
        if a.ndim != 1 or a.shape[0] != n:
            raise ValueError(f"{name} has shape {a.shape}; expected ({n},)")

    def wdot(self, rhoi, T):
        rhoi = _f64(rhoi)
        T = _f64(T)
        self._check_vector("T", T, len(T))
        self._check_states("rhoi", rhoi, self.model.ns, len(T))
        w = np.zeros_like(rhoi)
        self._chk(_chem_lib().pcfd_chem_wdot(self.h, len(T), _d(rhoi), _d(T), _d(w)), "pcfd_chem_wdot")
        return w

    def fr_aux(self, q, ref_pressure, ref_velocity):
        """[n, 2+2ns]: p, rho, cv_1..cv_ns, conc_1..conc_ns per FR state."""
        q = _f64(q)
        n = q.shape[0]
        self._check_states("q", q, self.model.ns + 4, n)
        a = np.zeros((n, 2 + 2 * self.model.ns))
        self._chk(_chem_lib().pcfd_chem_fr_aux(self.h, n, _d(q), float(ref_pressure), float(ref_velocity), _d(a)),
                  "pcfd_chem_fr_aux")
        return a

    def t_from_p(self, rhoi, p, Tinit, ref_pressure):
        """(T [n], Newton steps [n]) for n states."""
        rhoi, p, Tinit = _f64(rhoi), _f64(p), _f64(Tinit)
        n = len(p)
        self._check_vector("p", p, n)
        self._check_vector("Tinit", Tinit, n)
        self._check_states("rhoi", rhoi, self.model.ns, n)
        T = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        self._chk(_chem_lib().pcfd_chem_t_from_p(self.h, n, _d(rhoi), _d(p), _d(Tinit), float(ref_pressure), _d(T),
                                                 it.ctypes.data_as(C.POINTER(C.c_int))), "pcfd_chem_t_from_p")
        return T, it

    def source(self, q, vol, jacobian=True):
        q = _f64(q)
        vol = _f64(vol)
        neqn = self.model.ns + 4
        self._check_states("q", q, neqn)
        n = q.shape[0]
        self._check_vector("vol", vol, n)
        s = np.zeros((n, neqn))
        J = np.zeros((n, neqn, neqn)) if jacobian else None
        self._chk(_chem_lib().pcfd_chem_source(self.h, n, _d(q), _d(vol), _d(s), _d(J) if jacobian else None),
                  "pcfd_chem_source")
        return (s, J) if jacobian else s

    def source_device(self, q_ptr, vol_ptr, src_ptr, jac_ptr, n):
        """Device pointers (e.g. torch tensor data_ptr()); returns kernel ms."""
        self._chk(_chem_lib().pcfd_chem_source_device(self.h, n, q_ptr, vol_ptr, src_ptr, jac_ptr),
                  "pcfd_chem_source_device")
        return _chem_lib().pcfd_chem_last_kernel_ms(self.h)


# ---------------------------------------------------------------------------
# chemistry surrogate MLP (include/pcfd_chem.h, pcfd_mlp_*)

MLP_NPARAM = 6288
MLP_SYMBOLS = ["pcfd_mlp_create", "pcfd_mlp_destroy", "pcfd_mlp_last_error", "pcfd_mlp_last_kernel_ms",
               "pcfd_mlp_set_params", "pcfd_mlp_set_partner", "pcfd_mlp_get_params", "pcfd_mlp_forward",
               "pcfd_mlp_forward_device",
               "pcfd_mlp_loss_grad", "pcfd_mlp_train", "pcfd_chem_surrogate_wdot_device", "pcfd_chem_wdot_device",
               "pcfd_chem_surrogate_wdot", "pcfd_chem_surrogate_wdot_device_noalloc"]
SURROGATE_NOALLOC_MAX_CELLS = 1 << 22
CHEM_SYMBOLS += MLP_SYMBOLS

fp = C.POINTER(C.c_float)


class SurrogateDesc(C.Structure):
    _fields_ = [("lnT_mid", C.c_double), ("lnT_half", C.c_double), ("lnrho_mid", C.c_double),
                ("lnrho_half", C.c_double), ("y_mid", C.c_double), ("y_scale", C.c_double), ("y_eps", C.c_double),
                ("rate_scale", C.c_double), ("sigma", C.c_double * 5), ("T_min", C.c_double), ("T_max", C.c_double),
                ("rho_min", C.c_double), ("rho_max", C.c_double), ("tol", C.c_double), ("floor", C.c_double),
                ("atoms", (C.c_double * 5) * 2), ("err_tol", C.c_double), ("rate_floor", C.c_double),
                ("err_mid", C.c_double * 5), ("err_scale", C.c_double * 5)]


def _mlp_lib():
    L = _chem_lib()
    if not getattr(L, "_mlp_ready", False):
        vp = C.c_void_p
        L.pcfd_mlp_create.argtypes = [C.c_ulonglong, C.POINTER(vp)]
        L.pcfd_mlp_destroy.argtypes = [vp]
        L.pcfd_mlp_last_error.argtypes = [vp]
        L.pcfd_mlp_last_error.restype = C.c_char_p
        L.pcfd_mlp_last_kernel_ms.argtypes = [vp]
        L.pcfd_mlp_last_kernel_ms.restype = C.c_double
        L.pcfd_mlp_set_params.argtypes = [vp, fp]
        L.pcfd_mlp_set_partner.argtypes = [vp, vp]
        L.pcfd_mlp_get_params.argtypes = [vp, fp]
        L.pcfd_mlp_forward.argtypes = [vp, C.c_int, fp, fp]
        L.pcfd_mlp_forward_device.argtypes = [vp, C.c_int, vp, vp]
        L.pcfd_mlp_loss_grad.argtypes = [vp, C.c_int, fp, fp, fp, fp]
        L.pcfd_mlp_train.argtypes = [vp, C.c_int, fp, fp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp]
        L.pcfd_chem_surrogate_wdot_device.argtypes = [vp, vp, C.POINTER(SurrogateDesc), C.c_int, vp, vp, vp, ip,
                                                      dp]
        L.pcfd_chem_surrogate_wdot_device_noalloc.argtypes = [vp, vp, C.POINTER(SurrogateDesc), C.c_int, vp, vp,
                                                              vp, ip, dp]
        L.pcfd_chem_wdot_device.argtypes = [vp, C.c_int, vp, vp, vp]
        L.pcfd_chem_surrogate_wdot.argtypes = [vp, vp, C.POINTER(SurrogateDesc), C.c_int, dp, dp, dp, ip, dp]
        L._mlp_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """The surrogate network on the current HIP device."""

    def __init__(self, seed=0):
        L = _mlp_lib()
        h = C.c_void_p()
        rc = L.pcfd_mlp_create(seed, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"pcfd_mlp_create failed ({rc}): {L.pcfd_mlp_last_error(None).decode()}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {_mlp_lib().pcfd_mlp_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            _mlp_lib().pcfd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_partner(self, other):
        """Error network of the surrogate (None removes it); it is kept
        alive with this one."""
        self._chk(_mlp_lib().pcfd_mlp_set_partner(self.h, other.h if other is not None else None),
                  "pcfd_mlp_set_partner")
        self.partner = other

    @property
    def params(self):
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._chk(_mlp_lib().pcfd_mlp_get_params(self.h, p.ctypes.data_as(fp)), "pcfd_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        self._chk(_mlp_lib().pcfd_mlp_set_params(self.h, p.ctypes.data_as(fp)), "pcfd_mlp_set_params")

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], 5), dtype=np.float32)
        self._chk(_mlp_lib().pcfd_mlp_forward(self.h, x.shape[0], x.ctypes.data_as(fp), y.ctypes.data_as(fp)),
                  "pcfd_mlp_forward")
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        loss = C.c_float()
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._chk(_mlp_lib().pcfd_mlp_loss_grad(self.h, x.shape[0], x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                C.byref(loss), g.ctypes.data_as(fp)), "pcfd_mlp_loss_grad")
        return loss.value, g

    def train(self, x, t, steps, batch, lr, seed=1, history=True):
        x, t = _f32(x), _f32(t)
        hist = np.zeros(steps, dtype=np.float32) if history else None
        self._chk(_mlp_lib().pcfd_mlp_train(self.h, x.shape[0], x.ctypes.data_as(fp), t.ctypes.data_as(fp), steps,
                                            batch, lr, seed, hist.ctypes.data_as(fp) if history else None),
                  "pcfd_mlp_train")
        return hist

    @property
    def last_ms(self):
        return _mlp_lib().pcfd_mlp_last_kernel_ms(self.h)


# ---------------------------------------------------------------------------
# finite-rate reacting Euler (include/pcfd_fr.h, pcfd_fr_*)

FR_SYMBOLS = ["pcfd_fr_create", "pcfd_fr_destroy", "pcfd_fr_last_error", "pcfd_fr_set_params", "pcfd_fr_info",
              "pcfd_fr_upload_q", "pcfd_fr_download_q", "pcfd_fr_update_bcs", "pcfd_fr_residual", "pcfd_fr_timestep",
              "pcfd_fr_set_timestep", "pcfd_fr_jacobian", "pcfd_fr_download_jacobian", "pcfd_fr_pattern",
              "pcfd_fr_sgs", "pcfd_fr_sgs_order", "pcfd_fr_apply_dq", "pcfd_fr_init_iterations", "pcfd_fr_step",
              "pcfd_fr_solve", "pcfd_fr_phase_times", "pcfd_fr_set_surrogate", "pcfd_fr_surrogate_stats"]
FR_PHASES = ["jacobian", "timestep", "bcs", "residual", "sgs", "update", "k_jac_edge", "k_edge_flux"]


class FrParams(C.Structure):
    _fields_ = [("ref_pressure", C.c_double), ("ref_velocity", C.c_double), ("Pref", C.c_double),
                ("beta", C.c_double), ("symmetry2d", C.c_int), ("reactions", C.c_int), ("local_dt", C.c_int),
                ("dt_param", C.c_double), ("cfl", C.c_double), ("cfl_start", C.c_double), ("cfl_ramp", C.c_int),
                ("nsgs", C.c_int), ("jac_freq", C.c_int), ("precision", C.c_int), ("no_cvbc", C.c_int)]


def _fr_lib():
    L = lib()
    if not getattr(L, "_fr_ready", False):
        vp = C.c_void_p
        L.pcfd_fr_create.argtypes = [C.POINTER(MeshDesc), vp, dp, C.POINTER(vp)]
        L.pcfd_fr_destroy.argtypes = [vp]
        L.pcfd_fr_last_error.argtypes = [vp]
        L.pcfd_fr_last_error.restype = C.c_char_p
        L.pcfd_fr_set_params.argtypes = [vp, C.POINTER(FrParams)]
        L.pcfd_fr_info.argtypes = [vp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.pcfd_fr_upload_q.argtypes = [vp, dp, C.c_int]
        L.pcfd_fr_download_q.argtypes = [vp, dp, C.c_int]
        L.pcfd_fr_update_bcs.argtypes = [vp]
        L.pcfd_fr_residual.argtypes = [vp, dp, dp]
        L.pcfd_fr_timestep.argtypes = [vp, C.c_double, dp, dp]
        L.pcfd_fr_set_timestep.argtypes = [vp, dp]
        L.pcfd_fr_jacobian.argtypes = [vp]
        L.pcfd_fr_download_jacobian.argtypes = [vp, ip, ip, dp]
        L.pcfd_fr_pattern.argtypes = [vp, ip, ip]
        L.pcfd_fr_sgs.argtypes = [vp, C.c_int, dp, dp]
        L.pcfd_fr_sgs_order.argtypes = [vp, ip]
        L.pcfd_fr_apply_dq.argtypes = [vp, ip]
        L.pcfd_fr_init_iterations.argtypes = [vp]
        L.pcfd_fr_step.argtypes = [vp, dp, ip]
        L.pcfd_fr_solve.argtypes = [vp, C.c_int, C.c_double, dp, ip]
        L.pcfd_fr_phase_times.argtypes = [vp, dp, C.c_int]
        L.pcfd_fr_set_surrogate.argtypes = [vp, vp, C.c_void_p]
        L.pcfd_fr_surrogate_stats.argtypes = [vp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp]
        L._fr_ready = True
    return L


class FrContext:
    """One finite-rate (FR) context on the current HIP device.

    mesh: a proteuscfd_amd.mesh.Mesh-like object (edges, edge_a, bedges,
    bedge_a, bfactag, xyz, vol); bc_type [ntag]; qref [ntag, ns+4] native
    reference states; chem: a ChemContext (kept alive by this object)."""

    def __init__(self, mesh, bc_type, qref, chem: ChemContext):
        L = _fr_lib()
        self.chem = chem
        self.ns = chem.model.ns
        self.neqn, self.nvars = self.ns + 4, 3 * self.ns + 6
        self.nnode, self.nbnode, self.nedge = int(mesh.nnode), int(mesh.nbedge), int(mesh.nedge)
        bc_type = _i32(bc_type)
        qref = _f64(np.asarray(qref)[:, :self.neqn])
        if qref.shape != (len(bc_type), self.neqn):
            raise ValueError(f"qref has shape {qref.shape}; needs ({len(bc_type)}, {self.neqn})")
        self._keep = dict(edges=_i32(mesh.edges), edge_a=_f64(mesh.edge_a), bedges=_i32(mesh.bedges),
                          bedge_a=_f64(mesh.bedge_a), bfactag=_i32(mesh.bfactag), xyz=_f64(mesh.xyz),
                          vol=_f64(mesh.vol), bc_type=bc_type, qref5=np.zeros((len(bc_type), 5)),
                          twall=np.full(len(bc_type), -1.0), qref=qref)
        k = self._keep
        desc = MeshDesc(self.nnode, self.nbnode, self.nedge, _i(k["edges"]), _d(k["edge_a"]), _i(k["bedges"]),
                        _d(k["bedge_a"]), _i(k["bfactag"]), _d(k["xyz"]), _d(k["vol"]), len(bc_type),
                        _i(k["bc_type"]), _d(k["qref5"]), _d(k["twall"]))
        h = C.c_void_p()
        rc = L.pcfd_fr_create(C.byref(desc), chem.h, _d(k["qref"]), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"pcfd_fr_create failed ({rc}): {L.pcfd_fr_last_error(None).decode()}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {_fr_lib().pcfd_fr_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            _fr_lib().pcfd_fr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, ref_pressure, ref_velocity, Pref, beta, cfl=1.0, cfl_start=1.0, cfl_ramp=0, nsgs=4,
                   jac_freq=1, precision=PLAIN_FP64, reactions=True, local_dt=True, dt_param=-1.0, symmetry2d=0,
                   no_cvbc=False):
        p = FrParams(float(ref_pressure), float(ref_velocity), float(Pref), float(beta), int(symmetry2d),
                     int(reactions), int(local_dt), float(dt_param), float(cfl), float(cfl_start), int(cfl_ramp),
                     int(nsgs), int(jac_freq), int(precision), int(no_cvbc))
        self._chk(_fr_lib().pcfd_fr_set_params(self.h, C.byref(p)), "pcfd_fr_set_params")

    def info(self):
        nc, nb, by = C.c_int(), C.c_longlong(), C.c_longlong()
        self._chk(_fr_lib().pcfd_fr_info(self.h, C.byref(nc), C.byref(nb), C.byref(by)), "pcfd_fr_info")
        return dict(ncolor=nc.value, nblocks=nb.value, device_bytes=by.value)

    def upload_q(self, q):
        q = _f64(q)
        if q.ndim != 2 or q.shape[0] != self.nnode + self.nbnode or q.shape[1] < self.neqn:
            raise ValueError(f"q has shape {q.shape}; needs ({self.nnode + self.nbnode}, >= {self.neqn})")
        self._chk(_fr_lib().pcfd_fr_upload_q(self.h, _d(q), q.shape[1]), "pcfd_fr_upload_q")

    def download_q(self, full=True):
        q = np.zeros((self.nnode + self.nbnode, self.nvars if full else self.neqn))
        self._chk(_fr_lib().pcfd_fr_download_q(self.h, _d(q), q.shape[1]), "pcfd_fr_download_q")
        return q

    def update_bcs(self):
        self._chk(_fr_lib().pcfd_fr_update_bcs(self.h), "pcfd_fr_update_bcs")

    def residual(self, download=True):
        b = np.zeros((self.nnode, self.neqn)) if download else None
        norms = np.zeros(1 + self.neqn)
        self._chk(_fr_lib().pcfd_fr_residual(self.h, _d(b) if download else None, _d(norms)), "pcfd_fr_residual")
        return b, norms

    def timestep(self, cfl, download=True):
        dt = np.zeros(self.nnode) if download else None
        dtmin = C.c_double()
        self._chk(_fr_lib().pcfd_fr_timestep(self.h, float(cfl), _d(dt) if download else None, C.byref(dtmin)),
                  "pcfd_fr_timestep")
        return dt, dtmin.value

    def set_timestep(self, dt):
        dt = _f64(dt)
        if dt.shape != (self.nnode,):
            raise ValueError(f"dt has shape {dt.shape}; needs ({self.nnode},)")
        self._chk(_fr_lib().pcfd_fr_set_timestep(self.h, _d(dt)), "pcfd_fr_set_timestep")

    def jacobian(self):
        self._chk(_fr_lib().pcfd_fr_jacobian(self.h), "pcfd_fr_jacobian")

    def pattern(self):
        ia = np.zeros(self.nnode + 1, dtype=np.int32)
        ja = np.zeros(self.nnode + 2 * self.nedge, dtype=np.int32)
        self._chk(_fr_lib().pcfd_fr_pattern(self.h, _i(ia), _i(ja)), "pcfd_fr_pattern")
        return ia, ja

    def download_jacobian(self, ia, ja):
        ia, ja = _i32(ia), _i32(ja)
        if ia.shape != (self.nnode + 1,) or ja.shape != (self.nnode + 2 * self.nedge,) or ia[-1] != len(ja):
            raise ValueError("ia/ja do not describe this mesh's block pattern")
        M = np.zeros((len(ja), self.neqn, self.neqn))
        self._chk(_fr_lib().pcfd_fr_download_jacobian(self.h, _i(ia), _i(ja), _d(M)), "pcfd_fr_download_jacobian")
        return M

    def sgs(self, nsgs, download=True):
        x = np.zeros((self.nnode, self.neqn)) if download else None
        delta = C.c_double()
        self._chk(_fr_lib().pcfd_fr_sgs(self.h, int(nsgs), _d(x) if download else None, C.byref(delta)),
                  "pcfd_fr_sgs")
        return x, delta.value

    def sgs_order(self):
        o = np.zeros(self.nnode, dtype=np.int32)
        self._chk(_fr_lib().pcfd_fr_sgs_order(self.h, _i(o)), "pcfd_fr_sgs_order")
        return o

    def apply_dq(self):
        n = C.c_int()
        self._chk(_fr_lib().pcfd_fr_apply_dq(self.h, C.byref(n)), "pcfd_fr_apply_dq")
        return n.value

    def init_iterations(self):
        self._chk(_fr_lib().pcfd_fr_init_iterations(self.h), "pcfd_fr_init_iterations")

    def step(self):
        norms = np.zeros(1 + self.neqn)
        nz = C.c_int()
        self._chk(_fr_lib().pcfd_fr_step(self.h, _d(norms), C.byref(nz)), "pcfd_fr_step")
        return norms[0], norms, nz.value

    def solve(self, max_iter, drop=1e-8):
        hist = np.zeros(max_iter)
        n = C.c_int()
        self._chk(_fr_lib().pcfd_fr_solve(self.h, int(max_iter), float(drop), _d(hist), C.byref(n)), "pcfd_fr_solve")
        return hist[:n.value]

    def set_surrogate(self, mlp, desc):
        """mlp: a gpu.MLP (kept alive here) and its SurrogateDesc; None for the exact rates."""
        self._mlp, self._sdesc = mlp, desc
        self._chk(_fr_lib().pcfd_fr_set_surrogate(self.h, mlp.h if mlp is not None else None,
                                                  C.byref(desc) if mlp is not None else None),
                  "pcfd_fr_set_surrogate")

    def surrogate_stats(self):
        last, calls, tot, ms = C.c_int(), C.c_longlong(), C.c_longlong(), C.c_double()
        self._chk(_fr_lib().pcfd_fr_surrogate_stats(self.h, C.byref(last), C.byref(calls), C.byref(tot),
                                                    C.byref(ms)), "pcfd_fr_surrogate_stats")
        return dict(last_fallback=last.value, calls=calls.value, fallback_total=tot.value, ms=ms.value)

    def phase_times(self, reset=False):
        ms = np.zeros(len(FR_PHASES))
        self._chk(_fr_lib().pcfd_fr_phase_times(self.h, _d(ms), int(reset)), "pcfd_fr_phase_times")
        return dict(zip(FR_PHASES, ms))
