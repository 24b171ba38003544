"""ctypes binding of the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  The product package never imports it.
All arrays use the reference layouts (AoS q with 10 vars per node, etc.).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class OMesh(C.Structure):
    _fields_ = [("nnode", C.c_int), ("nbnode", C.c_int), ("nedge", C.c_int),
                ("edges", ip), ("edge_a", dp), ("bedges", ip), ("bedge_a", dp), ("bfactag", ip),
                ("xyz", dp), ("vol", dp), ("nfactag", C.c_int), ("bc_type", ip), ("qref", dp),
                ("twall", dp), ("bnormal", ip)]


class OParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("sorder", C.c_int), ("nfirst", C.c_int), ("chi", C.c_double),
                ("limiter", C.c_int), ("venkat_K", C.c_double), ("limiter_refresh", C.c_int),
                ("limiter_freeze", C.c_int), ("cfl", C.c_double), ("cfl_start", C.c_double),
                ("cfl_ramp", C.c_int), ("nsgs", C.c_int), ("symmetry2d", C.c_int), ("jac_freq", C.c_int),
                ("clip_jacobi", C.c_int), ("viscous", C.c_int), ("re_tilde", C.c_double), ("pr", C.c_double),
                ("prt", C.c_double), ("sutherland_s", C.c_double), ("gradient_type", C.c_int),
                ("global_dt", C.c_int), ("fixed_dt", C.c_double), ("enable_vnn", C.c_int), ("vnn", C.c_double),
                ("dynamic_cfl", C.c_int), ("jac_field_type", C.c_int), ("flux_type", C.c_int),
                ("jac_boundary_type", C.c_int), ("rk_stages", C.c_int)]


class OChem(C.Structure):
    _fields_ = [("ns", C.c_int), ("nr", C.c_int), ("mw", C.c_double * 8), ("lo", (C.c_double * 7) * 8),
                ("hi", (C.c_double * 7) * 8), ("type", C.c_int * 16), ("tb", C.c_int * 16),
                ("A", C.c_double * 16), ("EA", C.c_double * 16), ("n", C.c_double * 16), ("nloc", C.c_int * 16),
                ("loc", (C.c_int * 8) * 16), ("nup", (C.c_double * 8) * 16), ("nupp", (C.c_double * 8) * 16),
                ("eff", (C.c_double * 8) * 16), ("ref_density", C.c_double), ("ref_time", C.c_double),
                ("ref_temperature", C.c_double)]


class OCrs(C.Structure):
    _fields_ = [("nnode", C.c_int), ("nblocks", C.c_int), ("ia", ip), ("ja", ip), ("iau", ip), ("pv", ip),
                ("M", dp), ("e_lr", ip), ("e_rl", ip)]


class OSolver(C.Structure):
    _fields_ = [("m", C.POINTER(OMesh)), ("p", OParams), ("q", dp), ("grad", dp), ("l", dp), ("sw", dp),
                ("b", dp), ("x", dp), ("dt", dp), ("A", OCrs), ("iter", C.c_int), ("cfl", C.c_double),
                ("order", ip), ("t_jac", C.c_double), ("t_grad", C.c_double), ("t_res", C.c_double),
                ("t_lin", C.c_double), ("t_upd", C.c_double), ("res", C.c_double), ("res_nm1", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build()
        L = C.CDLL(path)
        L.o_aux.argtypes = [dp, C.c_double]
        L.o_roe_flux.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_int, dp]
        L.o_roe_flux.restype = C.c_int
        L.o_bc_state.argtypes = [C.c_int, dp, dp, dp, dp, C.c_double]
        L.o_bc_state_ex.argtypes = [C.c_int, dp, dp, dp, dp, C.c_double, C.c_double, dp]
        L.o_wall_normal_nodes.argtypes = [C.POINTER(OMesh), ip]
        L.o_viscous_flux.argtypes = [C.POINTER(OParams), dp, dp, dp, dp]
        L.o_viscous_jacobian.argtypes = [C.POINTER(OParams), dp, dp, dp, C.c_double, dp, dp, dp]
        L.o_update_bcs.argtypes = [C.POINTER(OMesh), C.POINTER(OParams), dp]
        L.o_lsq_moments.argtypes = [C.POINTER(OMesh), dp]
        L.o_gradient.argtypes = [C.POINTER(OMesh), dp, dp, dp]
        L.o_gradient_gg.argtypes = [C.POINTER(OMesh), dp, dp]
        L.o_limiter.argtypes = [C.POINTER(OMesh), C.POINTER(OParams), dp, dp, dp]
        L.o_residual.argtypes = [C.POINTER(OMesh), C.POINTER(OParams), C.c_int, dp, dp, dp, dp, dp]
        L.o_timestep.argtypes = [C.POINTER(OMesh), C.POINTER(OParams), dp, C.c_double, dp]
        L.o_timestep.restype = C.c_double
        L.o_cfl.argtypes = [C.POINTER(OParams), C.c_int]
        L.o_cfl.restype = C.c_double
        L.o_cfl_update.argtypes = [C.POINTER(OParams), C.c_int, C.c_double, C.c_double, C.c_double]
        L.o_cfl_update.restype = C.c_double
        L.o_crs_init.argtypes = [C.POINTER(OMesh), C.POINTER(OCrs)]
        L.o_crs_free.argtypes = [C.POINTER(OCrs)]
        L.o_jacobian.argtypes = [C.POINTER(OMesh), C.POINTER(OParams), dp, dp, C.POINTER(OCrs)]
        L.o_prepare_sgs.argtypes = [C.POINTER(OCrs)]
        L.o_sgs.argtypes = [C.POINTER(OCrs), C.c_int, dp, dp, ip]
        L.o_sgs.restype = C.c_double
        L.o_lu.argtypes = [dp, ip, C.c_int]
        L.o_lu.restype = C.c_int
        L.o_lu_solve.argtypes = [dp, dp, ip, dp, C.c_int]
        L.o_apply_dq.argtypes = [dp, dp, C.c_double]
        L.o_solver_init.argtypes = [C.POINTER(OSolver), C.POINTER(OMesh), C.POINTER(OParams), dp]
        L.o_solver_step.argtypes = [C.POINTER(OSolver), dp, ip]
        L.o_solver_step.restype = C.c_double
        L.o_solver_free.argtypes = [C.POINTER(OSolver)]
        L.o_chem_kf.argtypes = [C.POINTER(OChem), C.c_int, C.c_double]
        L.o_chem_kf.restype = C.c_double
        L.o_chem_kc.argtypes = [C.POINTER(OChem), C.c_int, C.c_double]
        L.o_chem_kc.restype = C.c_double
        L.o_chem_wdot.argtypes = [C.POINTER(OChem), dp, C.c_double, dp]
        L.o_chem_source.argtypes = [C.POINTER(OChem), dp, C.c_double, dp]
        L.o_chem_source_jacobian.argtypes = [C.POINTER(OChem), dp, C.c_double, dp]
        L.o_chem_cp.argtypes = [C.POINTER(OChem), C.c_int, C.c_double]
        L.o_chem_cp.restype = C.c_double
        L.o_fr_aux.argtypes = [C.POINTER(OChem), dp, C.c_double, C.c_double, dp]
        L.o_fr_t_from_p.argtypes = [C.POINTER(OChem), dp, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(C.c_int)]
        L.o_fr_t_from_p.restype = C.c_double
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


class Problem:
    """Mesh + BCs + parameters in oracle form (keeps the numpy buffers alive)."""

    def __init__(self, mesh, bc_type, qref, params, clip_jacobi=False, twall=None):
        nt = len(bc_type)
        tw = np.full(nt, -1.0) if twall is None else _f64(twall)
        self.keep = dict(edges=_i32(mesh.edges), edge_a=_f64(mesh.edge_a), bedges=_i32(mesh.bedges),
                         bedge_a=_f64(mesh.bedge_a), bfactag=_i32(mesh.bfactag), xyz=_f64(mesh.xyz),
                         vol=_f64(mesh.vol), bc_type=_i32(bc_type), qref=_f64(qref), twall=_f64(tw),
                         bnormal=np.full(mesh.nbedge, -1, dtype=np.int32))
        k = self.keep
        self.nnode, self.nbnode, self.nedge = mesh.nnode, mesh.nbedge, mesh.nedge
        self.m = OMesh(mesh.nnode, mesh.nbedge, mesh.nedge, _i(k["edges"]), _d(k["edge_a"]), _i(k["bedges"]),
                       _d(k["bedge_a"]), _i(k["bfactag"]), _d(k["xyz"]), _d(k["vol"]), nt,
                       _i(k["bc_type"]), _d(k["qref"]), _d(k["twall"]), _i(k["bnormal"]))
        lib().o_wall_normal_nodes(C.byref(self.m), _i(k["bnormal"]))
        self.params = params
        self.p = OParams(params.gamma, params.sorder, params.nfirst, params.chi, params.limiter, params.venkat_K,
                         params.limiter_refresh, params.limiter_freeze, params.cfl, params.cfl_start,
                         params.cfl_ramp, params.nsgs, params.symmetry2d, params.jac_freq, int(clip_jacobi),
                         int(params.viscous), params.re_tilde, params.prandtl, params.prandtl_t,
                         params.sutherland_s, params.gradient_type, params.global_dt, params.fixed_dt,
                         params.enable_vnn, params.vnn, params.dynamic_cfl, params.jac_field_type,
                         params.flux_type, params.jac_boundary_type, getattr(params, "rk_stages", 1))

    @property
    def bnormal(self):
        return self.keep["bnormal"]

    @classmethod
    def from_case(cls, case, clip_jacobi=False):
        return cls(case.mesh, case.bc_type, case.qref, case.params, clip_jacobi, twall=case.twall)

    def viscous_flux(self, Q, grad, avec):
        F = np.zeros(5)
        lib().o_viscous_flux(C.byref(self.p), _d(_f64(Q)), _d(_f64(grad)), _d(_f64(avec)), _d(F))
        return F

    def viscous_jacobian(self, QL, QR, dx, s2, avec):
        aL, aR = np.zeros(25), np.zeros(25)
        lib().o_viscous_jacobian(C.byref(self.p), _d(_f64(QL)), _d(_f64(QR)), _d(_f64(dx)), float(s2),
                                 _d(_f64(avec)), _d(aL), _d(aR))
        return aL.reshape(5, 5), aR.reshape(5, 5)

    # --- single phases (arrays in reference AoS layout) ---
    def update_bcs(self, q):
        q = _f64(q).copy()
        lib().o_update_bcs(C.byref(self.m), C.byref(self.p), _d(q))
        return q

    def lsq_moments(self):
        sw = np.zeros((self.nnode, 6))
        lib().o_lsq_moments(C.byref(self.m), _d(sw))
        return sw

    def gradient(self, q, sw=None):
        g = np.zeros((self.nnode, 27))
        if self.p.gradient_type == 1:
            lib().o_gradient_gg(C.byref(self.m), _d(_f64(q)), _d(g))
            return g
        if sw is None:
            sw = self.lsq_moments()
        lib().o_gradient(C.byref(self.m), _d(_f64(q)), _d(_f64(sw)), _d(g))
        return g

    def limiter(self, q, grad):
        l = np.ones((self.nnode, 5))
        lib().o_limiter(C.byref(self.m), C.byref(self.p), _d(_f64(q)), _d(_f64(grad)), _d(l))
        return l

    def residual(self, q, grad=None, l=None, it=1):
        if grad is None:
            grad = np.zeros((self.nnode, 27))
        if l is None:
            l = np.ones((self.nnode, 5))
        b = np.zeros((self.nnode, 5))
        norms = np.zeros(6)
        lib().o_residual(C.byref(self.m), C.byref(self.p), it, _d(_f64(q)), _d(_f64(grad)), _d(_f64(l)), _d(b),
                         _d(norms))
        return b, norms

    def timestep(self, q, cfl):
        dt = np.zeros(self.nnode)
        dtmin = lib().o_timestep(C.byref(self.m), C.byref(self.p), _d(_f64(q)), cfl, _d(dt))
        return dt, dtmin

    def cfl_update(self, it, res, res_nm1, cfl_prev):
        return lib().o_cfl_update(C.byref(self.p), it, res, res_nm1, cfl_prev)

    def jacobian(self, q, dt):
        """Returns (ia, ja, iau, M[nblocks,5,5], q_after) -- q_after holds the
        phantom states the boundary Jacobian re-evaluated."""
        A = OCrs()
        lib().o_crs_init(C.byref(self.m), C.byref(A))
        q = _f64(q).copy()
        lib().o_jacobian(C.byref(self.m), C.byref(self.p), _d(q), _d(_f64(dt)), C.byref(A))
        nb = A.nblocks
        ia = np.ctypeslib.as_array(A.ia, (self.nnode + 1,)).copy()
        ja = np.ctypeslib.as_array(A.ja, (nb,)).copy()
        iau = np.ctypeslib.as_array(A.iau, (self.nnode,)).copy()
        Mm = np.ctypeslib.as_array(A.M, (nb * 25,)).copy().reshape(nb, 5, 5)
        lib().o_crs_free(C.byref(A))
        return ia, ja, iau, Mm, q

    def sgs(self, q, dt, b, nsgs, order=None):
        """Jacobian + PrepareSGS + SGS from x0 = 0; returns x."""
        A = OCrs()
        lib().o_crs_init(C.byref(self.m), C.byref(A))
        qc = _f64(q).copy()
        lib().o_jacobian(C.byref(self.m), C.byref(self.p), _d(qc), _d(_f64(dt)), C.byref(A))
        lib().o_prepare_sgs(C.byref(A))
        x = np.zeros((self.nnode, 5))
        o = None if order is None else _i32(order)
        lib().o_sgs(C.byref(A), nsgs, _d(_f64(b)), _d(x), None if o is None else _i(o))
        lib().o_crs_free(C.byref(A))
        return x


def sgs_on_matrix(ia, ja, iau, M, b, nsgs, order=None):
    """PrepareSGS + SGS (x0 = 0) on a given block-CRS matrix (reference layout)."""
    ia, ja, iau = _i32(ia), _i32(ja), _i32(iau)
    Mc = _f64(M).copy().reshape(-1)
    nn = len(ia) - 1
    pv = np.zeros(5 * nn, dtype=np.int32)
    A = OCrs(nn, len(ja), _i(ia), _i(ja), _i(iau), _i(pv), _d(Mc), None, None)
    lib().o_prepare_sgs(C.byref(A))
    x = np.zeros((nn, 5))
    o = None if order is None else _i32(order)
    lib().o_sgs(C.byref(A), nsgs, _d(_f64(b)), _d(x), None if o is None else _i(o))
    return x


def roe_flux(QL, QR, avec, gamma=1.4, symmetry2d=-1):
    f = np.zeros(5)
    lib().o_roe_flux(_d(_f64(QL)), _d(_f64(QR)), _d(_f64(avec)), 0.0, gamma, symmetry2d, _d(f))
    return f


def lu(a):
    a = _f64(a).copy()
    p = np.zeros(a.shape[0], dtype=np.int32)
    bad = lib().o_lu(_d(a), _i(p), a.shape[0])
    return a, p, bad


def lu_solve(a, p, b):
    b = _f64(b).copy()
    x = np.zeros_like(b)
    lib().o_lu_solve(_d(_f64(a)), _d(b), _i(_i32(p)), _d(x), len(b))
    return b


class Solver:
    """The oracle's steady solve loop (PreTimeAdvance + NewtonIterate)."""

    def __init__(self, problem: Problem, q0, order=None):
        self.pb = problem
        self.s = OSolver()
        self._q0 = _f64(q0).copy()
        self._order = None if order is None else _i32(order)
        lib().o_solver_init(C.byref(self.s), C.byref(problem.m), C.byref(problem.p), _d(self._q0))
        if self._order is not None:
            self.s.order = _i(self._order)

    def step(self):
        norms = np.zeros(6)
        nz = C.c_int(0)
        r = lib().o_solver_step(C.byref(self.s), _d(norms), C.byref(nz))
        return r, norms, nz.value

    @property
    def q(self):
        n = self.pb.nnode + self.pb.nbnode
        return np.ctypeslib.as_array(self.s.q, (n * 10,)).copy().reshape(n, 10)

    @property
    def iter(self):
        return self.s.iter

    def timers(self):
        s = self.s
        return dict(jacobian=s.t_jac, gradient=s.t_grad, residual=s.t_res, linear=s.t_lin, update=s.t_upd)

    def close(self):
        if self.s is not None:
            lib().o_solver_free(C.byref(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Chem:
    """Oracle view of a proteuscfd_amd.chem.ChemModel."""

    def __init__(self, model):
        c = OChem()
        c.ns, c.nr = model.ns, model.nr
        for s in range(model.ns):
            c.mw[s] = model.mw[s]
            for k in range(7):
                c.lo[s][k] = model.lo[s, k]
                c.hi[s][k] = model.hi[s, k]
        for r in range(model.nr):
            c.type[r], c.tb[r] = int(model.rtype[r]), int(model.tb[r])
            c.A[r], c.EA[r], c.n[r] = model.A[r], model.EA[r], model.n[r]
            c.nloc[r] = int(model.nloc[r])
            for k in range(int(model.nloc[r])):
                c.loc[r][k] = int(model.loc[r, k])
                c.nup[r][k], c.nupp[r][k], c.eff[r][k] = model.nup[r, k], model.nupp[r, k], model.eff[r, k]
        c.ref_density, c.ref_time, c.ref_temperature = model.ref_density, model.ref_time, model.ref_temperature
        self.c = c
        self.model = model

    def kf(self, r, T):
        return lib().o_chem_kf(C.byref(self.c), r, T)

    def kc(self, r, T):
        return lib().o_chem_kc(C.byref(self.c), r, T)

    def wdot(self, rhoi, T):
        w = np.zeros(self.model.ns)
        lib().o_chem_wdot(C.byref(self.c), _d(_f64(rhoi)), float(T), _d(w))
        return w

    def source(self, Q, vol):
        s = np.zeros(self.model.ns + 4)
        lib().o_chem_source(C.byref(self.c), _d(_f64(Q)), float(vol), _d(s))
        return s

    def source_jacobian(self, Q, vol):
        n = self.model.ns + 4
        A = np.zeros(n * n)
        lib().o_chem_source_jacobian(C.byref(self.c), _d(_f64(Q)), float(vol), _d(A))
        return A.reshape(n, n)

    def cp(self, s, T):
        return lib().o_chem_cp(C.byref(self.c), s, T)

    def fr_aux(self, Q, ref_pressure, ref_velocity):
        """p, rho, cv_1..cv_ns, conc_1..conc_ns of one FR state."""
        a = np.zeros(2 + 2 * self.model.ns)
        lib().o_fr_aux(C.byref(self.c), _d(_f64(Q)), float(ref_pressure), float(ref_velocity), _d(a))
        return a

    def t_from_p(self, rhoi, p, Tinit, ref_pressure):
        """(T, Newton steps) for one state."""
        it = C.c_int(0)
        T = lib().o_fr_t_from_p(C.byref(self.c), _d(_f64(rhoi)), float(p), float(Tinit), float(ref_pressure),
                                C.byref(it))
        return T, it.value


# --- finite-rate (FR) equation set (oracle_fr.c) ---------------------------

class OFr(C.Structure):
    _fields_ = [("chem", C.POINTER(OChem)), ("ref_pressure", C.c_double), ("ref_velocity", C.c_double),
                ("Pref", C.c_double), ("beta", C.c_double), ("sym2d", C.c_int), ("rxn_on", C.c_int),
                ("local_dt", C.c_int), ("dt_param", C.c_double), ("no_cvbc", C.c_int)]


class OFrCrs(C.Structure):
    _fields_ = [("nnode", C.c_int), ("nblocks", C.c_int), ("neqn", C.c_int), ("ia", ip), ("ja", ip), ("iau", ip),
                ("e_lr", ip), ("e_rl", ip), ("pv", ip), ("M", dp), ("ludiag", C.c_int)]


class OFrSolver(C.Structure):
    _fields_ = [("m", C.POINTER(OMesh)), ("f", OFr), ("q", dp), ("qref", dp), ("b", dp), ("x", dp), ("dt", dp),
                ("A", OFrCrs), ("iter", C.c_int), ("nsgs", C.c_int), ("jac_freq", C.c_int), ("cfl_ramp", C.c_int),
                ("cfl", C.c_double), ("cfl_start", C.c_double), ("cfl_end", C.c_double), ("order", ip)]


_FR_BOUND = False


def _fr_lib():
    global _FR_BOUND
    L = lib()
    if not _FR_BOUND:
        F, M, R = C.POINTER(OFr), C.POINTER(OMesh), C.POINTER(OFrCrs)
        L.o_fr_aux_q.argtypes = [F, dp]
        L.o_fr_total_enthalpy.argtypes = [F, dp]
        L.o_fr_total_enthalpy.restype = C.c_double
        L.o_fr_total_energy.argtypes = [F, dp]
        L.o_fr_total_energy.restype = C.c_double
        L.o_fr_hllc.argtypes = [F, dp, dp, dp, C.c_double, dp]
        L.o_fr_hllc.restype = C.c_int
        L.o_fr_max_eig.argtypes = [F, dp, dp, C.c_double]
        L.o_fr_max_eig.restype = C.c_double
        L.o_fr_eigensystem.argtypes = [F, dp, dp, C.c_double, dp, dp, dp]
        L.o_fr_bc_state.argtypes = [F, C.c_int, dp, dp, dp, dp, C.c_double]
        L.o_fr_apply_dq.argtypes = [F, dp, dp]
        L.o_fr_temporal.argtypes = [F, dp, C.c_double, C.c_double, C.c_double, dp]
        L.o_fr_qinf.argtypes = [F, dp, C.c_double, C.c_double, dp, dp]
        L.o_fr_update_bcs.argtypes = [F, M, dp, dp]
        L.o_fr_residual.argtypes = [F, M, dp, dp, dp]
        L.o_fr_residual.restype = C.c_double
        L.o_fr_timestep.argtypes = [F, M, dp, C.c_double, dp]
        L.o_fr_timestep.restype = C.c_double
        L.o_fr_crs_init.argtypes = [M, C.c_int, R]
        L.o_fr_crs_free.argtypes = [R]
        L.o_fr_jacobian.argtypes = [F, M, dp, dp, dp, R]
        L.o_fr_sgs.argtypes = [R, C.c_int, dp, dp, ip]
        L.o_fr_sgs.restype = C.c_double
        L.o_fr_solver_init.argtypes = [C.POINTER(OFrSolver), M, F, dp, dp, C.c_double, C.c_double, C.c_int,
                                       C.c_int, C.c_int]
        L.o_fr_step.argtypes = [C.POINTER(OFrSolver), dp, ip]
        L.o_fr_step.restype = C.c_double
        L.o_fr_solver_free.argtypes = [C.POINTER(OFrSolver)]
        _FR_BOUND = True
    return L


class FR:
    """Oracle view of the finite-rate equation set on one chemistry model.

    ref: dict with ref_pressure, ref_velocity, ref_density, ref_time,
    ref_temperature (Param::PostCompute); Pref the freestream pressure
    (non-dimensional); beta the preconditioning parameter."""

    def __init__(self, model, ref, Pref=0.0, beta=1.0, sym2d=0, rxn_on=True, local_dt=True, dt_param=-1.0,
                 no_cvbc=False):
        self.model = model.with_reference(ref["ref_density"], ref["ref_time"], ref["ref_temperature"])
        self.chem = Chem(self.model)
        self.ns = self.model.ns
        self.neqn, self.nvars = self.ns + 4, 3 * self.ns + 6
        self.f = OFr(C.pointer(self.chem.c), float(ref["ref_pressure"]), float(ref["ref_velocity"]), float(Pref),
                     float(beta), int(sym2d), int(rxn_on), int(local_dt), float(dt_param), int(no_cvbc))
        _fr_lib()

    @property
    def beta(self):
        return self.f.beta

    def aux(self, Q):
        """Rows of nvars (or neqn) values -> rows of nvars with aux recomputed."""
        Q = np.atleast_2d(_f64(Q))
        out = np.zeros((Q.shape[0], self.nvars))
        out[:, :self.neqn] = Q[:, :self.neqn]
        for r in out:
            _fr_lib().o_fr_aux_q(C.byref(self.f), _d(r))
        return out

    def qinf(self, Y, p_dim, T_dim, vel):
        q = np.zeros(self.nvars)
        _fr_lib().o_fr_qinf(C.byref(self.f), _d(_f64(Y)), float(p_dim), float(T_dim), _d(_f64(vel)), _d(q))
        return q

    def hllc(self, QL, QR, avec, beta=None):
        F = np.zeros(self.neqn)
        z = _fr_lib().o_fr_hllc(C.byref(self.f), _d(_f64(QL)), _d(_f64(QR)), _d(_f64(avec)),
                                self.beta if beta is None else float(beta), _d(F))
        return F, z

    def max_eig(self, Q, avec, beta=None):
        return _fr_lib().o_fr_max_eig(C.byref(self.f), _d(_f64(Q)), _d(_f64(avec)),
                                      self.beta if beta is None else float(beta))

    def eigensystem(self, Q, avec, beta=None):
        n = self.neqn
        eig, T, Ti = np.zeros(n), np.zeros(n * n), np.zeros(n * n)
        _fr_lib().o_fr_eigensystem(C.byref(self.f), _d(_f64(Q)), _d(_f64(avec)),
                                   self.beta if beta is None else float(beta), _d(eig), _d(T), _d(Ti))
        return eig, T.reshape(n, n), Ti.reshape(n, n)

    def total_enthalpy(self, Q):
        return _fr_lib().o_fr_total_enthalpy(C.byref(self.f), _d(_f64(Q)))

    def total_energy(self, Q):
        return _fr_lib().o_fr_total_energy(C.byref(self.f), _d(_f64(Q)))

    def bc_state(self, bctype, QL, QR, Qref, avec, beta=None):
        QL, QR = _f64(QL).copy(), _f64(QR).copy()
        _fr_lib().o_fr_bc_state(C.byref(self.f), int(bctype), _d(QL), _d(QR), _d(_f64(Qref)), _d(_f64(avec)),
                                self.beta if beta is None else float(beta))
        return QL, QR

    def apply_dq(self, dQ, Q):
        Q = _f64(Q).copy()
        _fr_lib().o_fr_apply_dq(C.byref(self.f), _d(_f64(dQ)), _d(Q))
        return Q

    def temporal(self, Q, vol, dtau, beta=None):
        A = np.zeros(self.neqn * self.neqn)
        _fr_lib().o_fr_temporal(C.byref(self.f), _d(_f64(Q)), float(vol), float(dtau),
                                self.beta if beta is None else float(beta), _d(A))
        return A.reshape(self.neqn, self.neqn)


class FRProblem:
    """Mesh + BCs + FR equation set; qref holds nvars values per face tag."""

    def __init__(self, mesh, bc_type, qref, fr: FR):
        nt = len(bc_type)
        self.fr = fr
        self.keep = dict(edges=_i32(mesh.edges), edge_a=_f64(mesh.edge_a), bedges=_i32(mesh.bedges),
                         bedge_a=_f64(mesh.bedge_a), bfactag=_i32(mesh.bfactag), xyz=_f64(mesh.xyz),
                         vol=_f64(mesh.vol), bc_type=_i32(bc_type), qref10=np.zeros((nt, 10)),
                         twall=np.full(nt, -1.0), bnormal=np.full(mesh.nbedge, -1, dtype=np.int32),
                         qref=_f64(np.asarray(qref).reshape(nt, fr.nvars)))
        k = self.keep
        self.nnode, self.nbnode, self.nedge = mesh.nnode, mesh.nbedge, mesh.nedge
        self.m = OMesh(mesh.nnode, mesh.nbedge, mesh.nedge, _i(k["edges"]), _d(k["edge_a"]), _i(k["bedges"]),
                       _d(k["bedge_a"]), _i(k["bfactag"]), _d(k["xyz"]), _d(k["vol"]), nt,
                       _i(k["bc_type"]), _d(k["qref10"]), _d(k["twall"]), _i(k["bnormal"]))

    def update_bcs(self, q):
        q = _f64(q).copy()
        _fr_lib().o_fr_update_bcs(C.byref(self.fr.f), C.byref(self.m), _d(self.keep["qref"]), _d(q))
        return q

    def residual(self, q):
        b = np.zeros((self.nnode, self.fr.neqn))
        norms = np.zeros(1 + self.fr.neqn)
        _fr_lib().o_fr_residual(C.byref(self.fr.f), C.byref(self.m), _d(_f64(q)), _d(b), _d(norms))
        return b, norms

    def timestep(self, q, cfl):
        dt = np.zeros(self.nnode)
        dtmin = _fr_lib().o_fr_timestep(C.byref(self.fr.f), C.byref(self.m), _d(_f64(q)), float(cfl), _d(dt))
        return dt, dtmin

    def jacobian(self, q, dt):
        """(ia, ja, iau, M[nblocks, neqn, neqn], q after the boundary Jacobian)."""
        n = self.fr.neqn
        A = OFrCrs()
        _fr_lib().o_fr_crs_init(C.byref(self.m), n, C.byref(A))
        q = _f64(q).copy()
        _fr_lib().o_fr_jacobian(C.byref(self.fr.f), C.byref(self.m), _d(self.keep["qref"]), _d(q),
                                _d(_f64(dt)), C.byref(A))
        nb = A.nblocks
        ia = np.ctypeslib.as_array(A.ia, (self.nnode + 1,)).copy()
        ja = np.ctypeslib.as_array(A.ja, (nb,)).copy()
        iau = np.ctypeslib.as_array(A.iau, (self.nnode,)).copy()
        Mm = np.ctypeslib.as_array(A.M, (nb * n * n,)).copy().reshape(nb, n, n)
        _fr_lib().o_fr_crs_free(C.byref(A))
        return ia, ja, iau, Mm, q


def fr_sgs_on_matrix(ia, ja, iau, M, b, nsgs, order=None):
    """PrepareSGS + SGS (x0 = 0) on a given neqn x neqn block-CRS matrix;
    returns (x, |xNorm - xNormOld|)."""
    ia, ja, iau = _i32(ia), _i32(ja), _i32(iau)
    M = _f64(M)
    n = M.shape[-1]
    Mc = M.copy().reshape(-1)
    nn = len(ia) - 1
    pv = np.zeros(n * nn, dtype=np.int32)
    A = OFrCrs(nn, len(ja), n, _i(ia), _i(ja), _i(iau), None, None, _i(pv), _d(Mc))
    x = np.zeros((nn, n))
    o = None if order is None else _i32(order)
    d = _fr_lib().o_fr_sgs(C.byref(A), nsgs, _d(_f64(b)), _d(x), None if o is None else _i(o))
    return x, d


class FRSolver:
    """The oracle's first-order FR Newton loop (PreTimeAdvance + NewtonIterate)."""

    def __init__(self, pb: FRProblem, q0, cfl, cfl_start, cfl_ramp, nsgs, jac_freq=1, order=None, ja=None):
        """ja: optional block-column order per row (same pattern, diagonal
        first) -- the reference's CRS keeps each row's neighbours in psp
        order, and SGS sums a row's products in that order."""
        self.pb = pb
        self.s = OFrSolver()
        self._q0 = _f64(q0).copy()
        self._order = None if order is None else _i32(order)
        _fr_lib().o_fr_solver_init(C.byref(self.s), C.byref(pb.m), C.byref(pb.fr.f), _d(pb.keep["qref"]),
                                   _d(self._q0), float(cfl), float(cfl_start), int(cfl_ramp), int(nsgs),
                                   int(jac_freq))
        if self._order is not None:
            self.s.order = _i(self._order)
        if ja is not None:
            ja = _i32(ja)
            cur = np.ctypeslib.as_array(self.s.A.ja, (self.s.A.nblocks,))
            ia = np.ctypeslib.as_array(self.s.A.ia, (pb.nnode + 1,))
            for i in range(pb.nnode):
                a, b = ia[i], ia[i + 1]
                if ja[a] != i or sorted(ja[a:b]) != sorted(cur[a:b]):
                    raise ValueError(f"ja row {i} is not a reordering of the mesh's row")
            cur[:] = ja

    def step(self):
        norms = np.zeros(1 + self.pb.fr.neqn)
        nz = C.c_int(0)
        r = _fr_lib().o_fr_step(C.byref(self.s), _d(norms), C.byref(nz))
        return r, norms, nz.value

    @property
    def q(self):
        n, nv = self.pb.nnode + self.pb.nbnode, self.pb.fr.nvars
        return np.ctypeslib.as_array(self.s.q, (n * nv,)).copy().reshape(n, nv)

    @property
    def iter(self):
        return self.s.iter

    def close(self):
        if self.s is not None:
            _fr_lib().o_fr_solver_free(C.byref(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
