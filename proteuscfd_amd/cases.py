"""Benchmark cases: meshes, boundary conditions, freestream states, parameters.

Boundary-condition ids are the reference's BCTypes enum (ucs/bc_defines.h:4-30).
The freestream is the reference's non-dimensional Q_inf
(ucs/compressible.tcc:823-867): rho, rho*u, rho*v, rho*w, E = p/(g-1)+rho|u|^2/2
with the auxiliary set T = g*p/rho, p, u, v, w (ucs/compressible.tcc:1229-1243).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh as M

# ucs/bc_defines.h:4-30
BC_PARALLEL = 0
BC_DIRICHLET = 1
BC_NEUMANN = 2
BC_IMPERMEABLE = 3
BC_NOSLIP = 4
BC_FARFIELD = 6
BC_SONIC_INFLOW = 7
BC_SONIC_OUTFLOW = 8
BC_SYMMETRY = 9
BC_ISOTHERMAL = 18

NEQN = 5
NVARS = 10


def qinf(gamma: float, mach: float, rho: float = 1.0, p: float | None = None,
         direction=(1.0, 0.0, 0.0)) -> np.ndarray:
    """Freestream conservative + auxiliary state (10 values)."""
    if p is None:
        p = 1.0 / gamma
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    u, v, w = mach * d
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    q = np.zeros(NVARS)
    q[:5] = [rho, rho * u, rho * v, rho * w, E]
    return aux(q, gamma)


def aux(q: np.ndarray, gamma: float) -> np.ndarray:
    """ComputeAuxiliaryVariables on a (..., 10) array (in place, returned)."""
    q = np.asarray(q)
    r = q[..., 0]
    u, v, w = q[..., 1] / r, q[..., 2] / r, q[..., 3] / r
    P = (gamma - 1.0) * (q[..., 4] - 0.5 * r * (u * u + v * v + w * w))
    q[..., 6] = P
    q[..., 5] = gamma * P / r
    q[..., 7], q[..., 8], q[..., 9] = u, v, w
    return q


@dataclass
class Params:
    """Solver parameters (mirrors ucs/param.tcc keywords)."""
    gamma: float = 1.4
    sorder: int = 1               # spatialOrder
    nfirst: int = 0               # firstOrderSteps
    limiter: int = 0              # 0 off, 1 Barth, 2 Venkat, 3 modified Venkat
    venkat_K: float = 1.0
    limiter_refresh: int = 1
    limiter_freeze: int = 999999
    cfl: float = 5.0
    cfl_start: float = 0.5
    cfl_ramp: int = 0
    nsgs: int = 5                 # numberSGS (0 = explicit)
    symmetry2d: int = -1
    jac_freq: int = 1
    precision: int = 0            # 0 plain FP64, 1 FP32 Jacobian + SGS
    # laminar Navier-Stokes (CompressibleNS with the Laminar model, mut = 0)
    viscous: int = 0
    reynolds: float = 1.0         # Re (ucs/compressible.tcc:85-89)
    mach_ref: float = 1.0         # freestream velocity / c_inf (Param::GetVelocity)
    prandtl: float = 0.72         # prandtlNumber
    prandtl_t: float = 0.85       # PrT (ucs/param.tcc:89)
    sutherland_s: float = 110.4 / 300.0   # 110.4 K / T_ref (ucs/eqnset.h:251-268)
    # options off by default in the reference
    gradient_type: int = 0        # gradientType: 0 weighted LSQ, 1 Green-Gauss
    global_dt: int = 0            # 1: localTimeStepping off (one dt everywhere)
    fixed_dt: float = -1.0        # timeStep (used with global_dt when > 0)
    enable_vnn: int = 0           # enableVNN
    vnn: float = 20.0             # VNN
    dynamic_cfl: int = 0          # dynamicCFL
    jac_field_type: int = 0       # jacobianFieldType: 0 one-sided, 1 centred
    flux_type: int = 0            # 0 Roe (roeFlux), 1 algebraic (algebraicFlux)
    jac_boundary_type: int = 0    # jacobianBoundaryType: 0 one-sided, 1 centred
    rk_stages: int = 1            # explicit path: 1 the reference's single stage, 4 Jameson's four stages

    @property
    def re_tilde(self) -> float:  # Re / Mach (ucs/compressible.tcc:753-755)
        return self.reynolds / self.mach_ref if self.viscous else 1.0

    @property
    def chi(self) -> float:       # ucs/param.tcc:410-417
        return 0.5 if self.sorder >= 3 else 0.0


@dataclass
class Case:
    name: str
    mesh: "M.Mesh"
    bc_type: np.ndarray           # [nfactag] int32
    qref: np.ndarray              # [nfactag, 10]
    qinf: np.ndarray              # [10]
    params: Params
    extra: dict = field(default_factory=dict)
    twall: np.ndarray | None = None   # [nfactag] non-dimensional noSlip wall temperature (<0 adiabatic)

    def initial_state(self) -> np.ndarray:
        """Uniform freestream on interior and phantom nodes (AoS, 10 vars)."""
        n = self.mesh.nnode + self.mesh.nbedge
        return np.tile(self.qinf, (n, 1)).copy()


def wedge_case(nx=60, ny=30, nz=2, seed=1, jitter=0.05, mach=2.5, gamma=1.4, rcm=False,
               params: Params | None = None, z_bc=BC_SYMMETRY) -> Case:
    """Supersonic 15-degree wedge channel: sonic inflow (x=0), sonic outflow
    (x=3), slip walls (y), symmetry (z)."""
    m = M.Mesh.build(M.WEDGE, nx=nx, ny=ny, nz=nz, seed=seed, jitter=jitter, rcm=int(rcm))
    q0 = qinf(gamma, mach)
    bct = np.array([BC_PARALLEL, BC_SONIC_INFLOW, BC_SONIC_OUTFLOW, BC_IMPERMEABLE,
                    BC_IMPERMEABLE, z_bc, z_bc], dtype=np.int32)
    qref = np.tile(q0, (M.NFACTAG, 1))
    if params is None:
        params = Params(gamma=gamma, sorder=1, cfl=0.9, cfl_start=0.9, nsgs=0)
    return Case("wedge", m, bct, qref, q0, params)


def bump_case(n=100, seed=2, jitter=0.03, mach=0.5, gamma=1.4, rcm=True,
              params: Params | None = None) -> Case:
    """Subsonic bump in a unit cube: farfield on all sides, slip floor."""
    m = M.Mesh.build(M.BUMP, nx=n, ny=n, nz=n, seed=seed, jitter=jitter, rcm=int(rcm))
    q0 = qinf(gamma, mach)
    bct = np.array([BC_PARALLEL, BC_FARFIELD, BC_FARFIELD, BC_FARFIELD, BC_FARFIELD,
                    BC_IMPERMEABLE, BC_FARFIELD], dtype=np.int32)
    qref = np.tile(q0, (M.NFACTAG, 1))
    if params is None:
        params = Params(gamma=gamma, sorder=2, limiter=3, venkat_K=1.0, cfl=50.0, cfl_start=1.0,
                        cfl_ramp=50, nsgs=8)
    return Case("bump", m, bct, qref, q0, params)


def plate_case(nx=100, ny=20, nz=40, nlayers=20, first_h=1e-4, growth=1.15, seed=3, jitter=0.0, mach=0.3,
               reynolds=1.0e4, twall=1.0, gamma=1.4, rcm=True, incidence_deg=0.0, outflow_bc=BC_FARFIELD,
               top_bc=BC_FARFIELD, params: Params | None = None) -> Case:
    """Laminar flat-plate channel: prism layers on a noSlip isothermal floor
    (twall = T_wall / T_ref, 1 = freestream temperature) under a Kuhn-tet
    core; farfield inflow, outflow and top, symmetry side planes.  Re is
    per unit length; T_ref = 300 K sets Sutherland's constant.
    incidence_deg tilts the freestream towards +z: with 0 the flow is exactly
    tangential to the top farfield face, where the characteristic branch
    (inflow / outflow, theta == 0) is decided by the sign of round-off.
    outflow_bc sets the x-max face and top_bc the z-max face; the reference farfield returns early
    when the averaged normal velocity is exactly 0, which happens at the
    outflow corner of the no-slip floor (zero wall velocity against a
    phantom whose velocity the fixed point drove to 0) unless round-off
    leaves a few 1e-17 -- sonic outflow avoids that branch point."""
    m = M.Mesh.build(M.PLATE, nx=nx, ny=ny, nz=nz, nlayers=nlayers, first_h=first_h, growth=growth, seed=seed,
                     jitter=jitter, rcm=int(rcm))
    a = np.deg2rad(incidence_deg)
    q0 = qinf(gamma, mach, direction=(np.cos(a), 0.0, np.sin(a)))
    bct = np.array([BC_PARALLEL, BC_FARFIELD, outflow_bc, BC_SYMMETRY, BC_SYMMETRY, BC_NOSLIP, top_bc],
                   dtype=np.int32)
    qref = np.tile(q0, (M.NFACTAG, 1))
    tw = np.full(M.NFACTAG, -1.0)
    tw[5] = twall
    if params is None:
        params = Params(gamma=gamma, sorder=2, limiter=0, cfl=20.0, cfl_start=1.0, cfl_ramp=50, nsgs=8,
                        viscous=1, reynolds=reynolds, mach_ref=mach)
    return Case("plate", m, bct, qref, q0, params, twall=tw)


def file_case(path: str, bc: dict, mach: float = 0.5, alpha_deg: float = 0.0, beta_deg: float = 0.0,
              gamma: float = 1.4, rcm: bool = True, params: Params | None = None, twall: dict | None = None,
              name: str | None = None) -> Case:
    """A case on a mesh file (GMSH / UGRID / FieldView / SU2, read by
    pcfd_mesh_read).  bc maps every face tag in the file to a BC type id
    (BCTypes, ucs/bc_defines.h); the freestream has angle of attack alpha
    (x-z plane) and sideslip beta; twall maps noSlip tags to T_wall/T_ref."""
    m = M.Mesh.read(path, rcm=rcm)
    tags = set(int(t) for t in np.unique(m.bface_tag))
    missing = tags - set(bc)
    if missing:
        raise ValueError(f"{path}: no boundary condition for face tags {sorted(missing)}")
    a, b = np.deg2rad(alpha_deg), np.deg2rad(beta_deg)
    q0 = qinf(gamma, mach, direction=(np.cos(a) * np.cos(b), np.sin(b), np.sin(a) * np.cos(b)))
    bct = np.full(m.nfactag, BC_FARFIELD, dtype=np.int32)
    bct[0] = BC_PARALLEL
    for t, v in bc.items():
        if not 0 < t < m.nfactag:
            raise ValueError(f"face tag {t} is outside 1..{m.nfactag - 1}")
        bct[t] = v
    qref = np.tile(q0, (m.nfactag, 1))
    tw = np.full(m.nfactag, -1.0)
    for t, v in (twall or {}).items():
        tw[t] = v
    if params is None:
        params = Params(gamma=gamma, sorder=2, limiter=3, cfl=20.0, cfl_start=1.0, cfl_ramp=20, nsgs=8)
    return Case(name or path, m, bct, qref, q0, params, twall=tw)


def perturbed_state(case: Case, amp: float = 0.05, seed: int = 7) -> np.ndarray:
    """Freestream plus smooth and random perturbations (for parity tests)."""
    rng = np.random.default_rng(seed)
    m = case.mesh
    g = case.params.gamma
    q = case.initial_state()
    x = np.concatenate([m.xyz, np.zeros((m.nbedge, 3))], axis=0)
    n = q.shape[0]
    smooth = np.sin(2.1 * x[:, 0] + 0.3) * np.cos(1.7 * x[:, 1]) + 0.5 * np.sin(3.3 * x[:, 2])
    rho = q[:, 0] * (1.0 + amp * smooth + 0.2 * amp * rng.uniform(-1, 1, n))
    vel = q[:, 1:4] / q[:, :1] * (1.0 + amp * rng.uniform(-1, 1, (n, 3)))
    vel[:, 1] += amp * 0.3 * smooth
    p = q[:, 6] * (1.0 + amp * np.cos(1.3 * x[:, 0] - 0.4 * x[:, 2]) + 0.2 * amp * rng.uniform(-1, 1, n))
    out = np.zeros_like(q)
    out[:, 0] = rho
    out[:, 1:4] = rho[:, None] * vel
    out[:, 4] = p / (g - 1.0) + 0.5 * rho * (vel ** 2).sum(axis=1)
    return aux(out, g)
