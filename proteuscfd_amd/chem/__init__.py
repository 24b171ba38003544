"""Finite-rate air chemistry model data (the model ucs.x builds from
chemModels/5speciesAir.rxn and its Burcat-derived thermo database).

air5.json holds the species (MW in g/mol as in the database, NASA-7
low/high coefficients) and the reactions with their local species lists in
the reference parser's order; make_air5.py regenerates it from the
reference's data files.  ChemModel turns it into the flat arrays of
pcfd_chem_desc (include/pcfd_chem.h).
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
UNIV_R = 8.31447215  # J/(mol K), ucs/chem_constants.h:5
MAXLOC = 8


@dataclass
class ChemModel:
    symbols: list
    mw: np.ndarray        # [ns] kg/mol
    lo: np.ndarray        # [ns, 7]
    hi: np.ndarray        # [ns, 7]
    rtype: np.ndarray     # [nr]
    tb: np.ndarray        # [nr]
    A: np.ndarray
    EA: np.ndarray
    n: np.ndarray
    nloc: np.ndarray      # [nr]
    loc: np.ndarray       # [nr, MAXLOC] global species index (-1 unused)
    nup: np.ndarray       # [nr, MAXLOC]
    nupp: np.ndarray
    eff: np.ndarray
    ref_density: float = 1.0
    ref_time: float = 1.0
    ref_temperature: float = 1.0

    @property
    def ns(self):
        return len(self.symbols)

    @property
    def nr(self):
        return len(self.A)

    def with_reference(self, ref_density, ref_time, ref_temperature) -> "ChemModel":
        d = dict(self.__dict__)
        d.update(ref_density=ref_density, ref_time=ref_time, ref_temperature=ref_temperature)
        return ChemModel(**d)


def load_air5() -> ChemModel:
    with open(os.path.join(HERE, "air5.json")) as f:
        d = json.load(f)
    sp, rx = d["species"], d["reactions"]
    nr = len(rx)
    loc = -np.ones((nr, MAXLOC), dtype=np.int32)
    nup = np.zeros((nr, MAXLOC))
    nupp = np.zeros((nr, MAXLOC))
    eff = np.ones((nr, MAXLOC))
    for r, x in enumerate(rx):
        k = len(x["species"])
        loc[r, :k] = x["species"]
        nup[r, :k] = x["nup"]
        nupp[r, :k] = x["nupp"]
        eff[r, :k] = x["tb_eff"]
    return ChemModel(
        symbols=[s["symbol"] for s in sp],
        mw=np.array([s["MW"] for s in sp]) / 1000.0,  # g/mol -> kg/mol (ucs/species.tcc:167-171)
        lo=np.array([s["lo"] for s in sp]), hi=np.array([s["hi"] for s in sp]),
        rtype=np.array([x["type"] for x in rx], dtype=np.int32),
        tb=np.array([x["third_body"] for x in rx], dtype=np.int32),
        A=np.array([x["A"] for x in rx]), EA=np.array([x["EA"] for x in rx]), n=np.array([x["n"] for x in rx]),
        nloc=np.array([len(x["species"]) for x in rx], dtype=np.int32), loc=loc, nup=nup, nupp=nupp, eff=eff)


def mass_fractions(model: ChemModel, **Y) -> np.ndarray:
    """Mass-fraction vector in the model's species numbering, e.g.
    mass_fractions(m, N2=0.767, O2=0.233); species not named get 0."""
    out = np.zeros(model.ns)
    for sym, y in Y.items():
        out[list(model.symbols).index(sym)] = y
    return out


def freestream_scales(model: ChemModel, T_inf: float, rho_inf: float, Y) -> tuple[float, float, float]:
    """Reference density, time and temperature of the Mach-based
    non-dimensionalisation (Param::PostCompute, ucs/param.tcc:354-400) for a
    freestream (T, rho, mass fractions), with the reference pressure set to
    the freestream pressure and length 1 m: u_ref = c_inf, rho_ref =
    p_inf / c_inf^2, t_ref = 1 / c_inf, T_ref = T_inf (so T_hat = 1)."""
    Y = np.asarray(Y, dtype=float)
    R = float(np.sum(Y * UNIV_R / model.mw))
    cp = 0.0
    for s in range(model.ns):
        a = model.hi[s] if T_inf > 1000.0 else model.lo[s]
        cp += Y[s] * (a[0] + T_inf * (a[1] + T_inf * (a[2] + T_inf * (a[3] + T_inf * a[4])))) * UNIV_R / model.mw[s]
    gamma = cp / (cp - R)
    p = rho_inf * R * T_inf
    c = np.sqrt(gamma * R * T_inf)
    return p / (c * c), 1.0 / c, T_inf


R_UNIV_PARAM = 8.314459848  # J/(mol K), the gas constant of the case parameters (ucs/param.tcc:13)


def fr_case_scales(p0: float, T0: float, ref_pressure: float, MW: float = 28.966, Cp: float = 1006.43,
                   ref_length: float = 1.0) -> dict:
    """Reference scales of an FR case as Param::PostCompute derives them
    (ucs/param.tcc:355-400) from initialPressure p0 (Pa), initialTemperature
    T0 (K), refPressure and the parameter gas (MW g/mol, Cp J/(kg K)):
    velocity by that gas's sound speed, density by p_ref / u_ref^2, time by
    L / u_ref, and the temperature for which the gas's non-dimensional c^2
    equals its non-dimensional T."""
    Rs = R_UNIV_PARAM / (MW / 1000.0)
    Cv = Cp - Rs
    gamma = Cp / Cv
    ref_velocity = np.sqrt(gamma * Rs * T0)
    ref_density = ref_pressure / (ref_velocity * ref_velocity)
    ref_time = ref_length / ref_velocity
    rho_d = p0 / (Rs * T0)
    P_nd = p0 / ref_pressure
    rho_nd = rho_d / ref_density
    c2_nd = gamma * P_nd / rho_nd
    return dict(ref_pressure=float(ref_pressure), ref_velocity=float(ref_velocity), ref_density=float(ref_density),
                ref_time=float(ref_time), ref_temperature=float(T0 / c2_nd))


def fr_freestream(model: ChemModel, scales: dict, Y, p0: float, T0: float, velocity) -> tuple[np.ndarray, float]:
    """Qinf (ns + 4 native values) and Pref = P(Qinf) as UpdateQinf sets them
    (ucs/compressibleFR.tcc:1465-1518): the pressure split by mass fraction,
    rho_i = Y_i sum_j (p0 Y_j / (R_j T0)) / rho_ref, velocity non-dimensional
    (flow direction times the Mach number), T = T0 / T_ref."""
    Y = np.asarray(Y, dtype=float)
    ns = model.ns
    R = UNIV_R / model.mw
    rDim = 0.0
    for i in range(ns):
        rDim += p0 * Y[i] / (R[i] * T0)
    r = rDim / scales["ref_density"]
    q = np.zeros(ns + 4)
    q[:ns] = Y * r
    q[ns:ns + 3] = velocity
    q[ns + 3] = T0 / scales["ref_temperature"]
    TDim = q[ns + 3] * scales["ref_temperature"]
    P = 0.0
    for i in range(ns):
        P += (q[i] * scales["ref_density"]) * R[i] * TDim
    return q, P / scales["ref_pressure"]
