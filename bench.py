"""Benchmark: time to an 8-order residual drop, chosen form (FP32 Jacobian +
SGS, FP64 residual/gradients/norms) vs the plain FP64 form, on the 8M-cell
laminar flat plate (the headline line: the case the chosen form's 2x target
is quoted on), followed by the same two solves of the 1.03M-node
second-order Euler bump as a "bump_euler" record (--no-extra skips it;
--case bump makes the bump the headline) and a finite-rate reacting Mach 8
wedge record (--no-fr skips it; --only-fr runs it alone), on one MI355X.

A "step" here is one complete steady solve from freestream until
residGlobal <= 1e-8 * residGlobal(iteration 1); --steps K times K solves of
each form, --warmup W runs W untimed implicit iterations first.  Prints one
JSON line (rank 0).  Per-kernel device times come from HIP events on the
context stream (pcfd_phase_times); the CPU baseline times the reference
solver itself (its checker build under oracle/_ref; the oracle where that
is absent) on the host's cores on a bounded sample (see
cpu_baseline.sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to 8-order residual drop, chosen vs plain FP64 (s); edge fluxes/s; HBM%"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(m, nblocks, prec_bytes):
    nn, ne, nb = m.nnode, m.nedge, m.nbedge
    return {
        # compulsory bytes per launch (indices int32, state FP64)
        "residual_2nd": 264 * nn + 40 * ne + 84 * nb,
        "residual_1st": 80 * nn + 40 * ne + 84 * nb,
        # the viscous pass that continues the residual's node sums: q, the
        # T, u, v, w gradient terms, xyz and b read, b written per node
        "viscous": 240 * nn + 40 * ne + 84 * nb,
        "sgs_sweep": nblocks * 25 * prec_bytes + 8 * 15 * nn,
        "jacobian": nblocks * 25 * prec_bytes + 80 * nn + 40 * ne,
        "gradient9": nn * 8 * (10 + 6 + 3 + 27) + 8 * ne,
        "limiter": nn * 8 * (5 + 15 + 3 + 1 + 5) + 8 * ne,
    }


def sgs_pass_bytes(m, nsgs, prec_bytes):
    """Bytes the colour passes of one pcfd_sgs call must move, and the pass
    counts.  A pass over colour c reads its rows' blocks (diagonal + one per
    neighbour, 25 values each) and moves 15 values per row (b, the
    neighbours' x, the new x), all in the storage precision.  The library
    skips a pass that directly follows a pass over the same colour (it would
    rewrite the same x), so nsgs symmetric sweeps over C colours run
    2*nsgs*C - (2*nsgs - 1) passes."""
    col, nc = m.colour()
    deg = np.bincount(np.asarray(m.edges).ravel(), minlength=m.nnode)
    blocks = np.bincount(col, weights=deg + 1.0, minlength=nc)
    rows = np.bincount(col, minlength=nc).astype(float)
    seq = []
    for _ in range(nsgs):
        seq += list(range(nc)) + list(range(nc - 1, -1, -1))
    run = [c for k, c in enumerate(seq) if k == 0 or seq[k - 1] != c]
    nbytes = sum(blocks[c] * 25 * prec_bytes + rows[c] * 15 * prec_bytes for c in run)
    return float(nbytes), len(run), len(seq)


def phase_profile(ctx, case, iters, jac_freq=1):
    """Run `iters` implicit iterations with per-phase HIP-event timing: ms
    per iteration, and for the Jacobian also ms per assembly (iterations
    0, 1 and the multiples of jac_freq assemble it)."""
    ctx.upload_q(case.initial_state())
    ctx.reset_iteration()
    ctx.enable_phase_timing(True)
    ctx.phase_times(reset=True)
    for _ in range(iters):
        ctx.iterate()
    pt = ctx.phase_times(reset=True)
    ctx.enable_phase_timing(False)
    out = {k: v / iters for k, v in pt.items()}
    builds = sum(1 for it in range(iters) if it % jac_freq == 0 or it == 1)
    out["jacobian_builds"] = builds
    out["jacobian_per_build"] = pt["jacobian"] / builds
    out["k_jacobian_per_build"] = pt["k_jacobian"] / builds
    return out


def timed_solves(ctx, case, k, max_iter, drop):
    times, iters, hists = [], [], []
    for _ in range(k):
        ctx.upload_q(case.initial_state())
        ctx.reset_iteration()
        ctx.synchronize()
        hist, nit, secs = ctx.solve(max_iter, drop)
        times.append(secs)
        iters.append(nit)
        hists.append(hist)
    return times, iters, hists


def measured_traffic(kernel, tag):
    """Average HBM bytes per dispatch of `kernel` on workload `tag`, measured
    by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size) and recorded in profiles/k_sgs_color_<tag>_traffic.txt;
    None when absent."""
    path = os.path.join(ROOT, "profiles", f"k_sgs_color_{tag}_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] == f"traffic_bytes_per_launch[{kernel}]":
                return float(parts[1])
    return None


def copy_calibration(nbytes):
    """A plain device copy of known size (torch's copy kernel), run first so
    a counter pass can scale FETCH_SIZE/WRITE_SIZE to bytes."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {nbytes} bytes read + {nbytes} bytes written per copy")
    del a, b


def cpu_iteration(case):
    """One representative implicit iteration of the oracle (single-threaded
    FP64 restatement of the reference) on `case`: the second iteration (the
    first runs at first order), timed per phase."""
    import oracle as O
    pb = O.Problem.from_case(case)
    s = O.Solver(pb, case.initial_state())
    s.step()
    t0 = s.timers()
    t = time.time()
    s.step()
    it_s = time.time() - t
    t1 = s.timers()
    s.close()
    return it_s, {k: t1[k] - t0[k] for k in t1}


def reference_keys(kind, base):
    """The workload as the reference's .param keywords (ucs/param.tcc) and
    .bc surfaces: BC types per face tag, and the noSlip wall temperature."""
    k = {"timeOrder": 0, "reorderMesh": 0, "refPressure": 101325, "flowDirection": "[1.0, 0.0, 0.0]",
         "spatialOrder": base["sorder"], "limiter": base["limiter"], "numberSGS": base["nsgs"], "CFL": base["cfl"],
         "startingCFL": base["cfl_start"], "rampCFL": base["cfl_ramp"]}
    if kind == "plate":
        # refViscosity gives Re ~ 1e4 per unit length at Mach 0.3 and 300 K
        k.update(equationSet="compressibleNS", initialTemperature=300, refViscosity=0.01225, velocity=0.3)
        return k, {1: 6, 2: 6, 3: 9, 4: 9, 5: 4, 6: 9}, {5: " twall = [300.0]"}
    k.update(equationSet="compressibleEuler", velocity=0.5)
    return k, {1: 6, 2: 6, 3: 6, 4: 6, 5: 3, 6: 6}, {}


def cpu_worker(args):
    """--cpu-worker: one rank's share of the CPU baseline (a partition-sized
    mesh of the workload), printed as one JSON line: the reference solver
    itself (oracle/_ref/ref_harness, iterations of its own steady loop timed
    on the wall clock) or the oracle."""
    case, base, _ = make_case(args.case, args, parts=args.cpu_parts)
    if args.cpu_kind == "reference":
        from oracle import ref_case
        keys, bc, extra = reference_keys(args.case, base)
        d = ref_case.write_case(case.mesh, keys, bc, nsteps=args.cpu_iters, extra=extra)
        ref_case.run(d, "c", "time", args.cpu_iters)
        secs = ref_case.read(d, "iter_seconds")
        print(json.dumps({"s": float(secs[-1]), "iter_s": [float(x) for x in secs], "nnode": case.mesh.nnode}))
        return
    it_s, tm = cpu_iteration(case)
    print(json.dumps({"s": it_s, "phases": tm, "nnode": case.mesh.nnode}))


def cpu_baseline(args, n_iter_plain):
    """The reference's CPU layout: one MPI rank per core, each iterating its
    own partition.  Estimated by running one process per core at once (so
    they share the host's memory bandwidth), each on a mesh of the workload
    cut to 1/cores of its size in the two directions along the wall: the
    reference solver itself where its checker build (oracle/_ref) is present
    (the last of a few iterations of its own loop timed), else the oracle
    (one iteration); the slowest rank sets the iteration time, times the
    plain GPU path's iteration count.  Halo exchanges are not modelled."""
    import subprocess
    cores = max(1, min(16, os.cpu_count() or 1))

    def run_workers(kind):
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", "--case", args.case, "--n", str(args.n),
               "--plate", args.plate, "--cpu-parts", str(cores), "--cpu-kind", kind]
        procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
                 for _ in range(cores)]
        outs = []
        for pr in procs:
            o, _ = pr.communicate()
            if pr.returncode != 0:
                raise RuntimeError(f"cpu baseline worker ({kind}) failed")
            outs.append(json.loads(o.strip().splitlines()[-1]))
        return outs

    t0 = time.time()
    kind = "reference"
    try:
        from oracle import ref_case
        if not os.path.exists(ref_case.harness_path()):
            raise RuntimeError("the reference harness is not built")
        outs = run_workers("reference")
    except Exception as e:  # the reference checker is optional: time the oracle instead
        log(f"[bench] reference CPU baseline unavailable ({e}); timing the oracle")
        kind = "port"
        outs = run_workers("port")
    wall = time.time() - t0
    worst = max(outs, key=lambda r: r["s"])
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    if kind == "reference":
        sample = (f"{cores} concurrent processes of the reference solver itself (ProteusCFD ucs, built from its own "
                  f"sources as oracle/_ref/ref_harness; one per core, as MPI ranks), each running "
                  f"{args.cpu_iters} iterations of its own steady loop on a {worst['nnode']}-node partition-sized "
                  f"mesh of the workload; the last iteration (second order) timed: slowest rank "
                  f"{worst['s']:.2f} s (its iterations {', '.join(f'{x:.2f}' for x in worst['iter_s'])} s)")
    else:
        ph = worst["phases"]
        sample = (f"{cores} concurrent oracle processes (one per core, as MPI ranks), each one implicit iteration "
                  f"on a {worst['nnode']}-node partition-sized mesh of the workload: slowest {worst['s']:.2f} s "
                  f"(jacobian {ph['jacobian']:.2f} s, residual {ph['residual']:.2f} s, gradient+limiter "
                  f"{ph['gradient']:.2f} s, SGS {ph['linear']:.2f} s)")
    return {
        "value": worst["s"] * n_iter_plain,
        "unit": "s (estimated: slowest rank's s/iteration x the plain GPU path's iteration count)",
        "cores": cores, "kind": kind,
        "sample": sample + f"; {wall:.0f} s wall incl. mesh set-up; host {model}; halo exchange not modelled",
        "s_per_iteration": worst["s"],
        "nproc_host": os.cpu_count(),
    }


def make_case(kind, args, parts=1):
    """The workload's mesh, parameters and description; parts > 1 cuts the
    mesh to about 1/parts of its size (the two directions along the wall or
    floor) for one rank of the CPU baseline."""
    from proteuscfd_amd import cases
    from proteuscfd_amd.cases import Params
    px = int(round(parts ** 0.5)) if parts > 1 else 1
    py = max(1, parts // px)
    if kind == "bump":
        base = dict(sorder=2, limiter=3, venkat_K=args.venkat_k, cfl=args.cfl or 50.0, cfl_start=1.0, cfl_ramp=20,
                    nsgs=8)
        n = args.n
        if parts > 1:
            from proteuscfd_amd import mesh as M
            case = cases.bump_case(n=2, params=Params(**base, precision=0))
            case.mesh = M.Mesh.build(M.BUMP, nx=max(2, n // px), ny=max(2, n // py), nz=n, seed=2, jitter=0.03,
                                     rcm=1)
            return case, base, ""
        case = cases.bump_case(n=n, params=Params(**base, precision=0))
        m = case.mesh
        workload = (f"second-order Euler, sin^2 bump in a unit cube, {n}^3 hexes -> 6 tets each "
                    f"({m.nnode} nodes, {m.nedge} edges, {m.nbedge} half-edges), RCM, jitter 3%, "
                    f"M=0.5 farfield, weighted LSQ (9 terms), modified Venkat K={args.venkat_k:g}, "
                    f"first-order FD Jacobian, 8 SGS, CFL 1->{base['cfl']:g} over 20 iterations")
    else:
        nx, ny, nz = (int(v) for v in args.plate.split(","))
        cfl = (args.cfl if kind == args.case else 0.0) or 100.0
        base = dict(sorder=2, limiter=0, cfl=cfl, cfl_start=1.0, cfl_ramp=100, nsgs=8, viscous=1,
                    reynolds=1.0e4, mach_ref=0.3)
        nx, ny = max(2, nx // px), max(2, ny // py)
        case = cases.plate_case(nx=nx, ny=ny, nz=nz, nlayers=20, first_h=1e-4, growth=1.15, twall=1.0,
                                top_bc=cases.BC_SYMMETRY, params=Params(**base, precision=0))
        m = case.mesh
        ncell = nx * ny * (20 * 2 + (nz - 20) * 6)
        workload = (f"second-order laminar Navier-Stokes flat-plate channel, Re=1e4 per unit length, Pr=0.72, "
                    f"Sutherland, M=0.3, isothermal noSlip floor T=1; {nx}x{ny}x{nz} hexes: 20 prism "
                    f"layers (first 1e-4, growth 1.15) under a Kuhn-tet core = {ncell} cells, {m.nnode} nodes, "
                    f"{m.nedge} edges; farfield inflow/outflow, symmetry top and sides; unlimited MUSCL, "
                    f"first-order FD Jacobian + analytic viscous Jacobian, 8 SGS, CFL 1->{base['cfl']:g} over 100 "
                    f"iterations")
    return case, base, workload


def run_forms(case, base, warmup, profile_iters, iters_only, steps, max_iter, drop_goal, chosen_jac_freq=1,
              forms=((0, "plain_fp64"), (1, "chosen_fp32"))):
    """Warm up, profile and solve the case in the plain and the chosen form.
    The plain form is the reference's method as configured (Jacobian every
    iteration); the chosen form stores the Jacobian and SGS in FP32 and
    refreshes the Jacobian every chosen_jac_freq iterations (the reference's
    jacobianUpdateFrequency): both change only the preconditioner."""
    from proteuscfd_amd import gpu
    from proteuscfd_amd.cases import Params
    results = {}
    for prec, name in forms:
        jf = chosen_jac_freq if (prec == 1 or name.endswith("same_refresh")) else 1
        p = Params(**base, precision=prec, jac_freq=jf)
        ctx = gpu.Context(case, p, gradient_terms=9)
        info = ctx.info()
        ctx.upload_q(case.initial_state())
        ctx.reset_iteration()
        for _ in range(warmup):
            ctx.iterate()
        ctx.synchronize()
        prof = phase_profile(ctx, case, profile_iters, jf)
        log(f"[bench] {name} phase ms/iter: " + ", ".join(f"{k}={v:.3f}" for k, v in prof.items()))
        if iters_only:
            ctx.upload_q(case.initial_state())
            ctx.reset_iteration()
            ctx.synchronize()
            t = time.time()
            for _ in range(iters_only):
                r = ctx.iterate()
            ctx.synchronize()
            times, iters, hists = [time.time() - t], [iters_only], [np.array([r[0]])]
        else:
            times, iters, hists = timed_solves(ctx, case, steps, max_iter, drop_goal)
        h = hists[-1]
        drop = float(h[-1] / h[0]) if len(h) > 1 else 1.0
        log(f"[bench] {name}: {iters} iterations, {['%.3f' % x for x in times]} s, drop {drop:.2e}")
        qfinal = ctx.download_q()
        results[name] = dict(times=times, iters=iters, prof=prof, info=info, drop=drop, q=qfinal, jac_freq=jf)
        ctx.close()
    return results


def field_error(results, nnode):
    """Per-variable max error of the chosen form's converged field against the plain FP64 path."""
    qp, qc = results["plain_fp64"]["q"][:nnode, :5], results["chosen_fp32"]["q"][:nnode, :5]
    return {v: float(np.abs(qc[:, k] - qp[:, k]).max() / np.abs(qp[:, k]).max())
            for k, v in enumerate(["rho", "rhou", "rhov", "rhow", "E"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed full solves per precision form")
    ap.add_argument("--warmup", type=int, default=5, help="untimed implicit iterations before timing")
    ap.add_argument("--n", type=int, default=100, help="hex cells per direction of the bump mesh")
    ap.add_argument("--max-iter", type=int, default=4000)
    ap.add_argument("--drop", type=float, default=1e-8)
    ap.add_argument("--profile-iters", type=int, default=10)
    ap.add_argument("--iters-only", type=int, default=0,
                    help="profiling mode: run this many iterations of each form instead of full solves")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run a device copy of BYTES first (for counter scaling)")
    ap.add_argument("--venkat-k", type=float, default=5.0, help="modified-Venkat K (the workload's K=5)")
    ap.add_argument("--case", choices=["bump", "plate"], default="plate",
                    help="plate: 8M-cell laminar flat plate (the headline workload); bump: 1.03M-node Euler bump")
    ap.add_argument("--plate", default="200,40,180", help="plate hexes nx,ny,nz (the lowest 20 become prism layers)")
    ap.add_argument("--cfl", type=float, default=0.0, help="final CFL (default 50 bump, 100 plate)")
    ap.add_argument("--chosen-jac-freq", type=int, default=4,
                    help="chosen form: Jacobian refreshed every N iterations (jacobianUpdateFrequency)")
    ap.add_argument("--no-attribution", action="store_true",
                    help="skip the plain FP64 solve with the chosen form's Jacobian refresh")
    ap.add_argument("--no-fr", action="store_true", help="skip the finite-rate reacting wedge record")
    ap.add_argument("--only-fr", action="store_true", help="run the finite-rate reacting wedge record alone")
    ap.add_argument("--fr-wedge", default="200,100,17", help="FR wedge hexes nx,ny,nz (6 tets each)")
    ap.add_argument("--fr-steps", type=int, default=300)
    ap.add_argument("--fr-err-tol", type=float, default=0.05,
                    help="surrogate error-network bound (fraction of the cell's rate scale)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the record of the other workload that follows the headline")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-parts", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-kind", default="reference", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-iters", type=int, default=2, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_worker:
        cpu_worker(args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world > 1:
        print("bench.py runs on one GPU only: the solver has no multi-GPU path "
              "(a node-partitioned design with halo exchange is not built). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.only_fr:
        print(json.dumps({"reacting_wedge_fr": fr_record(args)}))
        return

    t0 = time.time()
    case, base, workload = make_case(args.case, args)
    m = case.mesh
    log(f"[bench] mesh {args.case}: {m.nnode} nodes {m.nedge} edges {m.nbedge} half-edges "
        f"({time.time() - t0:.1f}s)")
    results = run_forms(case, base, args.warmup, args.profile_iters, args.iters_only, args.steps, args.max_iter,
                        args.drop, args.chosen_jac_freq)
    same = None
    if not args.iters_only and args.chosen_jac_freq > 1 and not args.no_attribution:
        # attribution: the plain FP64 path with the chosen form's Jacobian
        # refresh, one solve (how much of the speed-up is the FP32 storage)
        r = run_forms(case, base, 2, 2, 0, 1, args.max_iter, args.drop, args.chosen_jac_freq,
                      forms=((0, "plain_fp64_same_refresh"),))["plain_fp64_same_refresh"]
        same = {"s": float(np.mean(r["times"])), "iterations": r["iters"][0], "jac_freq": r["jac_freq"],
                "field_max_rel_err_vs_plain": field_error({"plain_fp64": results["plain_fp64"],
                                                           "chosen_fp32": r}, m.nnode)}

    plain, chosen = results["plain_fp64"], results["chosen_fp32"]
    t_plain = float(np.mean(plain["times"]))
    t_chosen = float(np.mean(chosen["times"]))
    nblocks = plain["info"]["nblocks"]
    ab64 = algorithmic_bytes(m, nblocks, 8)
    ab32 = algorithmic_bytes(m, nblocks, 4)

    def gbs(nbytes, ms):
        return nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0

    pp, pc = plain["prof"], chosen["prof"]
    nsgs = base["nsgs"]
    sweeps = 2 * nsgs
    sgs64, npass, npass_full = sgs_pass_bytes(m, nsgs, 8)
    sgs32, _, _ = sgs_pass_bytes(m, nsgs, 4)
    # kernel-only times (HIP events on the library stream around the launches)
    res_key = "residual_2nd" if base["sorder"] > 1 else "residual_1st"
    kern = {
        # k_residual's window covers the viscous pass too when the case is viscous
        "residual_fp64": dict(ms=pp["k_residual"],
                              gbs=gbs(ab64[res_key] + (ab64["viscous"] if base.get("viscous") else 0),
                                      pp["k_residual"])),
        # per symmetric sweep (2 per SGS iteration); bytes of the passes actually run
        "sgs_sweep_fp64": dict(ms=pp["k_sgs_sweeps"] / sweeps, gbs=gbs(sgs64, pp["k_sgs_sweeps"])),
        "sgs_sweep_fp32": dict(ms=pc["k_sgs_sweeps"] / sweeps, gbs=gbs(sgs32, pc["k_sgs_sweeps"])),
        # per assembly
        "jacobian_fp64": dict(ms=pp["k_jacobian_per_build"], gbs=gbs(ab64["jacobian"], pp["k_jacobian_per_build"])),
        "jacobian_fp32": dict(ms=pc["k_jacobian_per_build"], gbs=gbs(ab32["jacobian"], pc["k_jacobian_per_build"])),
        "gradient9": dict(ms=pp["gradient"], gbs=gbs(ab64["gradient9"], pp["gradient"])),
    }
    if base.get("limiter"):
        kern["limiter"] = dict(ms=pp["limiter"], gbs=gbs(ab64["limiter"], pp["limiter"]))
    for v in kern.values():
        v["frac"] = v["gbs"] / HBM_PEAK_GBS
    # wall time per iteration of the timed solves; the phase times come from a
    # separate pass whose kernels run one after another (the timed solves run
    # the Jacobian rows beside the edge passes)
    per_iter_ms = {k: 1e3 * float(np.mean(r["times"])) / max(float(np.mean(r["iters"])), 1.0)
                   for k, r in [("plain", plain), ("chosen", chosen)]}
    # dominant kernel of the chosen form: the FP32 SGS colour pass.  Per
    # launch: the passes' algorithmic bytes / passes run, over the average
    # pass duration (both from the same HIP-event window).
    dom = kern["sgs_sweep_fp32"]
    sgs_launch_bytes = sgs32 / npass
    sgs_launch_us = pc["k_sgs_sweeps"] * 1e3 / npass
    tag = {"bump": f"bump{args.n}", "plate": "plate8M" if args.plate == "200,40,180" else None}[args.case]
    traffic = measured_traffic("k_sgs_color<float>", tag) if tag else None
    edge_flux = (m.nedge + m.nbedge) / (pp["k_residual"] * 1e-3)
    # converged-field error of the chosen form against the plain FP64 path
    field_err = field_error(results, m.nnode)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline: one reference (or oracle) process per core, each on a partition-sized mesh ...")
        cpu = cpu_baseline(args, int(np.mean(plain["iters"])))
        log(f"[bench] cpu: {cpu['s_per_iteration']:.2f} s/iteration on {cpu['cores']} cores")

    speedup = t_plain / t_chosen if t_chosen > 0 else None
    out = {
        "metric": METRIC,
        "value": t_chosen,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps if not args.iters_only else 1,
        "warmup": args.warmup,
        "ms_per_step": t_chosen * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 residual/gradients/norms, f32 Jacobian+SGS (plain path all f64)",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "step": "one full steady solve from freestream to an 8-order residual drop",
            "parallelism": "single GPU",
        },
        "plain_fp64_s": t_plain,
        "chosen_fp32_s": t_chosen,
        "speedup_chosen_vs_plain": speedup,
        "speedup_target": 2.0,
        "jacobian_update_frequency": {"plain": 1, "chosen": args.chosen_jac_freq},
        "plain_fp64_same_refresh": same,
        "iterations": {"plain": plain["iters"], "chosen": chosen["iters"]},
        "residual_drop": {"plain": plain["drop"], "chosen": chosen["drop"]},
        "edge_fluxes_per_s": edge_flux,
        "cell_iterations_per_s": m.nnode * float(np.mean(chosen["iters"])) / t_chosen,
        "chosen_vs_plain_field_max_rel_err": field_err,
        "per_iteration_ms": per_iter_ms,
        "phase_ms": {"plain": pp, "chosen": pc},
        "kernels": kern,
        "sgs_colour_passes": {"run": npass, "of": npass_full, "ncolor": plain["info"]["ncolor"]},
        "roofline": {"bound": "hbm", "kernel": "k_sgs_color<float>", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": traffic,
                     "algorithmic_bytes_per_launch": sgs_launch_bytes, "avg_launch_us": sgs_launch_us,
                     "launches_per_sgs_call": npass},
        "roofline_residual": {"bound": "hbm", "kernel": "k_residual + k_viscous" if base.get("viscous") else
                              "k_residual", "achieved": kern["residual_fp64"]["gbs"], "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": kern["residual_fp64"]["frac"]},
        "cpu_baseline": cpu,
    }
    if not args.iters_only and not args.no_extra:
        other = "bump" if args.case == "plate" else "plate"
        out["bump_euler" if other == "bump" else "laminar_plate"] = extra_record(other, args)
    if not args.iters_only and not args.no_fr:
        out["reacting_wedge_fr"] = fr_record(args)
    print(json.dumps(out))


def extra_record(kind, args):
    """The other workload solved in both forms, reported beside the headline."""
    t0 = time.time()
    case, base, workload = make_case(kind, args)
    m = case.mesh
    log(f"[bench] mesh {kind}: {m.nnode} nodes {m.nedge} edges ({time.time() - t0:.1f}s)")
    res = run_forms(case, base, 2, 5, 0, 1, 4000, args.drop, args.chosen_jac_freq)
    tp = float(np.mean(res["plain_fp64"]["times"]))
    tc = float(np.mean(res["chosen_fp32"]["times"]))
    ip, ic = res["plain_fp64"]["iters"][0], res["chosen_fp32"]["iters"][0]
    return {
        "workload": workload,
        "plain_fp64_s": tp,
        "chosen_fp32_s": tc,
        "speedup_chosen_vs_plain": tp / tc if tc > 0 else None,
        "jacobian_update_frequency": {"plain": 1, "chosen": args.chosen_jac_freq},
        "iterations": {"plain": ip, "chosen": ic},
        "residual_drop": {"plain": res["plain_fp64"]["drop"], "chosen": res["chosen_fp32"]["drop"]},
        "wall_ms_per_iteration": {"plain": 1e3 * tp / max(ip, 1), "chosen": 1e3 * tc / max(ic, 1)},
        "chosen_vs_plain_field_max_rel_err": field_error(res, m.nnode),
        "phase_ms": {"plain": res["plain_fp64"]["prof"], "chosen": res["chosen_fp32"]["prof"]},
        "edge_fluxes_per_s": (m.nedge + m.nbedge) / (res["plain_fp64"]["prof"]["k_residual"] * 1e-3),
        "cell_iterations_per_s": m.nnode * ic / tc if tc > 0 else None,
    }


def fr_record(args):
    """Finite-rate reacting Euler (5-species air) on the Mach 8 wedge
    channel: a fixed number of implicit steps from freestream in the plain
    form (FP64, exact chemistry), the chosen form (FP32 Jacobian and SGS) and
    the chosen form with the surrogate production rates (rate network + error
    network, exact fallback) in the residual.  The first-order solve with
    the noCVBC boundary states reaches a plateau rather than an 8-order drop
    (the reference's loop behaves the same on the small wedge), so the
    record is time per step and the residual ratio after the steps."""
    from proteuscfd_amd import cases, chem, gpu, surrogate as S
    nx, ny, nz = (int(v) for v in args.fr_wedge.split(","))
    t0 = time.time()
    case = cases.wedge_case(nx=nx, ny=ny, nz=nz, rcm=True)
    m = case.mesh
    log(f"[bench] FR wedge {nx}x{ny}x{nz}: {m.nnode} nodes {m.nedge} edges ({time.time() - t0:.1f}s)")
    model = chem.load_air5()
    Y = chem.mass_fractions(model, N2=0.767, O2=0.233)
    p0 = 1.0e-3 * float(np.sum(Y * chem.UNIV_R / model.mw)) * 300.0
    sc = chem.fr_case_scales(p0, 300.0, p0)
    qinf, Pref = chem.fr_freestream(model, sc, Y, p0, 300.0, [8.0, 0.0, 0.0])
    cc = gpu.ChemContext(model.with_reference(sc["ref_density"], sc["ref_time"], sc["ref_temperature"]))
    bct = np.array([0, 7, 8, 3, 3, 9, 9], dtype=np.int32)
    t0 = time.time()
    net, enet, tr, gm, gs, _ = S.train_with_error_net(cc, n=2_000_000)
    # rates below rho / t_ref cannot move the mass fractions within a flow
    # time: that is the floor of both fallback rules (the element-balance
    # floor in mol/(kg s), over the freestream's mean molar mass)
    rate_floor = 1.0 / sc["ref_time"]
    mbar = 1.0 / float(np.sum(Y / model.mw))
    desc = S.desc(tr, tol=0.1, floor=rate_floor / mbar, err_tol=args.fr_err_tol, rate_floor=rate_floor,
                  err_mid=gm, err_scale=gs)
    t_train = time.time() - t0
    log(f"[bench] surrogate trained in {t_train:.1f}s")
    runs, fields = {}, {}
    # the chosen form refreshes the Jacobian every step here: with the stiff
    # source Jacobian lagged over four steps the reacting solve diverges
    for name, prec, sur, jfreq in [("plain_fp64", gpu.PLAIN_FP64, False, 1),
                                   ("chosen_fp32", gpu.CHOSEN_FP32, False, 1),
                                   ("chosen_fp32_surrogate", gpu.CHOSEN_FP32, True, 1)]:
        f = gpu.FrContext(m, bct, np.tile(qinf, (len(bct), 1)), cc)
        f.set_params(sc["ref_pressure"], sc["ref_velocity"], Pref, 1.0, cfl=5.0, cfl_start=1.0, cfl_ramp=50, nsgs=6,
                     precision=prec, no_cvbc=True, jac_freq=jfreq)
        if sur:
            f.set_surrogate(net, desc)
        q0 = np.tile(qinf, (m.nnode + m.nbedge, 1))
        f.upload_q(q0)
        f.init_iterations()
        for _ in range(3):
            f.step()
        f.upload_q(q0)
        f.init_iterations()
        f.phase_times(reset=True)
        h = []
        t = time.time()
        for _ in range(args.fr_steps):
            h.append(f.step()[0])
        el = time.time() - t
        pt = f.phase_times()
        info = f.info()
        r = {"s": el, "ms_per_step": 1e3 * el / args.fr_steps, "residual_ratio": h[-1] / h[0], "jac_freq": jfreq,
             "phase_ms_per_step": {k: v / args.fr_steps for k, v in pt.items()}}
        # SGS against the HBM roofline: every pass reads its rows' blocks
        # (81 values each) and moves b, the neighbours' x and the new x (27
        # values per row); 2 nsgs C - (2 nsgs - 1) passes of the C colours
        nc, nsgs, sz = info["ncolor"], 6, 4 if prec == gpu.CHOSEN_FP32 else 8
        npass = 2 * nsgs * nc - (2 * nsgs - 1)
        sgs_bytes = npass * (info["nblocks"] * 81 * sz + m.nnode * 27 * sz) / nc
        sgs_gbs = sgs_bytes / (r["phase_ms_per_step"]["sgs"] * 1e-3) / 1e9
        r["sgs_roofline"] = {"bound": "hbm", "passes_per_step": npass, "algorithmic_bytes_per_step": sgs_bytes,
                             "achieved": sgs_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sgs_gbs / HBM_PEAK_GBS,
                             "note": "the SGS phase time includes PrepareSGS (LU of the diagonal blocks)"}
        if sur:
            st = f.surrogate_stats()
            r["surrogate_fallback_share_last_step"] = st["last_fallback"] / m.nnode
            r["surrogate_fallback_share_mean"] = st["fallback_total"] / max(st["calls"] * m.nnode, 1)
            r["surrogate_ms_per_step"] = st["ms"] / max(st["calls"], 1)
        runs[name] = r
        fields[name] = f.download_q()[:m.nnode, :9]
        log(f"[bench] FR {name}: {r['ms_per_step']:.2f} ms/step, residual ratio {r['residual_ratio']:.3e}")
        f.close()
    base = fields["plain_fp64"]
    rho = base[:, :5].sum(axis=1).max()

    def ferr(q, ref):
        return [float(np.abs(q[:, j] - ref[:, j]).max() / (rho if j < 5 else np.abs(ref[:, j]).max()))
                for j in range(9)]
    names = ["rho_O2", "rho_O", "rho_N", "rho_N2", "rho_NO", "u", "v", "w", "T"]
    enet.close()
    net.close()
    return {
        "workload": (f"5-species air (O2, O, N, N2, NO; 5 reactions) finite-rate reacting Euler, first order, 9 "
                     f"unknowns/node (9x9 blocks); 15-degree wedge channel {nx}x{ny}x{nz} hexes -> "
                     f"{6 * nx * ny * nz} tets ({m.nnode} nodes, {m.nedge} edges), RCM, Mach 8, freestream 300 K, "
                     f"1e-3 kg/m^3, Y_N2 0.767, Y_O2 0.233; noCVBC (upwind farfield in/out, mirrored walls and "
                     f"symmetry planes), HLLC, FD Jacobian, 6 SGS, CFL 1->5 over 50 steps; {args.fr_steps} steps "
                     f"from freestream"),
        "steps": args.fr_steps,
        "forms": runs,
        "speedup_chosen_vs_plain": runs["plain_fp64"]["s"] / runs["chosen_fp32"]["s"],
        "field_max_rel_err_vs_plain": {
            "chosen_fp32": dict(zip(names, ferr(fields["chosen_fp32"], base))),
            "chosen_fp32_surrogate": dict(zip(names, ferr(fields["chosen_fp32_surrogate"], base)))},
        "surrogate_vs_exact_field_max_rel_err": dict(zip(names, ferr(fields["chosen_fp32_surrogate"],
                                                                     fields["chosen_fp32"]))),
        "surrogate": {"train_s": t_train, "samples": 2_000_000, "err_tol": args.fr_err_tol,
                      "rate_floor_per_s": rate_floor,
                      "note": "the source Jacobian in the preconditioner stays exact; species errors over the "
                              "largest mixture density"},
    }


if __name__ == "__main__":
    main()
