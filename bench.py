#!/usr/bin/env python3
"""Throughput of the LDDRK6 time-advance hot path on one MI355X.

Workload (default): 2-D Navier-Stokes, 4096 x 4096 uniform grid, uniform mean
flow M=0.5, Gaussian acoustic pulse + isentropic vortex, 13-point filter
sigma=0.2 in every stage, 40-point buffer zones, FP64.  A "step" is one full
LDDRK6 step (6 stages of RHS + update + buffer + x/y filter + edges).

`--workload batch`: the batched sweep -- 256 independent 2-D Euler Gaussian
pulse cases on square grids drawn from {128,192,256,384,512}^2 (seeded; M in
[0,0.6], amplitude log-uniform in [1e-4,1e-2] p0, half-width 2-6), advanced
together by the batched path (one launch per kernel class, graph per step),
FP32-perturbation form by default; reports the aggregate rate and the p'
error against the FP64 batch after the same steps.

`--workload adjoint`: dJ/d(initial p) of the 2-D Euler equations on 1024^2
(M=0.3, 16 seeded Gaussians), J = time-integrated p'^2 on a radius-400
observer arc, 500 steps with checkpoints every 50, FP64 and the FP32
perturbation form; reports forward and backward times and checks the FP64
gradient against central differences on 32 seeded directions.

`--workload mixing`: 2-D compressible mixing layer, 2048 x 1024, uniform x,
tanh-stretched y (spacing ratio 10) with metrics everywhere, u = 0.25
tanh(2y/delta), Re_delta = 1000, Pr = 0.72, 60-point buffers, seeded v';
FP64 and the FP32 perturbation form for the same steps (default 20000);
reports both rates and the momentum-thickness growth-rate error of FP32.

Prints one JSON line: Mpts*steps/s over the timed steps, the roofline of the
dominant kernel (stage kernel, HBM bound) from HIP events on the library's
stream, the whole-step roofline against the 704 B/pt/step compulsory
traffic, and the CPU oracle timed on the host as a reported baseline (an
OpenMP leg on the workload itself and a one-core leg on the same workload at
half the grid size, with the host's CPU model and core count).  For the
ns and pulse workloads in FP64 the same run also times the FP32
perturbation form on the same workload and steps (`fp32_perturbation`: its
rate, speed-up over FP64 and relative p' = p - p_b L2 error against the
FP64 state after the same steps).
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

METRIC = "Mpts·steps/s + % HBM roofline; FP64 vs FP32-pert at ≤1e-4 p' error"
HBM_PEAK_GBS = 8000.0


def alg_bytes_per_point_step(nk: int, elem: int) -> float:
    # 2N-storage LDDRK6: U read 6x + written 6x, W read 5x (stages 2-6) and
    # written 5x (stages 1-5) -> 22 field accesses per point per step
    return 22.0 * nk * elem


def make_case(args):
    from navierstokes_amd import config
    if args.workload == "pulse":
        case = config.acoustic_pulse_case(args.size or 201)
        desc = f"euler-acoustic-pulse {case.imax}x{case.jmax} fp64 lddrk6 sigma=0.2 D=30"
    else:
        case = config.pulse_vortex_case(args.size or 4096)
        desc = (f"ns-pulse-vortex {case.imax}x{case.jmax} M=0.5 Re=1e4(pulse) fp64 lddrk6 "
                f"sigma=0.2 D=40")
    return case, desc


def profiled_traffic(workload: str, size: int, precision: str, variant: int, same_launch: bool = False):
    """HBM bytes per stage launch of the given stage variant (same_launch: the
    form whose launch also covers the edge tiles), measured with rocprofv3
    --pmc (FETCH_SIZE x2 + WRITE_SIZE, calibrated on a plain copy), from the
    committed summary profiles/hbm_traffic_ns4096.md; None when that
    configuration has none."""
    key = f"v{variant}" + ("e" if same_launch else "")
    path = os.path.join(ROOT, "profiles", "hbm_traffic_ns4096.md")
    try:
        for line in open(path):
            f = line.split()
            if (len(f) == 6 and f[0] == "traffic" and f[1] == workload and int(f[2]) == size
                    and f[3] == precision and f[4] == key):
                return int(f[5])
    except OSError:
        pass
    return None


def host_cpu():
    """CPU model and core counts of the host the baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def cpu_baseline(case, steps: int, serial_case=None, serial_steps: int = 1):
    """The oracle (C restatement of the reference) timed on the host: an
    OpenMP build over the host's threads on `steps` steps of the workload
    itself, and the serial build (one core, like the reference's serial
    Fortran) on `serial_steps` steps of `serial_case` (the same workload on
    a smaller grid: a step of the full grid takes about a minute on one
    core)."""
    from oracle.oracle import Oracle
    o = Oracle(case, parallel=True)
    t0 = time.perf_counter()
    o.run(case.U0, 0, steps)
    dt = time.perf_counter() - t0
    out = {"value": case.npoints * steps / dt / 1e6, "unit": "Mpts·steps/s",
           "cores": int(o.threads()), "kind": "port",
           "sample": f"{steps} LDDRK6 step(s) of the same {case.imax}x{case.jmax} workload, "
                     f"OpenMP build of oracle/ns_oracle.c on {o.threads()} threads ({dt:.1f} s)",
           **host_cpu()}
    if serial_case is not None:
        so = Oracle(serial_case, parallel=False)
        t0 = time.perf_counter()
        so.run(serial_case.U0, 0, serial_steps)
        ds = time.perf_counter() - t0
        out["serial"] = {"value": serial_case.npoints * serial_steps / ds / 1e6, "unit": "Mpts·steps/s",
                         "cores": 1,
                         "sample": f"{serial_steps} LDDRK6 step(s) of the same workload on a "
                                   f"{serial_case.imax}x{serial_case.jmax} grid, serial build of "
                                   f"oracle/ns_oracle.c ({ds:.1f} s)"}
    return out


def copy_probe(mib: int):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE
    of this kernel calibrate the counters against bytes actually moved, and
    its rate (read + write bytes over time) is the practical HBM ceiling a
    streaming kernel is compared with."""
    import torch
    n = mib * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_probe_MiB": mib, "bytes_read": n * 8, "bytes_written": n * 8,
                      "ms": round(ms, 4), "GB_per_s": round(2 * n * 8 / ms / 1e6, 1)}), flush=True)


def bench_batch(args):
    from navierstokes_amd import config
    from navierstokes_amd.solver import BatchSolver

    args.precision = args.precision or "fp32pert"
    prec = 1 if args.precision == "fp32pert" else 0
    cases = config.batch_sweep_cases(args.ncases, seed=2024)
    npts = sum(c.npoints for c in cases)
    sizes = sorted({c.imax for c in cases})
    desc = (f"euler-pulse batched sweep: {len(cases)} cases, square grids {sizes}, "
            f"{npts / 1e6:.2f} Mpts, M~U[0,0.6], amp~logU[1e-4,1e-2], hw~U[2,6], D=20, "
            f"{'fp32-perturbation' if prec else 'fp64'} lddrk6 sigma=0.2")
    print(f"[bench] {desc}: set-up", file=sys.stderr, flush=True)
    b = BatchSolver(cases, precision=prec)
    variant, tune_ms = b.autotune()
    b.run(0, args.warmup)
    print(f"[bench] warm-up {args.warmup} steps done; timing {args.steps}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b.run(args.warmup, args.steps)
    t = time.perf_counter() - t0
    ms_step = t / args.steps * 1e3
    value = npts * args.steps / t / 1e6
    nk, elem = cases[0].nk, (4 if prec else 8)
    B_step = alg_bytes_per_point_step(nk, elem)
    ach = B_step * npts / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpts·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec else "f64",
        "data": "synthetic",
        "config": {"workload": desc, "ncases": len(cases), "points": npts, "nk": nk,
                   "precision": args.precision, "parallelism": "single-gpu",
                   "stage_block_variant": variant, "stage_variant_ms": [round(x, 4) for x in tune_ms]},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batched step (graph of stage, filter and edge launches)",
                     "alg_bytes_per_launch": B_step * npts, "avg_launch_ms": round(ms_step, 4)},
    }
    if prec:
        # p' error of every case against the FP64 batch (bitwise the FP64 oracle)
        nsteps = args.warmup + args.steps
        ref = BatchSolver(cases, precision=0)
        ref.run(0, nsteps)
        errs = []
        for i, c in enumerate(cases):
            a, r = b.get_state(i), ref.get_state(i)
            pa, pr = a[..., 3] - c.Ub[..., 3], r[..., 3] - c.Ub[..., 3]
            errs.append(float(np.linalg.norm(pa - pr) / np.linalg.norm(pr)))
        ref.close()
        out["pprime_rel_l2_vs_fp64"] = {"max": max(errs), "median": float(np.median(errs)),
                                        "steps": nsteps}
    b.close()
    if not args.no_cpu_baseline:
        from oracle.oracle import Oracle
        sample = cases[:16]
        t0 = time.perf_counter()
        threads = 1
        for c in sample:
            o = Oracle(c, parallel=True)
            threads = o.threads()
            o.run(c.U0, 0, args.cpu_steps)
        dt = time.perf_counter() - t0
        spts = sum(c.npoints for c in sample)
        out["cpu_baseline"] = {"value": spts * args.cpu_steps / dt / 1e6, "unit": "Mpts·steps/s",
                               "cores": int(threads), "kind": "port",
                               "sample": f"{args.cpu_steps} LDDRK6 step(s) of the first {len(sample)} cases "
                                         f"({spts / 1e6:.2f} Mpts), OpenMP build of oracle/ns_oracle.c "
                                         f"({dt:.1f} s incl. per-case set-up)"}
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    from navierstokes_amd import config
    from navierstokes_amd.solver import Solver

    N = args.size or 1024
    nsteps = args.steps if args.steps != 50 else 500
    every = args.every
    case = config.adjoint_case(N, D=40, M=0.3, ngauss=16, seed=5)
    arc = config.observer_arc(N, 0.390625 * N)  # radius 400 at 1024^2
    dp = case.U0[..., 3] - case.Ub[..., 3]
    desc = (f"euler-adjoint {N}x{N} M=0.3 16 gaussians, observer arc r={0.390625 * N:.0f} "
            f"({len(arc)} pts), {nsteps} steps, checkpoint every {every}")
    print(f"[bench] {desc}: set-up", file=sys.stderr, flush=True)

    def J_of(s, x, chk):
        U0 = case.Ub.copy()
        U0[..., 3] += x
        s.put_state(U0)
        return s.adj_forward(0, nsteps, chk)

    def J_at(s, p, chk):  # J from the initial pressure field p (the rest at the base state)
        U0 = case.Ub.copy()
        U0[..., 3] = p
        s.put_state(U0)
        return s.adj_forward(0, nsteps, chk)

    res = {}
    grads = {}
    for prec, name in ((0, "fp64"), (1, "fp32pert")):
        s = Solver(case, precision=prec)
        s.adj_set_arc(arc)
        s.autotune()
        J_of(s, dp, every)  # warm-up (kernels, buffers)
        s.synchronize()
        t0 = time.perf_counter()
        J = J_of(s, dp, every)
        t1 = time.perf_counter()
        g = s.adj_backward()
        t2 = time.perf_counter()
        grads[name] = g
        res[name] = {"J": J, "forward_s": t1 - t0, "backward_s": t2 - t1,
                     "forward_ms_per_step": (t1 - t0) / nsteps * 1e3,
                     "backward_ms_per_step": (t2 - t1) / nsteps * 1e3}
        print(f"[bench] {name}: J={J:.6e} forward {t1 - t0:.2f} s backward {t2 - t1:.2f} s",
              file=sys.stderr, flush=True)
        if prec == 0 and args.fd_directions > 0:
            rng = np.random.default_rng(7)
            g3 = g[..., 3]
            errs = []
            for t in range(args.fd_directions):
                d = rng.standard_normal(dp.shape)
                d *= np.linalg.norm(dp) / np.linalg.norm(d)
                eps = args.fd_eps
                # the perturbed initial pressures as stored: p = pb + (dp +- eps d) rounds the
                # perturbation to the resolution of p (~1e-16 of pb), so the direction applied is
                # d_eff, not d; the adjoint is compared along d_eff
                pp = case.Ub[..., 3] + (dp + eps * d)
                pm = case.Ub[..., 3] + (dp - eps * d)
                d_eff = (pp - pm) / (2 * eps)
                fd = (J_at(s, pp, nsteps) - J_at(s, pm, nsteps)) / (2 * eps)
                ad = float(np.sum(g3 * d_eff))
                errs.append(abs(fd - ad) / abs(fd))
            res[name]["fd_check"] = {"directions": args.fd_directions, "eps_rel": args.fd_eps,
                                     "max_rel_err": max(errs), "median_rel_err": float(np.median(errs))}
            print(f"[bench] fd check: max rel err {max(errs):.2e}", file=sys.stderr, flush=True)
        s.close()
    g64, g32 = grads["fp64"], grads["fp32pert"]
    res["fp32pert"]["grad_p_rel_l2_vs_fp64"] = float(np.linalg.norm(g32[..., 3] - g64[..., 3])
                                                      / np.linalg.norm(g64[..., 3]))
    prec = args.precision or "fp64"
    r = res[prec]
    t = r["forward_s"] + r["backward_s"]
    npts = case.npoints
    out = {
        "metric": METRIC,
        "value": round(npts * nsteps / t / 1e6, 3),
        "unit": "Mpts·steps/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": 1,
        "ms_per_step": round(t / nsteps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "fp32pert" else "f64",
        "data": "synthetic",
        "config": {"workload": desc, "grid": [N, N], "nk": 4, "precision": prec, "parallelism": "single-gpu",
                   "step": "forward + reverse (gradient) LDDRK6 step"},
        "roofline": None,
        "adjoint": res,
    }
    print(json.dumps(out), flush=True)


def bench_mixing(args):
    from navierstokes_amd import config
    from navierstokes_amd.solver import Solver

    nx, ny = (args.size, args.size // 2) if args.size else (2048, 1024)
    nsteps = args.steps if args.steps != 50 else 20000
    every = args.sample_every
    case = config.mixing_layer_case(nx, ny, D=60)
    y = config.mixing_layer_rows(case)
    desc = (f"temporal mixing layer {nx} (periodic x) x {ny} (tanh-stretched y, ratio 10, metrics everywhere), "
            f"M=0.5, Re_delta=1000, Pr=0.72, 60-row y sponges, {nsteps} steps, theta every {every}")
    print(f"[bench] {desc}: set-up", file=sys.stderr, flush=True)
    res = {}
    for prec, name in ((0, "fp64"), (1, "fp32pert")):
        s = Solver(case, precision=prec)
        s.autotune()
        t_steps = 0.0
        times, theta = [], []
        n = 0
        while True:
            ub = s.row_average(1, *config.mixing_layer_columns(case))
            theta.append(config.momentum_thickness(ub, y, 0.25, -0.25))
            times.append(n * case.dt)
            if n >= nsteps:
                break
            k = min(every, nsteps - n)
            t0 = time.perf_counter()
            s.run(n, k)
            t_steps += time.perf_counter() - t0
            n += k
            if n % (every * 10) == 0 or n == nsteps:
                print(f"[bench] {name}: step {n}, theta {theta[-1]:.6f}", file=sys.stderr, flush=True)
        s.check_finite()
        s.close()
        t = np.array(times)
        th = np.array(theta)
        half = len(t) // 2
        rate = float(np.polyfit(t[half:], th[half:], 1)[0])
        res[name] = {"ms_per_step": t_steps / nsteps * 1e3, "Mpts_steps_per_s": case.npoints * nsteps / t_steps / 1e6,
                     "theta0": float(th[0]), "theta_end": float(th[-1]), "dtheta_dt": rate}
    r64, r32 = res["fp64"]["dtheta_dt"], res["fp32pert"]["dtheta_dt"]
    res["fp32pert"]["growth_rate_rel_err_vs_fp64"] = abs(r32 - r64) / abs(r64) if r64 != 0 else None
    prec = args.precision or "fp32pert"
    r = res[prec]
    out = {
        "metric": METRIC,
        "value": round(r["Mpts_steps_per_s"], 3),
        "unit": "Mpts·steps/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": 0,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "fp32pert" else "f64",
        "data": "synthetic",
        "config": {"workload": desc, "grid": [nx, ny], "nk": 4, "precision": prec, "parallelism": "single-gpu"},
        "roofline": None,
        "mixing": res,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["ns", "pulse", "batch", "adjoint", "mixing"], default="ns")
    ap.add_argument("--sample-every", type=int, default=200, help="mixing: steps between theta samples")
    ap.add_argument("--fd-directions", type=int, default=32)
    ap.add_argument("--fd-eps", type=float, default=1e-2,
                    help="adjoint: central-difference step, relative to the perturbation's norm "
                         "(below 1e-2 the FP64 forward's rounding in J, ~4e-9 of J, divided by 2 eps "
                         "dominates the check's error; above it the O(eps^2) truncation)")
    ap.add_argument("--every", type=int, default=50, help="adjoint checkpoint interval")
    ap.add_argument("--ncases", type=int, default=256)
    ap.add_argument("--size", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serial-baseline", action="store_true", help="skip the one-core leg of the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="ns/pulse: skip the FP32-perturbation leg")
    ap.add_argument("--timing-steps", type=int, default=3)
    ap.add_argument("--variant", type=int, default=-1, help="stage variant 0..2 (default: the autotuned one)")
    ap.add_argument("--filter-mode", type=int, default=-1,
                    help="0: x and y filter sweeps as two kernels, 1: one fused sweep (default: autotuned)")
    ap.add_argument("--precision", choices=["fp64", "fp32pert"], default=None,
                    help="default fp64 (ns, pulse) / fp32pert (batch)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only, not built)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "mixing":
        bench_mixing(args)
        return
    args.precision = args.precision or "fp64"

    case, desc = make_case(args)
    prec = 1 if args.precision == "fp32pert" else 0
    if prec:
        desc = desc.replace("fp64", "fp32-perturbation")
    r = run_form(case, desc, prec, args)
    ms_step, value, kt, npts_int, variant, tune_ms = (r[k] for k in ("ms_step", "value", "kt", "npts_int",
                                                                     "variant", "tune_ms"))
    nk, elem = case.nk, (4 if prec else 8)
    B_step = alg_bytes_per_point_step(nk, elem)
    B_stage = B_step / 6.0  # per point per stage-kernel launch (average over the 6 stages)
    # the dominant kernel: the stage's interior kernel, timed with HIP events
    # on its own stream, charged the algorithmic bytes of the points it covers
    t_int = kt["interior"] * 1e-3
    ach = B_stage * npts_int / t_int / 1e9
    step_ach = B_step * case.npoints / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpts·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec else "f64",
        "data": "synthetic",
        "config": {"workload": desc, "grid": [case.imax, case.jmax], "nk": nk, "precision": args.precision,
                   "viscous": not case.inviscid, "parallelism": "single-gpu", "stage_variant": variant,
                   "stage_variant_ms": [round(x, 4) for x in tune_ms], "filter_mode": r["filter_mode"],
                   "filter_tune_ms": r["filter_tune_ms"]},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(args.workload, case.imax, args.precision, variant,
                                                 npts_int == case.npoints),
                     "kernel": ("stage kernel (RHS + LDDRK update + buffer; marching blocks over the interior "
                                "region and the edge tiles in one launch, whole grid)" if npts_int == case.npoints
                                else "stage interior kernel (RHS + LDDRK update + buffer over the interior region; "
                                "the edge tiles run beside it on a second stream)"),
                     "interior_points": npts_int,
                     "alg_bytes_per_launch": B_stage * npts_int,
                     "avg_launch_ms": round(kt["interior"], 4)},
        "step_roofline": {"alg_bytes_per_point_step": B_step, "achieved_GBs": round(step_ach, 1),
                          "frac": round(step_ach / HBM_PEAK_GBS, 4)},
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
    }
    if not prec and not args.no_fp32:
        # the metric's second half: the FP32 perturbation form on the same
        # workload and steps, its speed-up and its p' error against FP64
        r32 = run_form(case, desc.replace("fp64", "fp32-perturbation"), 1, args)
        kt32 = r32["kt"]
        ach32 = alg_bytes_per_point_step(nk, 4) / 6.0 * r32["npts_int"] / (kt32["interior"] * 1e-3) / 1e9
        out["fp32_perturbation"] = {
            "ms_per_step": round(r32["ms_step"], 4),
            "value": round(r32["value"], 3),
            "speedup_vs_fp64": round(ms_step / r32["ms_step"], 3),
            "pprime_rel_l2_vs_fp64": float(np.linalg.norm(r32["pprime"] - r["pprime"]) / np.linalg.norm(r["pprime"])),
            "steps_compared": r["nsteps"],
            "stage_variant": r32["variant"],
            "stage_variant_ms": [round(x, 4) for x in r32["tune_ms"]],
            "filter_mode": r32["filter_mode"],
            "filter_tune_ms": r32["filter_tune_ms"],
            "kernel_ms": {k: round(v, 4) for k, v in kt32.items()},
            "interior_roofline_frac": round(ach32 / HBM_PEAK_GBS, 4),
        }
    if not args.no_cpu_baseline:
        print(f"[bench] cpu baseline ({args.cpu_steps} steps)", file=sys.stderr, flush=True)
        serial = make_case(argparse.Namespace(workload=args.workload, size=max(64, case.imax // 2)))[0] \
            if not args.no_serial_baseline else None
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps, serial)
    print(json.dumps(out), flush=True)


def run_form(case, desc, prec, args):
    """Warm-up, timed steps (wall clock) and HIP-event kernel times of one
    precision form; returns the rates and the final p - p_b (host)."""
    from navierstokes_amd.solver import Solver
    print(f"[bench] {desc}: set-up", file=sys.stderr, flush=True)
    s = Solver(case, precision=prec)
    variant, tune_ms = s.autotune()
    if args.variant >= 0:
        s.set_variant(args.variant)
        variant = args.variant
    if args.filter_mode >= 0:
        s.set_filter_mode(args.filter_mode)
    fmode, fms = s.filter_mode()
    s.run(0, args.warmup)
    s.synchronize()
    print(f"[bench] warm-up {args.warmup} steps done; timing {args.steps}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    s.run(args.warmup, args.steps)
    s.synchronize()
    t = time.perf_counter() - t0
    s.check_finite()
    # per-kernel device time (HIP events on the library's stream)
    s.enable_timing(True)
    s.run(args.warmup + args.steps, args.timing_steps)
    kt = {name: s.kernel_time(name)[0] for name in ("rhs", "interior", "filter_x", "filter_y", "bc")}
    npts_int = s.stage_interior_points()
    s.enable_timing(False)
    pprime = s.get_state()[..., 3] - case.Ub[..., 3]
    s.close()
    return {"ms_step": t / args.steps * 1e3, "value": case.npoints * args.steps / t / 1e6, "kt": kt,
            "npts_int": npts_int, "variant": variant, "tune_ms": tune_ms, "pprime": pprime,
            "nsteps": args.warmup + args.steps + args.timing_steps, "filter_mode": fmode,
            "filter_tune_ms": [round(float(x), 4) for x in fms]}


if __name__ == "__main__":
    main()
