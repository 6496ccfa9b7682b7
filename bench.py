#!/usr/bin/env python3
"""Benchmark: cell-timesteps/s of the ECFV immiscible two-phase Newton step on one MI355X.

Workload: the 3D heterogeneous DNAPL case, 128 x 128 x 64 cells (1,048,576 cells, 2x2
BCRS blocks on a 7-point stencil), dt = 500 s, Newton tolerance 1e-6, BiCGStab + ILU(0)
with the reference's default linear tolerances.

Forms (each timed the same way):
  fp64           FP64 throughout (the plain path);
  fp32_precond   FP32 ILU(0) factors and sweeps inside the FP64 BiCGStab;
  fp32_flux      FP32 face fluxes accumulated into the FP64 residual and Jacobian;
  mixed          both of the above;
  learned_mixed  mixed, with each Newton solve started from the warm-start network's
                 prediction (BF16 matrix-core convolutions); the network is trained
                 before, outside every timed region, on pairs harvested from other
                 realizations;
  mixed_op, learned_mixed_op (not run by default)
                 the same with BiCGStab's operator products reading the Jacobian's
                 blocks rounded to FP32.
`value` is the metric's form, learned_mixed, when it runs (else fp64); the plain FP64
value and the ratio are reported beside it. For each reduced-precision form a second,
untimed pass steps it beside the FP64 path and reports the pressure and saturation
deviation and the Newton and BiCGStab iterations of every time step.

A step is one implicit-Euler time step driven to Newton convergence (linearization,
ILU(0) factorization, BiCGStab, update, norms; with the warm start also the prediction).
W warm-up steps run first, the state is reset to the initial condition, and then exactly
K steps from t = 0 are timed with the state resident in HBM (default K = 40, the whole
configured run). value = cells*K/seconds.

The CPU baseline is the C++ oracle timed on a bounded sample of the same workload on
this host (the first --cpu-sample-steps steps from t = 0): single thread, the reference's
default ThreadsPerProcess = 1, and with the linearization's element loop over every usable
core; the headline form's GPU rate over the same first steps is reported beside it. The
CPU model and core counts are recorded.

--kernel NAME times one kernel alone (for profiler and counter runs).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ewoms-numerics_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
EWG_COPY_BYTES = 1 << 30   # include/ewg_capi.h: the copy probe moves this many bytes each way


def algorithmic_bytes(P, nnzb, fp32_precond=False, fp32_operator=False):
    """Bytes each kernel must move at least once per launch (FP64 values unless stated,
    implicit structured column indices; intermediate buffers of an implementation do
    not count). fp32_operator: BiCGStab's SpMV reads the FP32 copy of the blocks."""
    N = P.N
    blk = 32
    oblk = 16 if fp32_operator else 32  # operator blocks read by the SpMV
    fblk = 16 if fp32_precond else 32  # preconditioner factor blocks
    fval = 4 if fp32_precond else 8    # the forward sweep's intermediate vector
    lower = (nnzb - N) // 2
    per_cell_in = 16 + 8 + 8 + 16 + (8 if P.pe is not None else 0) + \
        (4 if len(P.region_params) > 1 else 0)
    return {
        # x, K, phi, storage cache, [pe], [region] in; residual + all Jacobian blocks out
        "linearize": N * (per_cell_in + 16) + nnzb * blk,
        # all blocks + x in, y out
        "spmv": nnzb * oblk + 2 * 16 * N,
        # forward: L + d in, w out; backward: U + Dinv + w in, v out
        "ilu_apply": 2 * lower * fblk + N * fblk + (16 + 2 * fval + 2 * fval + 16) * N,
        # A (lower, diagonal, upper) in; L + Dinv out (FP32: and the U copy)
        "ilu_factor": nnzb * blk + lower * fblk + N * fblk + (lower * fblk if fp32_precond else 0),
    }


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (profiles/pmc_<kernel>_het3d_128x128x64.md, line 'traffic_bytes_per_launch: <n>'),
    or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_het3d_128x128x64.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        return None
    return None


def timed_run(ctx, P, warmup, steps):
    """W untimed warm-up steps, then the state is reset to the initial condition and the
    first K steps of the configured run are timed (the first step, where the pressure
    field is established, is the most expensive one and stays in the timed region)."""
    if warmup:
        ctx.run(0.0, warmup * P.dt, P.dt)
        ctx.set_solution(P.x0, 0)
        ctx.set_solution(P.x0, 1)
    ctx.reset_stats()
    t0 = time.perf_counter()
    n = ctx.run(0.0, steps * P.dt, P.dt)
    t1 = time.perf_counter()
    return n, t1 - t0, ctx.stats()


FORMS = {
    "fp64": {},
    "fp32_precond": {"precond_fp32": 1},
    "fp32_flux": {"linearize_fp32": 1},
    "mixed": {"precond_fp32": 1, "linearize_fp32": 1},
    "learned_mixed": {"precond_fp32": 1, "linearize_fp32": 1, "warm_start": 1},
    "learned": {"precond_fp32": 1, "warm_start": 1},
    "mixed_op": {"precond_fp32": 1, "linearize_fp32": 1, "operator_fp32": 1},
    "learned_mixed_op": {"precond_fp32": 1, "linearize_fp32": 1, "operator_fp32": 1,
                         "warm_start": 1},
}
FORM_DETAIL = {
    "fp64": "FP64 throughout",
    "fp32_precond": "FP32 ILU(0) factors and sweeps inside FP64 BiCGStab",
    "fp32_flux": "FP32 face fluxes accumulated into the FP64 residual and Jacobian",
    "mixed": "FP32 ILU(0) factors and sweeps + FP32 face fluxes; FP64 state, residual, "
             "Newton norms, Krylov recurrences and update",
    "learned_mixed": "mixed + Newton initial guess from the BF16 MFMA network (saturation "
                     "update), judged by the unchanged Newton test with fallback; FP64 state, "
                     "residual, Newton norms, Krylov recurrences and update",
    "learned": "FP32 ILU(0) factors and sweeps inside FP64 BiCGStab + Newton initial guess "
               "from the BF16 MFMA network (saturation update), judged by the unchanged Newton "
               "test with fallback; FP64 state, fluxes, residual, Jacobian, Newton norms, "
               "Krylov recurrences and update",
    "mixed_op": "mixed + BiCGStab's operator products read the Jacobian's blocks rounded to "
                "FP32 (FP64 products and sums)",
    "learned_mixed_op": "learned_mixed + BiCGStab's operator products read the Jacobian's "
                        "blocks rounded to FP32 (FP64 products and sums)",
}
# the arithmetic each form computes in: FP64 state, residual, Newton norms, Krylov
# recurrences and update everywhere; FP32 ILU(0) factors and sweeps, FP32 face fluxes and
# the BF16 network (FP32 accumulation) where the form uses them
DTYPE = {"fp64": "f64", "fp32_precond": "f64+f32", "fp32_flux": "f64+f32",
         "mixed": "f64+f32", "learned_mixed": "f64+f32+bf16", "learned": "f64+f32+bf16",
         "mixed_op": "f64+f32", "learned_mixed_op": "f64+f32+bf16"}
MFMA_BF16_PEAK_TFLOPS = 2500.0
CNN_FLOP_PER_CELL = 2 * 88992  # multiply-adds of the five 3x3x3 convolutions, x2


def form_params(form, **extra):
    import ewg
    return ewg.default_params(newton_tolerance=1e-6, linear_tolerance=1e-3, **FORMS[form],
                              **extra)


def converged_reference(P, steps, tol=1e-9):
    """Saturations per step of the FP64 path at a Newton tolerance far below the
    benchmark's (1e-9 instead of 1e-6, BiCGStab relative 1e-8): the converged solution the
    1e-6 trajectories, FP64 included, are measured against."""
    import ewg
    ref = ewg.GpuContext(P, ewg.default_params(newton_tolerance=tol, linear_tolerance=1e-8,
                                               max_time_step_size=P.dt, **FORMS["fp64"]))
    out = []
    for _ in range(steps):
        rc, it, err = ref.newton(P.dt)
        if rc != 0:
            raise RuntimeError(f"converged reference: Newton failed (rc {rc})")
        ref.advance_time_level()
        out.append(ref.solution(0)[1::2].copy())
    ref.close()
    return out


def per_step_deviation(P, form, steps, net=None, converged=None):
    """Steps the form and the FP64 path side by side from the initial condition: max |dS_w|,
    max |dp_w| / max |p_w| of the FP64 state, Newton and BiCGStab iterations per step
    (the two trajectories drift apart by where each Newton solve stops within its
    tolerance). max_dS_one_step: the form's step taken from the FP64 path's state of the
    step before, against the FP64 step (no drift). With `converged` (saturations per step
    of a tightly converged FP64 run): max_dS_vs_converged and max_dS_fp64_vs_converged,
    each trajectory's distance from the converged solution."""
    import numpy as np
    import ewg
    ref = ewg.GpuContext(P, form_params("fp64", max_time_step_size=P.dt))
    oth = ewg.GpuContext(P, form_params(form, max_time_step_size=P.dt))
    one = ewg.GpuContext(P, form_params(form, max_time_step_size=P.dt))
    if net is not None:
        from ewg import cnn
        cnn.attach(oth, net)
        cnn.attach(one, net)
    out = {"max_dS": [], "max_rel_dp": [], "newton": [], "newton_fp64": [],
           "krylov": [], "krylov_fp64": [], "max_dS_one_step": [],
           "max_dS_vs_converged": [], "max_dS_fp64_vs_converged": []}
    for _ in range(steps):
        its = []
        xprev = ref.solution(0)
        for c in (ref, oth):
            k0 = c.stats()["linear_iterations"]
            rc, it, err = c.newton(P.dt)
            if rc != 0:
                raise RuntimeError(f"{form}: Newton failed (rc {rc})")
            its.append((it, c.stats()["linear_iterations"] - k0))
            c.advance_time_level()
        one.set_solution(xprev, 1)
        one.set_solution(xprev, 0)
        rc, _, _ = one.newton(P.dt)
        if rc != 0:
            raise RuntimeError(f"{form}: Newton failed from the FP64 state (rc {rc})")
        xr, xo = ref.solution(0), oth.solution(0)
        out["max_dS"].append(float(np.abs(xo[1::2] - xr[1::2]).max()))
        out["max_rel_dp"].append(float(np.abs(xo[0::2] - xr[0::2]).max() /
                                       np.abs(xr[0::2]).max()))
        out["max_dS_one_step"].append(float(np.abs(one.solution(0)[1::2] - xr[1::2]).max()))
        if converged is not None:
            k = len(out["max_dS_vs_converged"])
            out["max_dS_vs_converged"].append(float(np.abs(xo[1::2] - converged[k]).max()))
            out["max_dS_fp64_vs_converged"].append(float(np.abs(xr[1::2] - converged[k]).max()))
        out["newton_fp64"].append(its[0][0])
        out["newton"].append(its[1][0])
        out["krylov_fp64"].append(its[0][1])
        out["krylov"].append(its[1][1])
    ref.close()
    oth.close()
    one.close()
    return out


def warm_start_network(P, args):
    """Train the warm-start network (outside every timed region) and measure its forward
    pass on the full grid: device time of the five convolutions and the matrix-core
    utilisation at algorithmic FLOPs."""
    import numpy as np
    from ewg import cnn
    net, info = cnn.train_warm_start(P.nx, P.ny, P.nz, n_pairs=args.train_pairs,
                                     epochs=args.train_epochs, seed=args.train_seed)
    feat = net.features(P.x0, P.K, P.phi, P.dt)
    ms = []
    for r in range(6):
        net.forward(feat, 1)
        if r:
            ms.append(net.last_forward_ms())
    fms = float(np.median(ms))
    tflops = CNN_FLOP_PER_CELL * P.N / (fms * 1e-3) / 1e12
    info["forward_ms"] = fms
    info["forward_TFLOPs"] = round(tflops, 1)
    info["mfma_frac"] = round(tflops / MFMA_BF16_PEAK_TFLOPS, 4)
    print(f"# warm-start network: {info}", file=sys.stderr, flush=True)
    return net, info


def host_cpu():
    """CPU model (from /proc/cpuinfo) and the cores this process may use."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, usable, os.cpu_count() or usable


def _oracle_steps(P, sample_steps, threads):
    from oracle.pyoracle import OracleContext
    o = OracleContext(P)
    o.set_threads(threads)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        rc, its, err = o.newton(P.dt)
        if rc != 0:
            raise RuntimeError(f"oracle Newton failed: rc {rc}")
        o.advance_time_level()
    el = time.perf_counter() - t0
    s = o.stats()
    o.close()
    return P.N * sample_steps / el, el, s


def cpu_baseline(P, sample_steps, gpu_same_sample=None):
    """The oracle on this host over the first `sample_steps` steps of the same case:
    (i) one thread, the reference default ThreadsPerProcess = 1; (ii) the element loop of
    the linearization over every usable core (ThreadsPerProcess = nproc; the linear
    solver stays serial, as in the reference)."""
    model, usable, total = host_cpu()
    v1, el1, s = _oracle_steps(P, sample_steps, 1)
    vn, eln, _ = _oracle_steps(P, sample_steps, usable)
    out = {"value": v1, "unit": "cell-timesteps/s", "cores": 1, "kind": "port",
           "sample": (f"first {sample_steps} time step(s) of the same 128x128x64 case "
                      f"({s['newton_iterations']} linearizations, "
                      f"{s['linear_iterations']} BiCGStab iterations) in {el1:.1f} s"),
           "cpu_model": model, "nproc": total, "usable_cores": usable,
           "threaded_linearization": {"value": vn, "cores": usable, "seconds": round(eln, 2)}}
    if gpu_same_sample is not None:
        out["gpu_same_sample"] = gpu_same_sample
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--forms", default="fp64,fp32_precond,mixed,learned_mixed")
    ap.add_argument("--train-pairs", type=int, default=2000)
    ap.add_argument("--train-epochs", type=int, default=30)
    ap.add_argument("--train-seed", type=int, default=2)
    ap.add_argument("--no-deviation", action="store_true",
                    help="skip the untimed per-step comparison of the reduced-precision forms")
    ap.add_argument("--cpu-sample-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble leg (config 3)")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--copy-reps", type=int, default=10)
    ap.add_argument("--kernel", choices=["linearize", "spmv", "ilu_apply", "ilu_factor", "copy"],
                    help="time one kernel only (the first form of --forms) and exit")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--cold", action="store_true",
                    help="with --kernel: read 1 GiB of other data before every launch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.kernel:
        kernel_mode(args)
        return

    import ewg
    from ewg import problems
    from ewg.capi import (EWG_KERNEL_COPY, EWG_KERNEL_ILU_APPLY, EWG_KERNEL_ILU_FACTOR,
                          EWG_KERNEL_LINEARIZE, EWG_KERNEL_SPMV)

    P = problems.heterogeneous_3d(args.nx, args.ny, args.nz)
    print(f"# device {ewg.device_info()}; case {P.name} {args.nx}x{args.ny}x{args.nz} "
          f"N={P.N}", file=sys.stderr, flush=True)
    forms = {}
    converged = None
    kernels = {}
    copy_gbs = None
    net, net_info = None, None
    for form in args.forms.split(","):
        if form not in FORMS:
            raise SystemExit(f"bench.py: unknown form {form!r} (known: {', '.join(FORMS)})")
        fp32 = bool(FORMS[form].get("precond_fp32"))
        params = form_params(form, max_time_step_size=P.dt, kernel_events=1)
        ctx = ewg.GpuContext(P, params)
        if FORMS[form].get("warm_start"):
            if net is None:
                net, net_info = warm_start_network(P, args)
            from ewg import cnn
            cnn.attach(ctx, net)
        n, el, st = timed_run(ctx, P, args.warmup, args.steps)
        # launches inside the timed run, bracketed by device events on the library's
        # stream (the first apply and SpMV of each batch of BiCGStab iterations, every
        # factorization and linearization); mean duration and algorithmic bandwidth.
        # Read before the sample run below, which resets them.
        ab = algorithmic_bytes(P, ctx.nnzb, fp32, bool(FORMS[form].get("operator_fp32")))
        ks = {}
        for name, kid in (("linearize", EWG_KERNEL_LINEARIZE), ("spmv", EWG_KERNEL_SPMV),
                          ("ilu_apply", EWG_KERNEL_ILU_APPLY),
                          ("ilu_factor", EWG_KERNEL_ILU_FACTOR)):
            cnt, tot = ctx.kernel_samples(kid)
            if cnt == 0:
                continue
            ms = tot / cnt
            gbs = ab[name] / (ms * 1e-3) / 1e9
            ks[name] = {"mean_ms": ms, "samples": cnt, "bytes": ab[name], "achieved_GBs": gbs,
                        "frac": gbs / HBM_PEAK_GBS}
        warm = None
        if FORMS[form].get("warm_start"):
            from ewg import cnn
            ws = cnn.warm_stats(ctx)
            warm = {"accepted": ws["accepted"], "rejected": ws["rejected"],
                    "predict_s": ws["predict_s"],
                    "last_region_cells": net.last_inference_cells(), "network": net_info}
        # the same form over the CPU baseline's sample (the first steps from t = 0)
        ctx.set_solution(P.x0, 0)
        ctx.set_solution(P.x0, 1)
        nf, elf, _ = timed_run(ctx, P, 0, args.cpu_sample_steps)
        forms[form] = {"value": P.N * n / el, "steps": n, "seconds": el,
                       "first_steps": {"steps": nf, "seconds": elf, "value": P.N * nf / elf},
                       "newton_iterations": st["newton_iterations"],
                       "linear_iterations": st["linear_iterations"],
                       "linear_solves": st["linear_solves"],
                       "failed_steps": st["failed_steps"],
                       "split_s": {"linearize": st["time_linearize_s"],
                                   "solve": st["time_solve_s"],
                                   "update": st["time_update_s"]},
                       "kernels": ks}
        if warm is not None:
            forms[form]["warm_start"] = warm
        print(f"# {form} kernels {ks}", file=sys.stderr, flush=True)
        if form == "fp64":
            ctx.time_kernel(EWG_KERNEL_COPY, 1)
            ms = ctx.time_kernel(EWG_KERNEL_COPY, args.copy_reps)
            copy_bytes = 2 * EWG_COPY_BYTES
            copy_gbs = copy_bytes / (ms * 1e-3) / 1e9
            print(f"# copy {copy_bytes} bytes {ms:.4f} ms -> {copy_gbs:.0f} GB/s",
                  file=sys.stderr, flush=True)
        ctx.close()
        if form != "fp64" and not args.no_deviation:
            if converged is None:
                converged = converged_reference(P, args.steps)
            dev = per_step_deviation(P, form, args.steps,
                                     net if FORMS[form].get("warm_start") else None, converged)
            forms[form]["vs_fp64"] = dev
            print(f"# {form} vs fp64: max dS {max(dev['max_dS']):.3e} (one step from the same "
                  f"state {max(dev['max_dS_one_step']):.3e}), max rel dp "
                  f"{max(dev['max_rel_dp']):.3e}, newton {sum(dev['newton'])} vs "
                  f"{sum(dev['newton_fp64'])}, krylov {sum(dev['krylov'])} vs "
                  f"{sum(dev['krylov_fp64'])}; from the converged solution: "
                  f"{max(dev['max_dS_vs_converged']):.3e} (fp64 at the same tolerance "
                  f"{max(dev['max_dS_fp64_vs_converged']):.3e})", file=sys.stderr, flush=True)

    ens_line = None
    if not args.no_ensemble:
        ens_line = ensemble_leg(args)
        print(f"# ensemble {ens_line['members']} members {ens_line['cells']} cells: "
              f"{ens_line['value'] / 1e6:.1f}M cell-timesteps/s, replay {ens_line['tuning']['replay']}"
              + ("; spot check " + ", ".join(
                  f"member {c['member']} max dS {c['max_dS_vs_fp64_oracle']:.1e} newton "
                  f"{c['newton']}/{c['newton_oracle']}"
                  for c in ens_line["cpu_baseline"]["spot_check"])
                 if ens_line.get("cpu_baseline") else ""), file=sys.stderr, flush=True)

    # the metric's form: learned warm start + FP32 preconditioner (+ FP32 face fluxes) when
    # it ran, compared with the plain FP64 path
    head_form = next((f for f in ("learned_mixed", "fp64") if f in forms), next(iter(forms)))
    head = forms[head_form]
    kernels = head["kernels"]
    # dominant kernel: the largest share of the timed run's device time (per BiCGStab
    # iteration 2 ILU applies and 2 SpMV, per Newton iteration 1 linearization)
    dom_name = None
    if kernels:
        share = {"ilu_apply": 2 * head["linear_iterations"],
                 "spmv": 2 * head["linear_iterations"],
                 "linearize": head["newton_iterations"],
                 "ilu_factor": head["linear_solves"]}
        dom_name = max(kernels, key=lambda k: kernels[k]["mean_ms"] * share[k])
    roofline = None
    if dom_name:
        k = kernels[dom_name]
        prec = "fp32" if FORMS[head_form].get("precond_fp32") and dom_name.startswith("ilu") \
            else "fp64"
        roofline = {"kernel": f"{dom_name} ({prec})", "bound": "hbm",
                    "achieved": round(k["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(k["frac"], 4),
                    "traffic": pmc_traffic(dom_name if prec == "fp64" else f"{dom_name}_fp32")}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(P, args.cpu_sample_steps,
                           {"form": head_form, **head["first_steps"]})
    line = {
        "metric": "cell-timesteps/s to Newton tol 1e-6 (learned+FP32-precond vs plain FP64); HBM GB/s",
        "value": head["value"],
        "unit": "cell-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["seconds"] / max(1, head["steps"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head_form],
        "data": "synthetic (seeded log-normal permeability, spectral synthesis)",
        "config": {"workload": f"3D heterogeneous two-phase DNAPL injection {args.nx}x{args.ny}x{args.nz}"
                               f" cells, dt {P.dt:g} s, Newton tol 1e-6, BiCGStab+ILU0",
                   "form": head_form, "form_detail": FORM_DETAIL[head_form],
                   "cells": P.N, "seed": P.meta["seed"]},
        "plain_fp64": ({"value": forms["fp64"]["value"],
                        "ms_per_step": 1e3 * forms["fp64"]["seconds"] / max(1, forms["fp64"]["steps"])}
                       if "fp64" in forms else None),
        "speedup_vs_plain_fp64": (round(head["value"] / forms["fp64"]["value"], 4)
                                  if "fp64" in forms else None),
        "forms": forms,
        "kernels": kernels,
        "copy_GBs": round(copy_gbs, 1) if copy_gbs else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "ensemble": ens_line,
    }
    print(json.dumps(line), flush=True)
    # one compact closing line on stderr (the tail a log keeps)
    summ = [f"{f} {forms[f]['value'] / 1e6:.1f}M" for f in forms]
    dev = forms.get(head_form, {}).get("vs_fp64")
    print(f"# summary: {', '.join(summ)} cell-timesteps/s; {head_form} / fp64 = "
          f"{line['speedup_vs_plain_fp64']}"
          + (f"; max |dS| vs fp64 {max(dev['max_dS']):.2e} (one step "
             f"{max(dev['max_dS_one_step']):.2e})" if dev else "")
          + (f"; roofline {roofline['kernel']} {roofline['frac']:.3f}" if roofline else "")
          + (f"; copy {line['copy_GBs']} GB/s" if line["copy_GBs"] else "")
          + (f"; ensemble {ens_line['value'] / 1e6:.1f}M" if ens_line else ""),
          file=sys.stderr, flush=True)


def ensemble_leg(args):
    """The config-3 ensemble: realizations of the same physics in the sizes of
    problems.ENSEMBLE_SIZES (seed 3), FP64 state with the FP32 preconditioner (a solve that
    fails with it is redone, and the member's later solves made, with FP64 factors),
    Newton tolerance 1e-6, dt 500 s, `steps` steps, all members time-stepped in the batched
    mode, each at its own pace; the tuner picks the batched mode's BiCGStab replay depth on
    one trial step first (the persistent mode's trials, one step of every member in one
    workgroup each, take seconds apiece and are left to `python -m ewg.ensemble --tune`)."""
    import numpy as np
    from ewg import ensemble, problems
    t0 = time.perf_counter()
    probs, ctxs = ensemble.build_members(args.ensemble_members, 3, args.steps)
    ens = ensemble.Ensemble(ctxs)
    t1 = time.perf_counter()
    dt = probs[0].dt
    tuning = ens.tune(dt, modes=ensemble.TUNE_BATCHED)
    t2 = time.perf_counter()
    sec = ens.run(dt, args.steps)
    st = [ens.stats(i) for i in range(len(probs))]
    cells = np.array([P.N for P in probs])
    done = np.array([x["steps"] for x in st])
    out = {"metric": "ensemble cell-timesteps/s (cells x completed steps over all members / wall time)",
           "value": float((cells * done).sum() / sec), "unit": "cell-timesteps/s",
           "members": len(probs), "cells": int(cells.sum()), "steps": args.steps,
           "seconds": sec, "completed_steps": int(done.sum()),
           "failed_members": int(sum(1 for x in st if x["status"] != 0)),
           "step_divisions": int(sum(x["step_divisions"] for x in st)),
           "precond_retries": int(sum(x["precond_retries"] for x in st)),
           "newton_iterations": int(sum(x["newton_iterations"] for x in st)),
           "linear_iterations": int(sum(x["linear_iterations"] for x in st)),
           "sizes": {f"{a}x{b}x{c}": int(sum(1 for P in probs if (P.nx, P.ny, P.nz) == (a, b, c)))
                     for (a, b, c) in problems.ENSEMBLE_SIZES},
           "mode": "batched", "tuning": tuning, "setup_s": round(t1 - t0, 2),
           "tune_s": round(t2 - t1, 2),
           "workload": f"{len(probs)} realizations, {int(cells.sum())} cells, FP64 state, FP32 "
                       f"ILU(0) preconditioner (a failed solve redone with FP64 factors), "
                       f"Newton tol 1e-6, dt {dt:g} s, seed 3"}
    states = {i: ctxs[i].solution(0) for i in range(len(probs))}
    ens.close()
    for c in ctxs:
        c.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = ensemble_cpu_baseline(probs, st, states, args.steps)
    return out


def ensemble_cpu_baseline(probs, st, states, steps):
    """The oracle (FP64 throughout, one thread) on the two smallest members that completed
    every step undivided: its rate over those members, and, as a spot check of the
    ensemble's results, each member's saturations and Newton iterations against it."""
    import numpy as np
    import ewg
    from oracle.pyoracle import OracleContext
    pick = sorted((i for i, x in enumerate(st) if x["status"] == 0 and x["step_divisions"] == 0
                   and x["steps"] == steps), key=lambda i: (probs[i].N, i))[:2]
    checks, cells, seconds = [], 0, 0.0
    for i in pick:
        P = probs[i]
        o = OracleContext(P, ewg.default_params(newton_tolerance=1e-6, linear_tolerance=1e-3))
        t0 = time.perf_counter()
        newton, ok = 0, True
        for _ in range(steps):
            rc, it, _ = o.newton(P.dt)
            newton += it
            if rc != 0:
                ok = False
                break
            o.advance_time_level()
        seconds += time.perf_counter() - t0
        cells += P.N * steps
        x = o.solution(0)
        o.close()
        checks.append({"member": i, "grid": f"{P.nx}x{P.ny}x{P.nz}", "oracle_ok": ok,
                       "max_dS_vs_fp64_oracle": float(np.abs(states[i][1::2] - x[1::2]).max()),
                       "max_rel_dp": float(np.abs(states[i][0::2] - x[0::2]).max() /
                                           np.abs(x[0::2]).max()),
                       "newton": int(st[i]["newton_iterations"]), "newton_oracle": newton})
    return {"value": cells / seconds if seconds else None, "unit": "cell-timesteps/s",
            "cores": 1, "kind": "port",
            "sample": f"{len(pick)} smallest members, all {steps} steps, in {seconds:.1f} s",
            "spot_check": checks}


def kernel_mode(args):
    """One kernel of the 128x128x64 case launched `reps` times at the state after two time
    steps of the chosen form (for profiler and counter runs)."""
    import ewg
    from ewg import problems
    from ewg.capi import (EWG_KERNEL_COPY, EWG_KERNEL_ILU_APPLY, EWG_KERNEL_ILU_FACTOR,
                          EWG_KERNEL_LINEARIZE, EWG_KERNEL_SPMV)
    kid = {"linearize": EWG_KERNEL_LINEARIZE, "spmv": EWG_KERNEL_SPMV,
           "ilu_apply": EWG_KERNEL_ILU_APPLY, "ilu_factor": EWG_KERNEL_ILU_FACTOR,
           "copy": EWG_KERNEL_COPY}[args.kernel]
    form = args.forms.split(",")[0]
    P = problems.heterogeneous_3d(args.nx, args.ny, args.nz)
    ctx = ewg.GpuContext(P, form_params(form, max_time_step_size=P.dt))
    ctx.run(0.0, 2 * P.dt, P.dt)
    ctx.linearize(P.dt, 1, want=False)
    if args.cold:
        from ewg.capi import EWG_KERNEL_COLD
        kid |= EWG_KERNEL_COLD
    ctx.time_kernel(kid, 1)
    ms = ctx.time_kernel(kid, args.reps)
    nbytes = (2 * EWG_COPY_BYTES if args.kernel == "copy" else
              algorithmic_bytes(P, ctx.nnzb, bool(FORMS[form].get("precond_fp32")),
                                bool(FORMS[form].get("operator_fp32")))[args.kernel])
    print(json.dumps({"kernel": args.kernel, "form": form, "reps": args.reps, "cold": args.cold,
                      "mean_ms": ms,
                      "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9}), flush=True)


if __name__ == "__main__":
    main()
