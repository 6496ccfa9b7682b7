"""ewg.ensemble -- binding of include/ewg_ensemble.h: realizations of uneven size
time-stepped together, either as batched launches over the cells of all members (each
BiCGStab iteration of all members replayed from a HIP graph) or in one persistent launch
(one workgroup per member), the mode and launch shape chosen by the library's tuner.

`python -m ewg.ensemble` measures the ensemble workload of config 3 (256 members of the
config-1 physics, sizes from {32^3, 48^3, 64x64x32, 96x96x48}, FP32 preconditioner):
cell-timesteps/s = sum of cells x completed steps over all members / wall time of the run.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import GpuContext, load, problems
from .capi import default_params

_vp = C.c_void_p


class MemberStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("newton_iterations", C.c_int64),
                ("linear_iterations", C.c_int64), ("linear_solves", C.c_int64),
                ("status", C.c_int32), ("step_divisions", C.c_int32),
                ("last_error", C.c_double), ("precond_retries", C.c_int32)]


MAXC = 8


PERSISTENT, BATCHED = 0, 1
MODES = {PERSISTENT: "persistent", BATCHED: "batched"}


class Tuning(C.Structure):
    _fields_ = [("mode", C.c_int32), ("threads", C.c_int32), ("largest_first", C.c_int32),
                ("replay", C.c_int32), ("ncandidates", C.c_int32),
                ("cand_mode", C.c_int32 * MAXC), ("cand_threads", C.c_int32 * MAXC),
                ("cand_largest_first", C.c_int32 * MAXC), ("cand_replay", C.c_int32 * MAXC),
                ("cand_ms", C.c_double * MAXC)]


TUNE_BATCHED = 1     # EWG_ENSEMBLE_TUNE_BATCHED
TUNE_PERSISTENT = 2  # EWG_ENSEMBLE_TUNE_PERSISTENT


_BOUND = False


def _lib():
    global _BOUND
    lib = load()
    if not _BOUND:
        sig = {
            "ewg_ensemble_create": (C.c_int, [C.POINTER(_vp), C.c_int32, C.POINTER(_vp)]),
            "ewg_ensemble_destroy": (None, [_vp]),
            "ewg_ensemble_last_error": (C.c_char_p, [_vp]),
            "ewg_ensemble_tune": (C.c_int, [_vp, C.c_double, C.POINTER(Tuning)]),
            "ewg_ensemble_tune_modes": (C.c_int, [_vp, C.c_double, C.c_int32, C.POINTER(Tuning)]),
            "ewg_ensemble_set_replay": (C.c_int, [_vp, C.c_int32]),
            "ewg_ensemble_set_graph": (C.c_int, [_vp, C.c_int32]),
            "ewg_ensemble_set_compact": (C.c_int, [_vp, C.c_int32]),
            "ewg_ensemble_set_mode": (C.c_int, [_vp, C.c_int32, C.c_int32]),
            "ewg_ensemble_run": (C.c_int, [_vp, C.c_double, C.c_int32,
                                           C.POINTER(C.c_double)]),
            "ewg_ensemble_stats": (C.c_int, [_vp, C.c_int32, C.POINTER(MemberStats)]),
            "ewg_ensemble_force_sweep_failure": (C.c_int, [_vp, C.c_int32, C.c_int32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _BOUND = True
    return lib


def member_params(**overrides):
    """Parameters of an ensemble member: FP32 preconditioner by default, with a failed
    solve redone, and the member's later solves made, with FP64 factors (precond_fp32 = 2);
    with ilu_schedule = 0 (default) the batched mode sweeps the members' tile-wavefront
    layouts, with ilu_schedule = 1 (no tile buffers) their hyperplanes."""
    kw = dict(newton_tolerance=1e-6, linear_tolerance=1e-3, precond_fp32=2, ilu_schedule=0)
    kw.update(overrides)
    return default_params(**kw)


class Ensemble:
    """Members are GpuContext objects (kept alive by the ensemble)."""

    def __init__(self, contexts):
        self.lib = _lib()
        self.contexts = list(contexts)
        arr = (_vp * len(self.contexts))(*[c.h.value for c in self.contexts])
        h = _vp()
        rc = self.lib.ewg_ensemble_create(arr, len(self.contexts), C.byref(h))
        if rc != 0 or not h.value:
            raise RuntimeError(f"ewg_ensemble_create failed: {rc}")
        self.h = h

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.ewg_ensemble_last_error(self.h)
            raise RuntimeError(f"ewg_ensemble: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ewg_ensemble_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, dt, modes=TUNE_BATCHED | TUNE_PERSISTENT):
        """Time the candidates on one step and keep the fastest (ewg_ensemble_tune_modes);
        candidates of a mode left out of `modes` report ms = -1."""
        t = Tuning()
        self._check(self.lib.ewg_ensemble_tune_modes(self.h, dt, modes, C.byref(t)))
        return {"mode": MODES[t.mode], "threads": t.threads,
                "largest_first": bool(t.largest_first), "replay": t.replay,
                "candidates": [{"mode": MODES[t.cand_mode[k]], "threads": t.cand_threads[k],
                                "largest_first": bool(t.cand_largest_first[k]),
                                "replay": t.cand_replay[k], "ms": t.cand_ms[k]}
                               for k in range(t.ncandidates)]}

    def force_sweep_failure(self, member, solves=1):
        """Test hook: the member's next `solves` batched solves start with its sweep error
        flag set (ewg_ensemble_force_sweep_failure)."""
        self._check(self.lib.ewg_ensemble_force_sweep_failure(self.h, member, solves))

    def set_mode(self, mode, threads=512):
        self._check(self.lib.ewg_ensemble_set_mode(self.h, mode, threads))

    def set_graph(self, on):
        """Batched mode: replay the BiCGStab iteration from a HIP graph (True, default) or
        launch its kernels each time (False)."""
        self._check(self.lib.ewg_ensemble_set_graph(self.h, 1 if on else 0))

    def set_compact(self, on):
        """Batched mode: a round whose running solves belong to members holding at most
        half of the cells launches its BiCGStab iterations over those members only (True or
        1, default), every round does ("always" or 2), or none does (False or 0)."""
        v = 2 if on == "always" else int(on)
        self._check(self.lib.ewg_ensemble_set_compact(self.h, v))

    def set_replay(self, depth):
        """Batched mode: most BiCGStab iterations queued between state reads."""
        self._check(self.lib.ewg_ensemble_set_replay(self.h, depth))

    def run(self, dt, steps):
        sec = C.c_double(0)
        self._check(self.lib.ewg_ensemble_run(self.h, dt, steps, C.byref(sec)))
        return sec.value

    def stats(self, i):
        s = MemberStats()
        self._check(self.lib.ewg_ensemble_stats(self.h, i, C.byref(s)))
        return {n: getattr(s, n) for n, _ in s._fields_}


def build_members(n, seed, steps=40, sizes=problems.ENSEMBLE_SIZES, **params):
    """n realizations of the config-3 family and their contexts."""
    probs = [problems.ensemble_member(k, seed, sizes=sizes, steps=steps) for k in range(n)]
    ctxs = [GpuContext(P, member_params(max_time_step_size=P.dt, **params)) for P in probs]
    return probs, ctxs


def main(argv=None):
    ap = argparse.ArgumentParser(description="ensemble throughput (config 3)")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--dt", type=float, default=500.0)
    ap.add_argument("--tune", action="store_true",
                    help="time every mode on one step first (the persistent candidates take "
                         "long on large members)")
    ap.add_argument("--mode", choices=("batched", "persistent"), default="batched")
    ap.add_argument("--threads", type=int, default=512)
    ap.add_argument("--replay", type=int, default=0,
                    help="batched mode: BiCGStab iterations queued per round (0: the default)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched mode: launch the BiCGStab iteration's kernels instead of "
                         "replaying its graph (for kernel tracers)")
    ap.add_argument("--no-compact", action="store_true",
                    help="batched mode: launch every round's BiCGStab iterations over all "
                         "members")
    ap.add_argument("--ilu-schedule", type=int, default=0, choices=(0, 1),
                    help="0: tile-wavefront sweeps in the batched mode; 1: hyperplane sweeps")
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    probs, ctxs = build_members(args.members, args.seed, args.steps,
                                ilu_schedule=args.ilu_schedule)
    ens = Ensemble(ctxs)
    t1 = time.perf_counter()
    if args.tune:
        tuning = ens.tune(args.dt)
    else:
        ens.set_mode(BATCHED if args.mode == "batched" else PERSISTENT, args.threads)
        if args.replay:
            ens.set_replay(args.replay)
        if args.no_graph:
            ens.set_graph(False)
        if args.no_compact:
            ens.set_compact(False)
        tuning = {"mode": args.mode, "threads": args.threads, "replay": args.replay,
                  "compact": not args.no_compact, "candidates": []}
    t2 = time.perf_counter()
    sec = ens.run(args.dt, args.steps)
    st = [ens.stats(i) for i in range(len(probs))]
    cells = np.array([P.N for P in probs])
    done = np.array([s["steps"] for s in st])
    out = {
        "metric": "ensemble cell-timesteps/s (sum of cells x steps over all members / wall time)",
        "value": float((cells * done).sum() / sec),
        "unit": "cell-timesteps/s",
        "members": len(probs), "cells": int(cells.sum()), "steps": args.steps,
        "seconds": sec, "completed_steps": int(done.sum()),
        "failed_members": int(sum(1 for s in st if s["status"] != 0)),
        "step_divisions": int(sum(s["step_divisions"] for s in st)),
        "newton_iterations": int(sum(s["newton_iterations"] for s in st)),
        "linear_iterations": int(sum(s["linear_iterations"] for s in st)),
        "sizes": {f"{a}x{b}x{c}": int(sum(1 for P in probs if (P.nx, P.ny, P.nz) == (a, b, c)))
                  for (a, b, c) in problems.ENSEMBLE_SIZES},
        "tuning": tuning, "ilu_schedule": args.ilu_schedule,
        "setup_s": t1 - t0, "tune_s": t2 - t1,
        "config": {"workload": "config-3 ensemble: FP64 state, FP32 ILU(0) preconditioner, "
                               f"dt {args.dt:g} s, Newton tol 1e-6", "seed": args.seed},
    }
    print(json.dumps(out))
    ens.close()
    for c in ctxs:
        c.close()


if __name__ == "__main__":
    sys.exit(main())
