"""Benchmark of the SWEHR coupled time step on one MI355X.

Default workload (the plain FP64 configuration): 2048 x 2048 cells, dx 1 m, fractal
topography, 7 size classes, 80 mm/h rain for 1800 s, per-cell Ks.  A "step" is one
coupled time step over the whole grid.  After ``--warmup`` untimed steps (the default
reaches ~600 s of simulated time, when runoff is established) exactly ``--steps`` steps
are timed with HIP events on the solver's stream, bracketed by device
synchronisations.  ``value`` is cell-class-steps/s = cells x size classes x steps / s.

Roofline: the fused step kernel's average duration is measured live (HIP events around
eagerly launched steps); its algorithmic HBM bytes per launch are cells x B_cell with
B_cell = 2 w (6 + 2 S) + 16 + w bytes (one read and one write of h, hu, hv, z, MAXDEPTH,
MAXVEL, ch_k, M_k at w bytes, VINF as FP64, one read of the Ks raster).  ``traffic`` is
the step kernel's measured HBM bytes per launch, read from profiles/traffic.json (made
with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a copy-kernel calibration).

CPU baseline: the FP64 C/OpenMP oracle on the host cores, same case, a bounded number
of steps from the initial state (the GPU is also timed on that window for reference).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime access for events and device synchronisation."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed: %d" % (what, rc))

    def event(self):
        e = C.c_void_p()
        self._ok(self.L.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e

    def record(self, e, stream):
        self._ok(self.L.hipEventRecord(e, C.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self._ok(self.L.hipEventSynchronize(b), "hipEventSynchronize")
        ms = C.c_float()
        self._ok(self.L.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
        return float(ms.value)

    def sync(self):
        self._ok(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")


def bytes_per_cell(S: int, w: int, ks_raster: bool = True) -> int:
    return 2 * w * (6 + 2 * S) + 16 + (w if ks_raster else 0)


def timed_run(hip, solver, steps):
    stream = solver.stream()
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    i0 = solver.info()
    hip.record(e0, stream)
    t0 = time.perf_counter()
    info = solver.run(steps)
    hip.record(e1, stream)
    hip.sync()
    wall = time.perf_counter() - t0
    ms = hip.elapsed_ms(e0, e1)
    return info["steps"] - i0["steps"], ms, wall, info


def cpu_baseline(case, budget_s: float, threads: int):
    """Oracle steps on the host cores from the case's initial state, ~budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    o = Oracle(case, threads=threads)
    o.run(1)  # first touch of the oracle's arrays, untimed
    t0 = time.perf_counter()
    o.run(1)
    one = time.perf_counter() - t0
    more = int(max(0, min(400, budget_s / max(one, 1e-6) - 1)))
    t0 = time.perf_counter()
    o.run(more)
    el = one + (time.perf_counter() - t0)
    o.close()
    return 1 + more, el


def traffic_for(key: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    v = d.get(key)
    return None if v is None else v.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)
    ap.add_argument("--warmup", type=int, default=2400)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--steps-per-graph", type=int, default=16)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch

        n = args.calib_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU sharding "
              "of the step is described in DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from swehr_amd import _abi
    from swehr_amd import solver as sv
    from swehr_amd.inputs import config_c2

    prec = _abi.FP64 if args.precision == "fp64" else _abi.FP32
    w = 8 if prec == _abi.FP64 else 4
    t0 = time.perf_counter()
    case = config_c2(n=args.grid)
    log("case %s: %d x %d, S=%d built in %.1f s" % (case.name, case.nx, case.ny, case.nclass, time.perf_counter() - t0))
    hip = Hip()
    S = case.nclass
    cells = case.nx * case.ny

    s = sv.Solver(case, precision=prec, steps_per_graph=args.steps_per_graph, tile=args.tile)
    # same window as the CPU baseline: the first steps from the initial state
    n_first, ms_first, _, _ = timed_run(hip, s, 20)
    s.run(max(0, args.warmup - n_first))
    log("warmup done: t=%.1f s after %d steps" % (s.info()["t"], s.info()["steps"]))
    n, ms, wall, info = timed_run(hip, s, args.steps)
    value = cells * S * n / (ms * 1e-3)
    log("timed %d steps: %.3f ms (wall %.3f s) -> %.4e cell-class-steps/s, t=%.1f" % (n, ms, wall, value, info["t"]))
    kms, taken = s.profile(args.profile_steps)
    bud = s.budget()
    s.close()

    bcell = bytes_per_cell(S, w)
    step_ms = float(kms[1])
    achieved = cells * bcell / (step_ms * 1e-3) / 1e9
    key = "c2_%d_%s" % (args.grid, args.precision)
    traffic = traffic_for(key)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "step_kernel", "kernel_ms": round(step_ms, 4), "bytes_per_cell": bcell,
                "cells_per_launch": cells}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ncpu, el = cpu_baseline(case, args.cpu_seconds, threads)
        cpu = {"value": cells * S * ncpu / el, "unit": "cell-class-steps/s", "cores": threads, "kind": "port",
               "sample": "FP64 C/OpenMP oracle, %d steps of the %dx%d case from its initial state (t=0), %.1f s; "
                         "GPU on the same first 20 steps: %.4e cell-class-steps/s"
                         % (ncpu, args.grid, args.grid, el, cells * S * n_first / (ms_first * 1e-3))}
        log("cpu baseline: %s" % cpu)

    wres = sv.water_residual(bud)
    wscale = max(abs(bud[k]) for k in (0, 1, 2, 3, 9, 10))
    out = {
        "metric": "cell-class-steps/s (cells x size classes x steps per second), plain FP64 step",
        "value": value,
        "unit": "cell-class-steps/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": ms / max(n, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == _abi.FP64 else "f32",
        "data": "synthetic (seeded diamond-square topography, log-normal Ks raster)",
        "config": {"workload": "plain %s step, %dx%d cells, dx 1 m, S=%d, 80 mm/h rain, per-cell Ks"
                               % (args.precision.upper(), args.grid, args.grid, S),
                   "grid": [case.nx, case.ny], "classes": S, "precision": args.precision,
                   "t_start_s": None, "t_end_s": info["t"], "tile": args.tile or "auto"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms": {"wave_speed": round(float(kms[0]), 4), "step": round(step_ms, 4),
                       "control": round(float(kms[2]), 4), "profiled_steps": taken},
        "water_budget_rel_residual": abs(wres) / wscale if wscale else 0.0,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
