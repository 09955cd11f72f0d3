"""Benchmark of the SWEHR coupled time step on one MI355X.

Default workload (the plain FP64 configuration): 2048 x 2048 cells, dx 1 m, fractal
topography, 7 size classes, 80 mm/h rain for 1800 s, per-cell Ks.  A "step" is one
coupled time step over the whole grid.  After at least ``--warmup`` untimed steps, and
in any case not before ``--t-start`` (600 s) of simulated time, when runoff is
established, exactly ``--steps`` steps are timed with HIP events on the solver's
stream, bracketed by device synchronisations; ``config.t_window_s`` gives the
simulated times the window spans.  ``value`` is cell-class-steps/s = cells x size
classes x steps / s.

Roofline: the fused step kernel's average duration is measured live over the same
steps as the timed window (the window's steps are launched one by one with HIP events
around each kernel, the state is put back from a device copy, then the window is timed
as graph launches); its algorithmic HBM bytes per launch are cells x B_cell with
B_cell = 2 w (6 + 2 S) + 16 + w bytes (one read and one write of h, hu, hv, z, MAXDEPTH,
MAXVEL, ch_k, M_k at w bytes, VINF as FP64, one read of the Ks raster).  The FP32 form
keeps z and the Ks raster in FP64, 12 B per cell more than this count credits it with.  ``traffic`` is
the step kernel's measured HBM bytes per launch, read from profiles/traffic.json (made
with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a copy-kernel calibration).

CPU baseline: the FP64 C/OpenMP oracle on the host cores, same case, a bounded number
of steps from the initial state (the GPU is also timed on that window for reference).
"""
from __future__ import annotations

import argparse
import copy
import ctypes as C
import json
import os
import subprocess
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

    def copy_bandwidth(self, mib: int = 1024, reps: int = 5) -> float:
        """Device-to-device copy of mib MiB, reps times, timed with events: GB/s of HBM
        traffic (bytes read + bytes written), the measured counterpart of the 8 TB/s spec."""
        n = mib << 20
        a, b = C.c_void_p(), C.c_void_p()
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemsetD8.argtypes = [C.c_void_p, C.c_ubyte, C.c_size_t]
        self.L.hipMemcpyDtoDAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        self._ok(self.L.hipMalloc(C.byref(a), n), "hipMalloc")
        self._ok(self.L.hipMalloc(C.byref(b), n), "hipMalloc")
        try:
            self._ok(self.L.hipMemsetD8(a, 1, n), "hipMemsetD8")
            self._ok(self.L.hipMemcpyDtoDAsync(b, a, n, None), "hipMemcpyDtoDAsync")  # warm
            e0, e1 = self.event(), self.event()
            self.sync()
            self.record(e0, 0)
            for _ in range(reps):
                self._ok(self.L.hipMemcpyDtoDAsync(b, a, n, None), "hipMemcpyDtoDAsync")
            self.record(e1, 0)
            ms = self.elapsed_ms(e0, e1)
        finally:
            self.L.hipFree(a)
            self.L.hipFree(b)
        return 2.0 * n * reps / (ms * 1e-3) / 1e9


def bytes_per_cell(S: int, w: int, ks_raster: bool = True) -> int:
    return 2 * w * (6 + 2 * S) + 16 + (w if ks_raster else 0)


def timed_run(hip, solver, steps):
    stream = solver.stream()
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    i0 = solver.info()
    hip.record(e0, stream)
    t0 = time.perf_counter()
    # one call runs the steps (the solver does not pause at print intervals here); run()
    # takes any that a pause or the end of the run left
    info = solver.advance(steps)
    if info["steps"] - i0["steps"] < steps and not info["done"]:
        info = solver.run(steps - (info["steps"] - i0["steps"]))
    hip.record(e1, stream)
    hip.sync()
    wall = time.perf_counter() - t0
    ms = hip.elapsed_ms(e0, e1)
    return info["steps"] - i0["steps"], ms, wall, info


def graph_length(steps: int, most: int) -> int:
    """Steps per captured graph for a window of `steps`: the largest divisor of steps
    not above max(most, 64), so the window is whole graphs."""
    top = max(1, max(most, 64))
    return max(d for d in range(1, min(steps, top) + 1) if steps % d == 0)


def advance_to(solver, t_target: float, maxstep: float):
    """Step until the simulated time reaches t_target.  A step is at most maxstep long,
    so ceil((t_target - t) / maxstep) steps never overshoot by more than one step."""
    import math

    info = solver.info()
    while info["t"] < t_target and not info["done"]:
        info = solver.run(max(1, int(math.ceil((t_target - info["t"]) / maxstep))))
    return info


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


def reference_binary_baseline(case, budget_s: float, step_s_estimate: float):
    """The compiled reference on the same case: SWEHR.c built from its source by
    oracle/Makefile with the parameter hooks of oracle/ref_hooks.h (inert unless set;
    here they carry the case's Ks raster, soil values, courant number and switches).
    It is single-threaded: its OpenMP pragmas are commented out (SWEHR.c:1091, 1416).
    One run stops after 1 step and one after 1 + n steps; the difference of their wall
    times is n steps, so parsing, set-up and the final raster writes cancel.  Returns
    (n, seconds, seconds of the 1-step run), or None where the build is absent."""
    import shutil
    import tempfile

    binary = os.path.join(ROOT, "oracle", "_ref", "swehr_ref_hooked")
    if not os.path.exists(binary):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from cases import hook_env
    from swehr_amd.inputs import write_case

    d = tempfile.mkdtemp(prefix="swehr_refbase_")
    try:
        write_case(case, d)

        def run(maxsteps):
            env = dict(os.environ)
            env.update(hook_env(case, d, maxsteps))
            t0 = time.perf_counter()
            p = subprocess.run([binary], cwd=d, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                               timeout=900)
            if p.returncode != 0:
                raise RuntimeError("reference build failed (%d): %s" % (p.returncode, p.stderr[-500:]))
            return time.perf_counter() - t0

        n = int(max(2, min(200, budget_s / max(step_s_estimate, 1e-6))))
        t_one = run(1)
        t_more = run(1 + n)
        return n, t_more - t_one, t_one
    finally:
        shutil.rmtree(d, ignore_errors=True)


def parity_c1(threads: int):
    """The reference's CPU-runnable configuration (100 x 40 plane, 1200 s) run in full by
    the FP64 oracle on the host cores (timed: the CPU baseline at this size) and by the
    GPU in FP64 and FP32, with the errors of each GPU result against the oracle's."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from swehr_amd import _abi
    from swehr_amd import solver as sv
    from swehr_amd.inputs import config_c1

    case = config_c1()
    S, cells = case.nclass, case.nx * case.ny

    def results(x):
        st, _ = x.stage()
        q, qs = x.discharge()
        return {"depth": x.field("depth"), "speed": x.field("speed"),
                "sed": [x.field("sed", k + 1) for k in range(S)], "stage_depth": st[1], "q": q, "qs": qs,
                "budget": x.budget(), "steps": x.info()["steps"]}

    o = Oracle(case, threads=threads)
    t0 = time.perf_counter()
    o.run()
    cpu_s = time.perf_counter() - t0
    ref = results(o)
    o.close()
    mid = copy.copy(case)
    mid.tend = 600.0  # the end of the rain: the plane is running off
    o = Oracle(mid, threads=threads)
    o.run()
    ref_mid = results(o)
    o.close()
    out = {"workload": "%dx%d plane, dx %.1f m, S=%d, %.0f s (%d steps), whole run" % (case.nx, case.ny, case.dx, S,
                                                                                   case.tend, ref["steps"]),
           "cpu_oracle": {"value": cells * S * ref["steps"] / cpu_s, "seconds": round(cpu_s, 3), "cores": threads},
           "mass_balance_rel_oracle": dict(zip(("water", "sediment"), budget_errors(ref["budget"])))}
    for name, prec in (("fp64", _abi.FP64), ("fp32", _abi.FP32)):
        with sv.Solver(case, precision=prec) as s:
            s.run()
            g = results(s)
        n = min(len(g["q"]), len(ref["q"]))
        out[name] = {
            "steps": g["steps"],
            "rel_l2_depth": rel_l2(g["depth"], ref["depth"]),
            "rel_l2_speed": rel_l2(g["speed"], ref["speed"]),
            "rel_l2_concentration_per_class": [rel_l2(a, b) for a, b in zip(g["sed"], ref["sed"])],
            "rel_l2_outlet_hydrograph": rel_l2(g["q"][:n], ref["q"][:n]),
            "rel_l2_outlet_sedigraph": rel_l2(g["qs"][:n], ref["qs"][:n]),
            "rel_l2_outlet_stage_depth": rel_l2(g["stage_depth"][:n], ref["stage_depth"][:n]),
            "mass_balance_rel": dict(zip(("water", "sediment"), budget_errors(g["budget"]))),
        }
        with sv.Solver(mid, precision=prec) as s:
            s.run()
            gm = results(s)
        out[name]["at_600s"] = {
            "rel_l2_depth": rel_l2(gm["depth"], ref_mid["depth"]),
            "rel_l2_speed": rel_l2(gm["speed"], ref_mid["speed"]),
            "rel_l2_concentration_per_class": [rel_l2(a, b) for a, b in zip(gm["sed"], ref_mid["sed"])]}
    return out


def host_description() -> str:
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return "%s, %d logical CPUs visible" % (model, os.cpu_count() or 0)


def traffic_for(key: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    v = d.get(key)
    return None if v is None else v.get("bytes_per_launch")


def rel_l2(a, b):
    import numpy as np

    m = np.isfinite(b) & np.isfinite(a)
    den = float(np.linalg.norm(b[m]))
    return float(np.linalg.norm(a[m] - b[m]) / den) if den > 0 else 0.0


def budget_errors(b):
    from swehr_amd import solver as sv

    wscale = max(abs(b[k]) for k in (0, 1, 2, 3, 9, 10))
    sscale = max(max(abs(b[k]) for k in (4, 5, 7, 8)), 1e-30)
    return abs(sv.water_residual(b)) / wscale, abs(sv.sediment_residual(b)) / sscale


def bench_precision(hip, grid: int, t_check: float, window: int):
    """Precision form on the channel-network case: FP64 and FP32 from the same initial
    state to t_check (errors of the FP32 result against FP64), then a timed window of
    each at that state (throughput and the step kernel's roofline)."""
    from swehr_amd import _abi
    from swehr_amd import solver as sv
    from swehr_amd.inputs import config_c3

    case = config_c3(n=grid)
    S, cells = case.nclass, case.nx * case.ny
    case.tend = t_check
    out = {"workload": "%dx%d channel network, S=%d, triangular hyetograph (peak 120 mm/h), per-cell Ks"
                       % (grid, grid, S), "t_check_s": t_check}
    res = {}
    for name, prec in (("fp64", _abi.FP64), ("fp32", _abi.FP32)):
        with sv.Solver(case, precision=prec) as s:
            s.run()
            st, _ = s.stage()
            S_fields = {"depth": s.field("depth"), "speed": s.field("speed"), "stage": st}
            S_fields["sed"] = [s.field("sed", k + 1) for k in range(S)]
            res[name] = (S_fields, s.info(), s.budget())
    # sensitivity of the FP64 solution itself at FP32 resolution: the same run with the
    # topography perturbed by 6e-8 relative (FP32's unit roundoff), and with the rain
    # rates lowered by 6e-8 (discrete events in the step amplify any change)
    floor = {"depth": 0.0, "speed": 0.0}
    b64 = res["fp64"][0]
    for pert in [perturbed_at_fp32_resolution(case, 7)[k] for k in (0, 3)]:
        with sv.Solver(pert, precision=_abi.FP64) as s:
            s.run()
            floor["depth"] = max(floor["depth"], rel_l2(s.field("depth"), b64["depth"]))
            floor["speed"] = max(floor["speed"], rel_l2(s.field("speed"), b64["speed"]))
    a, b = res["fp32"][0], res["fp64"][0]
    n = min(a["stage"].shape[1], b["stage"].shape[1])
    out["fp32_vs_fp64"] = {
        "rel_l2_depth": rel_l2(a["depth"], b["depth"]),
        "rel_l2_speed": rel_l2(a["speed"], b["speed"]),
        "rel_l2_concentration_per_class": [rel_l2(x, y) for x, y in zip(a["sed"], b["sed"])],
        "rel_l2_outlet_hydrograph": rel_l2(a["stage"][1, :n], b["stage"][1, :n]),
        "rel_l2_outlet_sedigraph": rel_l2(a["stage"][3, :n], b["stage"][3, :n]),
        "steps": [res["fp32"][1]["steps"], res["fp64"][1]["steps"]]}
    out["fp64_sensitivity"] = {
        "perturbation": "FP64 runs with the topography x (1 + 6e-8 N(0,1)), and with the rain x (1 - 6e-8) "
                        "(FP32 resolution): the larger rel L2",
        "rel_l2_depth": floor["depth"], "rel_l2_speed": floor["speed"]}
    for name in ("fp64", "fp32"):
        we, se = budget_errors(res[name][2])
        out["mass_balance_rel_%s" % name] = {"water": we, "sediment": se}
    # timed windows, continuing beyond t_check
    case.tend = 3600.0
    for name, prec, w in (("fp64", _abi.FP64, 8), ("fp32", _abi.FP32, 4)):
        with sv.Solver(case, precision=prec) as s:
            s.run(int(4 * t_check))
            best, _ = s.autotune(5)
            n1, ms, _, info = timed_run(hip, s, window)
            kms, _ = s.profile(10)
            bcell = bytes_per_cell(S, w)
            out[name] = {"value": cells * S * n1 / (ms * 1e-3), "ms_per_step": ms / max(n1, 1),
                         "step_kernel_ms": float(kms[1]),
                         "roofline_frac": cells * bcell / (float(kms[1]) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "bytes_per_cell": bcell, "t_s": info["t"], "tile": best}
    out["fp32_over_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    return out


def perturbed_at_fp32_resolution(case, seed: int):
    """The case with one input changed at FP32 resolution (6e-8 = 2^-24, FP32's unit
    roundoff), in four ways: the topography and the Ks raster (or scalar Ks) multiplied
    by (1 + 6e-8 N(0,1)), and the rain rates by (1 + 6e-8) and (1 - 6e-8)."""
    import numpy as np

    g = np.random.default_rng(seed).standard_normal(case.topo.shape)
    out = []
    p = copy.copy(case)
    p.topo = case.topo * (1 + 6e-8 * g)
    out.append(p)
    p = copy.copy(case)
    if case.ks_raster is not None:
        p.ks_raster = case.ks_raster * (1 + 6e-8 * g)
    else:
        p.sidecar = dict(case.sidecar, ks=case.side("ks") * (1 + 6e-8))
    out.append(p)
    for sign in (1, -1):
        p = copy.copy(case)
        p.rain = case.rain * (1 + sign * 6e-8)
        out.append(p)
    return out


def fp32_resolution_floor(case, depth64, seed: int) -> float:
    """How far the FP64 depth moves when an input changes at FP32 resolution: the
    largest rel L2 of the FP64 runs of perturbed_at_fp32_resolution.  Where the run
    turns at the model's thresholds (wet/dry, the film depth, ponding onset), this is how
    far any FP32 evaluation of it can be expected to land from FP64 (the FP32 form holds
    the rain rate, the topography and the state at that resolution)."""
    from swehr_amd import _abi
    from swehr_amd import solver as sv

    out = 0.0
    for p in perturbed_at_fp32_resolution(case, seed):
        with sv.Solver(p, precision=_abi.FP64) as s:
            s.run()
            out = max(out, rel_l2(s.field("depth"), depth64))
    return out


def bench_batched(hip, count: int, scale_max: float, tend: float, plain_every: int, floor_max: int,
                  persistent: bool = False):
    """Ensemble of hillslopes, each run for its configured duration (tend): the FP32 form
    in one ensemble (every stage one launch over all instances) against the plain FP64
    step run instance after instance.  Accuracy: each member's FP32 depth against its
    FP64 depth at tend (the FP64 reference comes from an FP64 ensemble, which steps every
    member exactly as it steps alone).  Where a member's FP32 error exceeds 1e-4, its
    FP32-resolution floor (fp32_resolution_floor) shows how far FP64 itself moves there
    under input changes of FP32's size (up to floor_max members).  The plain FP64
    throughput is timed on every plain_every-th member's whole run.  With persistent, the
    FP32 ensemble's whole run is timed again in the persistent mode (one cooperative
    launch per 16 steps) and in the graph mode, at the same 256-thread tile shape."""
    from swehr_amd import _abi
    from swehr_amd import solver as sv
    from swehr_amd.inputs import config_c4
    import numpy as np

    cases = config_c4(count=count, scale_max=scale_max, tend=tend)
    cells = sum(c.nx * c.ny for c in cases)
    S = cases[0].nclass

    def whole_run(b):
        hip.sync()
        t0 = time.perf_counter()
        left = b.advance()
        hip.sync()
        return time.perf_counter() - t0, left

    with sv.Batch(cases, precision=_abi.FP32) as b:
        best, tile_ms = b.autotune(3)
        batch_s, left = whole_run(b)
        if left:
            raise RuntimeError("%d ensemble members did not finish" % left)
        steps32 = [b.member(i).info()["steps"] for i in range(len(cases))]
        depth32 = [b.member(i).field("depth") for i in range(len(cases))]
        mb32 = [budget_errors(b.member(i).budget()) for i in range(len(cases))]
    work32 = sum(c.nx * c.ny * S * n for c, n in zip(cases, steps32))
    log("C4 FP32 ensemble: %d members to t = %.0f s in %.2f s" % (len(cases), tend, batch_s))
    pers = None
    if persistent:
        ptile = best if best in (1, 2, 3) else 1
        walls = {}
        for mode in (False, True):
            with sv.Batch(cases, precision=_abi.FP32, tile=ptile) as b:
                b.set_persistent(mode)
                walls[mode], left = whole_run(b)
                if mode:
                    for i in range(len(cases)):
                        if b.member(i).info()["steps"] != steps32[i] or not np.array_equal(b.member(i).field("depth"),
                                                                                             depth32[i]):
                            raise RuntimeError("member %d: persistent run differs from the graph run" % i)
        pers = {"tile": ptile, "graph_wall_s": walls[False], "persistent_wall_s": walls[True],
                "persistent_over_graph_time": walls[True] / walls[False]}
        log("C4 FP32 persistent mode: %.2f s vs %.2f s graph (tile %d)" % (walls[True], walls[False], ptile))
    with sv.Batch(cases, precision=_abi.FP64) as b:
        b64_s, left = whole_run(b)
        steps64 = [b.member(i).info()["steps"] for i in range(len(cases))]
        depth64 = [b.member(i).field("depth") for i in range(len(cases))]
    log("C4 FP64 ensemble (reference depths): %.2f s" % b64_s)
    plain_s, plain_cc, plain_n = 0.0, 0, 0
    for i in range(0, len(cases), max(1, plain_every)):
        with sv.Solver(cases[i], precision=_abi.FP64) as s:
            n, ms, wall, _ = timed_run(hip, s, 1 << 40)
            plain_s += wall
            plain_cc += cases[i].nx * cases[i].ny * S * n
            plain_n += 1
            if not np.array_equal(s.field("depth"), depth64[i]):
                raise RuntimeError("member %d: the FP64 ensemble differs from the plain FP64 run" % i)
    log("C4 FP64 plain: %d members in %.2f s" % (plain_n, plain_s))
    err = [rel_l2(a, r) for a, r in zip(depth32, depth64)]
    num = sum(float(np.sum((a - r) ** 2)) for a, r in zip(depth32, depth64))
    den = sum(float(np.sum(r ** 2)) for r in depth64)
    floors = {}
    for i in sorted(range(len(cases)), key=lambda k: -err[k]):
        if err[i] <= 1e-4 or len(floors) >= floor_max:
            break
        floors[i] = fp32_resolution_floor(cases[i], depth64[i], i)
    e = np.asarray(err)
    over = [i for i in range(len(cases)) if err[i] > 1e-4]
    return {"workload": "%d hillslopes, %d cells total, S=%d, each run for its configured %.0f s"
                        % (len(cases), cells, S, tend),
            "fp32_batched": {"value": work32 / batch_s, "wall_s": batch_s, "member_steps": int(sum(steps32)),
                             "tile": best, "autotune_ms_per_step": [round(float(x), 4) for x in tile_ms]},
            "fp64_plain_sequential": {"value": plain_cc / plain_s, "wall_s": plain_s,
                                      "members_timed": plain_n, "every": plain_every},
            "ratio": (work32 / batch_s) / (plain_cc / plain_s),
            "persistent_mode": pers,
            "fp32_vs_fp64_at_tend": {
                "rel_l2_depth_all_members": (num / den) ** 0.5 if den > 0 else 0.0,
                "rel_l2_depth_worst_member": float(e.max()),
                "rel_l2_depth_median_member": float(np.median(e)),
                "members_above_1e-4": int(np.sum(e > 1e-4)),
                "rel_l2_depth_per_member": [float("%.2g" % x) for x in err],
                "steps_fp32_minus_fp64_max": int(max(abs(a - b) for a, b in zip(steps32, steps64))),
                "fp32_resolution_floor": {
                    "what": "largest rel L2 of FP64 depth when one input changes at FP32 resolution (topography "
                            "or Ks raster x (1 + 6e-8 N(0,1)), rain x (1 +- 6e-8)), for the members above 1e-4",
                    "per_member": {str(k): float("%.3g" % v) for k, v in floors.items()},
                    "fp32_error_over_floor_max": max((err[k] / v for k, v in floors.items() if v > 0), default=None),
                    "members_above_1e-4_without_floor": [k for k in over if k not in floors]},
                "mass_balance_rel_worst_member": {"water": max(x[0] for x in mb32),
                                                  "sediment": max(x[1] for x in mb32)}}}


def serve(srv, reqs, arrivals):
    """Submit reqs at the given arrival times (seconds from now; all 0 for a closed batch)
    and pump the server until every request has finished.  Returns (latencies in
    seconds, completion minus arrival, per request; work in cell-class-steps; wall;
    (seconds, cell-class-steps) when the queue first stood empty after the last
    submission, from the server's work counter, which counts unfinished requests too)."""
    import numpy as np

    n = len(reqs)
    lat, work = {}, 0.0
    t0 = time.perf_counter()
    w0 = srv.work()
    sub, nxt, drained = {}, 0, None
    while len(lat) < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            rid = srv.submit(reqs[nxt])
            sub[rid] = (nxt, t0 + arrivals[nxt])
            nxt += 1
        running, queued, finished = srv.pump(16)
        if drained is None and nxt == n and queued == 0:
            drained = (time.perf_counter() - t0, srv.work() - w0)
        while True:
            d = srv.next_done()
            if d is None:
                break
            rid, info = d
            k, ta = sub[rid]
            lat[k] = time.perf_counter() - ta
            c = reqs[k]
            work += c.nx * c.ny * c.nclass * info["steps"]
            srv.release(rid)
        if running == 0 and queued == 0 and nxt < n:
            time.sleep(max(0.0, min(0.01, t0 + arrivals[nxt] - time.perf_counter())))
    return np.array([lat[k] for k in range(n)]), work, time.perf_counter() - t0, drained


def bench_serving(hip, nreq: int, load: float, cap_requests: int, mem_gib: float, slots: int, seed: int = 5):
    """Solve requests of the configured mix (square grids 128^2-2048^2, 3-10 classes,
    durations U[600, 3600] s of simulated time) through the continuous-batching server
    (FP32 state, FP64 budgets).  Capacity first: cap_requests requests of the same mix
    (more than the server's slots) submitted at once; the server runs full until its
    queue empties, and the capacity is the work it stepped over that window (its work
    counter, unfinished requests included) per second.  The makespan rate of the whole
    batch is reported beside it: its drain, when fewer requests are left than slots,
    lasts as long as the slowest request and says more about that one request than
    about the server.  Then an open loop: nreq requests arriving as a Poisson stream at load x capacity,
    in requests per second of the batch's mean work per request; latency = completion -
    arrival (wall clock)."""
    import numpy as np

    from swehr_amd import _abi
    from swehr_amd import solver as sv
    from swehr_amd.inputs import config_c5_request

    rng = np.random.default_rng(seed)
    cap_reqs = [config_c5_request(rng, 10000 + k) for k in range(cap_requests)]
    reqs = [config_c5_request(rng, k) for k in range(nreq)]
    if cap_requests <= slots:
        raise ValueError("the capacity batch must exceed the server's slots (%d)" % slots)
    with sv.Server(mem_bytes=int(mem_gib * 2 ** 30), max_slots=slots, precision=_abi.FP32) as srv:
        _, cap_work, cap_wall, (sat_s, sat_work) = serve(srv, cap_reqs, np.zeros(cap_requests))
        capacity = sat_work / sat_s
        per_req = cap_work / cap_requests
        rate = load * capacity / per_req
        log("C5 capacity: %.3e cell-class-steps/s over the first %.2f s (server full); the batch of %d "
            "requests took %.2f s; open loop at %.3f requests/s" % (capacity, sat_s, cap_requests, cap_wall, rate))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
        L, work, wall, _ = serve(srv, reqs, arrivals)
    sizes = sorted(set(c.nx for c in reqs))
    return {"workload": "%d requests, Poisson at %.3f/s = %.2f x measured capacity, sizes %s (weights "
                        ".4/.3/.15/.1/.05), S 3-10, durations U[600,3600] s (unscaled)"
                        % (nreq, rate, load, "/".join("%d^2" % n for n in sizes)),
            "capacity": {"requests": cap_requests, "saturated_s": sat_s, "cell_class_steps_per_s": capacity,
                         "batch_wall_s": cap_wall, "batch_cell_class_steps_per_s": cap_work / cap_wall,
                         "mean_work_per_request": per_req, "requests_per_s": capacity / per_req},
            "arrival_rate_per_s": rate, "load": load,
            "throughput_cell_class_steps_per_s": work / wall, "wall_s": wall,
            "mean_duration_s": float(np.mean([c.tend for c in reqs])),
            "latency_p50_s": float(np.percentile(L, 50)), "latency_p99_s": float(np.percentile(L, 99)),
            "latency_mean_s": float(L.mean()), "latency_max_s": float(L.max()),
            "mem_budget_gib": mem_gib, "slots": slots}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)
    ap.add_argument("--warmup", type=int, default=2400,
                    help="untimed steps before the timed window (a minimum: the window always starts at "
                         "--t-start simulated seconds or later)")
    ap.add_argument("--t-start", type=float, default=600.0,
                    help="simulated time at which the timed window starts at the earliest (runoff established)")
    ap.add_argument("--grid", type=int, default=0, help="grid side (default 2048, or 4096 with --case c3)")
    ap.add_argument("--case", choices=["c2", "c3"], default="c2",
                    help="headline workload: the 2048^2, S=7 fractal plane (c2, the metric's configuration) or "
                         "the 4096^2, S=10 channel network of the precision form (c3)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref-binary", action="store_true",
                    help="skip timing the compiled reference (oracle/_ref) beside the OpenMP port")
    ap.add_argument("--steps-per-graph", type=int, default=16)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--forms", default="precision,batched,serving",
                    help="extra measurements after the headline: comma list of precision, batched, serving, or none")
    ap.add_argument("--c3-grid", type=int, default=4096)
    ap.add_argument("--c3-check", type=float, default=900.0, help="simulated seconds for the FP32-vs-FP64 check")
    ap.add_argument("--c3-window", type=int, default=100)
    ap.add_argument("--c4-count", type=int, default=256)
    ap.add_argument("--c4-tend", type=float, default=1800.0, help="simulated seconds each C4 member runs")
    ap.add_argument("--c4-plain-every", type=int, default=8,
                    help="time the plain FP64 runs of every n-th C4 member")
    ap.add_argument("--c4-persistent", action="store_true",
                    help="also time the C4 FP32 ensemble in the persistent mode (one cooperative launch per 16 steps)")
    ap.add_argument("--c4-floor-max", type=int, default=32,
                    help="most C4 members whose FP32-resolution floor is measured (those with FP32 error > 1e-4)")
    ap.add_argument("--c4-scale-max", type=float, default=16.0)
    ap.add_argument("--c5-requests", type=int, default=40)
    ap.add_argument("--c5-load", type=float, default=0.7, help="C5 arrival rate as a fraction of measured capacity")
    ap.add_argument("--c5-capacity-requests", type=int, default=48,
                    help="requests submitted at once to measure the server's capacity (more than --c5-slots: "
                         "the server runs full until its queue empties)")
    ap.add_argument("--c5-mem-gib", type=float, default=64.0)
    ap.add_argument("--c5-slots", type=int, default=32)
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
              "of the step is a design note of this repository)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from swehr_amd import _abi
    from swehr_amd import solver as sv
    from swehr_amd.inputs import config_c2, config_c3

    prec = _abi.FP64 if args.precision == "fp64" else _abi.FP32
    w = 8 if prec == _abi.FP64 else 4
    t0 = time.perf_counter()
    if not args.grid:
        args.grid = 4096 if args.case == "c3" else 2048
    case = config_c2(n=args.grid) if args.case == "c2" else config_c3(n=args.grid)
    log("case %s: %d x %d, S=%d built in %.1f s" % (case.name, case.nx, case.ny, case.nclass, time.perf_counter() - t0))
    hip = Hip()
    S = case.nclass
    cells = case.nx * case.ny

    s = sv.Solver(case, precision=prec, steps_per_graph=args.steps_per_graph, tile=args.tile)
    # same window as the CPU baseline: the first steps from the initial state
    n_first, ms_first, _, _ = timed_run(hip, s, 20)
    # --warmup is a minimum: the timed window starts once runoff is established
    # (t >= --t-start simulated seconds), whatever the step count asked for
    s.run(max(0, args.warmup - n_first))
    advance_to(s, args.t_start, case.side("maxstep"))
    t_start = s.info()["t"]
    log("warmup done: t=%.1f s after %d steps" % (t_start, s.info()["steps"]))
    tile_ms = None
    if not args.tile:
        best, tile_ms = s.autotune(8)
        log("autotuned tile %d: %s ms per step kernel" % (best, [round(float(x), 4) for x in tile_ms]))
    # the step kernel's duration over exactly the steps of the timed window: the window's
    # steps are launched one by one with events around each kernel, the state is put
    # back from a device copy, and the same steps are timed as graph launches (captured
    # in graphs whose length divides the window, so no graph runs past its end)
    kms, taken = s.profile_window(args.steps)
    s.configure(steps_per_graph=graph_length(args.steps, args.steps_per_graph))
    n, ms, wall, info = timed_run(hip, s, args.steps)
    value = cells * S * n / (ms * 1e-3)
    log("timed %d steps: %.3f ms (wall %.3f s) -> %.4e cell-class-steps/s, t=%.1f..%.1f"
        % (n, ms, wall, value, t_start, info["t"]))
    bud = s.budget()
    s.close()

    bcell = bytes_per_cell(S, w)
    step_ms = float(kms[1])
    achieved = cells * bcell / (step_ms * 1e-3) / 1e9
    key = "%s_%d_%s" % (args.case, args.grid, args.precision)
    traffic = traffic_for(key)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "step_kernel", "kernel_ms": round(step_ms, 4), "bytes_per_cell": bcell,
                "cells_per_launch": cells}

    roofline["hbm_copy_measured_GBs"] = round(hip.copy_bandwidth(), 1)
    log("device copy: %.1f GB/s" % roofline["hbm_copy_measured_GBs"])
    # practical ceiling: the library's streaming kernel over the step's own planes (state
    # 6 + 2S read and written, z read and written, the Ks raster read; z and Ks are FP64)
    planes = 6 + 2 * S
    if w == 8:
        pin, pout = planes + 2, planes + 1
    else:
        pin, pout = planes + 4, planes + 2
    stream_gbs = sv.hbm_stream(case.nx, case.ny, w, pin, pout)
    roofline["hbm_stream_GBs"] = round(stream_gbs, 1)
    roofline["frac_of_stream"] = round(achieved / stream_gbs, 4)
    log("streaming kernel over the step's %d + %d planes: %.1f GB/s (step kernel at %.3f of it)"
        % (pin, pout, stream_gbs, achieved / stream_gbs))

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ncpu, el = cpu_baseline(case, args.cpu_seconds, threads)
        n1, el1 = cpu_baseline(case, args.cpu_seconds / 3, 1)
        cpu = {"value": cells * S * ncpu / el, "unit": "cell-class-steps/s", "cores": threads, "kind": "port",
               "sample": "FP64 C/OpenMP oracle, %d steps of the %dx%d case from its initial state (t=0), %.1f s; "
                         "GPU on the same first 20 steps: %.4e cell-class-steps/s"
                         % (ncpu, args.grid, args.grid, el, cells * S * n_first / (ms_first * 1e-3)),
               "one_thread": {"value": cells * S * n1 / el1, "steps": n1, "seconds": round(el1, 2)},
               "host": host_description(),
               "parity_c1": parity_c1(threads)}
        if not args.no_ref_binary:
            try:
                rb = reference_binary_baseline(case, args.cpu_seconds / 2, el1 / n1)
            except (OSError, RuntimeError, ValueError, subprocess.SubprocessError) as ex:  # a baseline, not the product
                rb = None
                cpu["reference_binary"] = {"error": str(ex)[:300]}
            if rb is not None:
                nr, elr, t_one = rb
                cpu["reference_binary"] = {
                    "value": cells * S * nr / elr, "unit": "cell-class-steps/s", "cores": 1, "kind": "reference",
                    "sample": "SWEHR.c compiled from source (oracle/_ref/swehr_ref_hooked, gcc -O2), %d steps of the "
                              "same case from t=0: wall time of a %d-step run minus a 1-step run (%.1f s, which "
                              "holds parsing, set-up and the final raster writes), %.1f s" % (nr, nr + 1, t_one, elr),
                    "port_one_thread_over_reference": round(cells * S * n1 / el1 / (cells * S * nr / elr), 3)}
        log("cpu baseline: %s" % cpu)

    forms = {}
    wanted = [f for f in args.forms.split(",") if f and f != "none"]
    if "precision" in wanted:
        forms["precision_c3"] = bench_precision(hip, args.c3_grid, args.c3_check, args.c3_window)
        log("precision: %s" % json.dumps(forms["precision_c3"]))
    if "batched" in wanted:
        forms["batched_c4"] = bench_batched(hip, args.c4_count, args.c4_scale_max, args.c4_tend, args.c4_plain_every,
                                            args.c4_floor_max, args.c4_persistent)
        log("batched: %s" % json.dumps(forms["batched_c4"]))
    if "serving" in wanted:
        forms["serving_c5"] = bench_serving(hip, args.c5_requests, args.c5_load, args.c5_capacity_requests,
                                            args.c5_mem_gib, args.c5_slots)
        log("serving: %s" % json.dumps(forms["serving_c5"]))

    wres = sv.water_residual(bud)
    wscale = max(abs(bud[k]) for k in (0, 1, 2, 3, 9, 10))
    out = {
        "metric": "cell-class-steps/s (cells x size classes x steps per second), plain %s step"
                  % args.precision.upper(),
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
        "config": {"workload": ("plain %s step, %dx%d cells, dx 1 m, S=%d, 80 mm/h rain, per-cell Ks"
                                if args.case == "c2" else
                                "plain %s step, %dx%d cells, dx 1 m, channel network, S=%d, triangular hyetograph, "
                                "per-cell Ks") % (args.precision.upper(), args.grid, args.grid, S),
                   "grid": [case.nx, case.ny], "classes": S, "precision": args.precision,
                   "t_window_s": [round(t_start, 3), round(info["t"], 3)], "tile": args.tile or "autotuned",
                   "autotune_step_kernel_ms": None if tile_ms is None else [round(float(x), 4) for x in tile_ms]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms": {"wave_speed": round(float(kms[0]), 4), "step": round(step_ms, 4),
                       "control": round(float(kms[2]), 4), "profiled_steps": taken},
        "water_budget_rel_residual": abs(wres) / wscale if wscale else 0.0,
        "forms": forms,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
