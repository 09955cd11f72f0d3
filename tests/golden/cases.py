"""Golden cases: inputs whose reference outputs are committed under tests/golden/<name>/.

Each case names the reference build that produced its outputs:
  "plain"   SWEHR.c compiled unmodified (only the input files can vary),
  "hooked"  the same source with the parameter hooks of oracle/ref_hooks.h, for the
            values the reference hard-codes (Ks, soil, Manning n, courant number,
            boundary types, inflow switch, movie output).
``maxsteps`` > 0 stops both the reference and the oracle after that many steps.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from swehr_amd.inputs import MM_PER_H, Case, config_c1, read_case

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE_DIR = "/root/reference"


@dataclasses.dataclass
class GoldenCase:
    name: str
    build: str  # "plain" | "hooked"
    make: object  # () -> Case
    maxsteps: int = 0
    dump: bool = False
    note: str = ""


def _shipped() -> Case:
    """The sample input set that ships with the reference (3 x 51 flume, 600 s)."""
    return read_case(os.path.join(HERE, "shipped", "inputs"), name="shipped")


def _inflow_mix() -> Case:
    rng = np.random.default_rng(11)
    nx, ny, dx = 4, 52, 0.25
    topo = 0.7 * dx * np.arange(ny)[None, :] + 0.002 * rng.standard_normal((nx, ny))
    rain = np.array([30 * MM_PER_H] * 3 + [0.0] * 3)
    return Case(nx=nx, ny=ny, dx=dx, tend=60.0, topo=topo, rain=rain, rint=10.0, ds=np.array([4e-3, 1e-3, 5e-5]),
                pfrac=np.array([0.4, 0.35, 0.25]), a=500.0, ad=1000.0, F=0.004, J=30.0, name="inflow_mix")


def _narrow() -> Case:
    rng = np.random.default_rng(12)
    ny = 55
    topo = 0.3 * 0.5 * np.arange(ny)[None, :] + 0.003 * rng.standard_normal((2, ny))
    return Case(nx=2, ny=ny, dx=0.5, tend=30.0, topo=topo, rain=np.array([20 * MM_PER_H] * 2), rint=20.0,
                ds=np.array([1e-4, 2.5e-3]), pfrac=np.array([0.5, 0.5]), a=300.0, ad=600.0, name="narrow")


def _rain_plane() -> Case:
    rng = np.random.default_rng(13)
    nx, ny, dx = 10, 60, 0.5
    topo = 0.10 * dx * np.arange(ny)[None, :] + 0.005 * rng.standard_normal((nx, ny))
    topo -= topo.min()
    rain = np.array([50 * MM_PER_H] * 2 + [0.0] * 2)
    return Case(nx=nx, ny=ny, dx=dx, tend=240.0, topo=topo, rain=rain, rint=60.0,
                ds=np.array([2e-5, 1e-4, 5e-4]), pfrac=np.array([0.3, 0.4, 0.3]), a=2000.0, ad=4000.0,
                sidecar=dict(cstable=0.5, inflow=0, ks=5 * MM_PER_H, hf=0.1, theta0=0.1, thetas=0.4,
                             roughness=0.03), movies=0, name="rain_plane")


def _transmissive() -> Case:
    rng = np.random.default_rng(14)
    nx, ny, dx = 8, 55, 1.0
    j = np.arange(ny)[None, :]
    i = np.arange(nx)[:, None]
    topo = 0.08 * dx * j + 0.03 * dx * np.abs(i - nx / 2) + 0.01 * rng.standard_normal((nx, ny))
    rain = np.array([90 * MM_PER_H] * 3 + [10 * MM_PER_H] * 3)
    ks = 6 * MM_PER_H * np.exp(0.5 * rng.standard_normal((nx, ny)))
    return Case(nx=nx, ny=ny, dx=dx, tend=150.0, topo=topo, rain=rain, rint=25.0, ds=np.array([1e-5, 3e-4]),
                pfrac=np.array([0.6, 0.4]), a=1500.0, ad=3000.0, F=0.02, J=25.0, ks_raster=ks,
                sidecar=dict(bndry=(0, 0, 0, 0), inflow=0, cstable=0.3, hf=0.05, theta0=0.15, thetas=0.45,
                             roughness=0.05), name="transmissive")


def _closed_basin() -> Case:
    rng = np.random.default_rng(15)
    nx, ny, dx = 6, 50, 0.5
    j = np.arange(ny)[None, :]
    topo = 0.02 * (j - ny / 3.0) ** 2 * dx + 0.004 * rng.standard_normal((nx, ny))
    rain = np.array([120 * MM_PER_H] * 4)
    return Case(nx=nx, ny=ny, dx=dx, tend=100.0, topo=topo, rain=rain, rint=30.0, ds=np.array([4e-5]),
                pfrac=np.array([1.0]), a=800.0, ad=900.0,
                sidecar=dict(bndry=(1, 1, 1, 1), inflow=0, cstable=0.4, ks=1 * MM_PER_H, roughness=0.04),
                movies=0, name="closed_basin")


def _cfl_abort() -> Case:
    rng = np.random.default_rng(16)
    nx, ny, dx = 3, 51, 0.25
    topo = 0.5 * dx * np.arange(ny)[None, :] + 0.001 * rng.standard_normal((nx, ny))
    return Case(nx=nx, ny=ny, dx=dx, tend=200.0, topo=topo, rain=np.zeros(1), rint=60.0, ds=np.array([2e-4]),
                pfrac=np.array([1.0]), sidecar=dict(cstable=1.5), name="cfl_abort")


def _param_rasters() -> Case:
    """Per-cell Ks, suction head, soil moisture and Manning n0, each varying by +-25%
    around its mean, on a grid with nx != ny (flipping any one raster along i moves the
    final state's most affected field by 7e-2 to 4.5e-1).  theta_0 and theta_s are multiples of 2^-10, so
    theta_s - theta_0 is exact and the case's dtheta raster equals the reference's
    THETAS - THETA0 bit for bit.  Rain exceeds the infiltration capacity early, so no
    cell recedes through the film-depth thresholds: the run's rounding floor is 3e-14."""
    rng = np.random.default_rng(17)
    nx, ny, dx = 7, 52, 0.5
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    topo = 0.12 * dx * j + 0.02 * dx * np.abs(i - 2) + 0.004 * rng.standard_normal((nx, ny))
    rain = np.array([100 * MM_PER_H] * 5)

    def around(mean):
        return mean * (1 + 0.5 * (rng.random((nx, ny)) - 0.5))

    ks = around(5 * MM_PER_H)
    hf = around(0.01)
    theta0 = np.round((0.1 + 0.025 * (rng.random((nx, ny)) - 0.5)) * 1024) / 1024
    thetas = np.round((0.4 + 0.05 * (rng.random((nx, ny)) - 0.5)) * 1024) / 1024
    n0 = around(0.05)
    c = Case(nx=nx, ny=ny, dx=dx, tend=150.0, topo=topo, rain=rain, rint=30.0, ds=np.array([3e-5, 6e-4]),
             pfrac=np.array([0.55, 0.45]), a=1200.0, ad=2500.0, F=0.01, J=25.0, ks_raster=ks, hf_raster=hf,
             dtheta_raster=thetas - theta0, roughness_raster=n0,
             sidecar=dict(cstable=0.4, inflow=0), movies=0, name="param_rasters")
    c.theta_rasters = (theta0, thetas)
    return c


def _c1() -> Case:
    c = config_c1(seed=1)
    c.name = "c1_plane_100x40"
    return c


CASES = [
    GoldenCase("shipped", "plain", _shipped, dump=True, note="the reference's own sample inputs"),
    GoldenCase("inflow_mix", "plain", _inflow_mix, dump=True,
               note="inflow, rain, 3 classes incl. 3 mm, entrainment, slope failures"),
    GoldenCase("inflow_mix_400", "hooked", _inflow_mix, maxsteps=400, dump=True, note="first 400 steps"),
    GoldenCase("narrow", "plain", _narrow, dump=True, note="two cells across (nx = 2, one-sided slopes on both sides; the reference needs nx >= 2)"),
    GoldenCase("rain_plane", "hooked", _rain_plane, dump=True,
               note="rain on a rough plane, Green-Ampt Ks > 0, Manning 0.03, courant 0.5, no inflow"),
    GoldenCase("transmissive", "hooked", _transmissive, dump=True,
               note="all sides transmissive, per-cell Ks raster, rain table"),
    GoldenCase("closed_basin", "hooked", _closed_basin, dump=True, note="all four sides walls"),
    GoldenCase("cfl_abort", "hooked", _cfl_abort, dump=True, note="courant number 1.5 -> CFL stop"),
    GoldenCase("c1_plane_100x40", "hooked", _c1, note="the 100 x 40 plane of the CPU configuration, 1200 s"),
    GoldenCase("param_rasters", "hooked", _param_rasters, dump=True,
               note="per-cell Ks, HF, THETA0, THETAS and Manning n0 rasters on a 7 x 52 grid"),
]

BY_NAME = {c.name: c for c in CASES}


def hook_env(case: Case, d: str, maxsteps: int = 0, dump_path: str = "") -> dict:
    """Environment of the hooked reference build that reproduces ``case``'s sidecar."""
    env = {}
    sc = case.sidecar
    if case.ks_raster is not None:
        env["REFH_KS_FILE"] = os.path.join(d, "ks.txt")
    elif "ks" in sc:
        env["REFH_KS"] = "%.17g" % sc["ks"]
    # the per-cell rasters: the sidecar's hf.txt and roughness.txt are in the reference
    # hooks' raster format; theta_s - theta_0 needs both rasters, written here from the
    # pair the case was built from (their difference must be the case's dtheta exactly)
    if case.hf_raster is not None:
        env["REFH_HF_FILE"] = os.path.join(d, "hf.txt")
    if case.roughness_raster is not None:
        env["REFH_ROUGHNESS_FILE"] = os.path.join(d, "roughness.txt")
    if case.dtheta_raster is not None:
        theta0, thetas = case.theta_rasters
        if not np.array_equal(thetas - theta0, case.dtheta_raster):
            raise ValueError("theta_s - theta_0 differs from the case's dtheta raster")
        for fn, r in (("theta0.txt", theta0), ("thetas.txt", thetas)):
            with open(os.path.join(d, fn), "w") as fh:
                fh.write("\t".join("%.17g" % v for v in np.asarray(r, dtype=np.float64).reshape(-1)) + "\n")
        env["REFH_THETA0_FILE"] = os.path.join(d, "theta0.txt")
        env["REFH_THETAS_FILE"] = os.path.join(d, "thetas.txt")
    for key, var in (("theta0", "REFH_THETA0"), ("thetas", "REFH_THETAS"), ("hf", "REFH_HF"),
                     ("roughness", "REFH_ROUGHNESS"), ("cstable", "REFH_CSTABLE")):
        if var + "_FILE" in env or (key in ("theta0", "thetas") and case.dtheta_raster is not None):
            continue
        if key in sc:
            env[var] = "%.17g" % sc[key]
    if "inflow" in sc:
        env["REFH_INFLOW"] = "%d" % int(sc["inflow"])
    if "bndry" in sc:
        env["REFH_BNDRY"] = "%d,%d,%d,%d" % tuple(sc["bndry"])
    for key in ("vinf0", "maxstep"):
        if key in sc:
            raise ValueError("the hooked reference has no hook for %s" % key)
    if not case.movies:
        env["REFH_MOVIES"] = "0"
    if maxsteps:
        env["REFH_MAXSTEPS"] = str(int(maxsteps))
    if dump_path:
        env["REFH_DUMP"] = dump_path
    return env


OUTPUT_FILES = ["topoout.txt", "depth.txt", "uh.txt", "vh.txt", "ch.txt", "m.txt", "c.txt", "vel.txt",
                "stage.txt", "maxvel.txt", "maxdepth.txt", "saveflow.txt", "topomovie.txt", "depthmovie.txt",
                "velocitymovie.txt", "cmovie.txt", "Mmovie.txt"]


def stdout_lines(text: str):
    """stdout without the wall-clock line, which differs between runs."""
    return [l for l in text.splitlines() if not l.startswith("Elapsed wall clock time")]


def read_dump(path: str) -> dict:
    """Parse the full-precision dump written by the hooked reference (format in
    oracle/ref_hooks.h)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    assert raw[:8] == b"SWEHRDMP"
    nx, ny, nn, cnt = np.frombuffer(raw, dtype="<i4", count=4, offset=8)
    steps = int(np.frombuffer(raw, dtype="<i8", count=1, offset=24)[0])
    t = float(np.frombuffer(raw, dtype="<f8", count=1, offset=32)[0])
    S = int(nn) - 3
    names = ["depth", "uh", "vh"] + ["ch%d" % (k + 1) for k in range(S)] + ["m%d" % (k + 1) for k in range(S)] + \
            ["topo", "vinf", "totm", "dep", "totsed", "maxvel", "maxdepth", "velx", "vely"]
    off = 40
    n = int(nx) * int(ny)
    out = {"steps": steps, "t": t, "stagecnt": int(cnt)}
    for name in names:
        out[name] = np.frombuffer(raw, dtype="<f8", count=n, offset=off).reshape(int(nx), int(ny)).copy()
        off += 8 * n
    out["stage"] = np.frombuffer(raw, dtype="<f8", count=24 * int(cnt), offset=off).reshape(24, int(cnt)).copy()
    off += 8 * 24 * int(cnt)
    out["saveflow"] = np.frombuffer(raw, dtype="<f8", count=18 * int(cnt), offset=off).reshape(18, int(cnt)).copy()
    return out
