"""TEST INFRASTRUCTURE: comparison helpers for HIP-path vs oracle parity tests."""
from __future__ import annotations

import numpy as np

STATE_FIELDS = ("depth", "uh", "vh", "topo", "vinf", "totm", "dep", "totsed", "maxvel", "maxdepth", "velx", "vely")


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b|| / ||b|| over the finite entries of b (NaN positions must agree)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb, na = np.isnan(b), np.isnan(a)
    if not np.array_equal(na, nb):
        return float("inf")
    m = ~nb
    den = np.linalg.norm(b[m])
    num = np.linalg.norm(a[m] - b[m])
    if den == 0:
        return 0.0 if num == 0 else float("inf")
    return float(num / den)


def compare_states(gpu: dict, ora: dict, nclass: int):
    """Relative L2 error per field (per class for ch and m)."""
    out = {}
    for k in STATE_FIELDS:
        out[k] = rel_l2(gpu[k], ora[k])
    for c in range(nclass):
        out["ch%d" % (c + 1)] = rel_l2(gpu["ch"][c], ora["ch"][c])
        out["m%d" % (c + 1)] = rel_l2(gpu["m"][c], ora["m"][c])
    return out


def worst(errs: dict):
    k = max(errs, key=lambda x: errs[x])
    return k, errs[k]


def outlet_case(nx: int = 20, ny: int = 30, tend: float = 300.0, seed: int = 5):
    """Tilted rough plane, dx 10 m, constant 108 mm/h rain, no infiltration, three
    classes with rainsplash: walls on three sides and free outflow through the j = 1/2
    face (the reference's boundary types, inflow off).  cstable 0.9 and maxstep 1 s give
    dt = 1 s in every step (wave speeds stay below 9 m/s), so every step ends past a
    whole second and takes a stage record: the discharge series then holds every step."""
    from swehr_amd.inputs import Case, planar_rough

    rng = np.random.default_rng(seed)
    dx = 10.0
    topo = planar_rough(nx, ny, dx, 0.10, 0.02, rng)
    topo -= topo.min()
    rain = np.full(int(tend // 60) + 2, 3e-5)
    return Case(nx=nx, ny=ny, dx=dx, tend=tend, topo=topo, rain=rain, rint=60.0,
                ds=np.array([2e-5, 1e-4, 5e-4]), pfrac=np.array([0.3, 0.4, 0.3]), a=2000.0, ad=4000.0,
                sidecar=dict(cstable=0.9, maxstep=1.0, inflow=0, ks=0.0, roughness=0.03), movies=0,
                name="outlet_plane")


def many_classes_case(S: int, nx: int, ny: int, seed: int):
    """A rough plane with S size classes 0.002-2 mm (S up to the maximum of 16), rain
    with rainsplash and entrainment, a grid whose sizes are not multiples of any tile,
    free outflow at j = 1/2 and the sediment-laden inflow at j = ny."""
    from swehr_amd.inputs import Case, log_spaced_classes, planar_rough

    rng = np.random.default_rng(seed)
    topo = planar_rough(nx, ny, 1.0, 0.08, 0.01, rng)
    topo -= topo.min()
    return Case(nx=nx, ny=ny, dx=1.0, tend=600.0, topo=topo, rain=np.array([4e-5] * 6 + [0.0] * 6), rint=60.0,
                ds=log_spaced_classes(S), pfrac=rng.dirichlet(np.full(S, 2.0)), a=2000.0, ad=4000.0, F=0.01,
                sidecar=dict(cstable=0.5, ks=2e-6, hf=0.1, theta0=0.1, thetas=0.4, roughness=0.03), movies=0,
                name="classes%d_%dx%d" % (S, nx, ny))
