"""Inputs in the reference's file contract, and seeded synthetic cases.

The reference reads five whitespace-separated text files from its working directory
(SWEHR.c:468-522, 218-227):

* ``input.txt``  14 numbers: nx ny dx F manning0 tend a ad h0 mtstar0 rnum rint S J
* ``rain.txt``   rnum rain rates (m/s), entry k used for t in ((k-1)*rint, k*rint]
  shifted 0.01*rint earlier (SWEHR.c:813)
* ``particlesizein.txt`` / ``particlepercentin.txt``  S diameters (m) / fractions
* ``topoin.txt`` nx*ny bed elevations, i (across) outer, j (downslope) inner

Values the reference hard-codes go into the optional sidecar ``swehr_ext.txt``
(``key = value``); a ``Case`` whose sidecar equals the reference's values writes no
sidecar at all.  Numbers are written with 17 significant digits so the files carry
the exact doubles.

Grid orientation: ``topo[i, j]`` with i = 0..nx-1 across the slope and j = 0..ny-1
downslope; j = 0 is the outlet row and elevation rises with j.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

#: values SWEHR.c hard-codes (SWEHR.c:236-243, 617-633, 1010-1035)
REFERENCE_SIDECAR = {
    "cstable": 0.05,
    "maxstep": 0.25,
    "bndry": (1, 0, 1, 1),
    "inflow": 1,
    "roughness": 0.1,
    "theta0": 0.05,
    "thetas": 0.4,
    "hf": 0.001,
    "ks": 0.0,
    "vinf0": 1e-4,
}

MM_PER_H = 1.0 / 1000.0 / 3600.0  # m/s in one mm/h


@dataclasses.dataclass
class Case:
    """One run: the 14 input.txt fields, rasters, tables and sidecar values."""

    nx: int
    ny: int
    dx: float
    tend: float
    topo: np.ndarray  # (nx, ny) float64
    rain: np.ndarray  # (rnum,) m/s
    ds: np.ndarray  # (S,) m
    pfrac: np.ndarray  # (S,)
    rint: float = 60.0
    F: float = 0.004
    manning0: float = 1.0
    a: float = 0.0
    ad: float = 0.0
    h0: float = 0.00099
    mtstar0: float = 3.0
    J: float = 30.0
    ks_raster: Optional[np.ndarray] = None  # (nx, ny) m/s, or None for sidecar["ks"]
    # per-cell suction head (m), theta_s - theta_0, Manning n0, or None for the scalars
    hf_raster: Optional[np.ndarray] = None
    dtheta_raster: Optional[np.ndarray] = None
    roughness_raster: Optional[np.ndarray] = None
    sidecar: dict = dataclasses.field(default_factory=dict)
    movies: int = 1
    name: str = "case"

    @property
    def nclass(self) -> int:
        return int(len(self.ds))

    @property
    def rnum(self) -> int:
        return int(len(self.rain))

    def side(self, key):
        return self.sidecar.get(key, REFERENCE_SIDECAR[key])

    def params(self) -> dict:
        """All run parameters as a flat dict (swehr_params field names)."""
        d = dict(nx=self.nx, ny=self.ny, dx=self.dx, F=self.F, manning0=self.manning0, tend=self.tend,
                 a=self.a, ad=self.ad, h0=self.h0, mtstar0=self.mtstar0, rnum=self.rnum, rint=self.rint,
                 nclass=self.nclass, J=self.J)
        for k in REFERENCE_SIDECAR:
            d[k] = self.side(k)
        return d

    def cells(self) -> int:
        return self.nx * self.ny


def _fmt(v) -> str:
    return repr(float(v)) if not float(v).is_integer() or abs(v) >= 1e16 else "%d" % int(v)


def _fmt17(v) -> str:
    return "%.17g" % float(v)


def write_case(case: Case, d: str) -> None:
    """Write the five input files (and the sidecar when it differs from the reference)."""
    os.makedirs(d, exist_ok=True)
    f = [case.nx, case.ny, case.dx, case.F, case.manning0, case.tend, case.a, case.ad, case.h0,
         case.mtstar0, case.rnum, case.rint, case.nclass, case.J]
    with open(os.path.join(d, "input.txt"), "w") as fh:
        fh.write("\t".join(_fmt17(v) for v in f) + "\n")
    with open(os.path.join(d, "rain.txt"), "w") as fh:
        fh.write("\t".join(_fmt17(v) for v in case.rain) + "\n")
    with open(os.path.join(d, "particlesizein.txt"), "w") as fh:
        fh.write("\t".join(_fmt17(v) for v in case.ds) + "\n")
    with open(os.path.join(d, "particlepercentin.txt"), "w") as fh:
        fh.write("\t".join(_fmt17(v) for v in case.pfrac) + "\n")
    topo = np.asarray(case.topo, dtype=np.float64)
    assert topo.shape == (case.nx, case.ny)
    with open(os.path.join(d, "topoin.txt"), "w") as fh:
        fh.write("\t".join(_fmt17(v) for v in topo.reshape(-1)) + "\n")
    lines = []
    for k, ref in REFERENCE_SIDECAR.items():
        v = case.side(k)
        if k == "bndry":
            if tuple(v) != tuple(ref):
                lines.append("bndry = %d %d %d %d" % tuple(v))
        elif float(v) != float(ref):
            lines.append("%s = %s" % (k, _fmt17(v) if k not in ("inflow",) else "%d" % int(v)))
    for attr, key, fn in (("ks_raster", "ks_file", "ks.txt"), ("hf_raster", "hf_file", "hf.txt"),
                          ("dtheta_raster", "dtheta_file", "dtheta.txt"),
                          ("roughness_raster", "roughness_file", "roughness.txt")):
        r = getattr(case, attr)
        if r is not None:
            r = np.asarray(r, dtype=np.float64)
            assert r.shape == (case.nx, case.ny)
            with open(os.path.join(d, fn), "w") as fh:
                fh.write("\t".join(_fmt17(v) for v in r.reshape(-1)) + "\n")
            lines.append("%s = %s" % (key, fn))
    if not case.movies:
        lines.append("movies = 0")
    path = os.path.join(d, "swehr_ext.txt")
    if lines:
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
    elif os.path.exists(path):
        os.remove(path)


def read_case(d: str, name: str = "case") -> Case:
    """Read a directory in the reference's format (plus an optional sidecar)."""
    def nums(fn):
        p = os.path.join(d, fn)
        if not os.path.exists(p):
            return np.zeros(0)
        with open(p) as fh:
            return np.array([float(x) for x in fh.read().split()], dtype=np.float64)

    v = nums("input.txt")
    v = np.concatenate([v, np.zeros(max(0, 14 - len(v)))])
    nx, ny, S, rnum = int(v[0]), int(v[1]), int(v[12]), int(v[10])

    def padded(a, n):
        a = a[:n]
        return np.concatenate([a, np.zeros(n - len(a))])

    case = Case(nx=nx, ny=ny, dx=v[2], F=v[3], manning0=v[4], tend=v[5], a=v[6], ad=v[7], h0=v[8],
                mtstar0=v[9], rint=v[11], J=v[13], rain=padded(nums("rain.txt"), rnum),
                ds=padded(nums("particlesizein.txt"), S), pfrac=padded(nums("particlepercentin.txt"), S),
                topo=padded(nums("topoin.txt"), nx * ny).reshape(nx, ny), name=name)
    side = os.path.join(d, "swehr_ext.txt")
    if os.path.exists(side):
        with open(side) as fh:
            for line in fh:
                line = line.split("#")[0].strip()
                if not line:
                    continue
                k, val = [s.strip() for s in line.split("=", 1)]
                if k == "bndry":
                    case.sidecar[k] = tuple(int(x) for x in val.replace(",", " ").split())
                elif k in ("ks_file", "hf_file", "dtheta_file", "roughness_file"):
                    setattr(case, k.replace("_file", "_raster"), padded(nums(val), nx * ny).reshape(nx, ny))
                elif k == "movies":
                    case.movies = int(val)
                elif k == "inflow":
                    case.sidecar[k] = int(val)
                elif k == "precision":
                    pass
                else:
                    case.sidecar[k] = float(val)
    return case


# ----------------------------------------------------------------- topography

def diamond_square(n_pow: int, H: float, rng: np.random.Generator) -> np.ndarray:
    """Fractal surface on a (2^n+1)^2 lattice by the diamond-square method.

    Midpoint displacement with amplitude halving by 2^-H per level (Fournier, Fussell
    & Carpenter 1982).  Returns an array with zero mean and unit standard deviation.
    """
    n = 2 ** n_pow + 1
    a = np.zeros((n, n), dtype=np.float64)
    a[0, 0], a[0, -1], a[-1, 0], a[-1, -1] = rng.standard_normal(4)
    step, amp = n - 1, 1.0
    while step > 1:
        half = step // 2
        # diamond step: centres of squares
        c = (a[0:-1:step, 0:-1:step] + a[step::step, 0:-1:step] + a[0:-1:step, step::step] + a[step::step, step::step])
        a[half::step, half::step] = c / 4 + amp * rng.standard_normal(c.shape)
        # square step: edge midpoints, averaging the (up to four) neighbours
        for (oi, oj) in ((half, 0), (0, half)):
            ii = np.arange(oi, n, step)
            jj = np.arange(oj, n, step)
            I, Jm = np.meshgrid(ii, jj, indexing="ij")
            s = np.zeros(I.shape)
            cnt = np.zeros(I.shape)
            for di, dj in ((-half, 0), (half, 0), (0, -half), (0, half)):
                ni, nj = I + di, Jm + dj
                ok = (ni >= 0) & (ni < n) & (nj >= 0) & (nj < n)
                s[ok] += a[ni[ok], nj[ok]]
                cnt[ok] += 1
            a[I, Jm] = s / cnt + amp * rng.standard_normal(I.shape)
        step = half
        amp *= 0.5 ** H
    a -= a.mean()
    sd = a.std()
    return a / sd if sd > 0 else a


def fractal_topography(nx: int, ny: int, dx: float, relief: float, tilt: float, H: float,
                       rng: np.random.Generator) -> np.ndarray:
    """Diamond-square surface cropped to (nx, ny), scaled to ``relief`` metres of
    peak-to-peak range, on a plane rising downslope with gradient ``tilt``."""
    n_pow = max(1, int(math.ceil(math.log2(max(nx, ny, 2)))))
    f = diamond_square(n_pow, H, rng)[:nx, :ny]
    f = (f - f.min()) / max(f.max() - f.min(), 1e-30) * relief
    j = np.arange(ny, dtype=np.float64)
    return f + tilt * dx * j[None, :]


def carve_channels(z: np.ndarray, dx: float, rng: np.random.Generator, n_channels: int = 0, depth: float = 1.5,
                   width: float = 0.0) -> np.ndarray:
    """Incise a seeded network of downslope channels: n_channels meandering centre
    lines i_c(j) (random phase, amplitude, wavelength and a random-walk drift), joined
    in pairs towards the outlet row, each carving a Gaussian valley of the given
    depth (m) and width (cells) that deepens downslope."""
    nx, ny = z.shape
    if n_channels <= 0:
        n_channels = max(2, nx // 128)
    if width <= 0:
        width = max(2.0, nx / (6.0 * n_channels))
    j = np.arange(ny, dtype=np.float64)
    i = np.arange(nx, dtype=np.float64)[:, None]
    out = z.copy()
    centres = (np.arange(n_channels) + 0.5) * nx / n_channels
    for k in range(n_channels):
        amp = rng.uniform(0.02, 0.08) * nx / n_channels * 4
        lam = rng.uniform(0.2, 0.6) * ny
        phase = rng.uniform(0, 2 * np.pi)
        walk = np.cumsum(rng.normal(0.0, 0.3, ny))
        # pairs of channels converge towards the outlet (j = 0)
        target = centres[(k // 2) * 2] if k % 2 else centres[k]
        blend = np.clip(1 - j / (0.5 * ny), 0, 1)
        ic = (1 - blend) * centres[k] + blend * 0.5 * (centres[k] + target)
        ic = ic + amp * np.sin(2 * np.pi * j / lam + phase) + walk
        dep = depth * (0.4 + 0.6 * (1 - j / ny))
        out -= dep[None, :] * np.exp(-((i - ic[None, :]) / width) ** 2)
    return out


def planar_rough(nx: int, ny: int, dx: float, slope: float, sigma: float, rng: np.random.Generator) -> np.ndarray:
    """Plane rising downslope with gradient ``slope`` plus iid Gaussian roughness."""
    j = np.arange(ny, dtype=np.float64)
    return slope * dx * j[None, :] + sigma * rng.standard_normal((nx, ny))


def log_spaced_classes(n: int, dmin: float = 2e-6, dmax: float = 2e-3) -> np.ndarray:
    if n == 1:
        return np.array([dmin])
    return dmin * (dmax / dmin) ** (np.arange(n) / (n - 1))


# ----------------------------------------------------------------- named cases

def config_c1(seed: int = 1) -> Case:
    """100 x 40 cells (downslope x across), dx 0.5 m, 10 % plane with 5 mm roughness;
    50 mm/h for 600 s then none; Ks 5 mm/h, suction head 0.1 m, dtheta 0.3; three
    classes 0.02/0.1/0.5 mm in fractions 0.3/0.4/0.3; Manning n 0.03; free outflow at
    j=1 (inflow off); 1200 s at courant 0.5."""
    rng = np.random.default_rng(seed)
    nx, ny, dx = 40, 100, 0.5
    topo = planar_rough(nx, ny, dx, 0.10, 0.005, rng)
    topo -= topo.min()
    rain = np.array([50 * MM_PER_H] * 10 + [0.0] * 10)
    return Case(nx=nx, ny=ny, dx=dx, tend=1200.0, topo=topo, rain=rain, rint=60.0,
                ds=np.array([2e-5, 1e-4, 5e-4]), pfrac=np.array([0.3, 0.4, 0.3]),
                a=2000.0, ad=4000.0, F=0.004, J=30.0, h0=0.00099, mtstar0=3.0,
                sidecar=dict(cstable=0.5, inflow=0, ks=5 * MM_PER_H, hf=0.1, theta0=0.1, thetas=0.4,
                             roughness=0.03), movies=0, name="c1_plane_100x40")


def config_c2(seed: int = 2, n: int = 2048, tend: float = 3600.0) -> Case:
    """n x n cells, dx 1 m, diamond-square surface (H 0.8, 20 m relief) on a 5 % tilt;
    7 classes 0.002-2 mm log-spaced in equal fractions; 80 mm/h for 1800 s then none;
    per-cell Ks log-normal (median 10 mm/h, sigma_ln 0.5); 3600 s."""
    rng = np.random.default_rng(seed)
    topo = fractal_topography(n, n, 1.0, 20.0, 0.05, 0.8, rng)
    rain = np.array([80 * MM_PER_H] * 30 + [0.0] * 30)
    ks = 10 * MM_PER_H * np.exp(0.5 * rng.standard_normal((n, n)))
    S = 7
    return Case(nx=n, ny=n, dx=1.0, tend=tend, topo=topo, rain=rain, rint=60.0,
                ds=log_spaced_classes(S), pfrac=np.full(S, 1.0 / S), a=2000.0, ad=4000.0,
                ks_raster=ks, sidecar=dict(cstable=0.5, inflow=0, hf=0.1, theta0=0.1, thetas=0.4,
                                           roughness=0.03), movies=0, name="c2_fractal_%d" % n)


def config_c3(seed: int = 3, n: int = 4096, tend: float = 3600.0, carve: bool = True) -> Case:
    """n x n cells, dx 1 m, fractal surface with a carved channel network; 10 classes
    0.002-2 mm log-spaced with Dirichlet(2) fractions; triangular hyetograph peaking at
    120 mm/h at 900 s and ending at 1800 s (10 s table); Ks as config_c2; 3600 s."""
    rng = np.random.default_rng(seed)
    topo = fractal_topography(n, n, 1.0, 20.0, 0.05, 0.8, rng)
    if carve:
        topo = carve_channels(topo, 1.0, rng)
    rint = 10.0
    tk = rint * (np.arange(int(tend / rint)) + 0.5)
    peak = 120 * MM_PER_H
    rain = np.where(tk < 900, peak * tk / 900, np.where(tk < 1800, peak * (1800 - tk) / 900, 0.0))
    ks = 10 * MM_PER_H * np.exp(0.5 * rng.standard_normal((n, n)))
    S = 10
    pf = rng.dirichlet(np.full(S, 2.0))
    return Case(nx=n, ny=n, dx=1.0, tend=tend, topo=topo, rain=rain, rint=rint, ds=log_spaced_classes(S),
                pfrac=pf, a=2000.0, ad=4000.0, ks_raster=ks,
                sidecar=dict(cstable=0.5, inflow=0, hf=0.1, theta0=0.1, thetas=0.4, roughness=0.03),
                movies=0, name="c3_channels_%d" % n)


def config_c4_member(rng: np.random.Generator, k: int, tend: float = 1800.0, scale_max: float = 16.0) -> Case:
    """One ensemble hillslope: (64 s) x (32 s) cells with s log-uniform in [1, scale_max],
    dx U[0.25, 2] m, slope U[2 %, 30 %], roughness sigma U[2, 20] mm, rain U[10, 150] mm/h
    constant, Ks log-normal (median 8 mm/h, sigma_ln 0.5), dtheta U[0.1, 0.4], 5 classes."""
    s = math.exp(rng.uniform(0.0, math.log(scale_max)))
    ny, nx = max(8, int(round(64 * s))), max(4, int(round(32 * s)))
    dx = rng.uniform(0.25, 2.0)
    slope = rng.uniform(0.02, 0.30)
    sigma = rng.uniform(0.002, 0.020)
    topo = planar_rough(nx, ny, dx, slope, sigma, rng)
    topo -= topo.min()
    R = rng.uniform(10, 150) * MM_PER_H
    ks = 8 * MM_PER_H * np.exp(0.5 * rng.standard_normal((nx, ny)))
    dth = rng.uniform(0.1, 0.4)
    S = 5
    return Case(nx=nx, ny=ny, dx=dx, tend=tend, topo=topo, rain=np.array([R]), rint=tend + 1.0,
                ds=log_spaced_classes(S), pfrac=np.full(S, 1.0 / S), a=2000.0, ad=4000.0, ks_raster=ks,
                sidecar=dict(cstable=0.5, inflow=0, hf=0.1, theta0=0.4 - dth, thetas=0.4, roughness=0.03),
                movies=0, name="c4_member_%03d" % k)


def config_c4(seed: int = 4, count: int = 256, tend: float = 1800.0, scale_max: float = 16.0):
    rng = np.random.default_rng(seed)
    return [config_c4_member(rng, k, tend, scale_max) for k in range(count)]


def config_c5_request(rng: np.random.Generator, k: int, sizes=(128, 256, 512, 1024, 2048),
                      weights=(0.4, 0.3, 0.15, 0.1, 0.05), tmin: float = 600.0, tmax: float = 3600.0) -> Case:
    """One serving request: square grid drawn from ``sizes`` with ``weights``, duration
    U[tmin, tmax] s, 3-10 classes with Dirichlet(2) fractions, triangular hyetograph
    (peak U[60, 150] mm/h at a quarter of the duration), Ks as config_c2.  The surface
    has config_c3's statistics: a fractal of H 0.8 on a 5 % tilt whose relief, U[5, 20]
    m over 4096 m, scales with the grid's extent as extent^H, as a window of that
    surface's would (a 128 m grid gets 0.3-1.3 m).  With the full 5-20 m on every grid,
    the small ones had slopes that the explicit scheme does not survive at CFL 0.5:
    speeds of 100-1000 m/s, depths of kilometres and 10^5-10^6 steps, in FP64 as in
    FP32."""
    n = int(rng.choice(np.asarray(sizes), p=np.asarray(weights) / np.sum(weights)))
    tend = float(rng.uniform(tmin, tmax))
    S = int(rng.integers(3, 11))
    relief = (5.0 + 15.0 * rng.random()) * (n / 4096.0) ** 0.8
    topo = fractal_topography(n, n, 1.0, relief, 0.05, 0.8, rng)
    rint = 10.0
    nr = int(math.ceil(tend / rint)) + 1
    tk = rint * (np.arange(nr) + 0.5)
    tp = tend / 4
    peak = rng.uniform(60, 150) * MM_PER_H
    rain = np.where(tk < tp, peak * tk / tp, np.where(tk < 2 * tp, peak * (2 * tp - tk) / tp, 0.0))
    ks = 10 * MM_PER_H * np.exp(0.5 * rng.standard_normal((n, n)))
    return Case(nx=n, ny=n, dx=1.0, tend=tend, topo=topo, rain=rain, rint=rint, ds=log_spaced_classes(S),
                pfrac=rng.dirichlet(np.full(S, 2.0)), a=2000.0, ad=4000.0, ks_raster=ks,
                sidecar=dict(cstable=0.5, inflow=0, hf=0.1, theta0=0.1, thetas=0.4, roughness=0.03),
                movies=0, name="c5_request_%04d" % k)
