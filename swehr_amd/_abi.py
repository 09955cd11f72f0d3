"""ctypes mirrors of the structs and enums in include/swehr.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

MAX_CLASSES = 16
STAGE_COLS = 24
FLOW_COLS = 18
STAGE_MAX = 30000
BUDGET_N = 12

OK, ERR_CFL, ERR_NONFINITE, ERR_OOM, ERR_BAD_INPUT, ERR_HIP, ERR_STAGE_FULL, ERR_NO_DEVICE, ERR_BUSY = range(9)
STATUS_NAMES = {0: "OK", 1: "CFL", 2: "NONFINITE", 3: "OOM", 4: "BAD_INPUT", 5: "HIP", 6: "STAGE_FULL",
                7: "NO_DEVICE", 8: "BUSY"}

FP64, FP32 = 0, 1

FIELDS = dict(depth=0, uh=1, vh=2, ch=3, m=4, topo=5, vinf=6, totm=7, dep=8, totsed=9, speed=10, maxvel=11,
              maxdepth=12, sed=13, velx=14, vely=15)
PER_CLASS = {"ch", "m", "sed"}


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("F", C.c_double), ("manning0", C.c_double),
        ("tend", C.c_double), ("a", C.c_double), ("ad", C.c_double), ("h0", C.c_double),
        ("mtstar0", C.c_double), ("rnum", C.c_int), ("rint", C.c_double), ("nclass", C.c_int),
        ("J", C.c_double), ("cstable", C.c_double), ("maxstep", C.c_double), ("bndry", C.c_int * 4),
        ("inflow", C.c_int), ("roughness", C.c_double), ("theta0", C.c_double), ("thetas", C.c_double),
        ("hf", C.c_double), ("ks", C.c_double), ("vinf0", C.c_double),
    ]


class StepInfo(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("courant", C.c_double), ("rain", C.c_double),
                ("t_prev", C.c_double), ("steps", C.c_int64), ("done", C.c_int32), ("print_due", C.c_int32),
                ("stage_count", C.c_int32), ("status", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def params_from_case(case) -> Params:
    d = case.params()
    p = Params()
    for k, v in d.items():
        if k == "bndry":
            for n in range(4):
                p.bndry[n] = int(v[n])
        elif k in ("nx", "ny", "rnum", "nclass", "inflow"):
            setattr(p, k, int(v))
        else:
            setattr(p, k, float(v))
    return p


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def case_arrays(case):
    """Contiguous float64 arrays for create(): topo, ks (or None), ds, pfrac, rain."""
    topo = np.ascontiguousarray(case.topo, dtype=np.float64).reshape(-1)
    ks = None if case.ks_raster is None else np.ascontiguousarray(case.ks_raster, dtype=np.float64).reshape(-1)
    ds = np.ascontiguousarray(case.ds, dtype=np.float64)
    pf = np.ascontiguousarray(case.pfrac, dtype=np.float64)
    rain = np.ascontiguousarray(case.rain, dtype=np.float64)
    return topo, ks, ds, pf, rain
