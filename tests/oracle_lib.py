"""TEST INFRASTRUCTURE: ctypes access to the CPU oracle (oracle/liboracle.so).

The oracle is the checker: tests compare the HIP path with it, never the other way
round, and the product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from swehr_amd import _abi
from swehr_amd._abi import dptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
#: the same oracle source compiled with FMA contraction: a second FP64 evaluation of the
#: same arithmetic that differs only in the rounding of a*b+c expressions
FMA_LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_fma.so")
RUN_PATH = os.path.join(ROOT, "oracle", "_bin", "swehr_oracle_run")

_libs = {}


def lib(variant: str = "plain"):
    if variant not in _libs:
        path = LIB_PATH if variant == "plain" else FMA_LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (run __graft_entry__.build())" % os.path.relpath(path, ROOT))
        L = C.CDLL(path)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(_abi.Params)] + [C.POINTER(C.c_double)] * 5
        L.oracle_step.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.oracle_run.argtypes = [C.c_void_p, C.c_long]
        L.oracle_info.argtypes = [C.c_void_p, C.POINTER(_abi.StepInfo)]
        L.oracle_download.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.oracle_stage.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                   C.POINTER(C.c_int)]
        L.oracle_set_rasters.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.oracle_discharge.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                       C.POINTER(C.c_int)]
        L.oracle_budget.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.oracle_fail_events.restype = C.c_long
        L.oracle_fail_events.argtypes = [C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        _libs[variant] = L
    return _libs[variant]


class Oracle:
    """One oracle instance for a swehr_amd.inputs.Case (FP64)."""

    def __init__(self, case, threads: int = 1, variant: str = "plain"):
        """variant "plain" (the checker) or "fma" (the rounding floor: FMA contraction)."""
        L = lib(variant)
        self._L = L
        L.oracle_set_threads(int(threads))
        self.case = case
        self._p = _abi.params_from_case(case)
        topo, ks, ds, pf, rain = _abi.case_arrays(case)
        self._keep = (topo, ks, ds, pf, rain)
        h = L.oracle_create(C.byref(self._p), dptr(topo), dptr(ks) if ks is not None else None, dptr(ds),
                            dptr(pf), dptr(rain))
        if not h:
            raise ValueError("oracle_create rejected the case")
        self._h = h
        from swehr_amd.solver import parameter_rasters

        pr = parameter_rasters(case)
        if any(r is not None for r in pr):
            self._keep = self._keep + tuple(pr)
            L.oracle_set_rasters(h, *[dptr(r) if r is not None else None for r in pr])

    def step(self) -> int:
        pd = C.c_int(0)
        return self._L.oracle_step(self._h, C.byref(pd))

    def run(self, n: int = 1 << 62) -> int:
        return self._L.oracle_run(self._h, int(n))

    def info(self) -> dict:
        i = _abi.StepInfo()
        self._L.oracle_info(self._h, C.byref(i))
        return i.as_dict()

    def field(self, name: str, k: int = 0) -> np.ndarray:
        out = np.empty(self.case.nx * self.case.ny, dtype=np.float64)
        rc = self._L.oracle_download(self._h, _abi.FIELDS[name], int(k), dptr(out))
        if rc != 0:
            raise ValueError("oracle_download(%s, %d) -> %d" % (name, k, rc))
        return out.reshape(self.case.nx, self.case.ny)

    def stage(self):
        cap = _abi.STAGE_MAX
        st = np.zeros(_abi.STAGE_COLS * cap)
        fl = np.zeros(_abi.FLOW_COLS * cap)
        cnt = C.c_int(0)
        self._L.oracle_stage(self._h, dptr(st), dptr(fl), cap, C.byref(cnt))
        n = cnt.value
        return st.reshape(_abi.STAGE_COLS, cap)[:, :n].copy(), fl.reshape(_abi.FLOW_COLS, cap)[:, :n].copy()

    def discharge(self):
        """(q, qs): outlet water (m^3/s) and sediment (kg/s) at each stage record."""
        cap = _abi.STAGE_MAX
        q, qs = np.zeros(cap), np.zeros(cap)
        cnt = C.c_int(0)
        self._L.oracle_discharge(self._h, dptr(q), dptr(qs), cap, C.byref(cnt))
        return q[:cnt.value].copy(), qs[:cnt.value].copy()

    def budget(self) -> np.ndarray:
        out = np.zeros(_abi.BUDGET_N)
        self._L.oracle_budget(self._h, dptr(out))
        return out

    def fail_events(self) -> int:
        return int(self._L.oracle_fail_events(self._h))

    def state(self) -> dict:
        """All persistent fields as (nx, ny) arrays; per-class fields as (S, nx, ny)."""
        S = self.case.nclass
        d = {k: self.field(k) for k in ("depth", "uh", "vh", "topo", "vinf", "totm", "dep", "totsed",
                                        "maxvel", "maxdepth", "velx", "vely")}
        d["ch"] = np.stack([self.field("ch", k + 1) for k in range(S)])
        d["m"] = np.stack([self.field("m", k + 1) for k in range(S)])
        return d

    def close(self):
        if self._h:
            self._L.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
