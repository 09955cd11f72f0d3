"""Python access to the HIP solver through its C ABI (include/swehr.h).

The library ``swehr_amd/libswehr_hip.so`` is built in-tree by ``__graft_entry__.build()``
(``make -C swehr_amd/csrc``).  There is no fallback: if the library is missing, or no
GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import _abi
from ._abi import dptr

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libswehr_hip.so")
HEADER = os.path.join(ROOT, "include", "swehr.h")
DRIVER = os.path.join(HERE, "bin", "swehr")

_lib = None


class SwehrError(RuntimeError):
    def __init__(self, where: str, code: int, msg: str = ""):
        self.code = code
        super().__init__("%s: %s%s" % (where, _abi.STATUS_NAMES.get(code, str(code)), (": " + msg) if msg else ""))


def header_functions(path: str = HEADER):
    """Names of the functions include/swehr.h declares."""
    with open(path) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s+\**\s*(swehr_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("%s is not built: run __graft_entry__.build() (make -C swehr_amd/csrc)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(C.c_double)
    L.swehr_params_init.argtypes = [C.POINTER(_abi.Params)]
    L.swehr_create.argtypes = [C.POINTER(_abi.Params), P, P, P, P, P, C.c_int, C.POINTER(C.c_void_p)]
    L.swehr_advance.argtypes = [C.c_void_p, C.c_int64, C.POINTER(_abi.StepInfo)]
    L.swehr_ack_print.argtypes = [C.c_void_p]
    L.swehr_download.argtypes = [C.c_void_p, C.c_int, C.c_int, P]
    L.swehr_stage.argtypes = [C.c_void_p, P, P, C.c_int, C.POINTER(C.c_int)]
    L.swehr_discharge.argtypes = [C.c_void_p, P, P, C.c_int, C.POINTER(C.c_int)]
    L.swehr_set_parameter_raster.argtypes = [C.c_void_p, C.c_int, P]
    L.swehr_budget.argtypes = [C.c_void_p, P]
    L.swehr_budget_noalloc.argtypes = [C.c_void_p, P]
    L.swehr_instance_bytes_for.restype = C.c_int64
    L.swehr_instance_bytes_for.argtypes = [C.POINTER(_abi.Params), C.c_int, C.c_int]
    L.swehr_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.swehr_server_work.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.swehr_server_evict.argtypes = [C.c_void_p, C.c_int64]
    L.swehr_server_set_quantum.argtypes = [C.c_void_p, C.c_int64]
    L.swehr_server_evictions.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.swehr_instance_bytes.restype = C.c_int64
    L.swehr_instance_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.swehr_destroy.argtypes = [C.c_void_p]
    L.swehr_strerror.restype = C.c_char_p
    L.swehr_strerror.argtypes = [C.c_int]
    L.swehr_last_error.restype = C.c_char_p
    L.swehr_configure.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.swehr_tile_shape.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.swehr_get_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.swehr_profile.argtypes = [C.c_void_p, C.c_int, P, C.POINTER(C.c_int)]
    L.swehr_profile_window.argtypes = [C.c_void_p, C.c_int, P, C.POINTER(C.c_int)]
    L.swehr_snapshot_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.swehr_snapshot.argtypes = [C.c_void_p, C.c_void_p]
    L.swehr_restore.argtypes = [C.c_void_p, C.c_void_p]
    L.swehr_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), P]
    L.swehr_hbm_stream.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P]
    PP = C.POINTER(C.POINTER(C.c_double))
    L.swehr_batch_create.argtypes = [C.c_int, C.POINTER(_abi.Params), PP, PP, PP, PP, PP, C.c_int, C.c_int,
                                     C.POINTER(C.c_void_p)]
    L.swehr_batch_advance.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int)]
    L.swehr_batch_handle.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
    L.swehr_batch_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), P]
    L.swehr_batch_profile.argtypes = [C.c_void_p, C.c_int, P]
    L.swehr_batch_set_persistent.argtypes = [C.c_void_p, C.c_int]
    L.swehr_batch_destroy.argtypes = [C.c_void_p]
    L.swehr_server_create.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.swehr_server_submit.argtypes = [C.c_void_p, C.POINTER(_abi.Params), P, P, P, P, P, C.POINTER(C.c_int64)]
    L.swehr_server_pump.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]
    L.swehr_server_next_done.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(_abi.StepInfo)]
    L.swehr_server_request.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]
    L.swehr_server_release.argtypes = [C.c_void_p, C.c_int64]
    L.swehr_server_destroy.argtypes = [C.c_void_p]
    _lib = L
    return L


def last_error() -> str:
    v = lib().swehr_last_error()
    return v.decode() if v else ""


def check(where: str, rc: int, allow=(0,)) -> int:
    if rc not in allow:
        raise SwehrError(where, rc, last_error())
    return rc


def default_params() -> _abi.Params:
    p = _abi.Params()
    lib().swehr_params_init(C.byref(p))
    return p


def instance_bytes(nx: int, ny: int, nclass: int, precision: int = _abi.FP64) -> int:
    return int(lib().swehr_instance_bytes(nx, ny, nclass, precision))


def instance_bytes_for(case, precision: int = _abi.FP64) -> int:
    """Exact device bytes swehr_create allocates for this case."""
    p = _abi.params_from_case(case)
    return int(lib().swehr_instance_bytes_for(C.byref(p), precision, 1 if case.ks_raster is not None else 0))


def tile_shapes():
    L = lib()
    out = []
    for k in range(1, L.swehr_tile_count() + 1):
        a, b = C.c_int(), C.c_int()
        L.swehr_tile_shape(k, C.byref(a), C.byref(b))
        out.append((a.value, b.value))
    return out


def hbm_stream(nx: int, ny: int, elem_bytes: int, planes_in: int, planes_out: int, reps: int = 5) -> float:
    """GB/s that a kernel streaming planes_in + planes_out state planes of nx x ny cells
    (the step kernel's layout and traversal) reaches on the current device."""
    out = C.c_double(0)
    check("swehr_hbm_stream", lib().swehr_hbm_stream(int(nx), int(ny), int(elem_bytes), int(planes_in),
                                                     int(planes_out), int(reps), C.byref(out)))
    return out.value


def parameter_rasters(case):
    """(hf, dtheta, roughness) rasters of a case as contiguous float64 arrays or None, in
    the order of enum swehr_param_raster."""
    out = []
    for attr in ("hf_raster", "dtheta_raster", "roughness_raster"):
        r = getattr(case, attr, None)
        out.append(None if r is None else np.ascontiguousarray(r, dtype=np.float64).reshape(-1))
    return out


class Solver:
    """One GPU instance of a ``swehr_amd.inputs.Case``."""

    def __init__(self, case, precision: int = _abi.FP64, pause_on_print: bool = False, steps_per_graph: int = 0,
                 tile: int = 0, _borrow=None):
        L = lib()
        self.case = case
        self.precision = precision
        self._owned = _borrow is None
        if _borrow is not None:
            self._h = _borrow
            return
        self._p = _abi.params_from_case(case)
        topo, ks, ds, pf, rain = _abi.case_arrays(case)
        h = C.c_void_p()
        rc = L.swehr_create(C.byref(self._p), dptr(topo), dptr(ks) if ks is not None else None, dptr(ds), dptr(pf),
                            dptr(rain), int(precision), C.byref(h))
        check("swehr_create", rc)
        self._h = h
        try:
            check("swehr_configure", L.swehr_configure(self._h, 1 if pause_on_print else 0, int(steps_per_graph),
                                                       int(tile)))
            for which, r in enumerate(parameter_rasters(case)):
                if r is not None:
                    check("swehr_set_parameter_raster", L.swehr_set_parameter_raster(self._h, which, dptr(r)))
        except SwehrError:
            L.swehr_destroy(self._h)
            self._h = None
            raise

    def configure(self, pause_on_print: int = -1, steps_per_graph: int = 0, tile: int = 0):
        check("swehr_configure", lib().swehr_configure(self._h, int(pause_on_print), int(steps_per_graph), int(tile)))

    def advance(self, n: int = 1 << 62) -> dict:
        info = _abi.StepInfo()
        rc = lib().swehr_advance(self._h, int(n), C.byref(info))
        check("swehr_advance", rc, allow=(_abi.OK, _abi.ERR_CFL))
        return info.as_dict()

    def run(self, n: int = 1 << 62) -> dict:
        """Advance n steps (or to the end), acknowledging print pauses."""
        left = int(n)
        info = self.info()
        while left > 0 and not info["done"]:
            before = info["steps"]
            info = self.advance(left)
            left -= info["steps"] - before
            if info["print_due"]:
                self.ack_print()
        return info

    def ack_print(self):
        check("swehr_ack_print", lib().swehr_ack_print(self._h))

    def info(self) -> dict:
        return self.advance(0)

    def field(self, name: str, k: int = 0) -> np.ndarray:
        out = np.empty(self.case.nx * self.case.ny, dtype=np.float64)
        check("swehr_download", lib().swehr_download(self._h, _abi.FIELDS[name], int(k), dptr(out)))
        return out.reshape(self.case.nx, self.case.ny)

    def state(self) -> dict:
        S = self.case.nclass
        d = {k: self.field(k) for k in ("depth", "uh", "vh", "topo", "vinf", "totm", "dep", "totsed", "maxvel",
                                        "maxdepth", "velx", "vely")}
        d["ch"] = np.stack([self.field("ch", k + 1) for k in range(S)])
        d["m"] = np.stack([self.field("m", k + 1) for k in range(S)])
        return d

    def stage(self):
        L = lib()
        cnt = C.c_int(0)
        check("swehr_stage", L.swehr_stage(self._h, None, None, 0, C.byref(cnt)))
        cap = max(1, cnt.value)
        st = np.zeros(_abi.STAGE_COLS * cap)
        fl = np.zeros(_abi.FLOW_COLS * cap)
        check("swehr_stage", L.swehr_stage(self._h, dptr(st), dptr(fl), cap, C.byref(cnt)))
        n = cnt.value
        return st.reshape(_abi.STAGE_COLS, cap)[:, :n].copy(), fl.reshape(_abi.FLOW_COLS, cap)[:, :n].copy()

    def discharge(self):
        """(q, qs): outlet water (m^3/s) and sediment (kg/s) through the j = 1/2 face at
        each stage record (swehr_discharge)."""
        L = lib()
        cnt = C.c_int(0)
        check("swehr_discharge", L.swehr_discharge(self._h, None, None, 0, C.byref(cnt)))
        cap = max(1, cnt.value)
        q, qs = np.zeros(cap), np.zeros(cap)
        check("swehr_discharge", L.swehr_discharge(self._h, dptr(q), dptr(qs), cap, C.byref(cnt)))
        return q[:cnt.value].copy(), qs[:cnt.value].copy()

    def budget(self) -> np.ndarray:
        out = np.zeros(_abi.BUDGET_N)
        check("swehr_budget", lib().swehr_budget(self._h, dptr(out)))
        return out

    def profile(self, nsteps: int):
        """Average kernel durations (ms) of wave-speed, step and control kernels over up
        to nsteps eagerly launched steps; returns (avg_ms[3], steps_taken)."""
        out = np.zeros(3)
        n = C.c_int(0)
        check("swehr_profile", lib().swehr_profile(self._h, int(nsteps), dptr(out), C.byref(n)))
        return out, n.value

    def profile_window(self, nsteps: int):
        """As profile(), then the state is put back on the device: the profiled steps
        are the ones the next advance() takes.  Returns (avg_ms[3], steps_profiled)."""
        out = np.zeros(3)
        n = C.c_int(0)
        check("swehr_profile_window", lib().swehr_profile_window(self._h, int(nsteps), dptr(out), C.byref(n)))
        return out, n.value

    def stream(self) -> int:
        s = C.c_void_p()
        check("swehr_get_stream", lib().swehr_get_stream(self._h, C.byref(s)))
        return int(s.value or 0)

    def snapshot(self) -> np.ndarray:
        n = C.c_int64(0)
        check("swehr_snapshot_bytes", lib().swehr_snapshot_bytes(self._h, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        check("swehr_snapshot", lib().swehr_snapshot(self._h, buf.ctypes.data_as(C.c_void_p)))
        return buf

    def restore(self, buf: np.ndarray):
        """Load a snapshot of an instance with the same layout (the C side checks the
        header; the size is checked here, where the buffer's length is known)."""
        buf = np.ascontiguousarray(buf, dtype=np.uint8).reshape(-1)
        n = C.c_int64(0)
        check("swehr_snapshot_bytes", lib().swehr_snapshot_bytes(self._h, C.byref(n)))
        if buf.nbytes != n.value:
            raise ValueError("snapshot of %d bytes, this instance's snapshots hold %d" % (buf.nbytes, n.value))
        check("swehr_restore", lib().swehr_restore(self._h, buf.ctypes.data_as(C.c_void_p)))

    def autotune(self, nsteps: int = 10):
        """Time every tile shape for nsteps steps (state restored after each), keep the
        fastest; returns (best_id, ms_per_tile)."""
        best = C.c_int(0)
        ms = np.zeros(lib().swehr_tile_count())
        check("swehr_autotune", lib().swehr_autotune(self._h, int(nsteps), C.byref(best), dptr(ms)))
        return best.value, ms

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().swehr_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def water_residual(b: np.ndarray) -> float:
    """Water budget residual (m^3): storage change minus the sum of all sources."""
    stored = b[0] + b[1]
    sources = b[2] - b[3] + b[9] + b[10]
    return float(stored - sources)


def sediment_residual(b: np.ndarray) -> float:
    """Sediment budget residual (kg): suspended + deposited minus detachment and inflow."""
    return float(b[4] + b[5] - (b[8] - b[7]))


def _ptr_array(arrays):
    arr = (C.POINTER(C.c_double) * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = dptr(a) if a is not None else C.cast(None, C.POINTER(C.c_double))
    return arr


class Batch:
    """An ensemble of cases stepped together on the GPU (one precision, shared tile
    shape, per-instance dt)."""

    def __init__(self, cases, precision: int = _abi.FP64, tile: int = 0):
        if any(r is not None for c in cases for r in parameter_rasters(c)):
            raise ValueError("parameter rasters (hf, dtheta, roughness) are supported on single instances")
        L = lib()
        self.cases = list(cases)
        self.precision = precision
        n = len(self.cases)
        self._params = (_abi.Params * n)(*[_abi.params_from_case(c) for c in self.cases])
        self._keep = [_abi.case_arrays(c) for c in self.cases]
        cols = list(zip(*self._keep))  # topo, ks, ds, pf, rain
        ptrs = [_ptr_array(col) for col in cols]
        h = C.c_void_p()
        check("swehr_batch_create", L.swehr_batch_create(n, self._params, ptrs[0], ptrs[1], ptrs[2], ptrs[3], ptrs[4],
                                                         int(precision), int(tile), C.byref(h)))
        self._b = h
        self._views = {}

    def advance(self, n: int = 1 << 62) -> int:
        left = C.c_int(0)
        check("swehr_batch_advance", lib().swehr_batch_advance(self._b, int(n), C.byref(left)))
        return left.value

    def member(self, i: int) -> Solver:
        if i not in self._views:
            h = C.c_void_p()
            check("swehr_batch_handle", lib().swehr_batch_handle(self._b, int(i), C.byref(h)))
            self._views[i] = Solver(self.cases[i], self.precision, _borrow=h)
        return self._views[i]

    def autotune(self, nsteps: int = 5):
        best = C.c_int(0)
        ms = np.zeros(lib().swehr_tile_count())
        check("swehr_batch_autotune", lib().swehr_batch_autotune(self._b, int(nsteps), C.byref(best), dptr(ms)))
        return best.value, ms

    def set_persistent(self, on: bool = True):
        """Step in one cooperative launch per 16 steps instead of graphs of three
        launches per step (bit-identical results)."""
        check("swehr_batch_set_persistent", lib().swehr_batch_set_persistent(self._b, int(bool(on))))

    def profile(self, nsteps: int = 10):
        """Average durations (ms) of the ensemble's wave-speed, step and control launches
        over nsteps eagerly launched steps (which advance the ensemble)."""
        out = np.zeros(3)
        check("swehr_batch_profile", lib().swehr_batch_profile(self._b, int(nsteps), dptr(out)))
        return out

    def close(self):
        if getattr(self, "_b", None):
            for v in self._views.values():
                v._h = None
            lib().swehr_batch_destroy(self._b)
            self._b = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching solve server (requests admitted by slots and memory)."""

    def __init__(self, mem_bytes: int, max_slots: int, precision: int = _abi.FP32, tile: int = 0):
        h = C.c_void_p()
        check("swehr_server_create", lib().swehr_server_create(int(mem_bytes), int(max_slots), int(precision),
                                                               int(tile), C.byref(h)))
        self._s = h
        self.precision = precision
        self.cases = {}

    def submit(self, case) -> int:
        if any(r is not None for r in parameter_rasters(case)):
            raise ValueError("parameter rasters (hf, dtheta, roughness) are supported on single instances")
        p = _abi.params_from_case(case)
        topo, ks, ds, pf, rain = _abi.case_arrays(case)
        rid = C.c_int64(0)
        rc = lib().swehr_server_submit(self._s, C.byref(p), dptr(topo), dptr(ks) if ks is not None else None,
                                       dptr(ds), dptr(pf), dptr(rain), C.byref(rid))
        check("swehr_server_submit", rc)
        self.cases[rid.value] = case
        return rid.value

    def pump(self, steps: int = 16):
        r, q, f = C.c_int(0), C.c_int(0), C.c_int(0)
        check("swehr_server_pump", lib().swehr_server_pump(self._s, int(steps), C.byref(r), C.byref(q), C.byref(f)))
        return r.value, q.value, f.value

    def next_done(self):
        rid = C.c_int64(0)
        info = _abi.StepInfo()
        rc = lib().swehr_server_next_done(self._s, C.byref(rid), C.byref(info))
        if rc == _abi.ERR_BUSY:
            return None
        check("swehr_server_next_done", rc)
        return rid.value, info.as_dict()

    def result(self, rid: int) -> Solver:
        h = C.c_void_p()
        check("swehr_server_request", lib().swehr_server_request(self._s, int(rid), C.byref(h)))
        return Solver(self.cases[rid], self.precision, _borrow=h)

    def stats(self):
        used, budget = C.c_int64(0), C.c_int64(0)
        check("swehr_server_stats", lib().swehr_server_stats(self._s, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def work(self) -> int:
        """Cell-class-steps stepped so far over every request, finished or running."""
        w = C.c_int64(0)
        check("swehr_server_work", lib().swehr_server_work(self._s, C.byref(w)))
        return w.value

    def release(self, rid: int):
        check("swehr_server_release", lib().swehr_server_release(self._s, int(rid)))
        self.cases.pop(rid, None)

    def evict(self, rid: int):
        """Checkpoint a running request to host memory and requeue it (it resumes exactly)."""
        check("swehr_server_evict", lib().swehr_server_evict(self._s, int(rid)))

    def set_quantum(self, steps: int):
        """Time slicing: evict requests that had `steps` steps when the queue head waits (0: off)."""
        check("swehr_server_set_quantum", lib().swehr_server_set_quantum(self._s, int(steps)))

    def evictions(self) -> int:
        n = C.c_int64(0)
        check("swehr_server_evictions", lib().swehr_server_evictions(self._s, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "_s", None):
            lib().swehr_server_destroy(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
