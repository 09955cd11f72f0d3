"""The C ABI library without a GPU: it loads, exports every function include/swehr.h
declares, reports the reference's defaults, sizes instances, and refuses to create an
instance when no GPU is present (no compute call is made here)."""
import ctypes as C
import os

import pytest

from swehr_amd import _abi
from swehr_amd import solver as sv


def test_library_exports_every_declared_function():
    names = sv.header_functions()
    assert len(names) >= 25
    L = sv.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_params_init_gives_reference_defaults():
    p = sv.default_params()
    assert p.cstable == 0.05 and p.maxstep == 0.25  # SWEHR.c:632-633
    assert list(p.bndry) == [1, 0, 1, 1]  # SWEHR.c:618-621
    assert p.inflow == 1 and p.roughness == 0.1 and p.theta0 == 0.05 and p.thetas == 0.4
    assert p.hf == 0.001 and p.ks == 0.0 and p.vinf0 == 1e-4  # SWEHR.c:236-243


def test_instance_bytes_scale_with_grid_and_precision():
    a = sv.instance_bytes(1024, 1024, 7, _abi.FP64)
    b = sv.instance_bytes(2048, 2048, 7, _abi.FP64)
    c = sv.instance_bytes(2048, 2048, 7, _abi.FP32)
    assert 3.9 < b / a < 4.1
    assert c < 0.6 * b
    # 4096^2, S=10 plain FP64 state fits easily in 288 GB
    assert sv.instance_bytes(4096, 4096, 10, _abi.FP64) < 8 * 2 ** 30


def test_tile_shapes_are_256_or_128_threads():
    for ti, tj in sv.tile_shapes():
        assert ti * tj in (128, 256)


def test_strerror_and_bad_input():
    L = sv.lib()
    assert L.swehr_strerror(_abi.ERR_CFL).decode().startswith("CFL")
    info = _abi.StepInfo()
    assert L.swehr_advance(None, 1, C.byref(info)) == _abi.ERR_BAD_INPUT
    assert L.swehr_download(None, 0, 0, None) == _abi.ERR_BAD_INPUT
    assert L.swehr_server_evict(None, 1) == _abi.ERR_BAD_INPUT
    assert L.swehr_server_set_quantum(None, 16) == _abi.ERR_BAD_INPUT


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_reports_no_device():
    from swehr_amd.inputs import read_case

    case = read_case(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shipped", "inputs"))
    with pytest.raises(sv.SwehrError) as ei:
        sv.Solver(case)
    assert ei.value.code == _abi.ERR_NO_DEVICE


def test_create_rejects_a_grid_beyond_32bit_plane_offsets():
    # 2^29 FP64 cells per plane is the limit of the kernels' 32-bit byte offsets; the
    # check precedes any device call, so it runs without a GPU
    L = sv.lib()
    p = sv.default_params()
    p.nx, p.ny, p.nclass, p.rnum, p.rint, p.tend, p.dx = 24000, 24000, 3, 1, 1.0, 1.0, 1.0
    one = (C.c_double * 3)(1e-4, 1e-4, 1e-4)
    rain = (C.c_double * 1)(0.0)
    topo = (C.c_double * 1)(0.0)  # never read: the parameters are refused first
    h = C.c_void_p()
    rc = L.swehr_create(C.byref(p), topo, None, one, one, rain, _abi.FP64, C.byref(h))
    assert rc == _abi.ERR_BAD_INPUT
    assert b"too large" in L.swehr_last_error()
