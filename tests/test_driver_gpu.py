"""The drop-in program (swehr_amd/bin/swehr) against the oracle's program on the
golden inputs: same stdout lines and the same 17 output files, numerically equal to
the print precision of the reference's formats (rasters "%10.10lf", series "%lf").
The oracle's program itself reproduces the compiled reference byte for byte
(test_oracle_golden.py), so this pins the drop-in outputs to the reference's."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from golden.cases import OUTPUT_FILES, stdout_lines
from oracle_lib import RUN_PATH
from swehr_amd.solver import DRIVER

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _numbers(path):
    with open(path) as fh:
        return np.array([float(x) for x in fh.read().split()])


def _run(prog, src, dst):
    os.makedirs(dst)
    for fn in os.listdir(src):
        shutil.copy(os.path.join(src, fn), dst)
    p = subprocess.run([prog, dst], capture_output=True, text=True, timeout=600, env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert p.returncode == 0, p.stderr
    return stdout_lines(p.stdout)


@pytest.mark.parametrize("name", ["shipped", "inflow_mix", "rain_plane", "transmissive", "cfl_abort"])
def test_driver_outputs_match_oracle_program(name, tmp_path):
    src = os.path.join(GOLD, name, "inputs")
    print("case", name, flush=True)
    a = _run(RUN_PATH, src, str(tmp_path / "oracle"))
    b = _run(DRIVER, src, str(tmp_path / "gpu"))
    assert len(a) == len(b)
    for la, lb in zip(a, b):
        if la == lb:
            continue
        fa = [float(x) for x in la.replace("!!! CFL Error at t=", "").split()]
        fb = [float(x) for x in lb.replace("!!! CFL Error at t=", "").split()]
        np.testing.assert_allclose(fb, fa, rtol=1e-6, atol=2e-6)
    for fn in OUTPUT_FILES:
        pa, pb = str(tmp_path / "oracle" / fn), str(tmp_path / "gpu" / fn)
        assert os.path.exists(pa) == os.path.exists(pb), fn
        va, vb = _numbers(pa), _numbers(pb)
        assert va.shape == vb.shape, fn
        assert np.array_equal(np.isnan(va), np.isnan(vb)), fn
        m = ~np.isnan(va)
        tol = 2e-6 if fn in ("stage.txt", "saveflow.txt") else 1e-8
        np.testing.assert_allclose(vb[m], va[m], rtol=1e-7, atol=tol, err_msg=fn)


def test_driver_writes_outlet_discharge_on_request(tmp_path):
    """Sidecar "discharge = 1" adds discharge.txt (time, outlet water and sediment
    discharge at each stage record); the values are the library's series."""
    from parity_util import outlet_case
    from swehr_amd import solver as sv
    from swehr_amd.inputs import write_case

    case = outlet_case(tend=120.0)
    d = str(tmp_path / "run")
    os.makedirs(d)
    write_case(case, d)
    with open(os.path.join(d, "swehr_ext.txt"), "a") as fh:
        fh.write("discharge = 1\n")
    p = subprocess.run([DRIVER, d], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr
    rows = np.loadtxt(os.path.join(d, "discharge.txt"))
    with sv.Solver(case) as s:
        s.run()
        q, qs = s.discharge()
        st, _ = s.stage()
    assert rows.shape == (len(q), 3)
    np.testing.assert_allclose(rows[:, 0], st[0], atol=1e-6)
    np.testing.assert_allclose(rows[:, 1], q, rtol=0, atol=1e-6)
    np.testing.assert_allclose(rows[:, 2], qs, rtol=0, atol=1e-6)


def test_driver_reads_parameter_raster_files(tmp_path):
    """hf_file, dtheta_file and roughness_file in the sidecar: the drop-in program and
    the oracle's program read them alike and agree to the print precision."""
    from swehr_amd.inputs import write_case
    from test_oracle_invariants import _raster_case

    case = _raster_case()
    case.tend = 120.0
    src = str(tmp_path / "in")
    os.makedirs(src)
    write_case(case, src)
    assert os.path.exists(os.path.join(src, "dtheta.txt"))
    _run(RUN_PATH, src, str(tmp_path / "oracle"))
    _run(DRIVER, src, str(tmp_path / "gpu"))
    for fn in ("depth.txt", "uh.txt", "vh.txt", "ch.txt"):
        va, vb = _numbers(str(tmp_path / "oracle" / fn)), _numbers(str(tmp_path / "gpu" / fn))
        np.testing.assert_allclose(vb, va, rtol=1e-7, atol=1e-8, err_msg=fn)


@pytest.mark.parametrize("name", ["transmissive", "closed_basin"])
def test_driver_fp32_precision_form(name, tmp_path):
    """Sidecar "precision = fp32" runs the drop-in program in the precision form: the
    same files and stdout lines as the oracle's FP64 program, with every raster within
    FP32 accuracy (the oracle's program reads the key and stays FP64)."""
    src = str(tmp_path / "in")
    shutil.copytree(os.path.join(GOLD, name, "inputs"), src)
    with open(os.path.join(src, "swehr_ext.txt"), "a") as fh:
        fh.write("precision = fp32\n")
    a = _run(RUN_PATH, src, str(tmp_path / "oracle"))
    b = _run(DRIVER, src, str(tmp_path / "gpu"))
    assert len(a) == len(b)
    for fn in ("topoout.txt", "depth.txt", "uh.txt", "vh.txt", "ch.txt", "m.txt", "c.txt", "vel.txt",
               "maxvel.txt", "maxdepth.txt"):
        va, vb = _numbers(str(tmp_path / "oracle" / fn)), _numbers(str(tmp_path / "gpu" / fn))
        assert va.shape == vb.shape, fn
        scale = max(float(np.max(np.abs(va))), 1e-12)
        err = float(np.max(np.abs(vb - va))) / scale
        print(fn, "%.2e" % err)
        assert err <= 1e-4, fn
