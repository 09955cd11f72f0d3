"""The CPU oracle against outputs of the compiled reference (tests/golden/).

Each golden case holds the inputs, the SHA-256 of the reference's 17 output files and
of its stdout, and (for most cases) its final state at full precision.  The oracle must
reproduce all of them exactly: the text files byte for byte and the state bit for bit.
Together the cases cover the inflow boundary, rain, rainsplash and stream-power
entrainment, coarse-class settling, slope failures, Green-Ampt with Ks > 0, every
boundary type, a per-cell Ks raster, the two-cell-wide grid and the CFL stop.
"""
import hashlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from golden.cases import CASES, OUTPUT_FILES, stdout_lines
from oracle_lib import RUN_PATH, Oracle
from swehr_amd.inputs import read_case

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _sha(path):
    if not os.path.exists(path):
        return None
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def _expected(name):
    with open(os.path.join(GOLD, name, "expected.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("gc", CASES, ids=[c.name for c in CASES])
def test_oracle_outputs_match_reference_bytes(gc, tmp_path):
    exp = _expected(gc.name)
    src = os.path.join(GOLD, gc.name, "inputs")
    for fn in os.listdir(src):
        shutil.copy(os.path.join(src, fn), tmp_path)
    env = dict(os.environ, OMP_NUM_THREADS="1" if gc.name != "c1_plane_100x40" else "4")
    args = [RUN_PATH, str(tmp_path)] + ([str(gc.maxsteps)] if gc.maxsteps else [])
    p = subprocess.run(args, env=env, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr
    lines = stdout_lines(p.stdout)
    assert lines[:20] == exp["stdout_head"]
    assert len(lines) == exp["stdout_lines"]
    assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == exp["stdout_sha256"]
    for fn in OUTPUT_FILES:
        assert _sha(os.path.join(tmp_path, fn)) == exp["files"][fn], fn


@pytest.mark.parametrize("gc", [c for c in CASES if c.dump], ids=[c.name for c in CASES if c.dump])
def test_oracle_state_matches_reference_bits(gc):
    exp = _expected(gc.name)
    ref = np.load(os.path.join(GOLD, gc.name, "dump.npz"))
    case = read_case(os.path.join(GOLD, gc.name, "inputs"))
    o = Oracle(case)
    o.run(gc.maxsteps if gc.maxsteps else 1 << 62)
    info = o.info()
    assert info["steps"] == exp["steps"]
    assert info["t"] == exp["t"]
    st = o.state()
    names = {"depth": "depth", "uh": "uh", "vh": "vh", "topo": "topo", "vinf": "vinf", "totm": "totm", "dep": "dep",
             "totsed": "totsed", "maxvel": "maxvel", "maxdepth": "maxdepth", "velx": "velx", "vely": "vely"}
    for mine, theirs in names.items():
        np.testing.assert_array_equal(st[mine], ref[theirs], err_msg=mine)
    for k in range(case.nclass):
        np.testing.assert_array_equal(st["ch"][k], ref["ch%d" % (k + 1)], err_msg="ch%d" % (k + 1))
        np.testing.assert_array_equal(st["m"][k], ref["m%d" % (k + 1)], err_msg="m%d" % (k + 1))
    stage, flow = o.stage()
    np.testing.assert_array_equal(stage, ref["stage"])
    np.testing.assert_array_equal(flow, ref["saveflow"])


def test_golden_cases_exercise_the_branches():
    """The golden set reaches the branches only whole-run goldens can pin: slope
    failures, the dry-cell clamp with sediment and the CFL stop."""
    case = read_case(os.path.join(GOLD, "inflow_mix", "inputs"))
    o = Oracle(case)
    o.run()
    assert o.fail_events() > 0
    assert "!!! CFL Error" in "\n".join(_expected("cfl_abort")["stdout_tail"])
