"""The reference's file contract (swehr_amd/csrc/host/swehr_io.c, used by the drop-in
program) and the Python case writer: inputs written by write_case read back exactly,
the sidecar is parsed and validated, and the writers format like SWEHR.c:1718-1761."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import RUN_PATH
from swehr_amd.inputs import (REFERENCE_SIDECAR, Case, config_c1, diamond_square, log_spaced_classes, read_case,
                              write_case)


def test_write_read_round_trip_is_exact(tmp_path):
    c = config_c1(seed=3)
    write_case(c, str(tmp_path))
    r = read_case(str(tmp_path))
    assert (r.nx, r.ny, r.nclass, r.rnum) == (c.nx, c.ny, c.nclass, c.rnum)
    np.testing.assert_array_equal(r.topo, c.topo)
    np.testing.assert_array_equal(r.rain, c.rain)
    np.testing.assert_array_equal(r.ds, c.ds)
    for k in REFERENCE_SIDECAR:
        assert r.side(k) == c.side(k), k


def test_reference_case_writes_no_sidecar(tmp_path):
    c = read_case(os.path.join(os.path.dirname(__file__), "golden", "shipped", "inputs"))
    write_case(c, str(tmp_path))
    assert not os.path.exists(tmp_path / "swehr_ext.txt")


def _tiny(**kw):
    topo = np.arange(2 * 50, dtype=float).reshape(2, 50) * 0.01
    return Case(nx=2, ny=50, dx=0.5, tend=2.0, topo=topo, rain=np.zeros(1), ds=np.array([1e-4]),
                pfrac=np.array([1.0]), **kw)


def test_unknown_sidecar_key_is_rejected(tmp_path):
    write_case(_tiny(), str(tmp_path))
    with open(tmp_path / "swehr_ext.txt", "w") as fh:
        fh.write("cstabel = 0.3\n")
    p = subprocess.run([RUN_PATH, str(tmp_path)], capture_output=True, text=True)
    assert p.returncode == 1 and "unknown key" in p.stderr


def test_raster_format_is_reference_format(tmp_path):
    c = _tiny(movies=0)
    write_case(c, str(tmp_path))
    p = subprocess.run([RUN_PATH, str(tmp_path)], capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert p.returncode == 0, p.stderr
    lines = open(tmp_path / "topoout.txt").read().split("\n")
    assert len(lines) == c.ny + 1 and lines[-1] == ""
    row = lines[0].split("\t")
    assert len(row) == c.nx
    assert row[0] == "%10.10f" % c.topo[0, 0] and row[1] == "%10.10f" % c.topo[1, 0]
    st = open(tmp_path / "stage.txt").read().split("\n")[0].split("\t")
    assert len(st) == 24 and st[0] == "0.250000"  # first record after the first 0.25 s step


def test_short_input_reads_zero_like_reference(tmp_path):
    write_case(_tiny(movies=0), str(tmp_path))
    with open(tmp_path / "topoin.txt", "w") as fh:
        fh.write("1 2 3\n")
    p = subprocess.run([RUN_PATH, str(tmp_path)], capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert p.returncode == 0
    v = np.array([float(x) for x in open(tmp_path / "topoout.txt").read().split()])
    assert v.size == 100 and np.count_nonzero(v[:]) >= 1


def test_generators_are_seeded_and_scaled():
    rng1, rng2 = np.random.default_rng(5), np.random.default_rng(5)
    a, b = diamond_square(5, 0.8, rng1), diamond_square(5, 0.8, rng2)
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean()) < 1e-12 and abs(a.std() - 1) < 1e-12
    d = log_spaced_classes(7)
    assert d[0] == pytest.approx(2e-6) and d[-1] == pytest.approx(2e-3)
