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


# ---------------------------------------------------------------- fast text I/O
import ctypes as C

from swehr_amd import solver as sv

_libc = C.CDLL(None)
_libc.snprintf.restype = C.c_int


def _printf(fmt: bytes, v: float) -> bytes:
    buf = C.create_string_buffer(512)
    _libc.snprintf(buf, 512, fmt, C.c_double(v))
    return buf.value


def _fixed(v: float, prec: int) -> bytes:
    L = sv.lib()
    f = L.swehr_io_format_fixed
    f.argtypes = [C.c_char_p, C.c_double, C.c_int]
    f.restype = C.c_int
    buf = C.create_string_buffer(512)
    n = f(buf, v, prec)
    assert n == len(buf.value)
    return buf.value


def _hard_values(rng):
    """Values where a formatter that is not exact goes wrong: dyadic ties at the last
    printed digit (round half to even), values one ulp either side of them, signed
    zeros and tiny negatives, denormals, large values (fallback path), nan / inf."""
    vals = [0.0, -0.0, 1e-300, -1e-300, 5e-324, -5e-324, 0.5e-10, 5e-11, 4.9999999999e-11, 1e9 - 0.5, 1.8e9,
            2e9, 1e12, -1e15, 1e300, float("nan"), -float("nan"), float("inf"), -float("inf"), 0.1, 0.2, 0.3,
            123.45678901235, 2.5, -2.5, 1 / 3]
    for k in range(1, 4096, 7):  # k / 2^11 has 11 decimals: a tie at the 10th
        x = k / 2048.0
        vals += [x, -x, np.nextafter(x, 0), np.nextafter(x, 1), x + 17]
    for k in range(1, 256, 3):  # k / 2^7: ties at the 6th decimal
        vals += [k / 128.0, -k / 128.0]
    vals += list(rng.standard_normal(4000) * 10.0 ** rng.integers(-12, 9, 4000))
    vals += list(rng.uniform(0, 400, 2000))
    return [float(v) for v in vals]


def test_fixed_formatter_matches_printf_exactly():
    rng = np.random.default_rng(11)
    for v in _hard_values(rng):
        for prec, fmt in ((10, b"%.10f"), (6, b"%.6f")):
            assert _fixed(v, prec) == _printf(fmt, v), (v, prec)


def test_raster_and_series_writers_are_byte_identical_to_printf(tmp_path):
    rng = np.random.default_rng(12)
    nx, ny = 37, 301  # more rows than one writer batch
    vals = np.array(_hard_values(rng))
    v = rng.choice(vals, size=nx * ny)
    vy = rng.standard_normal(nx * ny)
    L = sv.lib()
    _libc.fopen.restype = C.c_void_p
    _libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    _libc.fclose.argtypes = [C.c_void_p]
    for name, call in (("r", lambda f: L.swehr_io_write_raster(C.c_void_p(f), v.ctypes.data_as(C.POINTER(C.c_double)),
                                                                 nx, ny)),
                       ("s", lambda f: L.swehr_io_write_speed(C.c_void_p(f), v.ctypes.data_as(C.POINTER(C.c_double)),
                                                                vy.ctypes.data_as(C.POINTER(C.c_double)), nx, ny))):
        path = str(tmp_path / name)
        f = _libc.fopen(path.encode(), b"w")
        assert call(f) == 0
        _libc.fclose(f)
        got = open(path, "rb").read()
        exp = []
        _libc.pow.restype = C.c_double
        _libc.pow.argtypes = [C.c_double, C.c_double]
        for j in range(ny):
            for i in range(nx):
                x = v[i * ny + j]
                if name == "s":  # pow(vx^2 + vy^2, 0.5) as the reference (SWEHR.c:1735)
                    with np.errstate(over="ignore", invalid="ignore"):
                        q = np.float64(x) * np.float64(x) + np.float64(vy[i * ny + j]) * np.float64(vy[i * ny + j])
                    x = _libc.pow(float(q), 0.5)
                exp.append(_printf(b"%10.10lf" + (b"\t" if i != nx - 1 else b"\n"), x))
        assert got == b"".join(exp)
    cols = rng.choice(vals, size=24 * 50)
    path = str(tmp_path / "series")
    f = _libc.fopen(path.encode(), b"w")
    assert L.swehr_io_write_series(C.c_void_p(f), cols.ctypes.data_as(C.POINTER(C.c_double)), 24, 50, 50) == 0
    _libc.fclose(f)
    exp = b"".join(_printf(b"%lf" + (b"\t" if c != 23 else b"\n"), cols[c * 50 + r]) for r in range(50) for c in range(24))
    assert open(path, "rb").read() == exp


def test_parallel_parser_reads_like_fscanf():
    rng = np.random.default_rng(13)
    v = np.concatenate([rng.standard_normal(30000) * 10.0 ** rng.integers(-8, 6, 30000), [0.0, -0.0, 1e-310]])
    text = "\n".join(" \t".join(repr(float(x)) for x in v[k:k + 97]) for k in range(0, len(v), 97)).encode()
    L = sv.lib()
    p = L.swehr_io_parse_doubles
    p.argtypes = [C.c_char_p, C.c_long, C.POINTER(C.c_double), C.c_long]
    p.restype = C.c_long
    out = np.empty(len(v) + 5)
    miss = p(text, len(text), out.ctypes.data_as(C.POINTER(C.c_double)), len(out))
    assert miss == 5
    assert np.array_equal(out[:len(v)], v) and not out[len(v):].any()
    assert np.array_equal(np.signbit(out[:len(v)]), np.signbit(v))
    bad = b"1.0 2.0 3x 4.0"
    assert p(bad, len(bad), out.ctypes.data_as(C.POINTER(C.c_double)), 4) == -1


def test_driver_reports_a_missing_input_file(tmp_path):
    """The drop-in program stops with a message naming the file and a nonzero status
    where an input file is missing (it reads every input before it touches a GPU)."""
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "swehr_amd", "bin", "swehr")
    if not os.path.exists(exe):
        pytest.skip("drop-in program not built")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "input.txt" in (r.stderr + r.stdout)
