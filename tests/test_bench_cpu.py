"""The CPU legs of bench.py that need no GPU."""
import os

import pytest

import bench
from swehr_amd.inputs import config_c2

REF = os.path.join(bench.ROOT, "oracle", "_ref", "swehr_ref_hooked")


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref is built only where the reference sources are")
def test_reference_binary_baseline_times_steps():
    """The compiled reference runs the bench case through its parameter hooks (Ks
    raster, soil values, courant number, inflow and movie switches): a 1-step and a
    1 + n-step run both succeed and the difference is a positive time for n steps."""
    case = config_c2(n=64)
    n, seconds, t_one = bench.reference_binary_baseline(case, 0.5, 0.01)
    assert n == 50
    assert seconds > 0 and t_one > 0


def test_bytes_per_cell_matches_the_algorithmic_count():
    """B_cell = 2 w (6 + 2 S) + 16 + w: one read and one write of h, hu, hv, z, MAXDEPTH,
    MAXVEL, ch_k and M_k at w bytes, VINF as FP64, one read of the Ks raster."""
    assert bench.bytes_per_cell(7, 8) == 344
    assert bench.bytes_per_cell(10, 8) == 440
    assert bench.bytes_per_cell(10, 4) == 228
    assert bench.bytes_per_cell(5, 4) == 148
    assert bench.bytes_per_cell(7, 8, ks_raster=False) == 336


def test_c5_requests_relief_scales_with_extent():
    """Serving requests carry the relief of a window of the 4096 m fractal surface,
    (5-20 m) x (extent / 4096 m)^0.8, on the 5 % tilt: steep small grids (the full 5-20 m
    on 128 m) drove the explicit scheme to speeds of ~1000 m/s and 10^5-10^6 steps."""
    import numpy as np

    from swehr_amd.inputs import config_c5_request

    rng = np.random.default_rng(5)
    seen = set()
    for k in range(40):
        c = config_c5_request(rng, k, sizes=(128, 256, 512), weights=(1, 1, 1))
        j = np.arange(c.ny, dtype=np.float64)
        rough = c.topo - 0.05 * c.dx * j[None, :]  # the fractal part, tilt removed
        relief = float(rough.max() - rough.min())
        cap = 20.0 * (c.nx * c.dx / 4096.0) ** 0.8
        assert 0.25 * cap <= relief <= cap * (1 + 1e-12), (c.nx, relief, cap)
        seen.add(c.nx)
    assert seen == {128, 256, 512}


def test_bench_refuses_more_than_one_gpu():
    """The build runs on one GPU: --gpus above 1 exits with status 2 and says why,
    before it touches a device."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
