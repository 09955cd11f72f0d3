"""Invariants the oracle must satisfy beyond the golden outputs: its water and
sediment budgets close (every source and sink is accounted for), and its results do
not depend on the OpenMP thread count (the lagged-geometry restatement removes the
reference's loop-order dependence)."""
import os

import numpy as np
import pytest

from oracle_lib import Oracle
from swehr_amd import solver as sv
from swehr_amd.inputs import read_case

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["transmissive", "closed_basin", "rain_plane"])
def test_oracle_budgets_close(name):
    case = read_case(os.path.join(GOLD, name, "inputs"))
    o = Oracle(case)
    o.run()
    b = o.budget()
    wscale = max(abs(b[n]) for n in (0, 1, 2, 3, 9, 10))
    sscale = max(max(abs(b[n]) for n in (4, 5, 7, 8)), 1e-30)
    assert abs(sv.water_residual(b)) <= 1e-12 * wscale
    assert abs(sv.sediment_residual(b)) <= 1e-12 * sscale


def test_oracle_thread_count_invariant():
    case = read_case(os.path.join(GOLD, "transmissive", "inputs"))
    a = Oracle(case, threads=1)
    a.run(300)
    b = Oracle(case, threads=4)
    b.run(300)
    for k in ("depth", "uh", "vh", "topo", "totm"):
        np.testing.assert_array_equal(a.field(k), b.field(k))


def test_outlet_discharge_integrates_to_boundary_budget_and_reaches_rain_rate():
    """The oracle's outlet discharge series is pinned two ways.  (1) With dt = 1 s and a
    stage record every step, the series summed over the run equals the net boundary
    outflow of the compensated budget (the other three sides are walls, whose mass flux
    is exactly zero): a wrong sign, face or width factor fails.  (2) Without
    infiltration the outflow of a plane under constant rain tends to rain rate x area,
    a closed form."""
    from parity_util import outlet_case

    case = outlet_case(tend=3000.0)
    o = Oracle(case, threads=4)
    o.run()
    q, qs = o.discharge()
    info = o.info()
    b = o.budget()
    o.close()
    assert len(q) == info["steps"] == 3000
    assert abs(q.sum() - b[3]) <= 1e-12 * abs(b[3])
    assert abs(qs.sum() - b[7]) <= 1e-12 * max(abs(b[7]), 1e-30)
    area_rate = 3e-5 * case.nx * case.ny * case.dx ** 2
    assert abs(q[-1] / area_rate - 1) < 0.02
    assert np.all(q >= -1e-12) and np.all(qs >= -1e-12)


def _raster_case(seed=9):
    from parity_util import outlet_case

    case = outlet_case(nx=24, ny=40, tend=600.0, seed=seed)
    case.sidecar = dict(case.sidecar, ks=3e-6, hf=0.1, theta0=0.1, thetas=0.4)
    rng = np.random.default_rng(seed)
    sh = (case.nx, case.ny)
    case.hf_raster = 0.1 * np.exp(0.3 * rng.standard_normal(sh))
    case.dtheta_raster = rng.uniform(0.1, 0.4, sh)
    case.roughness_raster = rng.uniform(0.02, 0.08, sh)
    return case


def test_parameter_rasters_equal_to_the_scalars_reproduce_the_scalar_run():
    """Rasters holding the sidecar's scalars reproduce the scalar run bit for bit (the
    per-cell path does the same arithmetic), and non-uniform rasters change it."""
    import copy

    base = _raster_case()
    plain = copy.copy(base)
    plain.hf_raster = plain.dtheta_raster = plain.roughness_raster = None
    same = copy.copy(plain)
    sh = (base.nx, base.ny)
    same.hf_raster = np.full(sh, 0.1)
    same.dtheta_raster = np.full(sh, 0.4 - 0.1)
    same.roughness_raster = np.full(sh, plain.side("roughness"))
    out = []
    for c in (plain, same, base):
        o = Oracle(c, threads=2)
        o.run()
        out.append({f: o.field(f) for f in ("depth", "vinf", "uh")})
        o.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f
    assert not np.array_equal(out[0]["vinf"], out[2]["vinf"])


def test_parameter_raster_files_round_trip(tmp_path):
    from swehr_amd.inputs import read_case, write_case

    c = _raster_case()
    write_case(c, str(tmp_path))
    r = read_case(str(tmp_path))
    for a in ("hf_raster", "dtheta_raster", "roughness_raster"):
        np.testing.assert_array_equal(getattr(r, a), getattr(c, a))


def test_oracle_self_divergence_after_a_threshold_flip():
    """The long-run bound of the GPU parity tests rests on this property of the FP64
    model itself: on the channel-network inputs, the oracle and the oracle with Ks
    multiplied by (1 + 1e-15 N(0,1)) agree to rounding level for 300 steps, and some 40
    steps later a receding cell meets the film-depth test (SWEHR.c:1466-1471,
    1532-1551) on different sides in the two runs, after which they differ by orders of
    magnitude more than rounding."""
    from parity_util import channel_case, compare_states, perturbed_ks, worst

    case = channel_case()
    out = {}
    for steps in (300, 400):
        a = Oracle(case, threads=4)
        a.run(steps)
        b = Oracle(perturbed_ks(case, 1), threads=4)
        b.run(steps)
        out[steps] = worst(compare_states(b.state(), a.state(), case.nclass))[1]
        a.close()
        b.close()
    print(out)
    assert out[300] <= 1e-12
    assert out[400] >= 1e-6
