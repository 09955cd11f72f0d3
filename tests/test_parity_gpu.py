"""HIP path against the CPU oracle on the same inputs (needs an MI355X).

The FP64 kernel keeps the reference's operation order but replaces some libm calls by
identities that are exact in real arithmetic (sqrt for pow(x, .5), 1/sqrt(1+x^2) for
cos(atan x), one exp instead of two equal ones) and uses the GPU's own exp/pow, so it
agrees with the oracle to rounding level, not bit for bit.  Over a few hundred steps
the relative L2 error of every field stays below RTOL_FP64_SHORT (measured: at most
2.2e-14); whole runs of the golden cases stay below their RTOL_FP64_RUN entry, set 10x
above the error measured on the MI355X (the FP64 path is deterministic: every
reduction has a fixed order, and the one atomic is a max, which is exact).  Time and
step counts are compared as well.
"""
import json
import os

import numpy as np
import pytest

from golden.cases import BY_NAME
from oracle_lib import Oracle
from parity_util import compare_states, rel_l2, worst
from swehr_amd import _abi
from swehr_amd import solver as sv
from swehr_amd.inputs import read_case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

RTOL_FP64_SHORT = 1e-12
#: whole runs: 10x the worst field measured (inflow_mix uh 7.2e-13, narrow 2.0e-15,
#: rain_plane 0, closed_basin 1.8e-15, cfl_abort velx 1.0e-12)
RTOL_FP64_RUN = {"inflow_mix": 1e-11, "narrow": 2e-14, "rain_plane": 1e-15, "closed_basin": 2e-14,
                 "cfl_abort": 1e-11}


def _case(name):
    return read_case(os.path.join(GOLD, name, "inputs"), name=name)


def _pair(case, steps, precision=_abi.FP64):
    o = Oracle(case)
    o.run(steps)
    s = sv.Solver(case, precision=precision)
    s.run(steps)
    return o, s


SHORT = [("shipped", 1), ("shipped", 40), ("inflow_mix", 1), ("inflow_mix", 300), ("narrow", 200),
         ("rain_plane", 120), ("transmissive", 150), ("closed_basin", 150), ("cfl_abort", 60),
         ("param_rasters", 300)]


@pytest.mark.parametrize("name,steps", SHORT, ids=["%s-%d" % x for x in SHORT])
def test_fp64_steps_match_oracle(name, steps):
    case = _case(name)
    print("case", name, "steps", steps, flush=True)
    o, s = _pair(case, steps)
    oi, si = o.info(), s.info()
    assert si["steps"] == oi["steps"]
    assert si["t"] == pytest.approx(oi["t"], rel=1e-12, abs=1e-12)
    errs = compare_states(s.state(), o.state(), case.nclass)
    print({k: "%.2e" % v for k, v in errs.items()})
    k, e = worst(errs)
    assert e <= RTOL_FP64_SHORT, (k, e)
    st_g, fl_g = s.stage()
    st_o, fl_o = o.stage()
    assert st_g.shape == st_o.shape
    assert rel_l2(st_g, st_o) <= RTOL_FP64_SHORT
    assert rel_l2(fl_g, fl_o) <= RTOL_FP64_SHORT
    s.close()


@pytest.mark.parametrize("name", ["inflow_mix", "narrow", "rain_plane", "closed_basin", "cfl_abort"])
def test_fp64_whole_run_matches_oracle(name):
    case = _case(name)
    print("case", name, "whole run", flush=True)
    o, s = _pair(case, 1 << 40)
    oi, si = o.info(), s.info()
    assert si["done"] == 1 and oi["done"] == 1
    assert si["status"] == oi["status"]
    assert abs(si["steps"] - oi["steps"]) <= max(2, oi["steps"] // 1000)
    errs = compare_states(s.state(), o.state(), case.nclass)
    print({k: "%.2e" % v for k, v in errs.items()})
    k, e = worst(errs)
    assert e <= RTOL_FP64_RUN[name], (k, e)
    s.close()


@pytest.mark.parametrize("name", ["transmissive", "closed_basin"])
def test_fp64_mass_budgets_close(name):
    """Water and sediment budgets close to rounding, and every budget term agrees with
    the oracle's (terms are compared relative to the largest water / sediment term)."""
    case = _case(name)
    with sv.Solver(case) as s:
        s.run()
        b = s.budget()
    o = Oracle(case)
    o.run()
    bo = o.budget()
    wscale = max(abs(bo[n]) for n in (0, 1, 2, 3, 9, 10))
    sscale = max(max(abs(bo[n]) for n in (4, 5, 7, 8)), 1e-30)
    print("gpu", b, "\noracle", bo)
    assert abs(sv.water_residual(b)) <= 1e-11 * wscale
    assert abs(sv.sediment_residual(b)) <= 1e-11 * sscale
    for n in (0, 1, 2, 3, 9, 10):
        assert abs(b[n] - bo[n]) <= 1e-7 * wscale, n
    for n in (4, 5, 7, 8):
        assert abs(b[n] - bo[n]) <= 1e-7 * sscale, n
    assert abs(b[6] - bo[6]) <= 1e-7 * max(abs(bo[6]), 1e-9)
    assert b[11] == bo[11] or abs(b[11] - bo[11]) <= max(2, bo[11] // 1000)


@pytest.mark.parametrize("name", ["transmissive", "closed_basin", "rain_plane"])
def test_fp32_form_close_to_fp64_oracle(name):
    """Precision form: FP32 state against the FP64 oracle over a whole run; depth within
    1e-4 relative L2.  The water and sediment budget residuals are the random walk of
    the FP32 roundings, which the unbiased rounding of small increments keeps free of
    drift: 1e-8 to 4e-8 on these grids of a few hundred cells (bound 1e-6)."""
    case = _case(name)
    o, s = _pair(case, 1 << 40, precision=_abi.FP32)
    errs = compare_states(s.state(), o.state(), case.nclass)
    print({k: "%.2e" % v for k, v in errs.items()})
    assert errs["depth"] <= 1e-4
    b = s.budget()
    wscale = max(abs(b[n]) for n in (0, 1, 2, 3, 9, 10))
    sscale = max(max(abs(b[n]) for n in (4, 5, 7, 8)), 1e-30)
    print(sv.water_residual(b) / wscale, sv.sediment_residual(b) / sscale)
    assert abs(sv.water_residual(b)) <= 1e-6 * wscale
    assert abs(sv.sediment_residual(b)) <= 1e-6 * sscale
    s.close()


def test_fp32_mass_balance_large_grid():
    """FP32 form on a 512 x 512 fractal hillslope (7 classes, rain, per-cell Ks):
    relative water and sediment budget residuals below 1e-6 after 600 steps."""
    from swehr_amd.inputs import config_c2

    case = config_c2(seed=5, n=512, tend=3600.0)
    with sv.Solver(case, precision=_abi.FP32) as s:
        s.run(600)
        b = s.budget()
    wscale = max(abs(b[n]) for n in (0, 1, 2, 3, 9, 10))
    sscale = max(max(abs(b[n]) for n in (4, 5, 7, 8)), 1e-30)
    print(b, sv.water_residual(b) / wscale, sv.sediment_residual(b) / sscale)
    assert abs(sv.water_residual(b)) <= 1e-6 * wscale
    assert abs(sv.sediment_residual(b)) <= 1e-6 * sscale


def test_fp32_budgets_whole_rain_event():
    """FP32 form over a whole rain event (256 x 256 channel network, 10 classes down to
    2 um, triangular hyetograph to 1800 s, ~11700 steps): relative water and sediment
    budget residuals below 1e-6.  The exchange with the bed changes the depth by
    increments far below the float spacing at the depth, and the dry clamp dumps
    suspended mass far below the spacing at the deposited mass; rounded to nearest,
    those increments are lost every step with one sign (water 7.3e-6, sediment 1.4e-6
    here), which the unbiased rounding of the FP32 kernel prevents."""
    from swehr_amd.inputs import config_c3

    case = config_c3(seed=13, n=256, tend=1800.0)
    with sv.Solver(case, precision=_abi.FP32) as s:
        s.run()
        info = s.info()
        b = s.budget()
    assert info["done"] == 1
    wscale = max(abs(b[n]) for n in (0, 1, 2, 3, 9, 10))
    sscale = max(max(abs(b[n]) for n in (4, 5, 7, 8)), 1e-30)
    print(info["steps"], sv.water_residual(b) / wscale, sv.sediment_residual(b) / sscale)
    assert abs(sv.water_residual(b)) <= 1e-6 * wscale
    assert abs(sv.sediment_residual(b)) <= 1e-6 * sscale


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(_abi.FP64, 1e-12), (_abi.FP32, 1e-4)])
def test_outlet_discharge_matches_oracle(precision, tol):
    """The outlet hydrograph and sedigraph (water and sediment through the j = 1/2 face at
    each stage record) against the oracle's, and summed over a run with a record every
    step against the GPU's own boundary-outflow budget."""
    from parity_util import outlet_case

    case = outlet_case(tend=400.0)
    print("outlet_plane", case.nx, case.ny, "precision", precision, flush=True)
    o = Oracle(case, threads=4)
    o.run()
    oq, oqs = o.discharge()
    o.close()
    with sv.Solver(case, precision=precision) as s:
        s.run()
        q, qs = s.discharge()
        b = s.budget()
        steps = s.info()["steps"]
    assert len(q) == len(oq) == steps == 400
    print("hydrograph %.3e sedigraph %.3e" % (rel_l2(q, oq), rel_l2(qs, oqs)))
    assert rel_l2(q, oq) < tol
    assert rel_l2(qs, oqs) < tol
    assert abs(q.sum() - b[3]) <= 1e-12 * abs(b[3]) * (1 if precision == _abi.FP64 else 1e4)
    assert abs(qs.sum() - b[7]) <= 1e-12 * abs(b[7]) * (1 if precision == _abi.FP64 else 1e4)


@pytest.mark.parametrize("S,nx,ny", [(10, 37, 53), (16, 21, 70), (9, 8, 33)])
def test_fp64_many_classes_match_oracle(S, nx, ny):
    """More than 7 classes and the maximum of 16, on ragged grids.  The FP64 kernel keeps
    the first 7 classes of a cell in registers and LDS planes of their own; the rest are
    read from memory and staged in a second chunk over LDS planes that are dead by then."""
    from parity_util import many_classes_case

    case = many_classes_case(S, nx, ny, seed=S)
    print("case", case.name, flush=True)
    o, s = _pair(case, 400)
    errs = compare_states(s.state(), o.state(), case.nclass)
    k, e = worst(errs)
    print("worst %s %.3e" % (k, e))
    if e > RTOL_FP64_SHORT:
        # the finest classes' deposited mass: its step increments are differences of
        # nearly equal numbers (measured 6e-12..1.2e-11, the FMA-compiled oracle is
        # 6.9e-11 from the oracle on the 10-class case); per field within 4x the floor
        from parity_util import fp64_check

        bad, fl = fp64_check(case, 400, o.state(), errs, RTOL_FP64_SHORT, 4.0)
        print("rounding floor on %s %.3e" % (k, fl[k]))
        assert not bad, bad
    st_g, _ = s.stage()
    st_o, _ = o.stage()
    assert rel_l2(st_g, st_o) <= RTOL_FP64_SHORT
    q, qs = s.discharge()
    oq, oqs = o.discharge()
    assert rel_l2(q, oq) <= RTOL_FP64_SHORT and rel_l2(qs, oqs) <= RTOL_FP64_SHORT
    s.close()
    o.close()


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_every_tile_shape_gives_the_same_bits(precision):
    """Every tile shape the step kernel is built for computes each cell from the same
    inputs with the same code: after 150 steps of a ragged 7-class case every field,
    the stage series and the outlet discharge are bit-identical to the first shape's,
    which is also checked against the oracle."""
    from parity_util import many_classes_case

    case = many_classes_case(7, 45, 70, seed=7)
    steps = 150
    ref = None
    ran = []
    for tid in range(1, len(sv.tile_shapes()) + 1):
        print("tile id", tid, flush=True)
        try:
            s = sv.Solver(case, precision=precision, tile=tid)
        except sv.SwehrError as e:
            assert e.code == _abi.ERR_BAD_INPUT, str(e)
            continue
        s.run(steps)
        got = (s.state(), s.stage(), s.discharge(), s.info())
        s.close()
        ran.append(tid)
        if ref is None:
            ref = got
            o = Oracle(case)
            o.run(steps)
            errs = compare_states(got[0], o.state(), case.nclass)
            o.close()
            k, e = worst(errs)
            assert e <= (RTOL_FP64_SHORT if precision == _abi.FP64 else 1e-4), (k, e)
            continue
        assert got[3]["steps"] == ref[3]["steps"] and got[3]["t"] == ref[3]["t"]
        for f in ref[0]:
            np.testing.assert_array_equal(got[0][f], ref[0][f], err_msg="tile %d %s" % (tid, f))
        for a, b in zip(got[1] + got[2], ref[1] + ref[2]):
            np.testing.assert_array_equal(a, b, err_msg="tile %d series" % tid)
    print("tile ids run:", ran)
    assert ran == [1, 2, 3, 4]


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_step_commit_in_next_wave_speed_pass_equals_commit_per_step(precision):
    """In a step graph the wave-speed pass of each step commits the previous step's end
    (time, stage records, print pause, stop) and the graph ends with the control kernel;
    a graph of one step commits every step in its control kernel.  Runs cut at step
    counts that end inside a graph, with print pauses, must give the same bits both ways:
    every field, the stage series, the outlet discharge, the step count and time."""
    from parity_util import many_classes_case

    case = many_classes_case(5, 40, 66, seed=11)
    cuts = [1, 15, 16, 17, 31, 5, 40, 25]
    got = []
    for spg in (16, 1):
        s = sv.Solver(case, precision=precision, steps_per_graph=spg, pause_on_print=True)
        pauses = 0
        for n in cuts:
            left = n
            while left > 0:
                before = s.info()["steps"]
                info = s.advance(left)
                left -= info["steps"] - before
                if info["print_due"]:
                    pauses += 1
                    s.ack_print()
                if info["done"]:
                    break
        st = s.state()
        got.append((st, s.stage(), s.discharge(), s.info(), pauses))
        s.close()
    (a, sa, qa, ia, pa), (b, sb, qb, ib, pb) = got
    print("steps", ia["steps"], "t", ia["t"], "print pauses", pa, pb)
    assert ia["steps"] == sum(cuts) and ia["steps"] == ib["steps"] and ia["t"] == ib["t"] and pa == pb > 0
    for f in a:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    for x, y in zip(sa + qa, sb + qb):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("precision,tol", [(_abi.FP64, RTOL_FP64_SHORT), (_abi.FP32, 1e-4)], ids=["fp64", "fp32"])
def test_channel_network_case_matches_oracle(precision, tol):
    """The inputs of the 4096^2 precision-form case at 72 x 56 (ragged against the tiles):
    fractal surface with carved channels, a per-cell Ks raster, 10 classes with
    Dirichlet fractions and a rain table of 10 s entries (here falling from the
    hyetograph's peak, so that the surface runs off within the steps compared): FP64 to
    rounding level, FP32 within 1e-4 on every field, after 300 steps.  Some 40 steps
    later a receding cell meets the film-depth threshold at a last-bit difference: from
    there the oracle itself, with Ks perturbed by 1e-15 relative, moves by 8e-4 rel L2
    at 350 steps (the GPU by 6e-4), so the comparison stops before."""
    from swehr_amd.inputs import carve_channels, config_c3, fractal_topography

    case = config_c3(seed=21, n=64)
    rng = np.random.default_rng(21)
    case.nx, case.ny = 72, 56
    case.topo = carve_channels(fractal_topography(72, 56, 1.0, 20.0, 0.05, 0.8, rng), 1.0, rng)
    case.ks_raster = 2.7778e-6 * np.exp(0.5 * rng.standard_normal((72, 56)))
    case.rain = float(case.rain.max()) * (1.0 - np.arange(len(case.rain)) / len(case.rain))
    print("channel case", case.nx, case.ny, "precision", precision, flush=True)
    o, s = _pair(case, 300, precision=precision)
    assert float(o.state()["depth"].mean()) > 5e-5  # the surface is wet (film depth 1e-5) and running off
    errs = compare_states(s.state(), o.state(), case.nclass)
    print({k: "%.2e" % v for k, v in errs.items()})
    k, e = worst(errs)
    assert e <= tol, (k, e)
    st_g, _ = s.stage()
    st_o, _ = o.stage()
    assert rel_l2(st_g, st_o) <= tol
    s.close()
    o.close()


def test_fp32_sixteen_classes_close_to_oracle():
    """FP32 form at the maximum of 16 classes (all of them in LDS planes of the tile)."""
    from parity_util import many_classes_case

    case = many_classes_case(16, 37, 53, seed=7)
    o, s = _pair(case, 400, precision=_abi.FP32)
    errs = compare_states(s.state(), o.state(), case.nclass)
    print({k: "%.2e" % v for k, v in errs.items()})
    assert errs["depth"] <= 1e-4
    k, e = worst(errs)
    assert e <= 1e-3, (k, e)
    s.close()
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(_abi.FP64, RTOL_FP64_SHORT), (_abi.FP32, 1e-3)])
def test_parameter_rasters_match_oracle(precision, tol):
    """Per-cell suction head, theta_s - theta_0 and Manning n0 rasters (with a Ks
    raster-free scalar Ks): the GPU state against the oracle's with the same rasters."""
    from test_oracle_invariants import _raster_case

    case = _raster_case()
    print("raster case", case.nx, case.ny, "precision", precision, flush=True)
    o = Oracle(case, threads=4)
    o.run(300)
    with sv.Solver(case, precision=precision) as s:
        s.run(300)
        errs = compare_states(s.state(), o.state(), case.nclass)
    o.close()
    k, e = worst(errs)
    print("worst %s %.3e" % (k, e))
    assert e <= tol, (k, e)


def _dump_state(d, nclass):
    """The reference's full-precision dump (tests/golden/<case>/dump.npz) as a state dict."""
    st = {k: d[k] for k in ("depth", "uh", "vh", "topo", "vinf", "totm", "dep", "totsed", "maxvel", "maxdepth",
                            "velx", "vely")}
    st["ch"] = np.stack([d["ch%d" % (k + 1)] for k in range(nclass)])
    st["m"] = np.stack([d["m%d" % (k + 1)] for k in range(nclass)])
    return st


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_parameter_raster_golden_matches_reference(precision):
    """Non-uniform Ks, suction head, theta_0 / theta_s and Manning n0 rasters on a 7 x 52
    grid, whole run (600 steps), against the final state the reference itself wrote
    (SWEHR.c built with the raster hooks; the rasters are read per cell where
    SWEHR.c:237-242 sets them).  FP64: every field within RTOL_FP64_SHORT of the
    reference; FP32: depth within 1e-4."""
    name = "param_rasters"
    case = _case(name)
    with open(os.path.join(GOLD, name, "expected.json")) as fh:
        exp = json.load(fh)
    ref = _dump_state(np.load(os.path.join(GOLD, name, "dump.npz")), case.nclass)
    print("case", name, "precision", precision, flush=True)
    with sv.Solver(case, precision=precision) as s:
        s.run()
        info = s.info()
        errs = compare_states(s.state(), ref, case.nclass)
    print({k: "%.2e" % v for k, v in errs.items()})
    assert info["done"] == 1
    if precision == _abi.FP64:
        assert info["steps"] == exp["steps"]
        assert info["t"] == pytest.approx(exp["t"], rel=1e-12, abs=1e-12)
        k, e = worst(errs)
        assert e <= RTOL_FP64_SHORT, (k, e)
    else:
        assert abs(info["steps"] - exp["steps"]) <= max(2, exp["steps"] // 100)
        assert errs["depth"] <= 1e-4, errs["depth"]
