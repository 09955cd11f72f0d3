"""swehr_budget_noalloc (include/swehr.h) against swehr_budget, through the C ABI.

The allocation-free entry point keeps the budget kernel's partial sums in the library's
fixed device array and reduces them on the device in the order swehr_budget adds them
on the host, so both return the same bits.  Two calls in a row, the second on a smaller
instance (fewer partial blocks), check that nothing of the first call's partials is
picked up by the second.
"""
import ctypes as C

import numpy as np
import pytest

from swehr_amd import _abi
from swehr_amd import solver as sv
from swehr_amd._abi import dptr
from swehr_amd.inputs import config_c2

pytestmark = pytest.mark.gpu


def _both(s):
    L = sv.lib()
    a = np.zeros(_abi.BUDGET_N)
    b = np.full(_abi.BUDGET_N, np.nan)
    sv.check("swehr_budget", L.swehr_budget(s._h, dptr(a)))
    sv.check("swehr_budget_noalloc", L.swehr_budget_noalloc(s._h, dptr(b)))
    return a, b


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_budget_noalloc_equals_budget_large_then_small(precision):
    big = config_c2(seed=11, n=600, tend=600.0)  # 360000 cells: 1024 partial blocks (the maximum)
    small = config_c2(seed=12, n=40, tend=600.0)  # 1600 cells: 7 blocks
    print("noalloc budget: 600^2 then 40^2, precision", precision, flush=True)
    with sv.Solver(big, precision=precision) as s1, sv.Solver(small, precision=precision) as s2:
        s1.run(300)
        s2.run(300)
        a1, b1 = _both(s1)
        a2, b2 = _both(s2)
        assert a1[11] == 300 and a2[11] == 300
        assert np.all(np.isfinite(b1)) and np.all(np.isfinite(b2))
        np.testing.assert_array_equal(b1, a1)
        np.testing.assert_array_equal(b2, a2)
        # the small instance's totals differ from the large one's: no stale partials
        assert abs(a2[0] - a1[0]) > 0
        # and once more on the large one after the small one
        s1.run(10)
        a3, b3 = _both(s1)
        np.testing.assert_array_equal(b3, a3)
