"""The VALU issue-rate microbenchmark that checks the 32-bit vector peak the
VALU-bound kernels are priced against: at 8 and at 1 waves per SIMD every
instruction kind returns a sustained cycles-per-wave64-instruction figure
between the 2-cycle issue limit (a little slack for timing) and a loose upper
bound, simple ops are not slower than 32-bit multiplies, and out-of-range
arguments are rejected before anything is launched."""
import ctypes
import sys

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wps", [8, 1])
def test_valu_issue_rates(wps):
    from pqc_amd import _lib as L
    cyc = {}
    for op in range(7):
        c = ctypes.c_double()
        assert L.lib().pqc_valu_issue_rate(op, wps, ctypes.byref(c), None) == 0
        cyc[op] = c.value
    print("waves per SIMD", wps, "cycles per wave64 instruction:", cyc, flush=True)
    sys.stdout.flush()
    for op, c in cyc.items():
        assert 1.8 < c < 32.0, (op, c)
    assert min(cyc[0], cyc[1]) <= cyc[3] * 1.05


def test_valu_issue_rate_rejects_bad_arguments():
    from pqc_amd import _lib as L
    c = ctypes.c_double()
    assert L.lib().pqc_valu_issue_rate(7, 8, ctypes.byref(c), None) != 0
    assert L.lib().pqc_valu_issue_rate(0, 0, ctypes.byref(c), None) != 0
    assert L.lib().pqc_valu_issue_rate(0, 9, ctypes.byref(c), None) != 0
    assert L.lib().pqc_valu_issue_rate(0, 8, None, None) != 0
