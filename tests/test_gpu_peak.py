"""The VALU issue-rate microbenchmark that checks the 32-bit vector peak the
VALU-bound kernels are priced against: every instruction kind returns a
sustained cycles-per-wave64-instruction figure between the 2-cycle issue limit
(a little slack for timing) and a loose upper bound, and simple ops are not
slower than 32-bit multiplies."""
import ctypes
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_valu_issue_rates():
    from pqc_amd import _lib as L
    cyc = {}
    for op in range(4):
        c = ctypes.c_double()
        assert L.lib().pqc_valu_issue_rate(op, ctypes.byref(c), None) == 0
        cyc[op] = c.value
    print("cycles per wave64 instruction:", cyc, flush=True)
    sys.stdout.flush()
    for op, c in cyc.items():
        assert 1.8 < c < 32.0, (op, c)
    assert min(cyc[0], cyc[1]) <= cyc[3] * 1.05
    assert L.lib().pqc_valu_issue_rate(4, ctypes.byref(ctypes.c_double()), None) != 0
