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
    for op in range(16):
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
    assert L.lib().pqc_valu_issue_rate(16, 8, ctypes.byref(c), None) != 0
    assert L.lib().pqc_valu_issue_rate(0, 0, ctypes.byref(c), None) != 0
    assert L.lib().pqc_valu_issue_rate(0, 9, ctypes.byref(c), None) != 0
    assert L.lib().pqc_valu_issue_rate(0, 8, None, None) != 0


def test_keccak_rate_forms():
    """Keccak-f throughput of the three state layouts: every form gives a
    positive rate, a second wave per SIMD never makes the single-lane form
    slower, and bad arguments are rejected before a launch."""
    from pqc_amd import _lib as L
    rates = {}
    for form in range(3):
        for wps in (1, 2):
            r = ctypes.c_double()
            assert L.lib().pqc_keccak_rate(form, wps, ctypes.byref(r), None) == 0
            rates[(form, wps)] = r.value
    print("Keccak state permutations per second:", rates, flush=True)
    sys.stdout.flush()
    assert all(v > 1e8 for v in rates.values())
    assert rates[(0, 2)] >= 0.95 * rates[(0, 1)]
    r = ctypes.c_double()
    assert L.lib().pqc_keccak_rate(3, 1, ctypes.byref(r), None) != 0
    assert L.lib().pqc_keccak_rate(0, 9, ctypes.byref(r), None) != 0
    assert L.lib().pqc_keccak_rate(0, 1, None, None) != 0

