"""Constants the HIP kernels hard-code, checked on the CPU against what they
replace: the 24-bit multiply-high division by q in compress_d (kyber_arith.cuh)
must equal integer division for every numerator compress_d can form, i.e.
(x << d) + q/2 for x in [0, q) and d <= 11."""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
Q = 3329


def _div_q_literal():
    src = open(os.path.join(HERE, "..", "pqc_amd", "csrc", "kyber_arith.cuh")).read()
    m = re.search(r'v_mul_hi_u32_u24 %0, (0x[0-9a-fA-F]+), %1', src)
    assert m, "div_q literal not found"
    return int(m.group(1), 16)


def test_div_q_is_exact_for_every_compress_numerator():
    m = _div_q_literal()
    assert m < (1 << 24)
    n = np.arange(0, ((Q - 1) << 11) + Q // 2 + 1, dtype=np.uint64)
    got = (n * np.uint64(m)) >> np.uint64(33)  # v_mul_hi_u32_u24 gives bits 32..47; the kernel shifts once more
    assert (got == n // np.uint64(Q)).all()


def test_compress_matches_definition_for_all_widths():
    m = _div_q_literal()
    x = np.arange(Q, dtype=np.uint64)
    for d in range(1, 12):
        num = (x << np.uint64(d)) + np.uint64(Q // 2)
        want = (num // np.uint64(Q)) & np.uint64((1 << d) - 1)
        got = ((num * np.uint64(m)) >> np.uint64(33)) & np.uint64((1 << d) - 1)
        assert (got == want).all(), d


def test_newhope_div_q_is_exact_below_2_21():
    """newhope.hip's div_q: the high word of n * m (m = ceil(2^32 / 12289)) is
    floor(n / 12289) for every n < 2^21; its operands stay below 2^17."""
    src = open(os.path.join(HERE, "..", "pqc_amd", "csrc", "newhope.hip")).read()
    m = int(re.search(r'v_mul_hi_u32_u24 %0, (0x[0-9a-fA-F]+), %1', src).group(1), 16)
    assert m < (1 << 24)
    n = np.arange(0, 1 << 21, dtype=np.uint64)
    assert ((n * np.uint64(m)) >> np.uint64(32) == n // np.uint64(12289)).all()
