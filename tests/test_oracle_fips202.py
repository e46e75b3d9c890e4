"""Pins the oracle's FIPS 202 sponge against Python's hashlib (OpenSSL) and
the Keccak team's published zero-state permutation value."""
import hashlib
import random

import pytest

from tests import oracle_lib as O


def test_keccakf_zero_state_kat():
    # KeccakF-1600 applied to the all-zero state: first two lanes of the
    # published intermediate values (Keccak reference, "KeccakF-1600-IntermediateValues").
    st = O.keccakf([0] * 25)
    assert st[0] == 0xF1258F7940E1DDE7
    assert st[1] == 0x84D5CCF933C0478A


@pytest.mark.parametrize("n", [0, 1, 31, 32, 33, 34, 64, 71, 72, 73, 135, 136, 137, 167, 168, 169,
                               335, 336, 337, 1088, 1184, 1568, 15744])
def test_hashes_vs_hashlib(n):
    rnd = random.Random(n)
    data = bytes(rnd.getrandbits(8) for _ in range(n))
    assert O.sha3_256(data) == hashlib.sha3_256(data).digest()
    assert O.sha3_512(data) == hashlib.sha3_512(data).digest()
    for outlen in (1, 32, 128, 136, 168, 672, 1000):
        assert O.shake128(data, outlen) == hashlib.shake_128(data).digest(outlen)
        assert O.shake256(data, outlen) == hashlib.shake_256(data).digest(outlen)


def test_drbg_definition():
    import struct
    want = hashlib.shake_256(b"pqc-mi355x" + struct.pack("<QQQ", 7, 3, 11)).digest(100)
    assert O.drbg(100, seed=7, stream=3, index=11) == want
