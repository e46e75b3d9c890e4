"""GPU reproduction of the known-answer files in tests/golden (bit-exact), through
the batched path and through the NIST harness's own call sequence
(randombytes_init, crypto_kem_keypair, crypto_kem_enc, crypto_kem_dec).

Every test prints the case it runs before touching the GPU."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rsp(alg):
    return os.path.join(HERE, "golden", "kat_%s.rsp" % alg)


@pytest.mark.parametrize("alg", ["kyber512", "kyber768", "kyber1024", "frodo976", "newhope1024", "newhope512"])
def test_gpu_reproduces_kat_file(alg):
    from pqc_amd import kat
    say("kat batch", alg)
    assert kat.check_rsp(rsp(alg), alg) == []


@pytest.mark.parametrize("alg", ["kyber768", "kyber1024", "frodo976", "newhope1024", "newhope512"])
def test_nist_harness_sequence(alg):
    """randombytes_init(seed) then the single-instance drop-ins give the file's entry."""
    import pqc_amd
    from pqc_amd import kat
    say("kat harness sequence", alg)
    _, ents = kat.read_rsp(rsp(alg))
    e = ents[-1]
    try:
        kat.drbg_init(e["seed"])
        pk, sk = pqc_amd.crypto_kem_keypair(alg)
        ct, ss = pqc_amd.crypto_kem_enc(pk, alg)
        ss1 = pqc_amd.crypto_kem_dec(ct, sk, alg)
    finally:
        kat.drbg_release()
    assert (pk, sk, ct, ss, ss1) == (e["pk"], e["sk"], e["ct"], e["ss"], e["ss"])
