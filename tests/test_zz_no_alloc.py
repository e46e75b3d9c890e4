"""The allocation-free decapsulation entry point (pqc_kyber_dec_batch_noalloc),
called through the C ABI, against crypto_kem_dec_batch on the same inputs.

Two calls in a row on each parameter set, the second with a smaller batch, so
that the second runs on the scratch array the first left behind: both must
equal the context path bit for bit.  The batch sizes cover the tile kernel
(which uses the scratch array) and the per-request kernel.  The ciphertexts
include tampered ones (implicit rejection), and the sizes above the maximum
are rejected without a launch.

Every test prints the case it runs before touching the GPU."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SIZES = {"kyber512": (800, 1632, 736), "kyber768": (1184, 2400, 1088), "kyber1024": (1568, 3168, 1568)}
ALG_ID = {"kyber512": 1, "kyber768": 2, "kyber1024": 3}
MAX = 16384


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(pqc, ctx, alg, n, seed, stream):
    pkb, skb, ctb = SIZES[alg]
    rng = np.random.default_rng(seed)
    coins = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 64), dtype=np.uint8))
    ecoins = pqc.DeviceBuffer.from_array(rng.integers(0, 256, (n, 32), dtype=np.uint8))
    pk, sk, ct, ss = pqc.DeviceBuffer(n * pkb), pqc.DeviceBuffer(n * skb), pqc.DeviceBuffer(n * ctb), \
        pqc.DeviceBuffer(n * 32)
    ctx.keypair_derand(alg, n, pk, sk, coins, stream=stream)
    ctx.enc_derand(alg, n, ct, ss, pk, ecoins, stream=stream)
    stream.synchronize()
    # tamper with every 7th ciphertext (implicit rejection path)
    cth = ct.download((n, ctb))
    cth[::7, rng.integers(0, ctb)] ^= 1
    ct.upload(cth)
    return sk, ct, ss


@pytest.mark.parametrize("alg", ["kyber768", "kyber1024", "kyber512"])
def test_noalloc_equals_context_path(alg):
    import pqc_amd as pqc
    from pqc_amd import _lib as L
    stream = pqc.Stream()
    ctx = pqc.Context([alg], max_batch=MAX)
    status = pqc.DeviceBuffer(8)
    status.upload(np.zeros(8, np.uint8))
    for n, seed in ((MAX, 1), (9000, 2), (100, 3)):
        say("noalloc", alg, "n", n)
        sk, ct, ss_enc = _inputs(pqc, ctx, alg, n, seed, stream)
        want = pqc.DeviceBuffer(n * 32)
        got = pqc.DeviceBuffer(n * 32)
        got.upload(np.full(n * 32, 0xA5, np.uint8))
        ctx.dec(alg, n, want, ct, sk, stream=stream)
        rc = L.lib().pqc_kyber_dec_batch_noalloc(ALG_ID[alg], n, got.ptr, ct.ptr, sk.ptr, status.ptr, stream.handle)
        assert rc == 0, rc
        stream.synchronize()
        w, g, e = want.download((n, 32)), got.download((n, 32)), ss_enc.download((n, 32))
        assert (g == w).all(), (alg, n)
        honest = np.ones(n, bool)
        honest[::7] = False
        assert (g[honest] == e[honest]).all() and not (g[~honest] == e[~honest]).all(axis=1).any()
    assert int(status.download((2,), np.uint32)[0]) == 0
    ctx.close()


def test_noalloc_rejects_bad_calls():
    import pqc_amd as pqc
    from pqc_amd import _lib as L
    say("noalloc argument checks")
    buf = pqc.DeviceBuffer(64)
    lib = L.lib()
    assert lib.pqc_kyber_dec_batch_noalloc(2, MAX + 1, buf.ptr, buf.ptr, buf.ptr, None, None) == L.PQC_ERR_BATCH_TOO_LARGE
    assert lib.pqc_kyber_dec_batch_noalloc(5, 1, buf.ptr, buf.ptr, buf.ptr, None, None) == L.PQC_ERR_UNSUPPORTED
    assert lib.pqc_kyber_dec_batch_noalloc(2, 1, None, buf.ptr, buf.ptr, None, None) == L.PQC_ERR_INVALID_ARG
    assert lib.pqc_kyber_dec_batch_noalloc(2, 1, buf.ptr + 4, buf.ptr, buf.ptr, None, None) == L.PQC_ERR_INVALID_ARG
    assert lib.pqc_kyber_dec_batch_noalloc(2, 0, None, None, None, None, None) == 0
