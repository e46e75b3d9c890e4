"""GPU parity of the NewHope CPA batched kernels (N = 1024 and 512) and the
NewHope CCA-KEM against the CPU oracle, bit-exact, through both the fused
per-element kernels (the default) and the per-operation kernels (PQC_F_PLAIN)."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


PATHS = ["fused", "plain"]


@pytest.fixture(scope="module")
def ctx():
    import pqc_amd
    c = pqc_amd.Context(["newhope1024cpa"], max_batch=16384)
    yield c
    c.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n", [1, 63, 130])
def test_triples_vs_oracle(ctx, n, path):
    say("newhope triples n", n, path)
    seed = O.drbg_array(n, 32, 0, 700 + n)
    msg = O.drbg_array(n, 32, 0, 701 + n)
    coin = O.drbg_array(n, 32, 0, 702 + n)
    pk, sk, ct, m2 = ctx.newhope_triples_host(seed, msg, coin, plain=path == "plain")
    opk, osk, oct_, om2 = O.newhope_triples(seed, msg, coin)
    assert (pk == opk).all()
    assert (sk == osk).all()
    assert (ct == oct_).all()
    assert (m2 == om2).all() and (m2 == msg).all()


@pytest.mark.parametrize("path", PATHS)
def test_decrypt_garbage_ciphertexts_vs_oracle(ctx, path):
    say("newhope garbage ct", path)
    n = 64
    seed = O.drbg_array(n, 32, 0, 800)
    rng = np.random.default_rng(3)
    pk, sk, _, _ = O.newhope_triples(seed, O.drbg_array(n, 32, 0, 801), O.drbg_array(n, 32, 0, 802))
    ct = rng.integers(0, 256, (n, 2176), dtype=np.uint8)
    m = np.empty((n, 32), np.uint8)
    ctx.newhope_dec(n, m, ct, sk, host=True, plain=path == "plain")
    for i in range(n):
        assert m[i].tobytes() == O.newhope_dec(ct[i].tobytes(), sk[i].tobytes())


def test_full_size_16384_roundtrip_and_sample(ctx):
    import pqc_amd
    n = 16384
    say("newhope full size", n)
    rng = np.random.default_rng(11)
    seed = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    msg = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    coin = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    d_seed, d_msg, d_coin = (pqc_amd.DeviceBuffer.from_array(x) for x in (seed, msg, coin))
    d_pk, d_sk, d_ct, d_m2 = (pqc_amd.DeviceBuffer(n * b) for b in (1824, 1792, 2176, 32))
    ctx.newhope_keypair(n, d_pk, d_sk, d_seed)
    ctx.newhope_enc(n, d_ct, d_msg, d_pk, d_coin)
    ctx.newhope_dec(n, d_m2, d_ct, d_sk)
    m2 = d_m2.download((n, 32))
    assert (m2 == msg).all()
    idx = rng.choice(n, 24, replace=False)
    ct = d_ct.download((n, 2176))
    pk = d_pk.download((n, 1824))
    opk, osk, oct_, _ = O.newhope_triples(np.ascontiguousarray(seed[idx]), np.ascontiguousarray(msg[idx]),
                                         np.ascontiguousarray(coin[idx]))
    assert (pk[idx] == opk).all() and (ct[idx] == oct_).all()


@pytest.mark.parametrize("op", [0, 1])
def test_poly_transform_vs_oracle(op):
    """The wave-per-polynomial NTT / inverse NTT on their own, element by element,
    on in-range coefficients and on arbitrary uint16 values (the reference's
    16-bit wraparound of the lazy layers must be reproduced exactly)."""
    import pqc_amd
    say("newhope poly transform op", op)
    rng = np.random.default_rng(40 + op)
    polys = np.concatenate([rng.integers(0, 12289, (150, 1024)), rng.integers(0, 65536, (50, 1024))]).astype(np.uint16)
    got = pqc_amd.newhope_poly_transform(op, polys)
    ref = O.newhope_ntt if op == 0 else O.newhope_invntt
    for i in range(len(polys)):
        assert (got[i] == ref(polys[i])).all(), i


def test_poly_sample_vs_oracle():
    import pqc_amd
    say("newhope poly_sample")
    seeds = O.drbg_array(70, 32, 0, 810)
    for nonce in (0, 2, 255):
        got = pqc_amd.newhope_poly_sample(seeds, nonce)
        for i in range(len(seeds)):
            assert (got[i] == O.newhope_sample(seeds[i].tobytes(), nonce)).all(), (nonce, i)


# ---- NewHope-1024 CCA-KEM over the CPA kernels ---------------------------------------
@pytest.fixture(scope="module")
def cctx():
    import pqc_amd
    c = pqc_amd.Context(["newhope1024"], max_batch=16384)
    yield c
    c.close()


def _oracle_cca(kc, ec):
    outs = [O.newhope_cca_keypair(kc[i].tobytes()) for i in range(len(kc))]
    encs = [O.newhope_cca_enc(pk, ec[i].tobytes()) for i, (pk, _) in enumerate(outs)]
    return outs, encs


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n", [1, 63, 65, 200])
def test_cca_vs_oracle(cctx, n, path):
    say("newhope cca n", n, path)
    plain = path == "plain"
    kc = O.drbg_array(n, 64, 0, 900 + n)
    ec = O.drbg_array(n, 32, 0, 901 + n)
    pk, sk = cctx.keypair_host("newhope1024", kc, plain=plain)
    ct, ss = cctx.enc_host("newhope1024", pk, ec, plain=plain)
    ss2 = cctx.dec_host("newhope1024", ct, sk, plain=plain)
    keys, encs = _oracle_cca(kc, ec)
    for i in range(n):
        assert (pk[i].tobytes(), sk[i].tobytes()) == keys[i]
        assert (ct[i].tobytes(), ss[i].tobytes()) == encs[i]
    assert (ss2 == ss).all()


@pytest.mark.parametrize("path", PATHS)
def test_cca_implicit_rejection_vs_oracle(cctx, path):
    """Tampered bytes in u, v and d, and random-byte ciphertexts: the GPU's shared
    secrets equal the oracle's (H(s || H(ct')))."""
    say("newhope cca rejection", path)
    n = 96
    kc = O.drbg_array(n, 64, 0, 950)
    ec = O.drbg_array(n, 32, 0, 951)
    pk, sk = cctx.keypair_host("newhope1024", kc)
    ct, ss = cctx.enc_host("newhope1024", pk, ec)
    bad = ct.copy()
    pos = np.array([0, 1791, 1792, 2175, 2176, 2207])
    for i in range(n):
        bad[i, pos[i % len(pos)]] ^= 0x10 if i % 3 else 0x01
    bad[-8:] = np.random.default_rng(5).integers(0, 256, (8, 2208), dtype=np.uint8)
    got = cctx.dec_host("newhope1024", bad, sk, plain=path == "plain")
    for i in range(n):
        assert got[i].tobytes() == O.newhope_cca_dec(bad[i].tobytes(), sk[i].tobytes())
    assert (got != ss).any(axis=1).all()


def test_cca_full_size_16384_roundtrip_and_sample(cctx):
    import pqc_amd
    n = 16384
    say("newhope cca full size", n)
    rng = np.random.default_rng(12)
    kc = rng.integers(0, 256, (n, 64), dtype=np.uint8)
    ec = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    d_kc, d_ec = pqc_amd.DeviceBuffer.from_array(kc), pqc_amd.DeviceBuffer.from_array(ec)
    pk, sk = pqc_amd.DeviceBuffer(n * 1824), pqc_amd.DeviceBuffer(n * 3680)
    ct, ss, ss2 = pqc_amd.DeviceBuffer(n * 2208), pqc_amd.DeviceBuffer(n * 32), pqc_amd.DeviceBuffer(n * 32)
    cctx.keypair_derand("newhope1024", n, pk, sk, d_kc)
    cctx.enc_derand("newhope1024", n, ct, ss, pk, d_ec)
    h_ct = ct.download((n, 2208))
    h_ct[::7, 2200] ^= 0x80  # every 7th ciphertext rejected
    ct.upload(h_ct)
    cctx.dec("newhope1024", n, ss2, ct, sk)
    a, e = ss2.download((n, 32)), ss.download((n, 32))
    good = np.ones(n, bool)
    good[::7] = False
    assert (a[good] == e[good]).all() and (a[~good] != e[~good]).any(axis=1).all()
    h_pk, h_sk = pk.download((n, 1824)), sk.download((n, 3680))
    for i in rng.choice(n, 12, replace=False):
        assert (h_pk[i].tobytes(), h_sk[i].tobytes()) == O.newhope_cca_keypair(kc[i].tobytes())
        assert a[i].tobytes() == O.newhope_cca_dec(h_ct[i].tobytes(), h_sk[i].tobytes())


# ---- NewHope-512: CPA kernels and CCA-KEM -------------------------------------------
@pytest.fixture(scope="module")
def ctx512():
    import pqc_amd
    c = pqc_amd.Context(["newhope512cpa", "newhope512"], max_batch=16384)
    yield c
    c.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n", [1, 63, 130])
def test_newhope512_triples_vs_oracle(ctx512, n, path):
    say("newhope512 triples n", n, path)
    seed = O.drbg_array(n, 32, 0, 1700 + n)
    msg = O.drbg_array(n, 32, 0, 1701 + n)
    coin = O.drbg_array(n, 32, 0, 1702 + n)
    pk, sk, ct, m2 = ctx512.newhope_triples_host(seed, msg, coin, N=512, plain=path == "plain")
    opk, osk, oct_, om2 = O.newhope_triples(seed, msg, coin, N=512)
    assert (pk == opk).all()
    assert (sk == osk).all()
    assert (ct == oct_).all()
    assert (m2 == om2).all() and (m2 == msg).all()


@pytest.mark.parametrize("path", PATHS)
def test_newhope512_decrypt_garbage_vs_oracle(ctx512, path):
    say("newhope512 garbage ct", path)
    n = 64
    seed = O.drbg_array(n, 32, 0, 1800)
    pk, sk, _, _ = O.newhope_triples(seed, O.drbg_array(n, 32, 0, 1801), O.drbg_array(n, 32, 0, 1802), N=512)
    ct = np.random.default_rng(4).integers(0, 256, (n, 1088), dtype=np.uint8)
    m = np.empty((n, 32), np.uint8)
    ctx512.newhope_dec(n, m, ct, sk, host=True, N=512, plain=path == "plain")
    for i in range(n):
        assert m[i].tobytes() == O.newhope_dec(ct[i].tobytes(), sk[i].tobytes(), 512)


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("n", [1, 65, 200])
def test_newhope512_cca_vs_oracle(ctx512, n, path):
    say("newhope512 cca n", n, path)
    plain = path == "plain"
    kc = O.drbg_array(n, 64, 0, 1900 + n)
    ec = O.drbg_array(n, 32, 0, 1901 + n)
    pk, sk = ctx512.keypair_host("newhope512", kc, plain=plain)
    ct, ss = ctx512.enc_host("newhope512", pk, ec, plain=plain)
    bad = ct.copy()
    bad[::3, 7] ^= 0x20
    got = ctx512.dec_host("newhope512", bad, sk, plain=plain)
    for i in range(n):
        okp, osk = O.newhope_cca_keypair(kc[i].tobytes(), 512)
        assert (pk[i].tobytes(), sk[i].tobytes()) == (okp, osk)
        assert (ct[i].tobytes(), ss[i].tobytes()) == O.newhope_cca_enc(okp, ec[i].tobytes(), 512)
        assert got[i].tobytes() == O.newhope_cca_dec(bad[i].tobytes(), osk, 512)
    assert (got[1::3] == ss[1::3]).all() and (got[::3] != ss[::3]).any(axis=1).all()


def test_newhope512_full_size_16384(ctx512):
    import pqc_amd
    n = 16384
    say("newhope512 full size", n)
    rng = np.random.default_rng(13)
    seed, msg, coin = (rng.integers(0, 256, (n, 32), dtype=np.uint8) for _ in range(3))
    d_seed, d_msg, d_coin = (pqc_amd.DeviceBuffer.from_array(x) for x in (seed, msg, coin))
    pk, sk, ct, m2 = (pqc_amd.DeviceBuffer(n * b) for b in (928, 896, 1088, 32))
    ctx512.newhope_keypair(n, pk, sk, d_seed, N=512)
    ctx512.newhope_enc(n, ct, d_msg, pk, d_coin, N=512)
    ctx512.newhope_dec(n, m2, ct, sk, N=512)
    assert (m2.download((n, 32)) == msg).all()
    h_pk, h_ct = pk.download((n, 928)), ct.download((n, 1088))
    for i in rng.choice(n, 12, replace=False):
        opk, _ = O.newhope_keypair(seed[i].tobytes(), 512)
        assert h_pk[i].tobytes() == opk
        assert h_ct[i].tobytes() == O.newhope_enc(msg[i].tobytes(), opk, coin[i].tobytes(), 512)


@pytest.mark.parametrize("N", [1024, 512])
def test_fused_equals_plain_full_size_16384(ctx, ctx512, N):
    """Every element of a 16384-element batch: the fused kernels' pk, sk, ct and
    decrypted messages equal the per-operation kernels' (garbage ciphertexts in
    every 5th slot for the decryption)."""
    import pqc_amd
    n = 16384
    say("newhope fused == plain", N, n)
    c = ctx if N == 1024 else ctx512
    P = 14 * N // 8
    rng = np.random.default_rng(20 + N)
    seed, msg, coin = (rng.integers(0, 256, (n, 32), dtype=np.uint8) for _ in range(3))
    d_seed, d_msg, d_coin = (pqc_amd.DeviceBuffer.from_array(x) for x in (seed, msg, coin))
    out = {}
    for plain in (False, True):
        pk, sk, ct, m2 = (pqc_amd.DeviceBuffer(n * b) for b in (P + 32, P, P + 3 * N // 8, 32))
        c.newhope_keypair(n, pk, sk, d_seed, N=N, plain=plain)
        c.newhope_enc(n, ct, d_msg, pk, d_coin, N=N, plain=plain)
        h_ct = ct.download((n, P + 3 * N // 8))
        h_ct[::5] = rng.integers(0, 256, h_ct[::5].shape, dtype=np.uint8) if not plain else out[False][2][::5]
        ct.upload(h_ct)
        c.newhope_dec(n, m2, ct, sk, N=N, plain=plain)
        out[plain] = (pk.download((n, P + 32)), sk.download((n, P)), h_ct, m2.download((n, 32)))
    for a, b in zip(out[False], out[True]):
        assert (a == b).all()
    good = np.ones(n, bool)
    good[::5] = False
    assert (out[False][3][good] == msg[good]).all()
