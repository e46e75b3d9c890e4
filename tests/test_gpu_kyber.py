"""GPU parity of the Kyber kernels against the CPU oracle (bit-exact).

Every test prints the case it runs before touching the GPU."""
import hashlib
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

ALG = {2: "kyber512", 3: "kyber768", 4: "kyber1024"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pqc():
    import pqc_amd
    return pqc_amd


@pytest.fixture(scope="module")
def ctx(pqc):
    c = pqc.Context(["kyber512", "kyber768", "kyber1024"], max_batch=4096)
    yield c
    c.close()


@pytest.mark.parametrize("op,outlen", [(0, 32), (1, 64), (2, 168), (2, 700), (3, 136), (3, 300)])
@pytest.mark.parametrize("inlen", [0, 32, 64, 136, 168, 1088, 1184])
def test_keccak_batch_vs_hashlib(pqc, op, outlen, inlen):
    say("keccak op", op, "inlen", inlen, "outlen", outlen)
    rng = np.random.default_rng(op * 1000 + inlen)
    n = 70
    msgs = rng.integers(0, 256, (n, inlen), dtype=np.uint8)
    out = pqc.keccak_batch(op, msgs, outlen)
    fns = [lambda b: hashlib.sha3_256(b).digest(), lambda b: hashlib.sha3_512(b).digest(),
           lambda b: hashlib.shake_128(b).digest(outlen), lambda b: hashlib.shake_256(b).digest(outlen)]
    for i in range(n):
        assert out[i].tobytes() == fns[op](msgs[i].tobytes()), i


def test_poly_ntt_invntt_vs_oracle(pqc):
    rng = np.random.default_rng(7)
    cases = [rng.integers(-3328, 3329, (40, 256)), rng.integers(-2, 3, (40, 256)), rng.integers(0, 3329, (40, 256))]
    for ci, polys in enumerate(cases):
        say("ntt case", ci)
        polys = polys.astype(np.int16)
        got = pqc.kyber_poly_transform(0, polys)
        for i in range(len(polys)):
            assert (got[i] == O.kyber_ntt(polys[i])).all(), (ci, i)
        say("invntt case", ci)
        src = np.stack([O.kyber_ntt(p) for p in polys])
        got = pqc.kyber_poly_transform(1, src)
        for i in range(len(polys)):
            assert (got[i] == O.kyber_invntt(src[i])).all(), (ci, i)


def oracle_triples(k, n, seed_stream):
    coins = O.drbg_array(n, 64, seed=0, stream=seed_stream)
    ecoins = O.drbg_array(n, 32, seed=0, stream=seed_stream + 1)
    pk, sk = O.kyber_batch("keypair", k, [coins])
    ct, ss = O.kyber_batch("enc", k, [pk, ecoins])
    return coins, ecoins, pk, sk, ct, ss


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("plain", [True, False])
def test_kem_vs_oracle(ctx, k, plain):
    n = 130  # two full tiles plus a ragged one
    say("kem k", k, "plain", plain, "n", n)
    coins, ecoins, pk, sk, ct, ss = oracle_triples(k, n, 10 * k)
    gpk, gsk = ctx.keypair_host(ALG[k], coins, plain=plain)
    assert (gpk == pk).all() and (gsk == sk).all()
    gct, gss = ctx.enc_host(ALG[k], pk, ecoins, plain=plain)
    assert (gct == ct).all()
    assert (gss == ss).all()
    gss2 = ctx.dec_host(ALG[k], ct, sk, plain=plain)
    assert (gss2 == ss).all()


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("plain", [True, False])
def test_decaps_implicit_rejection_vs_oracle(ctx, k, plain):
    n = 96
    say("reject k", k, "plain", plain)
    coins, ecoins, pk, sk, ct, ss = oracle_triples(k, n, 100 + 10 * k)
    rng = np.random.default_rng(k)
    bad = ct.copy()
    for i in range(n):
        if i % 3 == 0:
            bad[i, rng.integers(0, ct.shape[1])] ^= 1 << rng.integers(0, 8)   # one flipped bit
        elif i % 3 == 1:
            bad[i] = rng.integers(0, 256, ct.shape[1], dtype=np.uint8)        # garbage
    want = O.kyber_batch("dec", k, [bad, sk])
    got = ctx.dec_host(ALG[k], bad, sk, plain=plain)
    assert (got == want).all()
    for i in range(0, n, 3):
        z = sk[i, -32:].tobytes()
        assert got[i].tobytes() == hashlib.shake_256(z + hashlib.sha3_256(bad[i].tobytes()).digest()).digest(32)


@pytest.mark.parametrize("path", ["lowlat", "tiles"])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 1000])
def test_ragged_batches_kyber768(ctx, n, path):
    """Partial and multi-tile batches through both kernel families: one
    workgroup per request (the default below the low-latency limits) and the
    64-request tiles (lowlat_max = -1)."""
    from pqc_amd import tuner as T
    say("ragged n", n, path)
    coins, ecoins, pk, sk, ct, ss = oracle_triples(3, n, 500 + n)
    T.put_tuning(ctx, lowlat_max=0 if path == "lowlat" else -1)
    try:
        gss = ctx.dec_host("kyber768", ct, sk)
        assert (gss == ss).all()
        gct, gss2 = ctx.enc_host("kyber768", pk, ecoins)
        assert (gct == ct).all() and (gss2 == ss).all()
    finally:
        T.put_tuning(ctx, lowlat_max=0)


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("n", [1, 63, 65])
def test_ragged_fused_keypair_vs_oracle(ctx, k, n):
    say("fused keypair k", k, "n", n)
    coins = O.drbg_array(n, 64, seed=0, stream=700 + 10 * k + n)
    pk, sk = O.kyber_batch("keypair", k, [coins])
    gpk, gsk = ctx.keypair_host(ALG[k], coins)
    assert (gpk == pk).all() and (gsk == sk).all()


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 95, 1000, 9001])
def test_keypair_tile32_equals_tile64_plain_and_oracle(pqc, k, n):
    """The fused key generation in 32-request tiles (two waves per SIMD, G on
    both lanes of a pair and the PRF outputs split between them, A through the
    per-wave scratch; the default) against the 64-request tiles, the plain
    per-op kernels and the oracle, on ragged batch sizes; the keys then
    round-trip through encapsulation and decapsulation."""
    from pqc_amd import tuner as T
    say("keypair tile32 k", k, "n", n)
    alg = ALG[k]
    ctx = pqc.Context([alg], max_batch=max(n, 64))
    try:
        kc = O.drbg_array(n, 64, 0, 7000 + 10 * k + n % 97)
        T.put_tuning(ctx, fused_key_tile=32)
        pk32, sk32 = ctx.keypair_host(alg, kc)
        T.put_tuning(ctx, fused_key_tile=64)
        pk64, sk64 = ctx.keypair_host(alg, kc)
        pkp, skp = ctx.keypair_host(alg, kc, plain=True)
        assert (pk32 == pk64).all() and (sk32 == sk64).all()
        assert (pk32 == pkp).all() and (sk32 == skp).all()
        idx = np.unique(np.array([0, n // 2, n - 1]))
        wpk, wsk = O.kyber_batch("keypair", k, [np.ascontiguousarray(kc[idx])])
        assert (pk32[idx] == wpk).all() and (sk32[idx] == wsk).all()
        ec = O.drbg_array(n, 32, 0, 7001 + 10 * k + n % 97)
        ct, ss = ctx.enc_host(alg, pk32, ec)
        assert (ctx.dec_host(alg, ct, sk32) == ss).all()
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_tile32_many_tiles_per_wave(pqc, k):
    """The 32-request tiles of key generation, encapsulation and decapsulation
    with a grid of 5 waves, so that each wave runs six or seven tiles one after
    another (the last one partial) through the same A^T scratch region, LDS
    staging rows and message slots: the bytes must equal the default grid's,
    where each wave runs one tile, and the oracle's."""
    say("tile32 many tiles per wave k", k)
    alg = ALG[k]
    n = 1000
    ctx = pqc.Context([alg], max_batch=n)
    try:
        from pqc_amd import tuner as T
        T.put_tuning(ctx, lowlat_max=-1)
        kc = O.drbg_array(n, 64, 0, 7100 + k)
        ec = O.drbg_array(n, 32, 0, 7101 + k)
        pk, sk = ctx.keypair_host(alg, kc)
        ct, ss = ctx.enc_host(alg, pk, ec)
        ct[::4, 17] ^= 0x01
        dec = ctx.dec_host(alg, ct, sk)
        ctx.set_tuning(grid=5)
        pk5, sk5 = ctx.keypair_host(alg, kc)
        ct5, ss5 = ctx.enc_host(alg, pk, ec)
        ct5[::4, 17] ^= 0x01
        dec5 = ctx.dec_host(alg, ct5, sk)
        assert (pk5 == pk).all() and (sk5 == sk).all()
        assert (ct5 == ct).all() and (ss5 == ss).all() and (dec5 == dec).all()
        honest = np.arange(n) % 4 != 0
        assert (dec[honest] == ss[honest]).all() and (dec[~honest] != ss[~honest]).any(axis=1).all()
        idx = np.array([0, 31, 32, 499, 999])
        wpk, wsk = O.kyber_batch("keypair", k, [np.ascontiguousarray(kc[idx])])
        assert (pk[idx] == wpk).all() and (sk[idx] == wsk).all()
        want = O.kyber_batch("dec", k, [np.ascontiguousarray(ct[idx]), np.ascontiguousarray(sk[idx])])
        assert (dec[idx] == want).all()
    finally:
        ctx.close()


def test_batch_of_one_equals_batch_of_n(ctx):
    say("batch-of-1 vs batch-of-n")
    coins, ecoins, pk, sk, ct, ss = oracle_triples(3, 40, 900)
    full = ctx.dec_host("kyber768", ct, sk)
    for i in range(0, 40, 7):
        one = ctx.dec_host("kyber768", ct[i:i + 1], sk[i:i + 1])
        assert (one[0] == full[i]).all()


def test_errors(pqc, ctx):
    from pqc_amd import _lib as L
    say("error paths")
    with pytest.raises(pqc.PQCError) as e:
        ctx.dec_host("kyber768", np.zeros((4097, 1088), np.uint8), np.zeros((4097, 2400), np.uint8))
    assert e.value.code == L.PQC_ERR_BATCH_TOO_LARGE
    small = pqc.Context(["kyber512"], max_batch=8)
    with pytest.raises(pqc.PQCError) as e:
        small.dec_host("kyber768", np.zeros((1, 1088), np.uint8), np.zeros((1, 2400), np.uint8))
    assert e.value.code == L.PQC_ERR_UNSUPPORTED
    small.close()
    with pytest.raises(pqc.PQCError) as e:
        pqc.Context(["kyber1024"], max_batch=65536, hbm_budget=1 << 20)
    assert e.value.code == L.PQC_ERR_NO_MEMORY


def test_nist_dropins_deterministic(pqc):
    say("NIST drop-ins")
    for k, alg in ALG.items():
        pqc.randombytes_seed(42)
        pk, sk = pqc.crypto_kem_keypair(alg)
        ct, ss = pqc.crypto_kem_enc(pk, alg)
        assert pqc.crypto_kem_dec(ct, sk, alg) == ss
        # the same DRBG draws through the oracle: keypair d then z, encaps m
        import struct

        def draw(i):
            return hashlib.shake_256(b"pqc-mi355x" + struct.pack("<QQQ", 42, 0xFFFFFFFF, i)).digest(32)
        d0 = draw(0) + draw(1)
        d1 = draw(2)
        opk, osk = O.kyber_keypair(k, d0)
        assert (pk, sk) == (opk, osk)
        oct_, oss = O.kyber_enc(k, opk, d1)
        assert (ct, ss) == (oct_, oss)
    pqc.randombytes_system()


def test_full_size_fused_equals_plain_kyber768(pqc):
    """65536 Kyber-768 decaps: fused == plain on every element, round trip holds,
    and a sample of elements matches the oracle."""
    n = 65536
    say("full size", n)
    ctx = pqc.Context(["kyber768"], max_batch=n)
    rng = np.random.default_rng(0)
    coins = rng.integers(0, 256, (n, 64), dtype=np.uint8)
    ecoins = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    dco = pqc.DeviceBuffer.from_array(coins)
    dec_ = pqc.DeviceBuffer.from_array(ecoins)
    pk, sk = pqc.DeviceBuffer(n * 1184), pqc.DeviceBuffer(n * 2400)
    ct, ss = pqc.DeviceBuffer(n * 1088), pqc.DeviceBuffer(n * 32)
    ssf, ssp = pqc.DeviceBuffer(n * 32), pqc.DeviceBuffer(n * 32)
    ctx.keypair_derand("kyber768", n, pk, sk, dco)
    # fused key generation == plain key generation on every element
    pk2, sk2 = pqc.DeviceBuffer(n * 1184), pqc.DeviceBuffer(n * 2400)
    ctx.keypair_derand("kyber768", n, pk2, sk2, dco, plain=True)
    assert (pk.download((n, 1184)) == pk2.download((n, 1184))).all()
    assert (sk.download((n, 2400)) == sk2.download((n, 2400))).all()
    ctx.enc_derand("kyber768", n, ct, ss, pk, dec_)
    # tamper 1/8 of the ciphertexts in place
    h_ct = ct.download((n, 1088))
    h_ct[::8, 5] ^= 0x10
    ct.upload(h_ct)
    ctx.dec("kyber768", n, ssf, ct, sk)
    ctx.dec("kyber768", n, ssp, ct, sk, plain=True)
    a, b, e = ssf.download((n, 32)), ssp.download((n, 32)), ss.download((n, 32))
    assert (a == b).all()
    good = np.ones(n, bool)
    good[::8] = False
    assert (a[good] == e[good]).all()
    assert (a[~good] != e[~good]).any(axis=1).all()
    h_sk = sk.download((n, 2400))
    idx = rng.choice(n, 48, replace=False)
    want = O.kyber_batch("dec", 3, [np.ascontiguousarray(h_ct[idx]), np.ascontiguousarray(h_sk[idx])])
    assert (a[idx] == want).all()
    ctx.close()


def test_host_staging_pageable_and_pinned(pqc):
    """Host-buffer calls larger than one 16 MiB staging chunk, from pageable numpy
    arrays and from pqc_host_alloc memory, give the same bytes as device calls."""
    n = 12288  # sk 29.5 MB and pk 14.5 MB: two chunks each way
    say("host staging", n)
    ctx = pqc.Context(["kyber768"], max_batch=n)
    rng = np.random.default_rng(7)
    coins = rng.integers(0, 256, (n, 64), dtype=np.uint8)
    ecoins = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    dco, dec_ = pqc.DeviceBuffer.from_array(coins), pqc.DeviceBuffer.from_array(ecoins)
    pk, sk = pqc.DeviceBuffer(n * 1184), pqc.DeviceBuffer(n * 2400)
    ct, ss = pqc.DeviceBuffer(n * 1088), pqc.DeviceBuffer(n * 32)
    ctx.keypair_derand("kyber768", n, pk, sk, dco)
    ctx.enc_derand("kyber768", n, ct, ss, pk, dec_)
    d_pk, d_sk = pk.download((n, 1184)), sk.download((n, 2400))
    d_ct, d_ss = ct.download((n, 1088)), ss.download((n, 32))
    d_ct[::5, 100] ^= 1  # some implicit rejections
    ss2 = pqc.DeviceBuffer(n * 32)
    ct.upload(d_ct)
    ctx.dec("kyber768", n, ss2, ct, sk)
    want_dec = ss2.download((n, 32))

    h_pk, h_sk = ctx.keypair_host("kyber768", coins)
    assert (h_pk == d_pk).all() and (h_sk == d_sk).all()
    h_ct, h_ss = ctx.enc_host("kyber768", d_pk, ecoins)
    h_ct[::5, 100] ^= 1
    assert (h_ct == d_ct).all()
    assert (ctx.dec_host("kyber768", h_ct, h_sk) == want_dec).all()

    say("host staging pinned", n)
    bufs = [pqc.PinnedBuffer(n * w) for w in (1184, 2400, 1088, 2400, 32)]
    p_pk, p_sk = bufs[0].view((n, 1184)), bufs[1].view((n, 2400))
    ctx.keypair_host("kyber768", coins, out=(p_pk, p_sk))
    assert (p_pk == d_pk).all() and (p_sk == d_sk).all()
    p_ct, p_sk2, p_ss = bufs[2].view((n, 1088)), bufs[3].view((n, 2400)), bufs[4].view((n, 32))
    p_ct[:] = d_ct
    p_sk2[:] = d_sk
    ctx.dec_host("kyber768", p_ct, p_sk2, out=p_ss)
    assert (p_ss == want_dec).all()
    for b in bufs:
        b.free()
    ctx.close()


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("n", [1, 7, 200])
def test_lowlat_decaps_equals_tile_and_oracle(pqc, k, n):
    """Small decapsulation batches run one four-wave workgroup per request
    (k_lowlat_dec); the bytes equal the 64-request tile kernel's (forced with
    lowlat_max = -1) and the oracle's, tampered ciphertexts included."""
    from pqc_amd import tuner as T
    say("lowlat decaps k", k, "n", n)
    alg = ALG[k]
    ctx = pqc.Context([alg], max_batch=max(n, 64))
    try:
        kc = O.drbg_array(n, 64, 0, 3000 + 10 * k + n)
        ec = O.drbg_array(n, 32, 0, 3001 + 10 * k + n)
        pk, sk = ctx.keypair_host(alg, kc)
        ct, ss = ctx.enc_host(alg, pk, ec)
        ct[::2, n % ct.shape[1]] ^= 0x40  # every other ciphertext rejected
        T.put_tuning(ctx, lowlat_max=0)   # default: batches up to 256 take the low-latency kernel
        low = ctx.dec_host(alg, ct, sk)
        T.put_tuning(ctx, lowlat_max=-1)  # tiles only
        tile = ctx.dec_host(alg, ct, sk)
        assert (low == tile).all()
        assert (low[1::2] == ss[1::2]).all() and (low[::2] != ss[::2]).any(axis=1).all()
        want = O.kyber_batch("dec", k, [np.ascontiguousarray(ct), np.ascontiguousarray(sk)])
        assert (low == want).all()
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 95, 1000, 9001])
def test_dec_tile32_equals_tile64_plain_and_oracle(pqc, k, n):
    """The fused decapsulation in 32-request tiles (two waves per SIMD, A^T in a
    global scratch; the default) against the 64-request tiles, the plain per-op
    kernels and the oracle, on ragged batch sizes (partial tiles, partial
    sampling rounds) and with every third ciphertext tampered (implicit
    rejection).  lowlat_max = -1 keeps small batches on the tile kernels."""
    from pqc_amd import tuner as T
    say("dec tile32 k", k, "n", n)
    alg = ALG[k]
    ctx = pqc.Context([alg], max_batch=max(n, 64))
    try:
        kc = O.drbg_array(n, 64, 0, 5000 + 10 * k + n % 97)
        ec = O.drbg_array(n, 32, 0, 5001 + 10 * k + n % 97)
        pk, sk = ctx.keypair_host(alg, kc)
        ct, ss = ctx.enc_host(alg, pk, ec)
        ct[::3, (7 * n) % ct.shape[1]] ^= 0x10
        T.put_tuning(ctx, lowlat_max=-1, fused_dec_tile=32)
        t32 = ctx.dec_host(alg, ct, sk)
        T.put_tuning(ctx, fused_dec_tile=64)
        t64 = ctx.dec_host(alg, ct, sk)
        plain = ctx.dec_host(alg, ct, sk, plain=True)
        assert (t32 == t64).all() and (t32 == plain).all()
        honest = np.arange(n) % 3 != 0
        assert (t32[honest] == ss[honest]).all() and (t32[~honest] != ss[~honest]).any(axis=1).all()
        idx = np.unique(np.array([0, n // 2, n - 1]))
        want = O.kyber_batch("dec", k, [np.ascontiguousarray(ct[idx]), np.ascontiguousarray(sk[idx])])
        assert (t32[idx] == want).all()
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 95, 1000, 9001])
def test_enc_tile32_equals_tile64_plain_and_oracle(pqc, k, n):
    """The fused encapsulation in 32-request tiles (two waves per SIMD, H(pk) on
    half the lanes beside H(coins), G and the PRF outputs split over lane pairs,
    A^T in a global scratch; the default above lowlat_max) against the 64-request
    tiles, the plain per-op kernels and the oracle, on ragged batch sizes, with
    the same coins: identical ciphertexts and shared secrets, which decapsulate
    back to the same secrets."""
    from pqc_amd import tuner as T
    say("enc tile32 k", k, "n", n)
    alg = ALG[k]
    ctx = pqc.Context([alg], max_batch=max(n, 64))
    try:
        kc = O.drbg_array(n, 64, 0, 6000 + 10 * k + n % 97)
        ec = O.drbg_array(n, 32, 0, 6001 + 10 * k + n % 97)
        pk, sk = ctx.keypair_host(alg, kc)
        T.put_tuning(ctx, lowlat_max=-1, fused_enc_tile=32)
        ct32, ss32 = ctx.enc_host(alg, pk, ec)
        T.put_tuning(ctx, fused_enc_tile=64)
        ct64, ss64 = ctx.enc_host(alg, pk, ec)
        ctp, ssp = ctx.enc_host(alg, pk, ec, plain=True)
        assert (ct32 == ct64).all() and (ss32 == ss64).all()
        assert (ct32 == ctp).all() and (ss32 == ssp).all()
        assert (ctx.dec_host(alg, ct32, sk) == ss32).all()
        idx = np.unique(np.array([0, n // 2, n - 1]))
        wct, wss = O.kyber_batch("enc", k, [np.ascontiguousarray(pk[idx]), np.ascontiguousarray(ec[idx])])
        assert (ct32[idx] == wct).all() and (ss32[idx] == wss).all()
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_enc_tile32_random_pk_vs_tile64_and_oracle(pqc, k):
    """Encapsulation under random-byte public keys: 12-bit key coefficients up to
    4095 (never reduced by a key generation) and matrix entries from arbitrary
    seeds.  The 32-request tiles, the 64-request tiles and the oracle must give
    the same ciphertexts and shared secrets."""
    from pqc_amd import tuner as T
    n = 200
    say("enc tile32 random pk k", k, "n", n)
    alg = ALG[k]
    ctx = pqc.Context([alg], max_batch=n)
    try:
        pkb = {2: 800, 3: 1184, 4: 1568}[k]
        rng = np.random.default_rng(950 + k)
        pk = rng.integers(0, 256, (n, pkb), dtype=np.uint8)
        pk[: n // 2, : 384 * k] = 0xFF  # every key coefficient 4095
        ec = rng.integers(0, 256, (n, 32), dtype=np.uint8)
        T.put_tuning(ctx, lowlat_max=-1, fused_enc_tile=32)
        ct32, ss32 = ctx.enc_host(alg, pk, ec)
        T.put_tuning(ctx, fused_enc_tile=64)
        ct64, ss64 = ctx.enc_host(alg, pk, ec)
        assert (ct32 == ct64).all() and (ss32 == ss64).all()
        wct, wss = O.kyber_batch("enc", k, [np.ascontiguousarray(pk), np.ascontiguousarray(ec)])
        assert (ct32 == wct).all() and (ss32 == wss).all()
    finally:
        ctx.close()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_dec_tile32_random_bytes_vs_tile64_and_oracle(pqc, k):
    """Decapsulation of random-byte ciphertexts under random-byte secret keys:
    12-bit key coefficients up to 4095 (never reduced by a key generation),
    matrix entries from arbitrary seeds (accepted candidates up to 19q - 1) and
    decompressed values over their whole range.  The 32-request tiles compute
    on exact FP32 values mod q and the 64-request tiles on int16 Montgomery
    arithmetic; both must give the oracle's bytes for every request."""
    from pqc_amd import tuner as T
    n = 200
    say("dec tile32 random bytes k", k, "n", n)
    alg = ALG[k]
    ctx = pqc.Context([alg], max_batch=n)
    try:
        ctb, skb = {2: (736, 1632), 3: (1088, 2400), 4: (1568, 3168)}[k]
        rng = np.random.default_rng(900 + k)
        ct = rng.integers(0, 256, (n, ctb), dtype=np.uint8)
        sk = rng.integers(0, 256, (n, skb), dtype=np.uint8)
        sk[: n // 2, : 384 * k] = 0xFF  # every secret coefficient 4095
        T.put_tuning(ctx, lowlat_max=-1, fused_dec_tile=32)
        t32 = ctx.dec_host(alg, ct, sk)
        T.put_tuning(ctx, fused_dec_tile=64)
        t64 = ctx.dec_host(alg, ct, sk)
        assert (t32 == t64).all()
        want = O.kyber_batch("dec", k, [np.ascontiguousarray(ct), np.ascontiguousarray(sk)])
        assert (t32 == want).all()
    finally:
        ctx.close()
