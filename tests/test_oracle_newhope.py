"""Pins the NewHope-1024 CPA oracle: tables against the reference's own
precomputed arrays, the NTT by the negacyclic product identity, the samplers
against hashlib and their definitions, and the key / encryption equations."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from tests import kyber_math as KM
from tests import oracle_lib as O

Q = 12289
N = 1024
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
rng = np.random.default_rng(99)


def negacyclic(a, b):
    a = np.asarray(a, dtype=np.int64) % Q
    b = np.asarray(b, dtype=np.int64) % Q
    full = np.zeros(2 * N, dtype=np.int64)
    for i in range(N):  # exact: 1024 x 1024 products, int64 safe (each < 2^28)
        full[i: i + N] += a[i] * b
    r = full[:N] - full[N:]
    return r % Q


def test_tables_match_reference():
    with open(os.path.join(GOLDEN, "newhope1024_tables.json")) as f:
        ref = json.load(f)
    o, g, gi = O.newhope_tables()
    assert o.tolist() == ref["omegas_inv_bitrev_montgomery"]
    assert g.tolist() == ref["gammas_bitrev_montgomery"]
    assert gi.tolist() == ref["gammas_inv_montgomery"]
    brv = [int("{:010b}".format(i)[::-1], 2) for i in range(N)]
    assert brv == ref["bitrev_table"]


def test_montgomery_and_barrett():
    L = O.lib()
    L.oracle_newhope_montgomery_reduce.restype = ctypes.c_uint16
    L.oracle_newhope_montgomery_reduce.argtypes = [ctypes.c_uint32]
    L.oracle_newhope_barrett_reduce.restype = ctypes.c_uint16
    L.oracle_newhope_barrett_reduce.argtypes = [ctypes.c_uint16]
    rinv = pow(2 ** 18, -1, Q)
    for a in list(range(0, 1073491968, 9973)) + [1073491968]:
        r = L.oracle_newhope_montgomery_reduce(a)
        assert r < 2 ** 14 and (r - a * rinv) % Q == 0
    for a in range(65536):
        r = L.oracle_newhope_barrett_reduce(a)
        assert (r - a) % Q == 0 and r < 2 * Q


def test_ntt_roundtrip_and_product_identity():
    # poly_ntt reads its input in bit-reversed order and poly_invntt returns
    # normal order (poly.cu:513-537), so invntt(ntt(a)) is the bit-reversal
    # permutation of a, and products multiply the bit-reversed polynomials.
    # poly_invntt feeds lazily added sums into barrett_reduce(uint16_t)
    # (ntt.cu:228), so its inputs must already be below 2^14, as every
    # caller's are (pointwise products are Montgomery outputs).
    brv = np.array([int("{:010b}".format(i)[::-1], 2) for i in range(N)])
    for _ in range(2):
        a = rng.integers(0, Q, N)
        ah = O.newhope_ntt(a).astype(np.int64) % Q
        assert (O.newhope_invntt(ah).astype(np.int64) % Q == a[brv]).all()
    a = rng.integers(0, Q, N)
    b = (rng.integers(-8, 9, N) % Q)
    ah, bh = O.newhope_ntt(a).astype(np.int64), O.newhope_ntt(b).astype(np.int64)
    prod = (ah * bh) % Q
    back = O.newhope_invntt(prod).astype(np.int64) % Q
    assert (back == negacyclic(a[brv], b[brv])).all()


def test_uniform_vs_hashlib():
    seed = bytes(range(32))
    a = O.newhope_uniform(seed)
    for i in range(16):
        stream = hashlib.shake_128(seed + bytes([i])).digest(168 * 2)
        vals = []
        for p in range(0, len(stream), 2):
            v = stream[p] | (stream[p + 1] << 8)
            if v < 5 * Q:
                vals.append(v)
            if len(vals) == 64:
                break
        assert a[64 * i: 64 * i + 64].tolist() == vals


def test_sample_vs_definition():
    seed = bytes(range(100, 132))
    for nonce in (0, 1, 2):
        r = O.newhope_sample(seed, nonce)
        for i in range(16):
            buf = hashlib.shake_256(seed + bytes([nonce, i])).digest(128)
            for j in range(64):
                want = bin(buf[2 * j]).count("1") + Q - bin(buf[2 * j + 1]).count("1")
                assert r[64 * i + j] == want


def unpack14(b):
    return KM.unpack_bits(b, N, 14)


def test_keypair_equation():
    seed = O.drbg(32, seed=0, stream=40, index=0)
    pk, sk = O.newhope_keypair(seed)
    z = hashlib.shake_256(seed).digest(64)
    assert pk[1792:] == z[:32]
    ahat = O.newhope_uniform(z[:32]).astype(np.int64)
    shat = O.newhope_ntt(O.newhope_sample(z[32:], 0)).astype(np.int64)
    ehat = O.newhope_ntt(O.newhope_sample(z[32:], 1)).astype(np.int64)
    assert (unpack14(sk) == shat % Q).all()
    assert (unpack14(pk[:1792]) == (ahat * shat + ehat) % Q).all()


def test_encrypt_structure():
    pk, sk = O.newhope_keypair(O.drbg(32, seed=0, stream=41, index=0))
    m = O.drbg(32, seed=0, stream=42, index=0)
    coin = O.drbg(32, seed=0, stream=43, index=0)
    c = O.newhope_enc(m, pk, coin)
    ahat = O.newhope_uniform(pk[1792:]).astype(np.int64)
    bhat = unpack14(pk[:1792])
    sp = O.newhope_ntt(O.newhope_sample(coin, 0)).astype(np.int64)
    ep = O.newhope_ntt(O.newhope_sample(coin, 1)).astype(np.int64)
    epp = O.newhope_sample(coin, 2).astype(np.int64)
    assert (unpack14(c[:1792]) == (ahat * sp + ep) % Q).all()
    vprime = O.newhope_invntt((bhat * sp) % Q).astype(np.int64)
    bits = np.unpackbits(np.frombuffer(m, np.uint8), bitorder="little").astype(np.int64)
    msgpoly = np.tile(bits, 4) * (Q // 2)
    v = (vprime + epp + msgpoly) % Q
    want = ((v * 8 + Q // 2) // Q) % 8
    assert (KM.unpack_bits(c[1792:], N, 3) == want).all()


@pytest.mark.parametrize("idx", range(6))
def test_cpa_roundtrip(idx):
    pk, sk = O.newhope_keypair(O.drbg(32, seed=0, stream=44, index=idx))
    m = O.drbg(32, seed=0, stream=45, index=idx)
    c = O.newhope_enc(m, pk, O.drbg(32, seed=0, stream=46, index=idx))
    assert O.newhope_dec(c, sk) == m


def test_triple_batch_driver():
    n = 8
    seed = O.drbg_array(n, 32, 0, 47)
    msg = O.drbg_array(n, 32, 0, 48)
    coin = O.drbg_array(n, 32, 0, 49)
    pk, sk, ct, m2 = O.newhope_triples(seed, msg, coin, nthreads=3)
    assert (m2 == msg).all()
    for i in range(n):
        assert O.newhope_keypair(seed[i].tobytes()) == (pk[i].tobytes(), sk[i].tobytes())


# ---- CCA-KEM: layout and equations of the specification's transform, with hashlib ----
def _shake(data, n):
    return hashlib.shake_256(bytes(data)).digest(n)


@pytest.mark.parametrize("idx", range(3))
def test_cca_keypair_layout(idx):
    coins = O.drbg(64, seed=0, stream=60, index=idx)
    pk, sk = O.newhope_cca_keypair(coins)
    cpk, csk = O.newhope_keypair(coins[:32])
    assert pk == cpk and len(sk) == 3680
    assert sk[:1792] == csk and sk[1792:3616] == pk
    assert sk[3616:3648] == _shake(pk, 32) and sk[3648:] == coins[32:]


@pytest.mark.parametrize("idx", range(3))
def test_cca_encaps_equations(idx):
    """ct_cpa decrypts to mu = H(coins); d and ss follow from hashlib alone:
    (K, coin', d) = H(mu || H(pk), 96), ct_cpa = Enc(pk, mu, coin'),
    ss = H(K || H(ct_cpa || d))."""
    pk, sk = O.newhope_cca_keypair(O.drbg(64, seed=0, stream=61, index=idx))
    coins = O.drbg(32, seed=0, stream=62, index=idx)
    ct, ss = O.newhope_cca_enc(pk, coins)
    mu = _shake(coins, 32)
    assert O.newhope_dec(ct[:2176], sk[:1792]) == mu
    kcd = _shake(mu + _shake(pk, 32), 96)
    assert ct[:2176] == O.newhope_enc(mu, pk, kcd[32:64])
    assert ct[2176:] == kcd[64:]
    assert ss == _shake(kcd[:32] + _shake(ct, 32), 32)
    assert O.newhope_cca_dec(ct, sk) == ss


@pytest.mark.parametrize("pos", [0, 1791, 1792, 2175, 2176, 2207])
def test_cca_implicit_rejection(pos):
    """A changed ciphertext byte anywhere (u, v, d) gives H(s || H(ct')), s from the secret key."""
    pk, sk = O.newhope_cca_keypair(O.drbg(64, seed=0, stream=63, index=pos))
    ct, ss = O.newhope_cca_enc(pk, O.drbg(32, seed=0, stream=64, index=pos))
    bad = bytearray(ct)
    bad[pos] ^= 0x04
    got = O.newhope_cca_dec(bytes(bad), sk)
    assert got != ss and got == _shake(sk[3648:] + _shake(bad, 32), 32)
