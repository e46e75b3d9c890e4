"""Pins the NewHope oracle at N = 1024 and N = 512: tables against the
reference's own precomputed arrays, the NTT by the negacyclic product identity,
the samplers against hashlib and their definitions, the key / encryption
equations, and the CCA-KEM's layout and hash equations."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from tests import kyber_math as KM
from tests import oracle_lib as O

Q = 12289
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
rng = np.random.default_rng(99)
NS = [1024, 512]


def negacyclic(a, b):
    n = len(a)
    a = np.asarray(a, dtype=np.int64) % Q
    b = np.asarray(b, dtype=np.int64) % Q
    full = np.zeros(2 * n, dtype=np.int64)
    for i in range(n):  # exact: n x n products, int64 safe (each < 2^28)
        full[i: i + n] += a[i] * b
    r = full[:n] - full[n:]
    return r % Q


def pack_bits(vals, d):
    """LSB-first concatenation of d-bit values (the layout unpack_bits reads)."""
    bits = ((np.asarray(vals, dtype=np.int64)[:, None] >> np.arange(d)) & 1).astype(np.uint8).reshape(-1)
    return np.packbits(bits, bitorder="little").tobytes()


def brv(n):
    bits = n.bit_length() - 1
    return np.array([int(format(i, "0%db" % bits)[::-1], 2) for i in range(n)])


@pytest.mark.parametrize("N", NS)
def test_tables_match_reference(N):
    with open(os.path.join(GOLDEN, "newhope%d_tables.json" % N)) as f:
        ref = json.load(f)
    o, g, gi = O.newhope_tables(N)
    assert o.tolist() == ref["omegas_inv_bitrev_montgomery"]
    assert g.tolist() == ref["gammas_bitrev_montgomery"]
    assert gi.tolist() == ref["gammas_inv_montgomery"]
    assert brv(N).tolist() == ref["bitrev_table"]


@pytest.mark.parametrize("prefix", ["oracle_newhope_", "oracle_newhope512_"])
def test_montgomery_and_barrett(prefix):
    L = O.lib()
    mont, barr = getattr(L, prefix + "montgomery_reduce"), getattr(L, prefix + "barrett_reduce")
    mont.restype, mont.argtypes = ctypes.c_uint16, [ctypes.c_uint32]
    barr.restype, barr.argtypes = ctypes.c_uint16, [ctypes.c_uint16]
    rinv = pow(2 ** 18, -1, Q)
    for a in list(range(0, 1073491968, 9973)) + [1073491968]:
        r = mont(a)
        assert r < 2 ** 14 and (r - a * rinv) % Q == 0
    for a in range(65536):
        r = barr(a)
        assert (r - a) % Q == 0 and r < 2 * Q


@pytest.mark.parametrize("N", NS)
def test_ntt_roundtrip_and_product_identity(N):
    # poly_ntt reads its input in bit-reversed order and poly_invntt returns
    # normal order (poly.cu:513-537), so invntt(ntt(a)) is the bit-reversal
    # permutation of a, and products multiply the bit-reversed polynomials.
    # poly_invntt feeds lazily added sums into a reduction on uint16
    # (ntt.cu:228), so its inputs must already be below 2^14, as every
    # caller's are (pointwise products are Montgomery outputs).
    p = brv(N)
    for _ in range(2):
        a = rng.integers(0, Q, N)
        ah = O.newhope_ntt(a, N).astype(np.int64) % Q
        assert (O.newhope_invntt(ah, N).astype(np.int64) % Q == a[p]).all()
    a = rng.integers(0, Q, N)
    b = (rng.integers(-8, 9, N) % Q)
    ah, bh = O.newhope_ntt(a, N).astype(np.int64), O.newhope_ntt(b, N).astype(np.int64)
    prod = (ah * bh) % Q
    back = O.newhope_invntt(prod, N).astype(np.int64) % Q
    assert (back == negacyclic(a[p], b[p])).all()


@pytest.mark.parametrize("N", NS)
def test_uniform_vs_hashlib(N):
    seed = bytes(range(32))
    a = O.newhope_uniform(seed, N)
    for i in range(N // 64):
        stream = hashlib.shake_128(seed + bytes([i])).digest(168 * 2)
        vals = []
        for p in range(0, len(stream), 2):
            v = stream[p] | (stream[p + 1] << 8)
            if v < 5 * Q:
                vals.append(v)
            if len(vals) == 64:
                break
        assert a[64 * i: 64 * i + 64].tolist() == vals


@pytest.mark.parametrize("N", NS)
def test_sample_vs_definition(N):
    seed = bytes(range(100, 132))
    for nonce in (0, 1, 2):
        r = O.newhope_sample(seed, nonce, N)
        for i in range(N // 64):
            buf = hashlib.shake_256(seed + bytes([nonce, i])).digest(128)
            for j in range(64):
                want = bin(buf[2 * j]).count("1") + Q - bin(buf[2 * j + 1]).count("1")
                assert r[64 * i + j] == want


@pytest.mark.parametrize("N", NS)
def test_keypair_equation(N):
    P = 14 * N // 8
    seed = O.drbg(32, seed=0, stream=40, index=0)
    pk, sk = O.newhope_keypair(seed, N)
    z = hashlib.shake_256(seed).digest(64)
    assert pk[P:] == z[:32]
    ahat = O.newhope_uniform(z[:32], N).astype(np.int64)
    shat = O.newhope_ntt(O.newhope_sample(z[32:], 0, N), N).astype(np.int64)
    ehat = O.newhope_ntt(O.newhope_sample(z[32:], 1, N), N).astype(np.int64)
    assert (KM.unpack_bits(sk, N, 14) == shat % Q).all()
    assert (KM.unpack_bits(pk[:P], N, 14) == (ahat * shat + ehat) % Q).all()


@pytest.mark.parametrize("N", NS)
def test_encrypt_structure(N):
    P = 14 * N // 8
    pk, sk = O.newhope_keypair(O.drbg(32, seed=0, stream=41, index=0), N)
    m = O.drbg(32, seed=0, stream=42, index=0)
    coin = O.drbg(32, seed=0, stream=43, index=0)
    c = O.newhope_enc(m, pk, coin, N)
    ahat = O.newhope_uniform(pk[P:], N).astype(np.int64)
    bhat = KM.unpack_bits(pk[:P], N, 14)
    sp = O.newhope_ntt(O.newhope_sample(coin, 0, N), N).astype(np.int64)
    ep = O.newhope_ntt(O.newhope_sample(coin, 1, N), N).astype(np.int64)
    epp = O.newhope_sample(coin, 2, N).astype(np.int64)
    assert (KM.unpack_bits(c[:P], N, 14) == (ahat * sp + ep) % Q).all()
    vprime = O.newhope_invntt((bhat * sp) % Q, N).astype(np.int64)
    bits = np.unpackbits(np.frombuffer(m, np.uint8), bitorder="little").astype(np.int64)
    msgpoly = np.tile(bits, N // 256) * (Q // 2)  # each bit on coefficients i + 256 q
    v = (vprime + epp + msgpoly) % Q
    want = ((v * 8 + Q // 2) // Q) % 8
    assert (KM.unpack_bits(c[P:], N, 3) == want).all()


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("idx", range(6))
def test_cpa_roundtrip(N, idx):
    pk, sk = O.newhope_keypair(O.drbg(32, seed=0, stream=44, index=idx), N)
    m = O.drbg(32, seed=0, stream=45, index=idx)
    c = O.newhope_enc(m, pk, O.drbg(32, seed=0, stream=46, index=idx), N)
    assert O.newhope_dec(c, sk, N) == m


@pytest.mark.parametrize("N", NS)
def test_tomsg_decodes_noisy_message(N):
    """Decoding sums |x - q/2| over the N/256 coefficients of each bit and
    compares with q (N = 1024) or q/2 (N = 512): a message still decodes after
    every third 3-bit v coefficient moves by one step (q/8)."""
    P = 14 * N // 8
    pk, sk = O.newhope_keypair(O.drbg(32, seed=0, stream=50, index=N), N)
    m = O.drbg(32, seed=0, stream=51, index=N)
    c = bytearray(O.newhope_enc(m, pk, O.drbg(32, seed=0, stream=52, index=N), N))
    v = KM.unpack_bits(bytes(c[P:]), N, 3)
    # moving a 3-bit v coefficient by one step (q/8) keeps every bit decodable
    v2 = (v + np.where(np.arange(N) % 3 == 0, 1, 0)) % 8
    c[P:] = pack_bits(v2, 3)
    assert O.newhope_dec(bytes(c), sk, N) == m


@pytest.mark.parametrize("N", NS)
def test_triple_batch_driver(N):
    n = 8
    seed = O.drbg_array(n, 32, 0, 47)
    msg = O.drbg_array(n, 32, 0, 48)
    coin = O.drbg_array(n, 32, 0, 49)
    pk, sk, ct, m2 = O.newhope_triples(seed, msg, coin, nthreads=3, N=N)
    assert (m2 == msg).all()
    for i in range(n):
        assert O.newhope_keypair(seed[i].tobytes(), N) == (pk[i].tobytes(), sk[i].tobytes())


# ---- CCA-KEM: layout and equations of the specification's transform, with hashlib ----
def _shake(data, n):
    return hashlib.shake_256(bytes(data)).digest(n)


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("idx", range(3))
def test_cca_keypair_layout(N, idx):
    PK, SK, _, CSK, _ = O.nh_sizes(N)
    coins = O.drbg(64, seed=0, stream=60, index=idx)
    pk, sk = O.newhope_cca_keypair(coins, N)
    cpk, csk = O.newhope_keypair(coins[:32], N)
    assert pk == cpk and len(sk) == CSK
    assert sk[:SK] == csk and sk[SK:SK + PK] == pk
    assert sk[SK + PK:SK + PK + 32] == _shake(pk, 32) and sk[SK + PK + 32:] == coins[32:]


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("idx", range(3))
def test_cca_encaps_equations(N, idx):
    """ct_cpa decrypts to mu = H(coins); d and ss follow from hashlib alone:
    (K, coin', d) = H(mu || H(pk), 96), ct_cpa = Enc(pk, mu, coin'),
    ss = H(K || H(ct_cpa || d))."""
    _, SK, CT, _, _ = O.nh_sizes(N)
    pk, sk = O.newhope_cca_keypair(O.drbg(64, seed=0, stream=61, index=idx), N)
    coins = O.drbg(32, seed=0, stream=62, index=idx)
    ct, ss = O.newhope_cca_enc(pk, coins, N)
    mu = _shake(coins, 32)
    assert O.newhope_dec(ct[:CT], sk[:SK], N) == mu
    kcd = _shake(mu + _shake(pk, 32), 96)
    assert ct[:CT] == O.newhope_enc(mu, pk, kcd[32:64], N)
    assert ct[CT:] == kcd[64:]
    assert ss == _shake(kcd[:32] + _shake(ct, 32), 32)
    assert O.newhope_cca_dec(ct, sk, N) == ss


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("where", ["u0", "u_end", "v0", "v_end", "d0", "d_end"])
def test_cca_implicit_rejection(N, where):
    """A changed ciphertext byte anywhere (u, v, d) gives H(s || H(ct')), s from the secret key."""
    P, CT = 14 * N // 8, O.nh_sizes(N)[2]
    pos = {"u0": 0, "u_end": P - 1, "v0": P, "v_end": CT - 1, "d0": CT, "d_end": CT + 31}[where]
    pk, sk = O.newhope_cca_keypair(O.drbg(64, seed=0, stream=63, index=pos), N)
    ct, ss = O.newhope_cca_enc(pk, O.drbg(32, seed=0, stream=64, index=pos), N)
    bad = bytearray(ct)
    bad[pos] ^= 0x04
    got = O.newhope_cca_dec(bytes(bad), sk, N)
    assert got != ss and got == _shake(sk[-32:] + _shake(bad, 32), 32)
