"""Pins the Kyber (round 2) oracle: against the reference's own twiddle table,
against closed-form algebra (tests/kyber_math.py), exhaustively for the
scalar reductions, compressions and CBD, and against hashlib for every hash
the KEM glue uses."""
import hashlib
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from tests import kyber_math as M
from tests import oracle_lib as O

Q = 3329
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
rng = np.random.default_rng(1234)


def test_zetas_match_reference_table():
    with open(os.path.join(GOLDEN, "kyber_zetas.json")) as f:
        ref = json.load(f)
    z, zi = O.kyber_zetas()
    assert z.tolist() == ref["zetas"]
    assert zi.tolist() == ref["zetas_inv"]


def test_montgomery_reduce_exhaustive_range():
    L = O.lib()
    f = L.oracle_kyber_montgomery_reduce
    f.restype = __import__("ctypes").c_int16
    rinv = pow(2 ** 16, -1, Q)
    for a in list(range(-Q * 2 ** 15, Q * 2 ** 15, 104729)) + [-Q * 2 ** 15, Q * 2 ** 15 - 1]:
        r = f(a)
        assert -Q < r < Q and (r - a * rinv) % Q == 0


def test_barrett_and_csubq_exhaustive():
    import ctypes
    L = O.lib()
    br, cs = L.oracle_kyber_barrett_reduce, L.oracle_kyber_csubq
    br.restype = cs.restype = ctypes.c_int16
    br.argtypes = cs.argtypes = [ctypes.c_int16]
    for a in range(-32768, 32768):
        r = br(a)
        assert 0 <= r <= Q and (r - a) % Q == 0
    for a in range(0, 2 * Q):
        assert cs(a) == (a - Q if a >= Q else a)


def test_ntt_is_evaluation_mod_quadratics():
    for _ in range(3):
        a = rng.integers(-Q + 1, Q, 256)
        got = O.kyber_ntt(a).astype(np.int64) % Q
        assert (got == M.ntt_true(a)).all()
        raw = O.kyber_ntt(a, reduce=False).astype(np.int64) % Q
        assert (raw == M.ntt_true(a)).all()


def test_invntt_inverts_with_montgomery_factor():
    for _ in range(3):
        a = rng.integers(0, Q, 256)
        back = O.kyber_invntt(O.kyber_ntt(a)).astype(np.int64) % Q
        assert (back == (a * 2 ** 16) % Q).all()


def test_basemul_is_negacyclic_product():
    for _ in range(3):
        a = rng.integers(-2, 3, 256)
        b = rng.integers(0, Q, 256)
        ah, bh = O.kyber_ntt(a), O.kyber_ntt(b)
        prod = O.kyber_basemul(ah, bh).astype(np.int64) % Q
        rinv = pow(2 ** 16, -1, Q)
        assert (prod == M.pair_mul(ah, bh) * rinv % Q).all()
        back = O.kyber_invntt(O.kyber_basemul(ah, bh)).astype(np.int64) % Q
        assert (back == M.negacyclic(a, b)).all()


def test_cbd_vs_popcount_definition():
    for seed in range(4):
        b = bytes(np.random.default_rng(seed).integers(0, 256, 128, dtype=np.uint8))
        assert (O.kyber_cbd(b).astype(np.int64) == M.cbd2(b)).all()
    # every byte value in every position class
    b = bytes(range(128))
    assert (O.kyber_cbd(b).astype(np.int64) == M.cbd2(b)).all()
    b = bytes(range(128, 256))
    assert (O.kyber_cbd(b).astype(np.int64) == M.cbd2(b)).all()


@pytest.mark.parametrize("d", [1, 3, 4, 5, 10, 11])
def test_compress_rounding_exhaustive(d):
    xs = np.arange(Q + 1)
    for lo in range(0, len(xs), 256):
        chunk = np.zeros(256, dtype=np.int64)
        part = xs[lo: lo + 256]
        chunk[: len(part)] = part
        got = M.unpack_bits(O.kyber_compress(chunk, d), 256, d)
        want = np.array([int(Fraction(int(x) % Q * 2 ** d, Q) + Fraction(1, 2)) % 2 ** d for x in chunk])
        assert (got == want).all()


@pytest.mark.parametrize("d", [3, 4, 5, 10, 11])
def test_decompress_rounding_exhaustive(d):
    vals = np.arange(2 ** d) % (2 ** d)
    for lo in range(0, len(vals), 256):
        t = np.zeros(256, dtype=np.int64)
        part = vals[lo: lo + 256]
        t[: len(part)] = part
        packed = np.packbits(((t[:, None] >> np.arange(d)) & 1).astype(np.uint8).reshape(-1), bitorder="little")
        got = O.kyber_decompress(packed.tobytes(), d).astype(np.int64)
        want = np.array([int(Fraction(int(x) * Q, 2 ** d) + Fraction(1, 2)) for x in t])
        assert (got == want).all()


def test_packing_matches_reference_unrolled_byte_formulas():
    # 11-bit packer of polyvec.cu:36-46 and 5-bit packer of poly.cu:36-40,
    # restated here byte by byte, against the oracle's generic LSB-first packer.
    x = rng.integers(0, Q, 256)
    t11 = M.compress_int(x, 11)
    t5 = M.compress_int(x, 5)
    r11 = bytearray()
    for j in range(32):
        t = [int(v) for v in t11[8 * j: 8 * j + 8]]
        r11 += bytes([t[0] & 0xFF, ((t[0] >> 8) | ((t[1] & 0x1F) << 3)) & 0xFF,
                      ((t[1] >> 5) | ((t[2] & 0x03) << 6)) & 0xFF, (t[2] >> 2) & 0xFF,
                      ((t[2] >> 10) | ((t[3] & 0x7F) << 1)) & 0xFF, ((t[3] >> 7) | ((t[4] & 0x0F) << 4)) & 0xFF,
                      ((t[4] >> 4) | ((t[5] & 0x01) << 7)) & 0xFF, (t[5] >> 1) & 0xFF,
                      ((t[5] >> 9) | ((t[6] & 0x3F) << 2)) & 0xFF, ((t[6] >> 6) | ((t[7] & 0x07) << 5)) & 0xFF,
                      (t[7] >> 3) & 0xFF])
    r5 = bytearray()
    for j in range(32):
        t = [int(v) for v in t5[8 * j: 8 * j + 8]]
        r5 += bytes([(t[0] | (t[1] << 5)) & 0xFF, ((t[1] >> 3) | (t[2] << 2) | (t[3] << 7)) & 0xFF,
                     ((t[3] >> 1) | (t[4] << 4)) & 0xFF, ((t[4] >> 4) | (t[5] << 1) | (t[6] << 6)) & 0xFF,
                     ((t[6] >> 2) | (t[7] << 3)) & 0xFF])
    assert O.kyber_compress(x, 11) == bytes(r11)
    assert O.kyber_compress(x, 5) == bytes(r5)


def test_rej_uniform_brute_force():
    for seed in range(3):
        stream = np.random.default_rng(seed).integers(0, 256, 700, dtype=np.uint8).tobytes()
        want = M.rej_uniform(stream)
        got = O.kyber_rej_uniform(stream)
        assert got.tolist() == want[: len(got)] and len(got) == len(want)
    # candidates straddling 19q and the v >> 12 reduction boundaries
    cands = [0, 4095, 4096, 19 * Q - 1, 19 * Q, 65535, 61440, 3328, 3329]
    stream = b"".join(v.to_bytes(2, "little") for v in cands)
    assert O.kyber_rej_uniform(stream).tolist() == M.rej_uniform(stream)


@pytest.mark.parametrize("transposed", [0, 1])
def test_gen_matrix_vs_hashlib(transposed):
    rho = bytes(range(32))
    a = O.kyber_gen_matrix(3, rho, transposed)
    for i in range(3):
        for j in range(3):
            x, y = (i, j) if transposed else (j, i)
            assert (a[i, j].astype(np.int64) == M.matrix_entry(rho, x, y)).all()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_keypair_layout_and_key_equation(k):
    p = M.params(k)
    coins = O.drbg(64, seed=0, stream=100 + k, index=0)
    pk, sk = O.kyber_keypair(k, coins)
    assert len(pk) == p["pk"] and len(sk) == p["sk"]
    g = hashlib.sha3_512(coins[:32]).digest()
    rho, sigma = g[:32], g[32:]
    assert pk[-32:] == rho
    assert sk[384 * k: 384 * k + p["pk"]] == pk
    assert sk[-64:-32] == hashlib.sha3_256(pk).digest()
    assert sk[-32:] == coins[32:]
    s_hat = [M.unpack_bits(sk[384 * i: 384 * (i + 1)], 256, 12) for i in range(k)]
    t_hat = [M.unpack_bits(pk[384 * i: 384 * (i + 1)], 256, 12) for i in range(k)]
    s = [M.cbd2(M.prf(sigma, i)) for i in range(k)]
    e = [M.cbd2(M.prf(sigma, k + i)) for i in range(k)]
    for i in range(k):
        assert (s_hat[i] == M.ntt_true(s[i])).all()
    for i in range(k):
        acc = M.ntt_true(e[i])
        for j in range(k):
            acc = (acc + M.pair_mul(M.matrix_entry(rho, j, i), M.ntt_true(s[j]))) % Q
        assert (t_hat[i] == acc).all()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_encaps_structure(k):
    """ct = (Compress(A^T r + e1, du), Compress(t^T r + e2 + Decompress(m), dv)),
    with (Kbar, coins) = G(H(seed) || H(pk)) and ss = KDF(Kbar || H(ct))."""
    p = M.params(k)
    pk, sk = O.kyber_keypair(k, O.drbg(64, seed=0, stream=200 + k, index=0))
    seed = O.drbg(32, seed=0, stream=300 + k, index=0)
    ct, ss = O.kyber_enc(k, pk, seed)
    m = hashlib.sha3_256(seed).digest()
    kr = hashlib.sha3_512(m + hashlib.sha3_256(pk).digest()).digest()
    coins = kr[32:]
    assert ss == hashlib.shake_256(kr[:32] + hashlib.sha3_256(ct).digest()).digest(32)
    rho = pk[-32:]
    r_hat = [M.ntt_true(M.cbd2(M.prf(coins, j))) for j in range(k)]
    e1 = [M.cbd2(M.prf(coins, k + i)) for i in range(k)]
    e2 = M.cbd2(M.prf(coins, 2 * k))
    t_hat = [M.unpack_bits(pk[384 * i: 384 * (i + 1)], 256, 12) for i in range(k)]
    # invert the NTT through the oracle's (separately pinned) invntt: it returns 2^16 * a
    rinv = pow(2 ** 16, -1, Q)

    def inv(x):
        # x is a canonical NTT vector; scale to the Montgomery convention first
        return O.kyber_invntt(np.asarray(x, dtype=np.int64)).astype(np.int64) * rinv % Q

    du, dv = p["du"], p["dv"]
    for i in range(k):
        acc = np.zeros(256, dtype=np.int64)
        for j in range(k):
            acc = (acc + M.pair_mul(M.matrix_entry(rho, i, j), r_hat[j])) % Q
        u = (inv(acc) + e1[i]) % Q
        got = M.unpack_bits(ct[32 * du * i: 32 * du * (i + 1)], 256, du)
        assert (got == M.compress_int(u, du)).all()
    acc = np.zeros(256, dtype=np.int64)
    for j in range(k):
        acc = (acc + M.pair_mul(t_hat[j], r_hat[j])) % Q
    mbits = np.unpackbits(np.frombuffer(m, np.uint8), bitorder="little").astype(np.int64)
    v = (inv(acc) + e2 + mbits * ((Q + 1) // 2)) % Q
    got = M.unpack_bits(ct[32 * du * k:], 256, dv)
    assert (got == M.compress_int(v, dv)).all()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kem_roundtrip_and_implicit_rejection(k):
    for idx in range(8):
        pk, sk = O.kyber_keypair(k, O.drbg(64, seed=0, stream=k, index=idx))
        ct, ss = O.kyber_enc(k, pk, O.drbg(32, seed=0, stream=10 + k, index=idx))
        assert O.kyber_dec(k, ct, sk) == ss
        bad = bytearray(ct)
        bad[idx * 37 % len(ct)] ^= 1 << (idx % 8)
        bad = bytes(bad)
        want = hashlib.shake_256(sk[-32:] + hashlib.sha3_256(bad).digest()).digest(32)
        assert O.kyber_dec(k, bad, sk) == want


def test_batch_driver_matches_single():
    k, n = 3, 16
    coins = O.drbg_array(n, 64, seed=0, stream=1)
    pk, sk = O.kyber_batch("keypair", k, [coins], nthreads=4)
    for i in range(n):
        p1, s1 = O.kyber_keypair(k, coins[i].tobytes())
        assert p1 == pk[i].tobytes() and s1 == sk[i].tobytes()
    ecoins = O.drbg_array(n, 32, seed=0, stream=2)
    ct, ss = O.kyber_batch("enc", k, [pk, ecoins], nthreads=3)
    ss2 = O.kyber_batch("dec", k, [ct, sk], nthreads=5)
    assert (ss == ss2).all()
