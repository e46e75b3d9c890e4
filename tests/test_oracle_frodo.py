"""Pins the FrodoKEM-976-SHAKE oracle: the sampler exhaustively against its CDF
definition, packing against big-endian struct packing, matrix A against
hashlib, S'A + E' against numpy uint16 matrix products, and the KEM against
the key / encryption equations recomputed from hashlib streams."""
import hashlib
import struct

import numpy as np
import pytest

from tests import oracle_lib as O

N, NB = 976, 8
CDF = [5638, 15915, 23689, 28571, 31116, 32217, 32613, 32731, 32760, 32766, 32767]


def sample_def(r):
    """Inversion sampling by the CDF (noise.c:10-29): the count of CDF[0..9]
    entries below r >> 1, negated when r is odd, as a 16-bit two's complement."""
    prnd, sign = r >> 1, r & 1
    v = sum(1 for c in CDF[:10] if c < prnd)
    return (-v if sign else v) & 0xFFFF


def test_sampler_exhaustive():
    allv = np.arange(65536, dtype=np.uint16)
    got = O.frodo_sample(allv)
    want = np.array([sample_def(r) for r in range(65536)], dtype=np.uint16)
    assert (got == want).all()
    signed = got.astype(np.int16)
    assert signed.min() == -10 and signed.max() == 10


def test_pack16_is_big_endian_and_roundtrips():
    rng = np.random.default_rng(1)
    v = rng.integers(0, 65536, 1000).astype(np.uint16)
    packed = O.frodo_pack(v, 16)
    assert packed == struct.pack(">%dH" % len(v), *v.tolist())
    assert (O.frodo_unpack(packed, len(v), 16) == v).all()
    for lsb in (3, 5, 11, 15):
        w = (rng.integers(0, 1 << lsb, 64)).astype(np.uint16)
        assert (O.frodo_unpack(O.frodo_pack(w, lsb), 64, lsb) == w).all()


def test_gen_a_vs_hashlib():
    seedA = bytes(range(16))
    A = O.frodo_gen_a(seedA)
    for i in (0, 1, 255, 256, 975):
        row = hashlib.shake_128(struct.pack("<H", i) + seedA).digest(2 * N)
        assert (A[i] == np.frombuffer(row, dtype="<u2")).all()


def test_sa_plus_e_vs_numpy():
    rng = np.random.default_rng(2)
    A = rng.integers(0, 65536, (N, N)).astype(np.uint16)
    S = (rng.integers(-10, 11, (NB, N)) & 0xFFFF).astype(np.uint16)
    E = (rng.integers(-10, 11, (NB, N)) & 0xFFFF).astype(np.uint16)
    got = O.frodo_sa_plus_e(A, S, E)
    want = ((S.astype(np.int64) @ A.astype(np.int64) + E.astype(np.int64)) & 0xFFFF).astype(np.uint16)
    assert (got == want).all()


def shake_u16(prefix, seed, n):
    raw = hashlib.shake_256(bytes([prefix]) + seed).digest(2 * n)
    return np.frombuffer(raw, dtype="<u2").astype(np.uint16)


def test_keypair_equation_and_layout():
    rnd = O.drbg(64, seed=0, stream=60, index=0)
    pk, sk = O.frodo_keypair(rnd)
    s, seedSE, z = rnd[:24], rnd[24:48], rnd[48:64]
    seedA = hashlib.shake_256(z).digest(16)
    assert pk[:16] == seedA
    SE = O.frodo_sample(shake_u16(0x5F, seedSE, 2 * N * NB))
    S, E = SE[: N * NB].reshape(NB, N), SE[N * NB:].reshape(N, NB)
    A = O.frodo_gen_a(seedA).astype(np.int64)
    B = ((A @ S.T.astype(np.int64) + E.astype(np.int64)) & 0xFFFF).astype(np.uint16)
    assert pk[16:] == struct.pack(">%dH" % (N * NB), *B.reshape(-1).tolist())
    assert sk[:24] == s and sk[24:24 + 15632] == pk
    assert sk[24 + 15632: 24 + 15632 + 2 * N * NB] == S.astype("<u2").tobytes()
    assert sk[-24:] == hashlib.shake_256(pk).digest(24)


def test_encaps_equation():
    pk, sk = O.frodo_keypair(O.drbg(64, seed=0, stream=61, index=0))
    mu = O.drbg(24, seed=0, stream=62, index=0)
    ct, ss = O.frodo_enc(pk, mu)
    pkh = hashlib.shake_256(pk).digest(24)
    g2 = hashlib.shake_256(pkh + mu).digest(48)
    seedSE, k = g2[:24], g2[24:]
    assert ss == hashlib.shake_256(ct + k).digest(24)
    raw = shake_u16(0x96, seedSE, (2 * N + NB) * NB)
    Sp = O.frodo_sample(raw[: N * NB]).reshape(NB, N).astype(np.int64)
    Ep = O.frodo_sample(raw[N * NB: 2 * N * NB]).reshape(NB, N).astype(np.int64)
    Epp = O.frodo_sample(raw[2 * N * NB:]).reshape(NB, NB).astype(np.int64)
    A = O.frodo_gen_a(pk[:16]).astype(np.int64)
    B = np.frombuffer(pk[16:], dtype=">u2").reshape(N, NB).astype(np.int64)
    Bp = (Sp @ A + Ep) & 0xFFFF
    assert ct[: 2 * N * NB] == struct.pack(">%dH" % (N * NB), *Bp.reshape(-1).tolist())
    bits = int.from_bytes(mu, "little")
    enc = np.array([((bits >> (3 * i)) & 7) << 13 for i in range(64)], dtype=np.int64).reshape(NB, NB)
    Cm = (Sp @ B + Epp + enc) & 0xFFFF
    assert ct[2 * N * NB:] == struct.pack(">64H", *Cm.reshape(-1).tolist())


@pytest.mark.parametrize("idx", range(2))
def test_kem_roundtrip_and_implicit_rejection(idx):
    pk, sk = O.frodo_keypair(O.drbg(64, seed=0, stream=63, index=idx))
    ct, ss = O.frodo_enc(pk, O.drbg(24, seed=0, stream=64, index=idx))
    assert O.frodo_dec(ct, sk) == ss
    bad = bytearray(ct)
    bad[100 + 7000 * idx] ^= 0x40
    bad = bytes(bad)
    assert O.frodo_dec(bad, sk) == hashlib.shake_256(bad + sk[:24]).digest(24)


def test_batch_drivers_share_a():
    pk, sk = O.frodo_keypair(O.drbg(64, seed=0, stream=65, index=0))
    mu = O.drbg_array(4, 24, 0, 66)
    ct, ss = O.frodo_enc_batch(pk, mu, nthreads=2)
    assert (O.frodo_dec_batch(ct, sk, nthreads=2) == ss).all()
    assert O.frodo_enc(pk, mu[1].tobytes()) == (ct[1].tobytes(), ss[1].tobytes())
