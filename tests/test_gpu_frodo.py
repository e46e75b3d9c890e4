"""GPU parity of the FrodoKEM-976-SHAKE engine (shared-key batches, int8-MFMA
products) against the CPU oracle, bit for bit."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

N, NB, CT, SK, PK = 976, 8, 15744, 31296, 15632


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import pqc_amd
    c = pqc_amd.Context(["frodo976"], max_batch=4096)
    yield c
    c.close()


@pytest.fixture(scope="module")
def keys():
    pk, sk = O.frodo_keypair(O.drbg(64, seed=0, stream=900, index=0))
    return pk, sk


def test_keypair_vs_oracle(ctx):
    say("frodo keypair n=3")
    rnd = O.drbg_array(3, 64, 0, 901)
    pk, sk = ctx.frodo_keypair_host(rnd)
    for i in range(3):
        opk, osk = O.frodo_keypair(rnd[i].tobytes())
        assert pk[i].tobytes() == opk
        assert sk[i].tobytes() == osk


@pytest.mark.parametrize("n", [1, 2, 17, 130])
def test_enc_vs_oracle(ctx, keys, n):
    say("frodo enc n", n)
    pk, _ = keys
    mu = O.drbg_array(n, 24, 0, 910 + n)
    ct, ss = ctx.frodo_enc_host(pk, mu)
    oct_, oss = O.frodo_enc_batch(pk, mu, nthreads=8)
    assert (ct == oct_).all()
    assert (ss == oss).all()
    # B'' of the batch is the B' the ciphertext carries (big-endian in ct)
    bpp, v = ctx.frodo_products_host(n)
    assert (bpp == ct[:, : 2 * N * NB].view(">u2").reshape(n, NB, N)).all()


def test_dec_vs_oracle_with_tampering(ctx, keys):
    n = 64
    say("frodo dec n", n)
    pk, sk = keys
    mu = O.drbg_array(n, 24, 0, 920)
    ct, ss = O.frodo_enc_batch(pk, mu, nthreads=8)
    ct = ct.copy()
    # tamper in B' (early and late), in C, and in the last F block's bytes
    for i, pos in ((3, 0), (9, 7000), (17, 15615), (30, 15616), (41, 15743), (50, 15650)):
        ct[i, pos] ^= 0x01
    got = ctx.frodo_dec_host(sk, ct)
    want = O.frodo_dec_batch(ct, sk, nthreads=8)
    assert (got == want).all()
    tampered = {3, 9, 17, 30, 41, 50}
    for i in range(n):
        assert (got[i] == ss[i]).all() == (i not in tampered)


def test_dec_random_ciphertexts_vs_oracle(ctx, keys):
    say("frodo dec random ct")
    _, sk = keys
    rng = np.random.default_rng(5)
    ct = rng.integers(0, 256, (5, CT), dtype=np.uint8)
    got = ctx.frodo_dec_host(sk, ct)
    assert (got == O.frodo_dec_batch(ct, sk, nthreads=5)).all()


def test_key_switch_and_error_paths(ctx, keys):
    import pqc_amd
    say("frodo key switch / errors")
    pk, sk = keys
    pk2, sk2 = O.frodo_keypair(O.drbg(64, seed=0, stream=930, index=0))
    mu = O.drbg_array(4, 24, 0, 931)
    ct2, ss2 = ctx.frodo_enc_host(pk2, mu)
    assert (ct2 == O.frodo_enc_batch(pk2, mu, nthreads=4)[0]).all()
    # a public key alone cannot decapsulate
    ss = np.empty((4, 24), np.uint8)
    with pytest.raises(pqc_amd.PQCError):
        ctx.frodo_dec(4, ss, ct2, host=True)
    assert (ctx.frodo_dec_host(sk2, ct2) == ss2).all()
    with pytest.raises(pqc_amd.PQCError):
        big = np.zeros((4097, 24), np.uint8)
        ctx.frodo_enc(4097, np.zeros((4097, CT), np.uint8), np.zeros((4097, 24), np.uint8), big, host=True)


def test_full_size_4096_roundtrip_and_sample(ctx, keys):
    import pqc_amd
    n = 4096
    say("frodo full size", n)
    pk, sk = keys
    rng = np.random.default_rng(17)
    mu = rng.integers(0, 256, (n, 24), dtype=np.uint8)
    ct, ss = ctx.frodo_enc_host(pk, mu)
    bad = ct.copy()
    bad[::16, 1234] ^= 0x80
    d_ct = pqc_amd.DeviceBuffer.from_array(bad)
    d_ss = pqc_amd.DeviceBuffer(n * 24)
    ctx.frodo_load_key(np.frombuffer(sk, np.uint8).copy(), True, host=True)
    ctx.frodo_dec(n, d_ss, d_ct)
    ss2 = d_ss.download((n, 24))
    same = (ss2 == ss).all(axis=1)
    assert (~same[::16]).all()
    assert same[np.arange(n) % 16 != 0].all()
    idx = np.array([0, 1, 15, 16, 2047, 2048, 4094, 4095])
    oct_, oss = O.frodo_enc_batch(pk, mu[idx], nthreads=8)
    assert (ct[idx] == oct_).all() and (ss[idx] == oss).all()
    assert (ss2[idx] == O.frodo_dec_batch(bad[idx], sk, nthreads=8)).all()


def test_per_element_keys_batch_api_vs_oracle(ctx):
    """crypto_kem_*_batch with alg FRODO976: one key per element, A expanded
    per element (n = 130 crosses the 128-element chunk)."""
    n = 130
    say("frodo per-element keys n", n)
    rnd = O.drbg_array(n, 64, 0, 950)
    pk, sk = ctx.keypair_host("frodo976", rnd)
    for i in (0, 1, 129):
        opk, osk = O.frodo_keypair(rnd[i].tobytes())
        assert pk[i].tobytes() == opk and sk[i].tobytes() == osk
    mu = O.drbg_array(n, 24, 0, 951)
    ct, ss = ctx.enc_host("frodo976", pk, mu)
    for i in range(n):
        oct_, oss = O.frodo_enc(pk[i].tobytes(), mu[i].tobytes())
        assert ct[i].tobytes() == oct_ and ss[i].tobytes() == oss, i
    bad = ct.copy()
    bad[::5, 321] ^= 0x02
    bad[3, 15700] ^= 0x80  # inside C
    ss2 = ctx.dec_host("frodo976", bad, sk)
    for i in range(n):
        assert ss2[i].tobytes() == O.frodo_dec(bad[i].tobytes(), sk[i].tobytes()), i
        assert (ss2[i] == ss[i]).all() == (i % 5 != 0 and i != 3), i


def test_nist_dropins_frodo976(ctx):
    import pqc_amd
    say("frodo NIST drop-ins")
    pqc_amd.randombytes_seed(77)
    pk, sk = pqc_amd.crypto_kem_keypair("frodo976")
    ct, ss = pqc_amd.crypto_kem_enc(pk, "frodo976")
    assert pqc_amd.crypto_kem_dec(ct, sk, "frodo976") == ss
    assert O.frodo_dec(ct, sk) == ss
    pqc_amd.randombytes_system()


def test_pipelined_chunks_equal_unsplit_and_oracle(keys):
    """With frodo_chunk = 4096, batches above 4096 under one key run in
    4096-request chunks on four streams.  9000 requests (chunks 4096, 4096, 808; the last partial) must
    give the bytes of 4096-request batches over the same inputs, which run
    unsplit on one stream, and the oracle's on samples from every chunk."""
    import pqc_amd
    n = 9000
    say("frodo pipelined n", n)
    pk, sk = keys
    rng = np.random.default_rng(23)
    mu = rng.integers(0, 256, (n, 24), dtype=np.uint8)
    from pqc_amd import tuner as T
    big = pqc_amd.Context(["frodo976"], max_batch=n)
    try:
        T.put_tuning(big, frodo_chunk=4096)  # chunks of 4096 on four streams
        big.frodo_load_key(np.frombuffer(pk, np.uint8).copy(), False, host=True)
        d_mu = pqc_amd.DeviceBuffer.from_array(mu)
        d_ct, d_ss = pqc_amd.DeviceBuffer(n * CT), pqc_amd.DeviceBuffer(n * 24)
        big.frodo_enc(n, d_ct, d_ss, d_mu)
        ct, ss = d_ct.download((n, CT)), d_ss.download((n, 24))
        bad = ct.copy()
        bad[::13, 777] ^= 0x01
        big.frodo_load_key(np.frombuffer(sk, np.uint8).copy(), True, host=True)
        d_bad, d_ss2 = pqc_amd.DeviceBuffer.from_array(bad), pqc_amd.DeviceBuffer(n * 24)
        big.frodo_dec(n, d_ss2, d_bad)
        ss2 = d_ss2.download((n, 24))
    finally:
        big.close()
    # unsplit 4096-request batches over the same inputs
    one = pqc_amd.Context(["frodo976"], max_batch=4096)
    try:
        for lo in (0, 4096, n - 4096):
            ct1, ss1 = one.frodo_enc_host(pk, mu[lo:lo + 4096])
            assert (ct1 == ct[lo:lo + 4096]).all() and (ss1 == ss[lo:lo + 4096]).all(), lo
            one.frodo_load_key(np.frombuffer(sk, np.uint8).copy(), True, host=True)
            assert (one.frodo_dec_host(sk, bad[lo:lo + 4096]) == ss2[lo:lo + 4096]).all(), lo
    finally:
        one.close()
    same = (ss2 == ss).all(axis=1)
    assert (~same[::13]).all() and same[np.arange(n) % 13 != 0].all()
    idx = np.array([0, 4095, 4096, 8191, 8192, 8999])
    oct_, oss = O.frodo_enc_batch(pk, mu[idx], nthreads=8)
    assert (ct[idx] == oct_).all() and (ss[idx] == oss).all()
    assert (ss2[idx] == O.frodo_dec_batch(bad[idx], sk, nthreads=8)).all()


@pytest.mark.gpu
def test_lane_stream_large_call_equals_pair_stream_and_oracle(keys):
    """A shared-key decapsulation call of 50 000 requests runs the noise stream and
    the F prefix with one lane per request (and samples plain stream words);
    4096-request calls over the same inputs run them on lane pairs.  Decapsulation
    (with tampered ciphertexts) must give the same bytes both ways, and the
    oracle's on samples across the batch; encapsulation (lane pairs at every
    size) likewise."""
    import pqc_amd
    n = 50000
    say("frodo lane stream n", n)
    pk, sk = keys
    rng = np.random.default_rng(29)
    mu = rng.integers(0, 256, (n, 24), dtype=np.uint8)
    big = pqc_amd.Context(["frodo976"], max_batch=n)
    try:
        big.frodo_load_key(np.frombuffer(pk, np.uint8).copy(), False, host=True)
        d_mu = pqc_amd.DeviceBuffer.from_array(mu)
        d_ct, d_ss = pqc_amd.DeviceBuffer(n * CT), pqc_amd.DeviceBuffer(n * 24)
        big.frodo_enc(n, d_ct, d_ss, d_mu)
        ct, ss = d_ct.download((n, CT)), d_ss.download((n, 24))
        del d_ct
        bad = ct.copy()
        bad[::11, 3] ^= 0x10          # in B'
        bad[5::17, CT - 30] ^= 0x02   # in C
        big.frodo_load_key(np.frombuffer(sk, np.uint8).copy(), True, host=True)
        d_bad, d_ss2 = pqc_amd.DeviceBuffer.from_array(bad), pqc_amd.DeviceBuffer(n * 24)
        big.frodo_dec(n, d_ss2, d_bad)
        ss2 = d_ss2.download((n, 24))
    finally:
        big.close()
    one = pqc_amd.Context(["frodo976"], max_batch=4096)
    try:
        for lo in (0, 24576, n - 4096):
            ct1, ss1 = one.frodo_enc_host(pk, mu[lo:lo + 4096])
            assert (ct1 == ct[lo:lo + 4096]).all() and (ss1 == ss[lo:lo + 4096]).all(), lo
            one.frodo_load_key(np.frombuffer(sk, np.uint8).copy(), True, host=True)
            assert (one.frodo_dec_host(sk, bad[lo:lo + 4096]) == ss2[lo:lo + 4096]).all(), lo
    finally:
        one.close()
    tampered = np.zeros(n, bool)
    tampered[::11] = True
    tampered[5::17] = True
    same = (ss2 == ss).all(axis=1)
    assert (~same[tampered]).all() and same[~tampered].all()
    idx = np.array([0, 5, 63, 64, 12345, 33000, 49151, 49152, n - 1])
    oct_, oss = O.frodo_enc_batch(pk, mu[idx], nthreads=8)
    assert (ct[idx] == oct_).all() and (ss[idx] == oss).all()
    assert (ss2[idx] == O.frodo_dec_batch(bad[idx], sk, nthreads=8)).all()


@pytest.mark.gpu
def test_autotuner_chooses_stream_lane_layout(keys):
    """The Frodo tuning family measures the noise stream's Keccak lane layout
    (lane pairs or one state per lane) beside the block size: at 4096 requests
    it must pick lane pairs (one wave per SIMD would idle three quarters of the
    GPU on single lanes), at 65 536 single lanes (lane pairs spend 2 x 107
    against 180 instructions per round once every SIMD holds several waves).
    Decapsulation under the pick equals the default's."""
    import pqc_amd
    from pqc_amd import tuner as T
    pk, sk = keys
    for n, want_lanes in ((4096, False), (65536, True)):
        say("frodo lane layout autotune n", n)
        rng = np.random.default_rng(31)
        ctx = pqc_amd.Context(["frodo976"], max_batch=n)
        try:
            st = pqc_amd.Stream()
            ctx.frodo_load_key(np.frombuffer(pk, np.uint8).copy(), False, host=True)
            d_mu = pqc_amd.DeviceBuffer.from_array(rng.integers(0, 256, (n, 24), dtype=np.uint8))
            d_ct, d_ss, d_ss2 = (pqc_amd.DeviceBuffer(n * b) for b in (CT, 24, 24))
            ctx.frodo_enc(n, d_ct, d_ss, d_mu)
            ctx.frodo_load_key(np.frombuffer(sk, np.uint8).copy(), True, host=True)
            res = T.autotune(ctx, T.request("frodo", n, op=2, budget=0, reps=3),
                             lambda s: ctx.frodo_dec(n, d_ss2, d_ct, stream=s), st)
            pick = res.best
            lanes = pick.frodo_lane_min > 0 and n >= pick.frodo_lane_min or (pick.frodo_lane_min == 0 and n >= 49152)
            say("  pick", pick.frodo_stream_block, pick.frodo_lane_min, "%.3f ms" % res.best_ms, "lanes", lanes)
            assert res.measured == 6
            assert lanes == want_lanes
            ctx.frodo_dec(n, d_ss2, d_ct, stream=st)
            st.synchronize()
            assert (d_ss2.download((n, 24)) == d_ss.download((n, 24))).all()
        finally:
            ctx.close()
