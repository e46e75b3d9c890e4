"""Mixed Kyber-512/768/1024 batches through the fused persistent kernels: parity
with the CPU oracle, with the single-parameter-set fused and plain calls on the
same arenas, descriptor validation, and graph capture/replay."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

K_OF = {1: 2, 2: 3, 3: 4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_algs(n, seed):
    """Kyber-512/768/1024 at 1:2:1 in a seeded shuffle."""
    base = np.array([1, 2, 2, 3] * ((n + 3) // 4), dtype=np.int32)[:n]
    return np.random.default_rng(seed).permutation(base)


@pytest.fixture(scope="module")
def pqc():
    import pqc_amd
    return pqc_amd


@pytest.fixture(scope="module")
def ctx(pqc):
    c = pqc.Context(["kyber512", "kyber768", "kyber1024"], max_batch=65536)
    yield c
    c.close()


def fill_keys_on_gpu(pqc, ctx, lay, ar, seed):
    """Key pairs for every request, made per parameter set by the dense keypair
    call straight into the arenas' regions."""
    rng = np.random.default_rng(seed)
    kc = rng.integers(0, 256, (lay.nbytes["kcoins"],), dtype=np.uint8)
    ar.buf["kcoins"].upload(kc)
    for a in (1, 2, 3):
        cnt = lay.count[a]
        if cnt:
            ctx.keypair_derand(a, cnt, ar.region_ptr("pk", a), ar.region_ptr("sk", a), ar.region_ptr("kcoins", a))
    coins = rng.integers(0, 256, (lay.nbytes["coins"],), dtype=np.uint8)
    ar.buf["coins"].upload(coins)


def test_mixed_small_vs_oracle(pqc, ctx):
    n = 200
    say("mixed small n", n)
    lay = pqc.MixedLayout(mixed_algs(n, 1), np.random.default_rng(2))
    ar = pqc.DeviceArenas(lay)
    fill_keys_on_gpu(pqc, ctx, lay, ar, 3)
    ctx.kem_enc_mixed(n, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"])
    pk, sk, coins = ar.download("pk"), ar.download("sk"), ar.download("coins")
    ct, ss = ar.download("ct"), ar.download("ss")
    for i in range(n):
        k = K_OF[int(lay.algs[i])]
        oct_, oss = O.kyber_enc(k, lay.element(pk, "pk", i).tobytes(), lay.element(coins, "coins", i).tobytes())
        assert lay.element(ct, "ct", i).tobytes() == oct_, i
        assert lay.element(ss, "ss", i).tobytes() == oss, i
    # decapsulation, with a tampered ciphertext every 9th request
    bad = ct.copy()
    for i in range(0, n, 9):
        off = int(lay.offsets("ct")[i])
        bad[off + 5] ^= 0x10
    ar.buf["ct"].upload(bad)
    ar.buf["ss"].upload(np.zeros(lay.nbytes["ss"], np.uint8))
    ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ar.buf["ss"])
    ss2 = ar.download("ss")
    for i in range(n):
        k = K_OF[int(lay.algs[i])]
        want = O.kyber_dec(k, lay.element(bad, "ct", i).tobytes(), lay.element(sk, "sk", i).tobytes())
        assert lay.element(ss2, "ss", i).tobytes() == want, i
        assert (lay.element(ss2, "ss", i).tobytes() == lay.element(ss, "ss", i).tobytes()) == (i % 9 != 0), i
    assert ctx.take_status() == 0


def dense_run(pqc, ctx, lay, ar, op, plain):
    """The same requests through the per-parameter-set calls on the arenas' regions."""
    for a in (1, 2, 3):
        cnt = lay.count[a]
        if not cnt:
            continue
        if op == "enc":
            ctx.enc_derand(a, cnt, ar.region_ptr("ct", a), ar.region_ptr("ss", a), ar.region_ptr("pk", a),
                           ar.region_ptr("coins", a), plain=plain)
        else:
            ctx.dec(a, cnt, ar.region_ptr("ss", a), ar.region_ptr("ct", a), ar.region_ptr("sk", a), plain=plain)


@pytest.mark.parametrize("n", [3000, 65536])
def test_mixed_equals_dense_fused_and_plain(pqc, ctx, n):
    say("mixed vs dense n", n)
    lay = pqc.MixedLayout(mixed_algs(n, 10 + n), np.random.default_rng(11))
    ar = pqc.DeviceArenas(lay)
    fill_keys_on_gpu(pqc, ctx, lay, ar, 12)
    ctx.kem_enc_mixed(n, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"])
    ct_m, ss_m = ar.download("ct"), ar.download("ss")
    for plain in (False, True):
        say(" dense enc plain", plain)
        dense_run(pqc, ctx, lay, ar, "enc", plain)
        assert (ar.download("ct") == ct_m).all() and (ar.download("ss") == ss_m).all()
    # tamper every 7th ciphertext, then decapsulate three ways
    bad = ct_m.copy()
    offs = lay.offsets("ct")
    for i in range(0, n, 7):
        bad[int(offs[i]) + 3] ^= 0x01
    ar.buf["ct"].upload(bad)
    ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ar.buf["ss"])
    ss_d = ar.download("ss")
    sso = lay.offsets("ss")
    same = np.array([(ss_d[int(o):int(o) + 32] == ss_m[int(o):int(o) + 32]).all() for o in sso])
    assert (same == (np.arange(n) % 7 != 0)).all()
    for plain in (False, True):
        say(" dense dec plain", plain)
        ar.buf["ss"].upload(np.zeros(lay.nbytes["ss"], np.uint8))
        dense_run(pqc, ctx, lay, ar, "dec", plain)
        assert (ar.download("ss") == ss_d).all()
    assert ctx.take_status() == 0
    # a sample against the oracle
    sk = ar.download("sk")
    for i in list(range(0, n, max(1, n // 12)))[:12]:
        k = K_OF[int(lay.algs[i])]
        want = O.kyber_dec(k, lay.element(bad, "ct", i).tobytes(), lay.element(sk, "sk", i).tobytes())
        assert lay.element(ss_d, "ss", i).tobytes() == want, i


@pytest.fixture(params=["lowlat", "tiles"])
def path(request, ctx):
    """Small mixed batches run one workgroup per descriptor by default; "tiles"
    forces the classify + bucket + tile kernels (lowlat_max = -1)."""
    from pqc_amd import tuner as T
    T.put_tuning(ctx, lowlat_max=0 if request.param == "lowlat" else -1)
    yield request.param
    T.put_tuning(ctx, lowlat_max=0)


def test_mixed_rejects_bad_descriptors(pqc, ctx, path):
    n = 100
    say("mixed bad descriptors", path)
    lay = pqc.MixedLayout(mixed_algs(n, 20), np.random.default_rng(21))
    ar = pqc.DeviceArenas(lay)
    fill_keys_on_gpu(pqc, ctx, lay, ar, 22)
    ctx.kem_enc_mixed(n, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"])
    ss_ok = ar.download("ss")
    assert ctx.take_status() == 0
    reqs = lay.requests("sk")
    reqs["alg"][3] = 4            # NewHope id: not a Kyber set
    reqs["alg"][4] = 9            # unknown id
    reqs["ct_off"][5] += 4        # unaligned
    reqs["key_off"][6] = 1 << 40  # outside the arena
    reqs["ss_off"][7] = 2**64 - 1024  # wraps past 2^64 to 1 KiB below the arena
    reqs["key_off"][8] = 2**64 - 8
    d_reqs = pqc.DeviceBuffer.from_array(reqs)
    ar.buf["ss"].upload(np.full(lay.nbytes["ss"], 0xAA, np.uint8))
    ctx.kem_dec_mixed(n, d_reqs, ar.buf["ct"], ar.buf["sk"], ar.buf["ss"])
    assert ctx.take_status() & 2
    ss = ar.download("ss")
    sso = lay.offsets("ss")
    for i in range(n):
        got = ss[int(sso[i]):int(sso[i]) + 32]
        if i in (3, 4, 5, 6, 7, 8):
            assert (got == 0xAA).all(), i
        else:
            assert (got == ss_ok[int(sso[i]):int(sso[i]) + 32]).all(), i
    assert ctx.take_status() == 0


def test_graph_replay_mixed_dec(pqc, ctx, path):
    n = 1000
    say("graph replay mixed dec", path)
    lay = pqc.MixedLayout(mixed_algs(n, 30), np.random.default_rng(31))
    ar = pqc.DeviceArenas(lay)
    fill_keys_on_gpu(pqc, ctx, lay, ar, 32)
    ctx.kem_enc_mixed(n, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"])
    ct0 = ar.download("ct")
    stream = pqc.Stream()
    out = pqc.DeviceBuffer(lay.nbytes["ss"])
    g = pqc.Graph.capture(stream, lambda: ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], out,
                                                            stream=stream))
    offs = lay.offsets("ct")
    for rep in range(4):
        bad = ct0.copy()
        for i in range(rep, n, 5 + rep):
            bad[int(offs[i]) + 7 * rep] ^= 0x04
        ar.buf["ct"].upload(bad)
        # eager reference into the ss arena, then the graph into `out`
        ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ar.buf["ss"], stream=stream)
        stream.synchronize()
        want = ar.download("ss")
        out.upload(np.zeros(lay.nbytes["ss"], np.uint8))
        g.launch(stream)
        stream.synchronize()
        assert (out.download((lay.nbytes["ss"],)) == want).all(), rep
    g.destroy()
    assert ctx.take_status() == 0


def test_mixed_tile32_many_tiles_per_wave(pqc, ctx):
    """Mixed encapsulation and decapsulation in 32-request tiles on a grid of 3
    waves, each running many tiles of all three parameter sets in turn through
    one A^T scratch region: the same bytes as the default grid."""
    from pqc_amd import tuner as T
    n = 700
    say("mixed tile32 grid 3 n", n)
    lay = pqc.MixedLayout(mixed_algs(n, 77), np.random.default_rng(78))
    ar = pqc.DeviceArenas(lay)
    fill_keys_on_gpu(pqc, ctx, lay, ar, 79)
    offs = lay.offsets("ct")
    outs = []
    try:
        T.put_tuning(ctx, lowlat_max=-1, mixed_order=0, fused_dec_tile=0, fused_enc_tile=0)
        for grid in (0, 3):
            ctx.set_tuning(grid=grid)
            ar.buf["ct"].upload(np.zeros(lay.nbytes["ct"], np.uint8))
            ctx.kem_enc_mixed(n, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"])
            ct, ss = ar.download("ct"), ar.download("ss")
            bad = ct.copy()
            for i in range(0, n, 3):
                bad[int(offs[i]) + 5] ^= 0x08
            ar.buf["ct"].upload(bad)
            ar.buf["ss"].upload(np.zeros(lay.nbytes["ss"], np.uint8))
            ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ar.buf["ss"])
            outs.append((ct, ss, ar.download("ss")))
            assert ctx.take_status() == 0
    finally:
        ctx.set_tuning(grid=0)
        T.put_tuning(ctx, lowlat_max=0, mixed_order=0, fused_dec_tile=0, fused_enc_tile=0)
    for a, b in zip(outs[0], outs[1]):
        assert (a == b).all()
    sso = lay.offsets("ss")
    same = np.array([(outs[0][2][int(o):int(o) + 32] == outs[0][1][int(o):int(o) + 32]).all() for o in sso])
    assert (same == (np.arange(n) % 3 != 0)).all()


@pytest.mark.parametrize("n", [5, 65, 3000, 65536])
def test_mixed_tile_orders_agree(pqc, ctx, n):
    """Every tile order of the 64-request tiles (0: K 4,3,2; 1: K 2,3,4; 2:
    round-robin; 3: K 4,3,2 with 32-request tiles for K = 4 and K = 2), and the
    encapsulation and decapsulation in 32-request tiles at two waves per SIMD
    (the defaults, fused_enc_tile and fused_dec_tile 0; the order then applies
    to the 64-request tiles only), alone and mixed with the other tile size,
    give the same ciphertexts and shared secrets, through the tile kernels
    (lowlat_max = -1), at ragged and full sizes; tampered ciphertexts included."""
    from pqc_amd import tuner as T
    say("mixed tile orders n", n)
    lay = pqc.MixedLayout(mixed_algs(n, 30 + n), np.random.default_rng(31))
    ar = pqc.DeviceArenas(lay)
    fill_keys_on_gpu(pqc, ctx, lay, ar, 32)
    offs = lay.offsets("ct")
    ref = None
    try:
        for order, dec_tile, enc_tile in ((0, 0, 0), (0, 64, 64), (1, 64, 64), (2, 64, 64), (3, 64, 64),
                                          (3, 0, 64), (1, 64, 0)):
            say(" order", order, "dec tile", dec_tile, "enc tile", enc_tile)
            T.put_tuning(ctx, lowlat_max=-1, mixed_order=order, fused_dec_tile=dec_tile, fused_enc_tile=enc_tile)
            ar.buf["ct"].upload(np.zeros(lay.nbytes["ct"], np.uint8))
            ctx.kem_enc_mixed(n, ar.enc_reqs, ar.buf["ct"], ar.buf["pk"], ar.buf["ss"], ar.buf["coins"])
            ct, ss = ar.download("ct"), ar.download("ss")
            bad = ct.copy()
            for i in range(0, n, 5):
                bad[int(offs[i]) + 11] ^= 0x40
            ar.buf["ct"].upload(bad)
            ar.buf["ss"].upload(np.zeros(lay.nbytes["ss"], np.uint8))
            ctx.kem_dec_mixed(n, ar.dec_reqs, ar.buf["ct"], ar.buf["sk"], ar.buf["ss"])
            got = (ct, ss, ar.download("ss"))
            assert ctx.take_status() == 0
            if ref is None:
                ref = got
            else:
                for a, b in zip(ref, got):
                    assert (a == b).all(), (order, dec_tile, enc_tile)
    finally:
        T.put_tuning(ctx, lowlat_max=0, mixed_order=0, fused_dec_tile=0, fused_enc_tile=0)
    sso = lay.offsets("ss")
    same = np.array([(ref[2][int(o):int(o) + 32] == ref[1][int(o):int(o) + 32]).all() for o in sso])
    assert (same == (np.arange(n) % 5 != 0)).all()
