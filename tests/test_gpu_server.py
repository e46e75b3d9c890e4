"""The serving loop: every result is the same alone and in the mix (continuous
batching with several batches in flight), equal to the oracle, equal between the
fused and the plain paths on one trace, and admission holds to the HBM budget."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

K_OF = {1: 2, 2: 3, 3: 4}
FRODO = 5
ALGS = ["kyber512", "kyber768", "kyber1024", "frodo976"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def srv():
    from pqc_amd.server import Server
    s = Server(ALGS, keys_per_alg=8, key_seed=3, max_batch=128, slots=3, arena_bytes=16 << 20, max_wait_us=100)
    yield s
    s.close()


def make_requests(srv, n, seed, tamper_every=0):
    """n requests: algs 1:2:1 Kyber + some Frodo, encaps and decaps; decaps
    ciphertexts made by the oracle under the server's public keys."""
    rng = np.random.default_rng(seed)
    algs = rng.choice([1, 2, 2, 3, 5], size=n).astype(np.int32)
    ops = rng.choice([1, 2], size=n).astype(np.int32)
    kid = rng.integers(0, 8, size=n).astype(np.uint32)
    inputs, expect = [], []
    for i in range(n):
        a, op, k = int(algs[i]), int(ops[i]), int(kid[i])
        pk, sk = srv.key(a, k)
        if op == 1:
            coins = rng.integers(0, 256, 24 if a == FRODO else 32, dtype=np.uint8).tobytes()
            inputs.append(coins)
            ct, ss = (O.frodo_enc(pk.tobytes(), coins) if a == FRODO else O.kyber_enc(K_OF[a], pk.tobytes(), coins))
            expect.append(ct + ss)
        else:
            coins = rng.integers(0, 256, 24 if a == FRODO else 32, dtype=np.uint8).tobytes()
            ct, _ = (O.frodo_enc(pk.tobytes(), coins) if a == FRODO else O.kyber_enc(K_OF[a], pk.tobytes(), coins))
            if tamper_every and i % tamper_every == 0:
                ct = bytearray(ct)
                ct[11] ^= 0x20
                ct = bytes(ct)
            inputs.append(ct)
            expect.append(O.frodo_dec(ct, sk.tobytes()) if a == FRODO else O.kyber_dec(K_OF[a], ct, sk.tobytes()))
    return algs, ops, kid, inputs, expect


def test_results_same_alone_and_in_the_mix(srv):
    from pqc_amd.server import RequestSet
    n = 48
    say("server: requests alone")
    algs, ops, kid, inputs, expect = make_requests(srv, n, 1, tamper_every=5)
    alone = RequestSet(algs, ops, kid, inputs)
    for i in range(n):  # one at a time: each is a batch of one
        srv.submit(alone.reqs[i:i + 1])
        srv.drain()
    for i in range(n):
        assert alone.out(i) == expect[i], i
    say("server: the same requests inside a mix of 600")
    balgs, bops, bkid, binputs, _ = make_requests(srv, 552, 2)
    allalgs = np.concatenate([algs, balgs])
    perm = np.random.default_rng(3).permutation(len(allalgs))
    mixed = RequestSet(allalgs[perm], np.concatenate([ops, bops])[perm], np.concatenate([kid, bkid])[perm],
                       [(inputs + binputs)[j] for j in perm])
    # submitted in waves while earlier batches run
    for w in range(0, len(perm), 100):
        srv.submit(mixed.reqs[w:w + 100])
    srv.drain()
    where = {int(j): pos for pos, j in enumerate(perm)}
    for i in range(n):
        assert mixed.out(where[i]) == alone.out(i), i
    st = srv.stats()
    say(" batches %d, mean batch %.1f" % (st.batches, st.mean_batch))
    assert st.completed == st.submitted


def trace(n, rate_per_s, seed):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1e6 / rate_per_s, size=n)
    return np.cumsum(gaps)


def test_trace_fused_equals_plain():
    from pqc_amd.server import RequestSet, Server
    n = 3000
    say("server trace fused vs plain, n", n)
    rng = np.random.default_rng(7)
    algs = rng.choice([1, 2, 2, 3, 1, 2, 2, 3, 2, 5], size=n).astype(np.int32)
    ops = rng.choice([1, 2], size=n).astype(np.int32)
    kid = rng.integers(0, 16, size=n).astype(np.uint32)
    inputs = []
    for a, op in zip(algs, ops):
        if op == 1:
            inputs.append(rng.integers(0, 256, 24 if a == FRODO else 32, dtype=np.uint8).tobytes())
        else:  # any bytes are a ciphertext (implicit rejection does the full work)
            inputs.append(rng.integers(0, 256, {1: 736, 2: 1088, 3: 1568, 5: 15744}[int(a)], dtype=np.uint8).tobytes())
    arrivals = trace(n, 100000, 8)
    outs = []
    for plain in (False, True):
        s = Server(ALGS, keys_per_alg=16, key_seed=9, max_batch=512, slots=3, arena_bytes=32 << 20,
                   max_wait_us=100, plain=plain)
        rs = RequestSet(algs, ops, kid, inputs)
        st, lat = s.run_trace(arrivals, rs.reqs)
        say(" plain=%s: %.0f ops/s p50 %.0f us p99 %.0f us mean batch %.1f errors %d" %
            (plain, st.ops_per_s, st.p50_us, st.p99_us, st.mean_batch, st.errors))
        assert st.completed == n and st.errors == 0
        assert (lat > 0).all()
        outs.append(rs.outbuf.copy())
        if plain:  # spot-check against the oracle
            for i in (0, 1, 2, n - 1):
                a, op, k = int(algs[i]), int(ops[i]), int(kid[i])
                pk, sk = s.key(a, k)
                if op == 2:
                    want = (O.frodo_dec(inputs[i], sk.tobytes()) if a == FRODO
                            else O.kyber_dec(K_OF[a], inputs[i], sk.tobytes()))
                else:
                    ct, ss = (O.frodo_enc(pk.tobytes(), inputs[i]) if a == FRODO
                              else O.kyber_enc(K_OF[a], pk.tobytes(), inputs[i]))
                    want = ct + ss
                assert rs.out(i) == want, i
        s.close()
    assert (outs[0] == outs[1]).all()


def test_admission_holds_the_budget():
    import pqc_amd
    from pqc_amd.server import RequestSet, Server
    say("server admission by HBM budget")
    budget = 16 << 10  # 16 KiB of request buffers in flight
    s = Server(["kyber768", "frodo976"], keys_per_alg=4, key_seed=1, max_batch=1024, slots=4,
               arena_bytes=4 << 20, max_wait_us=50, hbm_budget=budget)
    n = 300
    rng = np.random.default_rng(4)
    ops = rng.choice([1, 2], size=n).astype(np.int32)
    inputs = [rng.integers(0, 256, 32 if op == 1 else 1088, dtype=np.uint8).tobytes() for op in ops]
    rs = RequestSet(np.full(n, 2, np.int32), ops, rng.integers(0, 4, n).astype(np.uint32), inputs)
    s.submit(rs.reqs)
    s.drain()
    st = s.stats()
    # each Kyber-768 request holds 1120 bytes: at most 14 of them fit the budget at once
    say(" batches %d mean batch %.1f" % (st.batches, st.mean_batch))
    assert st.completed == n and st.mean_batch <= budget // 1120 + 1
    # one Frodo decapsulation needs 15768 bytes: it fits; a smaller budget refuses it
    big = RequestSet(np.array([5], np.int32), np.array([2], np.int32), np.array([0], np.uint32),
                     [bytes(15744)])
    s.submit(big.reqs)
    s.drain()
    s.close()
    s2 = Server(["frodo976"], keys_per_alg=1, key_seed=1, max_batch=16, slots=1, arena_bytes=4 << 20,
                hbm_budget=8 << 10)
    with pytest.raises(pqc_amd.PQCError):
        s2.submit(big.reqs)
    s2.close()


def _oracle_out(srv, a, op, k, inp):
    pk, sk = srv.key(a, k)
    if op == 2:
        return O.frodo_dec(inp, sk.tobytes()) if a == FRODO else O.kyber_dec(K_OF[a], inp, sk.tobytes())
    ct, ss = O.frodo_enc(pk.tobytes(), inp) if a == FRODO else O.kyber_enc(K_OF[a], pk.tobytes(), inp)
    return ct + ss


def test_graph_mode_equals_eager_and_oracle():
    """PQC_SERVER_GRAPH: every batch is captured into a graph and launched once;
    a slot's graph of one batch shape is updated in place for the next batch of
    that shape.  Over a trace of many batches (so every graph is replayed many
    times, with the eager host work between replays) the bytes must equal the
    eager launches' on every request, and the oracle on a sample."""
    from pqc_amd.server import RequestSet, Server
    n = 4000
    say("server graph vs eager, n", n)
    rng = np.random.default_rng(11)
    algs = rng.choice([1, 2, 2, 3, 1, 2, 2, 3, 2, 5], size=n).astype(np.int32)
    ops = rng.choice([1, 2], size=n).astype(np.int32)
    kid = rng.integers(0, 16, size=n).astype(np.uint32)
    arrivals = trace(n, 200000, 12)
    outs, sets = [], []
    for graph in (False, True):
        s = Server(ALGS, keys_per_alg=16, key_seed=13, max_batch=128, slots=3, arena_bytes=32 << 20,
                   max_wait_us=50, graph=graph)
        rs = RequestSet.random(algs, ops, kid, np.random.default_rng(14))
        st, lat = s.run_trace(arrivals, rs.reqs)
        say(" graph=%s: %.0f ops/s p50 %.0f us p99 %.0f us mean batch %.1f errors %d" %
            (graph, st.ops_per_s, st.p50_us, st.p99_us, st.mean_batch, st.errors))
        assert st.completed == n and st.errors == 0
        assert s.stats().batches >= 32  # at most 128 requests per batch
        outs.append(rs.outbuf.copy())
        sets.append(rs)
        if graph:
            for i in list(range(0, n, 97))[:40]:
                a, op, k = int(algs[i]), int(ops[i]), int(kid[i])
                inp = rs.inbuf[rs.in_off[i]:rs.in_off[i + 1]].tobytes()
                assert rs.out(i) == _oracle_out(s, a, op, k, inp), i
        s.close()
    assert (sets[0].inbuf == sets[1].inbuf).all()
    assert (outs[0] == outs[1]).all()


def test_config_shape_slice_of_2m_trace():
    """The serving configuration's shape: a 0.05 s slice of the Poisson trace at
    2 M requests/s (90% Kyber-512/768/1024 at 1:2:1, 10% FrodoKEM-976, encaps and
    decaps 50/50) against 1024 device-resident key pairs per set.  Fused (graph
    mode) and plain servers must complete every request without error and give
    the same bytes; 50 requests are checked against the oracle."""
    from pqc_amd.server import RequestSet, Server
    n = 100000  # 0.05 s at 2 M/s
    say("server config-shape slice, n", n)
    rng = np.random.default_rng(21)
    arrivals = trace(n, 2e6, 22)
    kyber = rng.choice(np.array([1, 2, 2, 3], np.int32), size=n)
    algs = np.where(rng.random(n) < 0.9, kyber, FRODO).astype(np.int32)
    ops = rng.choice(np.array([1, 2], np.int32), size=n)
    kid = rng.integers(0, 1024, size=n).astype(np.uint32)
    outs = []
    for plain in (False, True):
        s = Server(ALGS, keys_per_alg=1024, key_seed=0, max_batch=16384, slots=4, arena_bytes=96 << 20,
                   max_wait_us=200, plain=plain, graph=not plain)
        rs = RequestSet.random(algs, ops, kid, np.random.default_rng(23))
        st, lat = s.run_trace(arrivals, rs.reqs)
        say(" plain=%s: %.0f ops/s p50 %.0f us p99 %.0f us mean batch %.1f errors %d" %
            (plain, st.ops_per_s, st.p50_us, st.p99_us, st.mean_batch, st.errors))
        assert st.completed == n and st.errors == 0
        outs.append(rs.outbuf.copy())
        if not plain:
            pick = np.concatenate([np.flatnonzero(algs == a)[:12] for a in (1, 2, 3)] +
                                  [np.flatnonzero(algs == FRODO)[:14]])
            assert len(pick) == 50
            for i in pick:
                a, op, k = int(algs[i]), int(ops[i]), int(kid[i])
                inp = rs.inbuf[rs.in_off[i]:rs.in_off[i + 1]].tobytes()
                assert rs.out(i) == _oracle_out(s, a, op, k, inp), i
        s.close()
    assert (outs[0] == outs[1]).all()
