"""Python mirror of the serving loop (pqc_server_* in include/pqc_api.h).  The
scheduler, the key store and the trace replay run in the library."""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check
from .kem import alg_id, sizes
from .tuner import Tuning

PLAIN = 1
GRAPH = 2

REQ_DTYPE = np.dtype([("alg", "<i4"), ("op", "<i4"), ("key_id", "<u4"), ("reserved", "<u4"), ("in", "<u8"),
                      ("out", "<u8"), ("tag", "<u8")])
assert REQ_DTYPE.itemsize == 40


class Config(C.Structure):
    _fields_ = [("alg_mask", C.c_uint32), ("keys_per_alg", C.c_uint32), ("key_seed", C.c_uint64),
                ("max_batch", C.c_uint32), ("slots", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("arena_bytes", C.c_uint64), ("max_wait_us", C.c_uint32), ("flags", C.c_uint32),
                ("tuning", Tuning)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("batch", C.c_uint32), ("submit_us", C.c_double),
                ("done_us", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("mean_batch", C.c_double), ("inflight_bytes", C.c_uint64), ("pack_us", C.c_double),
                ("device_us", C.c_double), ("retire_us", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TraceStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("errors", C.c_uint64), ("duration_s", C.c_double),
                ("ops_per_s", C.c_double), ("p50_us", C.c_double), ("p90_us", C.c_double), ("p99_us", C.c_double),
                ("p999_us", C.c_double), ("max_us", C.c_double), ("mean_batch", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def in_out_sizes(alg, op):
    """(input bytes, output bytes) of one request: encaps coins -> ct || ss, decaps ct -> ss."""
    _, _, ct, ss = sizes(alg)
    if op == 1:
        return (24 if alg_id(alg) == L.ALG_FRODO976 else 32), ct + ss
    return ct, ss


class Server:
    def __init__(self, algs, keys_per_alg=1024, key_seed=0, max_batch=8192, slots=3, hbm_budget=0,
                 arena_bytes=64 << 20, max_wait_us=200, plain=False, tuning=None, graph=False):
        cfg = Config()
        for a in algs:
            cfg.alg_mask |= 1 << alg_id(a)
        cfg.keys_per_alg = keys_per_alg
        cfg.key_seed = key_seed
        cfg.max_batch = max_batch
        cfg.slots = slots
        cfg.hbm_budget = hbm_budget
        cfg.arena_bytes = arena_bytes
        cfg.max_wait_us = max_wait_us
        cfg.flags = (PLAIN if plain else 0) | (GRAPH if graph else 0)
        if tuning is not None:
            cfg.tuning = tuning
        h = C.c_void_p()
        check(L.lib().pqc_server_create(C.byref(h), C.byref(cfg)), "pqc_server_create")
        self.handle = h.value

    def key(self, alg, key_id):
        pkb, skb, _, _ = sizes(alg)
        pk, sk = np.zeros(pkb, np.uint8), np.zeros(skb, np.uint8)
        check(L.lib().pqc_server_key(self.handle, alg_id(alg), key_id, pk.ctypes.data_as(C.c_void_p),
                                     sk.ctypes.data_as(C.c_void_p)), "pqc_server_key")
        return pk, sk

    def submit(self, reqs):
        reqs = np.ascontiguousarray(reqs, dtype=REQ_DTYPE)
        check(L.lib().pqc_server_submit(self.handle, reqs.ctypes.data_as(C.c_void_p), len(reqs)), "pqc_server_submit")

    def poll(self, cap=4096, wait_us=0):
        buf = (Completion * cap)()
        got = C.c_size_t()
        check(L.lib().pqc_server_poll(self.handle, buf, cap, C.byref(got), wait_us), "pqc_server_poll")
        return [buf[i] for i in range(got.value)]

    def drain(self):
        check(L.lib().pqc_server_drain(self.handle), "pqc_server_drain")

    def stats(self):
        st = Stats()
        check(L.lib().pqc_server_get_stats(self.handle, C.byref(st)), "pqc_server_get_stats")
        return st

    def run_trace(self, arrival_us, reqs):
        arrival_us = np.ascontiguousarray(arrival_us, dtype=np.float64)
        reqs = np.ascontiguousarray(reqs, dtype=REQ_DTYPE)
        lat = np.zeros(len(reqs), np.float64)
        st = TraceStats()
        check(L.lib().pqc_server_run_trace(self.handle, len(reqs), arrival_us.ctypes.data_as(C.c_void_p),
                                           reqs.ctypes.data_as(C.c_void_p), lat.ctypes.data_as(C.c_void_p),
                                           C.byref(st)), "pqc_server_run_trace")
        return st, lat

    def close(self):
        if self.handle:
            L.lib().pqc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RequestSet:
    """Host buffers and descriptors for a list of (alg, op, key_id, input bytes).

    Inputs and outputs live in two big byte arrays that stay alive with the set;
    out(i) is request i's output slice."""

    def __init__(self, algs, ops, key_ids, inputs):
        n = len(algs)
        self.n = n
        sz = [in_out_sizes(int(a), int(o)) for a, o in zip(algs, ops)]
        self.in_off = np.zeros(n + 1, np.int64)
        self.out_off = np.zeros(n + 1, np.int64)
        self.in_off[1:] = np.cumsum([s[0] for s in sz])
        self.out_off[1:] = np.cumsum([s[1] for s in sz])
        self.inbuf = np.zeros(int(self.in_off[-1]) + 8, np.uint8)
        self.outbuf = np.empty(int(self.out_off[-1]) + 8, np.uint8)
        self.outbuf.fill(0)  # touch every page now, not while the server writes results
        for i in range(n):
            b = np.frombuffer(bytes(inputs[i]), np.uint8)
            assert len(b) == sz[i][0], (i, len(b), sz[i])
            self.inbuf[self.in_off[i]:self.in_off[i + 1]] = b
        self.reqs = np.zeros(n, REQ_DTYPE)
        self.reqs["alg"] = algs
        self.reqs["op"] = ops
        self.reqs["key_id"] = key_ids
        self.reqs["in"] = self.inbuf.ctypes.data + self.in_off[:-1]
        self.reqs["out"] = self.outbuf.ctypes.data + self.out_off[:-1]
        self.reqs["tag"] = np.arange(n)

    @classmethod
    def random(cls, algs, ops, key_ids, rng):
        """Requests with random inputs (coins, or ciphertexts: decapsulating any
        bytes does the full work, implicit rejection included), built without a
        per-request Python loop."""
        self = cls.__new__(cls)
        algs = np.asarray(algs, np.int32)
        ops = np.asarray(ops, np.int32)
        n = len(algs)
        self.n = n
        ct = np.zeros(8, np.int64)
        ss = np.zeros(8, np.int64)
        for a in (L.ALG_KYBER512, L.ALG_KYBER768, L.ALG_KYBER1024, L.ALG_FRODO976):
            _, _, ct[a], ss[a] = sizes(a)
        coins = np.where(algs == L.ALG_FRODO976, 24, 32)
        isz = np.where(ops == 1, coins, ct[algs])
        osz = np.where(ops == 1, ct[algs] + ss[algs], ss[algs])
        self.in_off = np.zeros(n + 1, np.int64)
        self.out_off = np.zeros(n + 1, np.int64)
        self.in_off[1:] = np.cumsum(isz)
        self.out_off[1:] = np.cumsum(osz)
        self.inbuf = rng.integers(0, 256, int(self.in_off[-1]) + 8, dtype=np.uint8)
        self.outbuf = np.empty(int(self.out_off[-1]) + 8, np.uint8)
        self.outbuf.fill(0)  # touch every page now, not while the server writes results
        self.reqs = np.zeros(n, REQ_DTYPE)
        self.reqs["alg"] = algs
        self.reqs["op"] = ops
        self.reqs["key_id"] = key_ids
        self.reqs["in"] = self.inbuf.ctypes.data + self.in_off[:-1]
        self.reqs["out"] = self.outbuf.ctypes.data + self.out_off[:-1]
        self.reqs["tag"] = np.arange(n)
        return self

    def out(self, i):
        return self.outbuf[self.out_off[i]:self.out_off[i + 1]].tobytes()
