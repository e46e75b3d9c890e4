"""Python mirror of the engine's KEM interface.

Names and argument meaning follow the reference's NIST API
(Kyber_GPU_Batched/api.h:18-22: crypto_kem_keypair(pk, sk),
crypto_kem_enc(ct, ss, pk), crypto_kem_dec(ss, ct, sk)) and its batched
internal API (indcpa.h:16-26: a COUNT plus device buffers and a stream).
Errors raise PQCError carrying the C ABI's status code.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check

ALGS = {
    "kyber512": L.ALG_KYBER512,
    "kyber768": L.ALG_KYBER768,
    "kyber1024": L.ALG_KYBER1024,
    "newhope1024cpa": L.ALG_NEWHOPE1024_CPA,
    "frodo976": L.ALG_FRODO976,
    "newhope1024": L.ALG_NEWHOPE1024,
    "newhope512": L.ALG_NEWHOPE512,
    "newhope512cpa": L.ALG_NEWHOPE512_CPA,
}
KYBER_K = {L.ALG_KYBER512: 2, L.ALG_KYBER768: 3, L.ALG_KYBER1024: 4}


def alg_id(alg):
    return ALGS[alg] if isinstance(alg, str) else int(alg)


def sizes(alg):
    """(pk, sk, ct, ss) byte sizes of one element."""
    pk, sk, ct, ss = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
    check(L.lib().pqc_alg_sizes(alg_id(alg), C.byref(pk), C.byref(sk), C.byref(ct), C.byref(ss)), "pqc_alg_sizes")
    return pk.value, sk.value, ct.value, ss.value


def _ptr(a):
    if isinstance(a, DeviceBuffer) or hasattr(a, "ptr"):  # a buffer or a slice of one
        return C.c_void_p(a.ptr)
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(C.c_void_p)
    if a is None:
        return None
    raise TypeError(type(a))


class DeviceBuffer:
    """A device allocation made through the library (hipMalloc)."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(L.lib().pqc_device_malloc(C.byref(p), nbytes), "pqc_device_malloc")
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr, stream)
        return b

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(L.lib().pqc_memcpy_htod(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes,
                                      stream.handle if stream else None), "htod")

    def download(self, shape, dtype=np.uint8, stream=None):
        out = np.empty(shape, dtype=dtype)
        check(L.lib().pqc_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes,
                                      stream.handle if stream else None), "dtoh")
        return out

    def free(self):
        if self.ptr:
            L.lib().pqc_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory from pqc_host_alloc; `array` is a uint8 numpy view.

    Host-pointer calls (PQC_F_HOST_PTRS) on views of it skip the staging copy."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(L.lib().pqc_host_alloc(nbytes, C.byref(p)), "pqc_host_alloc")
        self.ptr = p.value
        self.nbytes = nbytes
        self.array = np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(self.ptr))

    def view(self, shape, dtype=np.uint8):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        assert n <= self.nbytes
        return self.array[:n].view(dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            check(L.lib().pqc_host_free(self.ptr), "pqc_host_free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(L.lib().pqc_stream_create(C.byref(p)), "pqc_stream_create")
        self.handle = p.value

    def synchronize(self):
        check(L.lib().pqc_stream_synchronize(self.handle), "pqc_stream_synchronize")

    def __del__(self):
        try:
            L.lib().pqc_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(L.lib().pqc_event_create(C.byref(p)), "pqc_event_create")
        self.handle = p.value

    def record(self, stream=None):
        check(L.lib().pqc_event_record(self.handle, stream.handle if stream else None), "pqc_event_record")

    def synchronize(self):
        check(L.lib().pqc_event_synchronize(self.handle), "pqc_event_synchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(L.lib().pqc_event_elapsed_ms(C.byref(ms), self.handle, end.handle), "pqc_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            L.lib().pqc_event_destroy(self.handle)
        except Exception:
            pass


# descriptor of one request in a mixed Kyber batch (pqc_kem_req, 40 bytes)
KEM_REQ_DTYPE = np.dtype([("alg", "<i4"), ("reserved", "<u4"), ("key_off", "<u8"), ("ct_off", "<u8"),
                          ("ss_off", "<u8"), ("coins_off", "<u8")])
assert KEM_REQ_DTYPE.itemsize == 40


class Graph:
    """Device-pointer batch calls recorded on a stream and replayed with one launch."""

    def __init__(self, handle):
        self.handle = handle

    @staticmethod
    def capture(stream, fn):
        check(L.lib().pqc_graph_capture_begin(stream.handle), "pqc_graph_capture_begin")
        try:
            fn()
        finally:
            h = C.c_void_p()
            rc = L.lib().pqc_graph_capture_end(stream.handle, C.byref(h))
        check(rc, "pqc_graph_capture_end")
        return Graph(h.value)

    def launch(self, stream):
        check(L.lib().pqc_graph_launch(self.handle, stream.handle), "pqc_graph_launch")

    def destroy(self):
        if self.handle:
            L.lib().pqc_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Context:
    """Owns device scratch for up to max_batch elements (replaces poly_set4)."""

    def __init__(self, algs=("kyber512", "kyber768", "kyber1024"), max_batch=1024, hbm_budget=0):
        mask = 0
        for a in algs:
            mask |= 1 << alg_id(a)
        p = C.c_void_p()
        check(L.lib().pqc_ctx_create(C.byref(p), mask, max_batch, hbm_budget), "pqc_ctx_create")
        self.handle = p.value
        self.max_batch = max_batch

    def close(self):
        if self.handle:
            L.lib().pqc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return L.lib().pqc_ctx_device_bytes(self.handle)

    def set_tuning(self, waves_per_cu=0, grid=0):
        check(L.lib().pqc_ctx_set_tuning(self.handle, waves_per_cu, grid), "pqc_ctx_set_tuning")

    # ---- batched calls on device buffers (asynchronous on `stream`) ----------
    def keypair_derand(self, alg, n, pk, sk, coins, plain=False, stream=None):
        flags = L.F_PLAIN if plain else 0
        check(L.lib().crypto_kem_keypair_derand_batch(self.handle, alg_id(alg), n, _ptr(pk), _ptr(sk), _ptr(coins),
                                                      flags, stream.handle if stream else None), "keypair")

    def enc_derand(self, alg, n, ct, ss, pk, coins, plain=False, stream=None):
        flags = L.F_PLAIN if plain else 0
        check(L.lib().crypto_kem_enc_derand_batch(self.handle, alg_id(alg), n, _ptr(ct), _ptr(ss), _ptr(pk),
                                                  _ptr(coins), flags, stream.handle if stream else None), "enc")

    def dec(self, alg, n, ss, ct, sk, plain=False, stream=None):
        flags = L.F_PLAIN if plain else 0
        check(L.lib().crypto_kem_dec_batch(self.handle, alg_id(alg), n, _ptr(ss), _ptr(ct), _ptr(sk), flags,
                                           stream.handle if stream else None), "dec")

    # ---- host numpy arrays (staged, synchronous) -----------------------------
    def keypair_host(self, alg, coins, plain=False, out=None):
        """out: optional (pk, sk) arrays to write into (e.g. PinnedBuffer views)."""
        pkb, skb, _, _ = sizes(alg)
        coins = np.ascontiguousarray(coins, dtype=np.uint8)
        n = coins.shape[0]
        pk, sk = out if out is not None else (np.empty((n, pkb), np.uint8), np.empty((n, skb), np.uint8))
        assert pk.shape == (n, pkb) and sk.shape == (n, skb)
        flags = L.F_HOST_PTRS | (L.F_PLAIN if plain else 0)
        check(L.lib().crypto_kem_keypair_derand_batch(self.handle, alg_id(alg), n, _ptr(pk), _ptr(sk), _ptr(coins),
                                                      flags, None), "keypair")
        return pk, sk

    def enc_host(self, alg, pk, coins, plain=False):
        _, _, ctb, ssb = sizes(alg)
        pk = np.ascontiguousarray(pk, dtype=np.uint8)
        coins = np.ascontiguousarray(coins, dtype=np.uint8)
        n = pk.shape[0]
        ct = np.empty((n, ctb), np.uint8)
        ss = np.empty((n, ssb), np.uint8)
        flags = L.F_HOST_PTRS | (L.F_PLAIN if plain else 0)
        check(L.lib().crypto_kem_enc_derand_batch(self.handle, alg_id(alg), n, _ptr(ct), _ptr(ss), _ptr(pk),
                                                  _ptr(coins), flags, None), "enc")
        return ct, ss

    def dec_host(self, alg, ct, sk, plain=False, out=None):
        """out: optional ss array to write into (e.g. a PinnedBuffer view)."""
        _, _, _, ssb = sizes(alg)
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        sk = np.ascontiguousarray(sk, dtype=np.uint8)
        n = ct.shape[0]
        ss = out if out is not None else np.empty((n, ssb), np.uint8)
        assert ss.shape == (n, ssb)
        flags = L.F_HOST_PTRS | (L.F_PLAIN if plain else 0)
        check(L.lib().crypto_kem_dec_batch(self.handle, alg_id(alg), n, _ptr(ss), _ptr(ct), _ptr(sk), flags, None),
              "dec")
        return ss


    # ---- NewHope CPA-PKE, N = 1024 (default) or 512 -------------------------------
    def newhope_keypair(self, n, pk, sk, seeds, stream=None, host=False, N=1024, plain=False):
        flags = (L.F_HOST_PTRS if host else 0) | (L.F_PLAIN if plain else 0)
        check(getattr(L.lib(), "pqc_newhope%d_cpa_keypair_batch" % N)(
            self.handle, n, _ptr(pk), _ptr(sk), _ptr(seeds), flags, stream.handle if stream else None),
            "newhope keypair")

    def newhope_enc(self, n, ct, msgs, pk, coins, stream=None, host=False, N=1024, plain=False):
        flags = (L.F_HOST_PTRS if host else 0) | (L.F_PLAIN if plain else 0)
        check(getattr(L.lib(), "pqc_newhope%d_cpa_enc_batch" % N)(
            self.handle, n, _ptr(ct), _ptr(msgs), _ptr(pk), _ptr(coins), flags, stream.handle if stream else None),
            "newhope enc")

    def newhope_dec(self, n, msgs, ct, sk, stream=None, host=False, N=1024, plain=False):
        flags = (L.F_HOST_PTRS if host else 0) | (L.F_PLAIN if plain else 0)
        check(getattr(L.lib(), "pqc_newhope%d_cpa_dec_batch" % N)(
            self.handle, n, _ptr(msgs), _ptr(ct), _ptr(sk), flags, stream.handle if stream else None),
            "newhope dec")

    def newhope_triples_host(self, seeds, msgs, coins, N=1024, plain=False):
        """keypair + enc + dec over host arrays; returns (pk, sk, ct, m')."""
        n = seeds.shape[0]
        poly = 14 * N // 8
        pk = np.empty((n, poly + 32), np.uint8)
        sk = np.empty((n, poly), np.uint8)
        ct = np.empty((n, poly + 3 * N // 8), np.uint8)
        m2 = np.empty((n, 32), np.uint8)
        seeds, msgs, coins = (np.ascontiguousarray(x, dtype=np.uint8) for x in (seeds, msgs, coins))
        self.newhope_keypair(n, pk, sk, seeds, host=True, N=N, plain=plain)
        self.newhope_enc(n, ct, msgs, pk, coins, host=True, N=N, plain=plain)
        self.newhope_dec(n, m2, ct, sk, host=True, N=N, plain=plain)
        return pk, sk, ct, m2

    # ---- mixed Kyber-512/768/1024 batches (device arenas + descriptors) ----------
    def kem_enc_mixed(self, n, reqs, ct, pk, ss, coins, stream=None):
        check(L.lib().pqc_kem_enc_mixed(self.handle, n, _ptr(reqs), _ptr(ct), ct.nbytes, _ptr(pk), pk.nbytes,
                                        _ptr(ss), ss.nbytes, _ptr(coins), coins.nbytes,
                                        stream.handle if stream else None), "pqc_kem_enc_mixed")

    def kem_dec_mixed(self, n, reqs, ct, sk, ss, stream=None):
        check(L.lib().pqc_kem_dec_mixed(self.handle, n, _ptr(reqs), _ptr(ct), ct.nbytes, _ptr(sk), sk.nbytes,
                                        _ptr(ss), ss.nbytes, stream.handle if stream else None), "pqc_kem_dec_mixed")

    def take_status(self, stream=None):
        st = C.c_uint32()
        check(L.lib().pqc_ctx_take_status(self.handle, stream.handle if stream else None, C.byref(st)),
              "pqc_ctx_take_status")
        return st.value

    # ---- FrodoKEM-976: batches under one device-resident key ----------------------
    def frodo_keypair(self, n, pk, sk, randomness, stream=None, host=False):
        flags = L.F_HOST_PTRS if host else 0
        check(L.lib().pqc_frodo976_keypair_batch(self.handle, n, _ptr(pk), _ptr(sk), _ptr(randomness), flags,
                                                 stream.handle if stream else None), "frodo keypair")

    def frodo_load_key(self, key, is_sk, stream=None, host=False):
        flags = L.F_HOST_PTRS if host else 0
        check(L.lib().pqc_frodo976_load_key(self.handle, _ptr(key), 1 if is_sk else 0, flags,
                                            stream.handle if stream else None), "frodo load_key")

    def frodo_enc(self, n, ct, ss, mu, stream=None, host=False):
        flags = L.F_HOST_PTRS if host else 0
        check(L.lib().pqc_frodo976_enc_batch(self.handle, n, _ptr(ct), _ptr(ss), _ptr(mu), flags,
                                             stream.handle if stream else None), "frodo enc")

    def frodo_dec(self, n, ss, ct, stream=None, host=False):
        flags = L.F_HOST_PTRS if host else 0
        check(L.lib().pqc_frodo976_dec_batch(self.handle, n, _ptr(ss), _ptr(ct), flags,
                                             stream.handle if stream else None), "frodo dec")

    def frodo_products(self, n, bpp=None, v=None, stream=None, host=False):
        flags = L.F_HOST_PTRS if host else 0
        check(L.lib().pqc_frodo976_products(self.handle, n, _ptr(bpp), _ptr(v), flags,
                                            stream.handle if stream else None), "frodo products")

    def frodo_gemm(self, n, ct=None, stream=None):
        """Re-run only the shared-A product of the last batch (measurement).  With
        the last decapsulation batch's device ciphertexts, the decapsulation form."""
        check(L.lib().pqc_frodo976_gemm(self.handle, n, _ptr(ct), stream.handle if stream else None), "frodo gemm")

    def frodo_keypair_host(self, randomness):
        randomness = np.ascontiguousarray(randomness, dtype=np.uint8)
        n = randomness.shape[0]
        pk = np.empty((n, 15632), np.uint8)
        sk = np.empty((n, 31296), np.uint8)
        self.frodo_keypair(n, pk, sk, randomness, host=True)
        return pk, sk

    def frodo_enc_host(self, pk, mu):
        """Encapsulate every row of mu under one public key: (ct, ss, B'', V)."""
        mu = np.ascontiguousarray(mu, dtype=np.uint8)
        n = mu.shape[0]
        self.frodo_load_key(np.frombuffer(bytes(pk), np.uint8).copy(), False, host=True)
        ct = np.empty((n, 15744), np.uint8)
        ss = np.empty((n, 24), np.uint8)
        self.frodo_enc(n, ct, ss, mu, host=True)
        return ct, ss

    def frodo_dec_host(self, sk, ct):
        ct = np.ascontiguousarray(ct, dtype=np.uint8)
        n = ct.shape[0]
        self.frodo_load_key(np.frombuffer(bytes(sk), np.uint8).copy(), True, host=True)
        ss = np.empty((n, 24), np.uint8)
        self.frodo_dec(n, ss, ct, host=True)
        return ss

    def frodo_products_host(self, n):
        """B'' (n x 8 x 976) and V (n x 8 x 8) of the last Frodo batch, uint16."""
        bpp = np.empty((n, 8, 976), np.uint16)
        v = np.empty((n, 8, 8), np.uint16)
        self.frodo_products(n, bpp, v, host=True)
        return bpp, v


# ---- NIST single-instance drop-ins (host bytes in, bytes out) -------------------
def _nist(prefix):
    return prefix + "crypto_kem_"


def crypto_kem_keypair(alg="kyber768"):
    pkb, skb, _, _ = sizes(alg)
    pk, sk = (C.c_uint8 * pkb)(), (C.c_uint8 * skb)()
    check(getattr(L.lib(), "pqc_%s_crypto_kem_keypair" % alg)(pk, sk), "crypto_kem_keypair")
    return bytes(pk), bytes(sk)


def crypto_kem_enc(pk, alg="kyber768"):
    _, _, ctb, ssb = sizes(alg)
    ct, ss = (C.c_uint8 * ctb)(), (C.c_uint8 * ssb)()
    check(getattr(L.lib(), "pqc_%s_crypto_kem_enc" % alg)(ct, ss, bytes(pk)), "crypto_kem_enc")
    return bytes(ct), bytes(ss)


def crypto_kem_dec(ct, sk, alg="kyber768"):
    _, _, _, ssb = sizes(alg)
    ss = (C.c_uint8 * ssb)()
    check(getattr(L.lib(), "pqc_%s_crypto_kem_dec" % alg)(ss, bytes(ct), bytes(sk)), "crypto_kem_dec")
    return bytes(ss)


def randombytes_seed(seed):
    L.lib().pqc_randombytes_seed(seed)


def randombytes_system():
    L.lib().pqc_randombytes_system()


def keccak_batch(op, msgs, outlen):
    """op 0 SHA3-256, 1 SHA3-512, 2 SHAKE128, 3 SHAKE256 over n fixed-length messages (device)."""
    msgs = np.ascontiguousarray(msgs, dtype=np.uint8)
    n, inlen = msgs.shape
    din = DeviceBuffer.from_array(msgs) if msgs.size else DeviceBuffer(8)
    dout = DeviceBuffer(n * outlen)
    check(L.lib().pqc_keccak_batch(op, n, dout.ptr, outlen, din.ptr, inlen, None), "pqc_keccak_batch")
    return dout.download((n, outlen))


def newhope_poly_transform(op, polys):
    """op 0 = poly_ntt, 1 = poly_invntt on an (n, 1024) uint16 array (device)."""
    polys = np.ascontiguousarray(polys, dtype=np.uint16)
    d = DeviceBuffer.from_array(polys)
    check(L.lib().pqc_newhope1024_poly_transform(op, polys.shape[0], d.ptr, None), "pqc_newhope1024_poly_transform")
    return d.download(polys.shape, np.uint16)


def newhope_poly_sample(seeds, nonce):
    """poly_sample of each 32-byte seed row with one nonce -> (n, 1024) uint16."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint8)
    n = seeds.shape[0]
    d_seed, d_out = DeviceBuffer.from_array(seeds), DeviceBuffer(n * 2048)
    check(L.lib().pqc_newhope1024_poly_sample(n, d_out.ptr, d_seed.ptr, seeds.shape[1], nonce, None),
          "pqc_newhope1024_poly_sample")
    return d_out.download((n, 1024), np.uint16)


def kyber_poly_transform(op, polys):
    """op 0 = poly_ntt, 1 = poly_invntt on an (n, 256) int16 array (device)."""
    polys = np.ascontiguousarray(polys, dtype=np.int16)
    d = DeviceBuffer.from_array(polys)
    check(L.lib().pqc_kyber_poly_transform(op, polys.shape[0], d.ptr, None), "pqc_kyber_poly_transform")
    return d.download(polys.shape, np.int16)
