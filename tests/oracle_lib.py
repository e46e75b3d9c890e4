"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.  The product path never imports it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

KYBER_SIZES = {}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _LIB = C.CDLL(path)
    return _LIB


class KyberParams(C.Structure):
    _fields_ = [("k", C.c_int), ("du", C.c_int), ("dv", C.c_int),
                ("polyvec_bytes", C.c_size_t), ("polyvec_cbytes", C.c_size_t),
                ("poly_cbytes", C.c_size_t), ("indcpa_pk", C.c_size_t),
                ("indcpa_sk", C.c_size_t), ("ct", C.c_size_t), ("pk", C.c_size_t),
                ("sk", C.c_size_t)]


def kyber_params(k):
    if k not in KYBER_SIZES:
        p = KyberParams()
        assert lib().oracle_kyber_params_get(k, C.byref(p)) == 0
        KYBER_SIZES[k] = p
    return KYBER_SIZES[k]


def buf(n):
    return (C.c_uint8 * n)()


def ptr(a):
    """Pointer to a contiguous numpy array or ctypes buffer."""
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(C.c_void_p)
    return C.cast(a, C.c_void_p)


def drbg(nbytes, seed=0, stream=0, index=0):
    out = buf(nbytes)
    lib().oracle_drbg(out, C.c_size_t(nbytes), C.c_uint64(seed), C.c_uint64(stream), C.c_uint64(index))
    return bytes(out)


def drbg_array(n, nbytes, seed=0, stream=0):
    """n independent nbytes-long DRBG draws: element i uses index i."""
    out = np.empty((n, nbytes), dtype=np.uint8)
    for i in range(n):
        out[i] = np.frombuffer(drbg(nbytes, seed, stream, i), dtype=np.uint8)
    return out


def shake128(data, n):
    out = buf(n)
    lib().oracle_shake128(out, C.c_size_t(n), bytes(data), C.c_size_t(len(data)))
    return bytes(out)


def shake256(data, n):
    out = buf(n)
    lib().oracle_shake256(out, C.c_size_t(n), bytes(data), C.c_size_t(len(data)))
    return bytes(out)


def sha3_256(data):
    out = buf(32)
    lib().oracle_sha3_256(out, bytes(data), C.c_size_t(len(data)))
    return bytes(out)


def sha3_512(data):
    out = buf(64)
    lib().oracle_sha3_512(out, bytes(data), C.c_size_t(len(data)))
    return bytes(out)


def keccakf(state):
    st = (C.c_uint64 * 25)(*state)
    lib().oracle_keccakf1600(st)
    return list(st)


# ---- Kyber ---------------------------------------------------------------
def _i16(a):
    return np.ascontiguousarray(a, dtype=np.int16)


def kyber_ntt(a, reduce=True):
    a = _i16(a).copy()
    (lib().oracle_kyber_ntt if reduce else lib().oracle_kyber_ntt_raw)(ptr(a))
    return a


def kyber_invntt(a):
    a = _i16(a).copy()
    lib().oracle_kyber_invntt(ptr(a))
    return a


def kyber_basemul(a, b):
    r = np.zeros(256, dtype=np.int16)
    lib().oracle_kyber_basemul(ptr(r), ptr(_i16(a)), ptr(_i16(b)))
    return r


def kyber_cbd(b128):
    r = np.zeros(256, dtype=np.int16)
    lib().oracle_kyber_cbd(ptr(r), bytes(b128))
    return r


def kyber_zetas():
    z = np.zeros(128, dtype=np.int16)
    zi = np.zeros(128, dtype=np.int16)
    lib().oracle_kyber_zetas(ptr(z), ptr(zi))
    return z, zi


def kyber_compress(a, d):
    out = buf(32 * d)
    lib().oracle_kyber_poly_compress(out, ptr(_i16(a)), C.c_int(d))
    return bytes(out)


def kyber_decompress(b, d):
    r = np.zeros(256, dtype=np.int16)
    lib().oracle_kyber_poly_decompress(ptr(r), bytes(b), C.c_int(d))
    return r


def kyber_rej_uniform(stream, n=256):
    r = np.zeros(n, dtype=np.int16)
    cnt = lib().oracle_kyber_rej_uniform(ptr(r), C.c_uint(n), bytes(stream), C.c_uint(len(stream)))
    return r[:cnt]


def kyber_gen_matrix(k, seed, transposed):
    a = np.zeros((k, k, 256), dtype=np.int16)
    lib().oracle_kyber_gen_matrix(C.c_int(k), ptr(a), bytes(seed), C.c_int(transposed))
    return a


def kyber_keypair(k, coins64):
    p = kyber_params(k)
    pk, sk = buf(p.pk), buf(p.sk)
    assert lib().oracle_kyber_keypair_derand(C.c_int(k), pk, sk, bytes(coins64)) == 0
    return bytes(pk), bytes(sk)


def kyber_enc(k, pk, coins32):
    p = kyber_params(k)
    ct, ss = buf(p.ct), buf(32)
    assert lib().oracle_kyber_enc_derand(C.c_int(k), ct, ss, bytes(pk), bytes(coins32)) == 0
    return bytes(ct), bytes(ss)


def kyber_dec(k, ct, sk):
    ss = buf(32)
    assert lib().oracle_kyber_dec(C.c_int(k), ss, bytes(ct), bytes(sk)) == 0
    return bytes(ss)


def kyber_indcpa_dec(k, ct, sk):
    m = buf(32)
    assert lib().oracle_kyber_indcpa_dec(C.c_int(k), m, bytes(ct), bytes(sk)) == 0
    return bytes(m)


def kyber_batch(op, k, arrays, nthreads=8):
    """Threaded batch over element-contiguous numpy uint8 arrays."""
    p = kyber_params(k)
    L = lib()
    if op == "keypair":
        coins, = arrays
        n = coins.shape[0]
        pk = np.empty((n, p.pk), np.uint8)
        sk = np.empty((n, p.sk), np.uint8)
        assert L.oracle_kyber_keypair_batch(k, C.c_size_t(n), ptr(pk), ptr(sk), ptr(coins), nthreads) == 0
        return pk, sk
    if op == "enc":
        pk, coins = arrays
        n = pk.shape[0]
        ct = np.empty((n, p.ct), np.uint8)
        ss = np.empty((n, 32), np.uint8)
        assert L.oracle_kyber_enc_batch(k, C.c_size_t(n), ptr(ct), ptr(ss), ptr(pk), ptr(coins), nthreads) == 0
        return ct, ss
    if op == "dec":
        ct, sk = arrays
        n = ct.shape[0]
        ss = np.empty((n, 32), np.uint8)
        assert L.oracle_kyber_dec_batch(k, C.c_size_t(n), ptr(ss), ptr(ct), ptr(sk), nthreads) == 0
        return ss
    raise ValueError(op)


# ---- NewHope-1024 CPA ------------------------------------------------------------
NH_PK, NH_SK, NH_CT = 1824, 1792, 2176
NH_CCA_SK, NH_CCA_CT = 1792 + 1824 + 64, 2176 + 32


def nh_sizes(n=1024):
    """(pk, sk, ct, cca_sk, cca_ct) of NewHope-n: 14-bit polynomials, 3-bit v."""
    poly, cb = 14 * n // 8, 3 * n // 8
    pk, ct = poly + 32, poly + cb
    return pk, poly, ct, poly + pk + 64, ct + 32


def _nh(name, n):
    return getattr(lib(), ("oracle_newhope_" if n == 1024 else "oracle_newhope512_") + name)


def _u16(a):
    return np.ascontiguousarray(a, dtype=np.uint16)


def newhope_tables(n=1024):
    o = np.zeros(n // 2, np.uint16)
    g = np.zeros(n, np.uint16)
    gi = np.zeros(n, np.uint16)
    _nh("tables", n)(ptr(o), ptr(g), ptr(gi))
    return o, g, gi


def newhope_ntt(a, n=1024):
    a = _u16(a).copy()
    assert a.size == n
    _nh("ntt", n)(ptr(a))
    return a


def newhope_invntt(a, n=1024):
    a = _u16(a).copy()
    assert a.size == n
    _nh("invntt", n)(ptr(a))
    return a


def newhope_uniform(seed, n=1024):
    a = np.zeros(n, np.uint16)
    _nh("poly_uniform", n)(ptr(a), bytes(seed))
    return a


def newhope_sample(seed, nonce, n=1024):
    a = np.zeros(n, np.uint16)
    _nh("poly_sample", n)(ptr(a), bytes(seed), C.c_uint8(nonce))
    return a


def newhope_keypair(seed, n=1024):
    pkb, skb, _, _, _ = nh_sizes(n)
    pk, sk = buf(pkb), buf(skb)
    assert _nh("cpa_keypair", n)(pk, sk, bytes(seed)) == 0
    return bytes(pk), bytes(sk)


def newhope_enc(m, pk, coin, n=1024):
    c = buf(nh_sizes(n)[2])
    assert _nh("cpa_enc", n)(c, bytes(m), bytes(pk), bytes(coin)) == 0
    return bytes(c)


def newhope_dec(c, sk, n=1024):
    m = buf(32)
    assert _nh("cpa_dec", n)(m, bytes(c), bytes(sk)) == 0
    return bytes(m)


def newhope_cca_keypair(coins64, n=1024):
    pkb, _, _, skb, _ = nh_sizes(n)
    pk, sk = buf(pkb), buf(skb)
    assert _nh("cca_keypair", n)(pk, sk, bytes(coins64)) == 0
    return bytes(pk), bytes(sk)


def newhope_cca_enc(pk, coins32, n=1024):
    ct, ss = buf(nh_sizes(n)[4]), buf(32)
    assert _nh("cca_enc", n)(ct, ss, bytes(pk), bytes(coins32)) == 0
    return bytes(ct), bytes(ss)


def newhope_cca_dec(ct, sk, n=1024):
    ss = buf(32)
    assert _nh("cca_dec", n)(ss, bytes(ct), bytes(sk)) == 0
    return bytes(ss)


def newhope_triples(seed, msg, coin, nthreads=8, N=1024):
    n = seed.shape[0]
    pkb, skb, ctb, _, _ = nh_sizes(N)
    pk = np.empty((n, pkb), np.uint8)
    sk = np.empty((n, skb), np.uint8)
    ct = np.empty((n, ctb), np.uint8)
    m2 = np.empty((n, 32), np.uint8)
    _nh("triple_batch", N)(C.c_size_t(n), ptr(pk), ptr(sk), ptr(ct), ptr(m2), ptr(seed), ptr(msg),
                                      ptr(coin), nthreads)
    return pk, sk, ct, m2


# ---- FrodoKEM-976-SHAKE ------------------------------------------------------------
FR_PK, FR_SK, FR_CT, FR_SS = 15632, 31296, 15744, 24


def frodo_sample(raw_u16):
    s = np.ascontiguousarray(raw_u16, dtype=np.uint16).copy()
    lib().oracle_frodo_sample_n(ptr(s), C.c_size_t(s.size))
    return s


def frodo_pack(vals, lsb):
    vals = np.ascontiguousarray(vals, dtype=np.uint16)
    outlen = (vals.size * lsb + 7) // 8
    out = buf(outlen)
    lib().oracle_frodo_pack(out, C.c_size_t(outlen), ptr(vals), C.c_size_t(vals.size), C.c_uint(lsb))
    return bytes(out)


def frodo_unpack(data, n, lsb):
    out = np.zeros(n, np.uint16)
    lib().oracle_frodo_unpack(ptr(out), C.c_size_t(n), bytes(data), C.c_size_t(len(data)), C.c_uint(lsb))
    return out


def frodo_gen_a(seedA):
    A = np.zeros((976, 976), np.uint16)
    lib().oracle_frodo_gen_a(ptr(A), bytes(seedA))
    return A


def frodo_sa_plus_e(A, S, E):
    out = np.zeros((8, 976), np.uint16)
    lib().oracle_frodo_mul_add_sa_plus_e(ptr(out), ptr(np.ascontiguousarray(A, np.uint16)),
                                         ptr(np.ascontiguousarray(S, np.uint16)), ptr(np.ascontiguousarray(E, np.uint16)))
    return out


def frodo_keypair(randomness64):
    pk, sk = buf(FR_PK), buf(FR_SK)
    assert lib().oracle_frodo_keypair_derand(pk, sk, bytes(randomness64)) == 0
    return bytes(pk), bytes(sk)


def frodo_enc(pk, mu):
    ct, ss = buf(FR_CT), buf(FR_SS)
    assert lib().oracle_frodo_enc_derand(ct, ss, bytes(pk), bytes(mu)) == 0
    return bytes(ct), bytes(ss)


def frodo_dec(ct, sk):
    ss = buf(FR_SS)
    assert lib().oracle_frodo_dec(ss, bytes(ct), bytes(sk)) == 0
    return bytes(ss)


def frodo_enc_batch(pk, mu, nthreads=8):
    mu = np.ascontiguousarray(mu, np.uint8)
    n = mu.shape[0]
    ct = np.empty((n, FR_CT), np.uint8)
    ss = np.empty((n, FR_SS), np.uint8)
    assert lib().oracle_frodo_enc_batch(C.c_size_t(n), ptr(ct), ptr(ss), bytes(pk), ptr(mu), nthreads) == 0
    return ct, ss


def frodo_dec_batch(ct, sk, nthreads=8):
    ct = np.ascontiguousarray(ct, np.uint8)
    n = ct.shape[0]
    ss = np.empty((n, FR_SS), np.uint8)
    assert lib().oracle_frodo_dec_batch(C.c_size_t(n), ptr(ss), ptr(ct), bytes(sk), nthreads) == 0
    return ss
