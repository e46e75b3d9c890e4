"""Independent Kyber (round 2) mathematics for the pinning tests.

Nothing here calls the oracle.  It states the algebra the reference's
arithmetic must satisfy: the NTT of Kyber_GPU_Batched/ntt.cu is evaluation of
a(X) modulo the 128 quadratics X^2 - gamma_m (gamma_m = +/- 17^brv7(64+m/2)),
basemul is multiplication in Z_q[X]/(X^2 - gamma_m) times 2^-16, CBD(eta=2)
is a difference of popcounts, and rejection sampling keeps 16-bit little-
endian candidates below 19q and subtracts (v >> 12) * q."""
import hashlib

import numpy as np

Q = 3329
N = 256


def brv7(i):
    return int("{:07b}".format(i)[::-1], 2)


def gammas():
    g = np.zeros(128, dtype=np.int64)
    for m in range(128):
        z = pow(17, brv7(64 + m // 2), Q)
        g[m] = z if m % 2 == 0 else (Q - z) % Q
    return g


GAMMA = gammas()


def ntt_true(a):
    """a(X) mod (X^2 - gamma_m) for m = 0..127, as the 256-vector (r0_m, r1_m)."""
    a = np.asarray(a, dtype=np.int64) % Q
    ae, ao = a[0::2], a[1::2]
    out = np.zeros(256, dtype=np.int64)
    for m in range(128):
        pw = np.array([pow(int(GAMMA[m]), k, Q) for k in range(128)], dtype=np.int64)
        out[2 * m] = int((ae * pw).sum() % Q)
        out[2 * m + 1] = int((ao * pw).sum() % Q)
    return out


def pair_mul(x, y):
    """Multiply NTT-domain vectors pairwise in Z_q[X]/(X^2 - gamma_m)."""
    x = np.asarray(x, dtype=np.int64) % Q
    y = np.asarray(y, dtype=np.int64) % Q
    out = np.zeros(256, dtype=np.int64)
    for m in range(128):
        a0, a1, b0, b1 = x[2 * m], x[2 * m + 1], y[2 * m], y[2 * m + 1]
        out[2 * m] = (a0 * b0 + a1 * b1 % Q * GAMMA[m]) % Q
        out[2 * m + 1] = (a0 * b1 + a1 * b0) % Q
    return out


def negacyclic(a, b):
    a = np.asarray(a, dtype=np.int64) % Q
    b = np.asarray(b, dtype=np.int64) % Q
    full = np.convolve(a, b)
    r = full[:256].copy()
    r[: len(full) - 256] -= full[256:]
    return r % Q


def cbd2(buf128):
    r = np.zeros(256, dtype=np.int64)
    bits = np.unpackbits(np.frombuffer(bytes(buf128), dtype=np.uint8), bitorder="little")
    for i in range(256):
        b = bits[4 * i: 4 * i + 4]
        r[i] = int(b[0]) + int(b[1]) - int(b[2]) - int(b[3])
    return r


def prf(key, nonce):
    return hashlib.shake_256(bytes(key) + bytes([nonce])).digest(128)


def rej_uniform(stream, n=256):
    out = []
    for pos in range(0, len(stream) - 1, 2):
        if len(out) == n:
            break
        v = stream[pos] | (stream[pos + 1] << 8)
        if v < 19 * Q:
            out.append(v - (v >> 12) * Q)
    return out


def matrix_entry(rho, x, y):
    """Parse(SHAKE128(rho || x || y)): 4 blocks, then more only if needed."""
    nbytes = 672
    while True:
        vals = rej_uniform(hashlib.shake_128(bytes(rho) + bytes([x, y])).digest(nbytes))
        if len(vals) == 256:
            return np.array(vals, dtype=np.int64)
        nbytes += 168


def unpack_bits(b, n, d):
    bits = np.unpackbits(np.frombuffer(bytes(b), dtype=np.uint8), bitorder="little")
    bits = bits[: n * d].reshape(n, d).astype(np.int64)
    return (bits << np.arange(d)).sum(axis=1)


def compress_int(x, d):
    """round(x * 2^d / q) mod 2^d for x in [0, q) (q odd: no ties)."""
    x = np.asarray(x, dtype=np.int64) % Q
    return ((2 * x * (1 << d) + Q) // (2 * Q)) % (1 << d)


def params(k):
    du = 11 if k == 4 else 10
    dv = {2: 3, 3: 4, 4: 5}[k]
    pk = 384 * k + 32
    sk = 384 * k + pk + 64
    ct = 32 * du * k + 32 * dv
    return dict(k=k, du=du, dv=dv, pk=pk, sk=sk, ct=ct)
