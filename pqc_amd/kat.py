"""NIST PQC known-answer-test files (.rsp) for the KEMs of this library.

The NIST harness (PQCgenKAT_kem.c) seeds the AES-256 CTR_DRBG with the bytes
0..47, draws one 48-byte seed per test count, then for each count re-seeds the
DRBG with that seed and runs crypto_kem_keypair, crypto_kem_enc and
crypto_kem_dec, writing count, seed, pk, sk, ct and ss.  The DRBG is the
library's own (randombytes_init / randombytes, host code); `coins` replays the
draws each parameter set makes, so a whole file is three batched GPU calls.

Draw order per parameter set (what the single-instance entry points do too):
Kyber and NewHope keypair CPA seed (32) then z / s (32), encaps m (32, hashed
inside); FrodoKEM-976
keypair s || seedSE || z (64, FrodoKEM_GPU/kem.c:33), encaps mu (24, kem.c:90).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .kem import Context

RSP_NAMES = {"kyber512": "Kyber512", "kyber768": "Kyber768", "kyber1024": "Kyber1024",
             "frodo976": "FrodoKEM-976-SHAKE", "newhope1024": "NewHope1024-CCAKEM",
             "newhope512": "NewHope512-CCAKEM"}
FIELDS = ("seed", "pk", "sk", "ct", "ss")


def drbg_init(seed48, personalization=None):
    """randombytes_init(seed48, personalization, 256): the library's draws now come from the DRBG."""
    assert len(seed48) == 48 and (personalization is None or len(personalization) == 48)
    e = (C.c_uint8 * 48).from_buffer_copy(bytes(seed48))
    p = (C.c_uint8 * 48).from_buffer_copy(bytes(personalization)) if personalization is not None else None
    L.lib().randombytes_init(e, p, 256)


def drbg(n):
    """One randombytes(x, n) call."""
    out = (C.c_uint8 * max(n, 1))()
    if L.lib().randombytes(out, n) != 0:
        raise L.PQCError(-1, "randombytes")
    return bytes(out)[:n]


def drbg_release():
    """Back to the operating system's randomness."""
    L.lib().pqc_randombytes_system()


def kat_seeds(count):
    """The harness's per-count seeds: DRBG(bytes 0..47), then `count` draws of 48 bytes."""
    drbg_init(bytes(range(48)))
    try:
        return [drbg(48) for _ in range(count)]
    finally:
        drbg_release()


def coins(alg, seed):
    """(keypair coins, encaps coins) that the harness's draws give for one seed."""
    drbg_init(seed)
    try:
        if alg == "frodo976":
            return drbg(64), drbg(24)
        d, z = drbg(32), drbg(32)
        return d + z, drbg(32)
    finally:
        drbg_release()


def generate(alg, seeds, ctx=None):
    """Runs keypair, encaps and decaps for every seed as three batched calls on the
    GPU; returns entries (dicts of count, seed, pk, sk, ct, ss).  Raises if a
    decapsulation disagrees with its encapsulation."""
    n = len(seeds)
    own = ctx is None
    ctx = ctx or Context([alg], max_batch=max(n, 1))
    try:
        kc, ec = zip(*(coins(alg, s) for s in seeds))
        kc = np.frombuffer(b"".join(kc), np.uint8).reshape(n, -1)
        ec = np.frombuffer(b"".join(ec), np.uint8).reshape(n, -1)
        pk, sk = ctx.keypair_host(alg, kc)
        ct, ss = ctx.enc_host(alg, pk, ec)
        ss2 = ctx.dec_host(alg, ct, sk)
        if not (ss2 == ss).all():
            raise AssertionError("decapsulation disagrees with encapsulation")
    finally:
        if own:
            ctx.close()
    return [{"count": i, "seed": seeds[i], "pk": pk[i].tobytes(), "sk": sk[i].tobytes(), "ct": ct[i].tobytes(),
             "ss": ss[i].tobytes()} for i in range(n)]


def write_rsp(path, alg, entries):
    """The harness's layout: '# <name>', blank line, then per count the six fields
    in upper-case hex and a blank line."""
    with open(path, "w") as f:
        f.write("# %s\n\n" % RSP_NAMES[alg])
        for e in entries:
            f.write("count = %d\n" % e["count"])
            for k in FIELDS:
                f.write("%s = %s\n" % (k, e[k].hex().upper()))
            f.write("\n")


def read_rsp(path):
    """(name, entries) of an .rsp file; fields other than count are bytes."""
    name, entries, cur = None, [], None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("#"):
                name = name or line[1:].strip()
                continue
            if not line:
                continue
            key, _, val = (t.strip() for t in line.partition("="))
            if key == "count":
                cur = {"count": int(val)}
                entries.append(cur)
            elif cur is not None:
                cur[key] = bytes.fromhex(val)
    return name, entries


def check_rsp(path, alg, ctx=None):
    """Re-runs an .rsp file's seeds on the GPU; returns the counts whose pk, sk, ct
    or ss differ from the file (empty when the file is reproduced)."""
    name, entries = read_rsp(path)
    if name != RSP_NAMES[alg]:
        raise ValueError("%s holds %r, not %s" % (path, name, RSP_NAMES[alg]))
    got = generate(alg, [e["seed"] for e in entries], ctx)
    bad = []
    for e, g in zip(entries, got):
        if any(e[k] != g[k] for k in ("pk", "sk", "ct", "ss")):
            bad.append(e["count"])
    return bad
