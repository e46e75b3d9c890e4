"""The NIST known-answer-test DRBG (host code, no GPU) and the .rsp fixtures.

Pins: the first seeds of NIST's published PQC KAT files (every .rsp the NIST
harness writes starts with these, whatever the scheme); FIPS 197 appendix C.3
for AES-256; an independent pure-Python AES-256 / CTR_DRBG (SP 800-90A 10.2.1,
no derivation function) compared with the library's on random seeds, with and
without a personalization string, over draw lengths that cross block edges."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

# count = 0 and count = 1 seeds of the NIST PQC KAT files (PQCgenKAT_kem.c, entropy 0..47)
NIST_SEED0 = bytes.fromhex("061550234D158C5EC95595FE04EF7A25767F2E24CC2BC479D09D86DC9ABCFDE7"
                           "056A8C266F9EF97ED08541DBD2E1FFA1")
NIST_SEED1 = bytes.fromhex("D81C4D8D734FCBFBEADE3D3F8A039FAA2A2C9957E835AD55B22E75BF57BB556A"
                           "C81ADDE6AEEB4A5A875C3BFCADFA958F")


# ---- independent AES-256 (FIPS 197), table-free: S-box by brute-force inverse --------
def _gmul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a = ((a << 1) ^ (0x11B if a & 0x80 else 0)) & 0x1FF
        b >>= 1
    return r & 0xFF


def _sbox():
    box = []
    for x in range(256):
        inv = 0 if x == 0 else next(y for y in range(1, 256) if _gmul(x, y) == 1)
        s = 0x63
        for i in range(8):
            bit = ((inv >> i) ^ (inv >> ((i + 4) % 8)) ^ (inv >> ((i + 5) % 8)) ^ (inv >> ((i + 6) % 8))
                   ^ (inv >> ((i + 7) % 8))) & 1
            s ^= bit << i
        box.append(s)
    return box


SBOX = _sbox()


def aes256_encrypt(key, block):
    w = [list(key[4 * i:4 * i + 4]) for i in range(8)]
    rcon = 1
    for i in range(8, 60):
        t = list(w[i - 1])
        if i % 8 == 0:
            t = [SBOX[b] for b in t[1:] + t[:1]]
            t[0] ^= rcon
            rcon = _gmul(rcon, 2)
        elif i % 8 == 4:
            t = [SBOX[b] for b in t]
        w.append([a ^ b for a, b in zip(w[i - 8], t)])
    st = [[block[r + 4 * c] for c in range(4)] for r in range(4)]  # st[row][col]

    def add(rd):
        for c in range(4):
            for r in range(4):
                st[r][c] ^= w[4 * rd + c][r]

    add(0)
    for rd in range(1, 15):
        for r in range(4):
            st[r] = [SBOX[b] for b in st[r]]
            st[r] = st[r][r:] + st[r][:r]
        if rd < 14:
            for c in range(4):
                a = [st[r][c] for r in range(4)]
                for r in range(4):
                    st[r][c] = (_gmul(a[r], 2) ^ _gmul(a[(r + 1) % 4], 3) ^ a[(r + 2) % 4] ^ a[(r + 3) % 4])
        add(rd)
    return bytes(st[r][c] for c in range(4) for r in range(4))


class PyCtrDrbg:
    def __init__(self, entropy, pers=None):
        seed = bytes(a ^ b for a, b in zip(entropy, pers)) if pers else bytes(entropy)
        self.key, self.v = bytes(32), bytes(16)
        self._update(seed)

    def _inc(self):
        self.v = ((int.from_bytes(self.v, "big") + 1) % (1 << 128)).to_bytes(16, "big")

    def _update(self, data):
        tmp = b""
        for _ in range(3):
            self._inc()
            tmp += aes256_encrypt(self.key, self.v)
        if data is not None:
            tmp = bytes(a ^ b for a, b in zip(tmp, data))
        self.key, self.v = tmp[:32], tmp[32:]

    def generate(self, n):
        out = b""
        while len(out) < n:
            self._inc()
            out += aes256_encrypt(self.key, self.v)
        self._update(None)
        return out[:n]


def test_python_aes256_fips197_c3():
    key = bytes(range(32))
    pt = bytes.fromhex("00112233445566778899aabbccddeeff")
    assert aes256_encrypt(key, pt).hex() == "8ea2b7ca516745bfeafc49904b496089"


def test_python_drbg_gives_nist_seeds():
    d = PyCtrDrbg(bytes(range(48)))
    assert d.generate(48) == NIST_SEED0
    assert d.generate(48) == NIST_SEED1


def test_library_drbg_gives_nist_seeds():
    from pqc_amd import kat
    s = kat.kat_seeds(3)
    assert s[0] == NIST_SEED0 and s[1] == NIST_SEED1
    d = PyCtrDrbg(bytes(range(48)))
    assert [d.generate(48) for _ in range(3)] == s


@pytest.mark.parametrize("pers", [False, True])
def test_library_drbg_vs_python(pers):
    from pqc_amd import kat
    rng = np.random.default_rng(11 + pers)
    for _ in range(3):
        ent = rng.integers(0, 256, 48, dtype=np.uint8).tobytes()
        p = rng.integers(0, 256, 48, dtype=np.uint8).tobytes() if pers else None
        py = PyCtrDrbg(ent, p)
        kat.drbg_init(ent, p)
        try:
            for n in (1, 15, 16, 17, 32, 48, 100):
                assert kat.drbg(n) == py.generate(n), n
        finally:
            kat.drbg_release()


def test_library_randombytes_follows_drbg_until_released():
    """pqc_randombytes (the non-derand entry points' source) draws from the DRBG
    after randombytes_init, and from the system after pqc_randombytes_system."""
    import ctypes as C

    from pqc_amd import _lib as L
    from pqc_amd import kat
    kat.drbg_init(NIST_SEED0)
    out = (C.c_uint8 * 32)()
    assert L.lib().pqc_randombytes(out, 32) == 0
    kat.drbg_release()
    assert bytes(out) == PyCtrDrbg(NIST_SEED0).generate(32)
    L.lib().pqc_randombytes_seed(5)  # the SHAKE256 DRBG again
    import hashlib
    import struct
    assert L.lib().pqc_randombytes(out, 32) == 0
    assert bytes(out) == hashlib.shake_256(b"pqc-mi355x" + struct.pack("<QQQ", 5, 0xFFFFFFFF, 0)).digest(32)
    L.lib().pqc_randombytes_system()


def test_rsp_write_read_roundtrip(tmp_path):
    from pqc_amd import kat
    rng = np.random.default_rng(3)
    ents = [{"count": i, **{k: rng.integers(0, 256, n, dtype=np.uint8).tobytes()
                            for k, n in (("seed", 48), ("pk", 800), ("sk", 1632), ("ct", 736), ("ss", 32))}}
            for i in range(3)]
    p = tmp_path / "x.rsp"
    kat.write_rsp(str(p), "kyber512", ents)
    name, back = kat.read_rsp(str(p))
    assert name == "Kyber512" and back == ents
    assert p.read_text().startswith("# Kyber512\n\ncount = 0\nseed = ")


@pytest.mark.parametrize("alg", ["kyber512", "kyber768", "kyber1024", "frodo976", "newhope1024", "newhope512"])
def test_kat_fixture_matches_oracle(alg):
    """The committed .rsp: NIST seeds, sizes, and count 0 recomputed by the oracle
    from the DRBG's coins."""
    from pqc_amd import kat, sizes
    from tests import oracle_lib as O
    name, ents = kat.read_rsp(os.path.join(HERE, "golden", "kat_%s.rsp" % alg))
    assert name == kat.RSP_NAMES[alg]
    assert [e["seed"] for e in ents] == kat.kat_seeds(len(ents))
    pkb, skb, ctb, ssb = sizes(alg)
    for e in ents:
        assert (len(e["pk"]), len(e["sk"]), len(e["ct"]), len(e["ss"])) == (pkb, skb, ctb, ssb)
    e = ents[0]
    kc, ec = kat.coins(alg, e["seed"])
    if alg == "frodo976":
        pk, sk = O.frodo_keypair(kc)
        ct, ss = O.frodo_enc(pk, ec)
        assert O.frodo_dec(ct, sk) == ss
    elif alg.startswith("newhope"):
        N = int(alg[7:])
        pk, sk = O.newhope_cca_keypair(kc, N)
        ct, ss = O.newhope_cca_enc(pk, ec, N)
        assert O.newhope_cca_dec(ct, sk, N) == ss
    else:
        k = {"kyber512": 2, "kyber768": 3, "kyber1024": 4}[alg]
        pk, sk = O.kyber_keypair(k, kc)
        ct, ss = O.kyber_enc(k, pk, ec)
        assert O.kyber_dec(k, ct, sk) == ss
        assert sk[-32:] == kc[32:] and sk[-64:-32] == O.sha3_256(pk)  # z and H(pk) in place
    assert (pk, sk, ct, ss) == (e["pk"], e["sk"], e["ct"], e["ss"])
