"""C ABI contract, checked without a GPU: the library loads, exports every
function include/pqc_api.h declares, and fails cleanly when no device is
visible."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not hdr.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", hdr)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        # function declarations; typedefs (function-pointer types) are not symbols
        for m in re.finditer(r"^\s*(?!typedef\b)(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", src,
                             flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "sizeof", "return"):
                names.append(name)
    return sorted(set(names))


def test_header_parse_finds_api():
    names = declared_functions()
    for must in ("crypto_kem_keypair", "crypto_kem_enc", "crypto_kem_dec", "crypto_kem_dec_batch",
                 "pqc_ctx_create", "pqc_kyber768_crypto_kem_dec"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import pqc_amd._lib as L
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    unbound = [n for n in declared_functions() if n not in L.SIGNATURES]
    assert not unbound, unbound


def test_no_cpu_fallback_without_device():
    import pqc_amd
    from pqc_amd import _lib as L
    n = ctypes.c_int(-1)
    rc = L.lib().pqc_device_count(ctypes.byref(n))
    if rc == L.PQC_OK and n.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(pqc_amd.PQCError) as e:
        pqc_amd.Context(["kyber768"], max_batch=4)
    assert e.value.code == L.PQC_ERR_NO_DEVICE
    # the NIST drop-in reports the missing device instead of computing on the CPU
    pk = (ctypes.c_uint8 * 1184)()
    sk = (ctypes.c_uint8 * 2400)()
    assert L.lib().crypto_kem_keypair(pk, sk) == L.PQC_ERR_NO_DEVICE


def test_sizes_and_strings():
    import pqc_amd
    from pqc_amd import _lib as L
    assert pqc_amd.sizes("kyber512") == (800, 1632, 736, 32)
    assert pqc_amd.sizes("kyber768") == (1184, 2400, 1088, 32)
    assert pqc_amd.sizes("kyber1024") == (1568, 3168, 1568, 32)
    assert L.lib().pqc_strerror(L.PQC_ERR_NO_MEMORY)
    assert b"gfx950" in L.lib().pqc_version()


def test_deterministic_randombytes_matches_drbg_definition():
    import hashlib
    import struct
    from pqc_amd import _lib as L
    L.lib().pqc_randombytes_seed(5)
    out = (ctypes.c_uint8 * 48)()
    assert L.lib().pqc_randombytes(out, 48) == 0
    want = hashlib.shake_256(b"pqc-mi355x" + struct.pack("<QQQ", 5, 0xFFFFFFFF, 0)).digest(48)
    assert bytes(out) == want
    L.lib().pqc_randombytes_system()


def test_tuner_features_encode_the_frodo_lane_layout():
    """pqc_tune_features (host code, no device): for the Frodo family, slot 10 is 1
    exactly when the stream runs one Keccak state per lane at the request's n
    (tuning.frodo_lane_min: 0 = the library's 49152, -1 = lane pairs always);
    other families leave it to their own layout."""
    from pqc_amd import tuner as T
    t = T.Tuning(plain_block=256, newhope_block=128, fused_waves_per_cu=4, frodo_stream_block=64)
    cases = ((0, 4096, 0), (0, 49151, 0), (0, 49152, 1), (0, 65536, 1), (-1, 65536, 0), (1, 1, 1), (1, 4096, 1),
             (8192, 8191, 0), (8192, 8192, 1))
    for lane_min, n, want in cases:
        t.frodo_lane_min = lane_min
        f = T.features(T.request("frodo", n, op=2), t)
        assert len(f) == T.N_FEATURES == 14
        assert f[T.FAMILIES["frodo"]] == 1.0 and f[:5].sum() == 1.0
        assert f[10] == want, (lane_min, n)
        assert f[11:].sum() == 0.0
        assert abs(f[8] - np.log2(64) / 10) < 1e-6  # the stream's block size
    t.frodo_lane_min = 1
    f = T.features(T.request("plain_kyber", 4096, k=3, op=2), t)
    assert f[10:].sum() == 0.0
