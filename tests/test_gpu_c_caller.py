"""The drop-in boundary exercised from C and from several host threads.

* tests/c/kat_kem.c is a NIST-harness-shaped KEM program written against the
  reference's api.h names only (CRYPTO_* sizes, CRYPTO_ALGNAME,
  crypto_kem_keypair / enc / dec, randombytes_init / randombytes); it is built
  by plain gcc against include/pqc_api.h, once per parameter set with
  -DPQC_NIST_ALG=<id> (make -C tests/c kat, run by the build step).  Its .rsp
  output must equal the known-answer fixtures in tests/golden, whose pk, sk,
  ct and ss come from the CPU oracle under the same DRBG seeds.
* Four host threads call different drop-ins at once; decapsulations must give
  the oracle's shared secrets and every encapsulation must decapsulate to its
  own secret (the drop-ins check a library-owned context out per call).

Every test prints the case it runs before touching the GPU."""
import os
import subprocess
import sys
import tempfile
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "c", "_bin")
ALGS = {"kyber512": "KYBER512", "kyber768": "KYBER768", "kyber1024": "KYBER1024", "frodo976": "FRODO976",
        "newhope1024": "NEWHOPE1024", "newhope512": "NEWHOPE512"}
CRYPTO_ALGNAME = {"kyber512": "Kyber512", "kyber768": "Kyber768", "kyber1024": "Kyber1024",
                  "frodo976": "FrodoKEM-976", "newhope1024": "NewHope1024-CCAKEM",
                  "newhope512": "NewHope512-CCAKEM"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("alg", list(ALGS))
def test_c_caller_reproduces_kat(alg):
    from pqc_amd import kat
    exe = os.path.join(BIN, "kat_" + ALGS[alg])
    if not os.path.exists(exe):
        pytest.fail("%s is not built: run make -C tests/c kat (part of the build step)" % exe)
    _, want = kat.read_rsp(os.path.join(HERE, "golden", "kat_%s.rsp" % alg))
    say("c caller", alg, "counts", len(want))
    r = subprocess.run([exe, str(len(want))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with tempfile.NamedTemporaryFile("w", suffix=".rsp", delete=False) as f:
        f.write(r.stdout)
    try:
        name, got = kat.read_rsp(f.name)
    finally:
        os.unlink(f.name)
    assert name == CRYPTO_ALGNAME[alg]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        for key in ("count", "seed", "pk", "sk", "ct", "ss"):
            assert g[key] == w[key], (alg, w["count"], key)


def test_dropins_from_four_threads():
    import pqc_amd
    from tests import oracle_lib as O
    say("drop-ins from 4 threads")
    # per thread: a parameter set, an oracle key pair and ciphertext, and the
    # oracle's shared secret
    cases = []
    for t, (alg, k) in enumerate([("kyber512", 2), ("kyber768", 3), ("kyber1024", 4), ("frodo976", None)]):
        if k:
            pk, sk = O.kyber_keypair(k, O.drbg(64, seed=40, stream=t))
            ct, ss = O.kyber_enc(k, pk, O.drbg(32, seed=41, stream=t))
        else:
            pk, sk = O.frodo_keypair(O.drbg(64, seed=40, stream=t))
            ct, ss = O.frodo_enc(pk, O.drbg(24, seed=41, stream=t))
        cases.append((alg, pk, sk, ct, ss))
    rounds = {"kyber512": 12, "kyber768": 12, "kyber1024": 12, "frodo976": 3}
    errors = []

    def worker(alg, pk, sk, ct, ss):
        try:
            for _ in range(rounds[alg]):
                got = pqc_amd.crypto_kem_dec(ct, sk, alg)
                if got != ss:
                    errors.append((alg, "dec differs from the oracle"))
                ct2, ss2 = pqc_amd.crypto_kem_enc(pk, alg)
                if pqc_amd.crypto_kem_dec(ct2, sk, alg) != ss2:
                    errors.append((alg, "enc/dec round trip"))
        except Exception as e:  # reported below
            errors.append((alg, repr(e)))

    threads = [threading.Thread(target=worker, args=c) for c in cases]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=110)
    assert not any(th.is_alive() for th in threads), "a drop-in thread did not finish"
    assert errors == []
