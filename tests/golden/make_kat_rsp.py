"""Writes tests/golden/kat_<alg>.rsp: NIST-harness-style known-answer files for
Kyber-512/768/1024 (counts 0..3), FrodoKEM-976 (count 0) and the NewHope-1024
and NewHope-512 CCA-KEMs (counts 0..2), computed by the
CPU oracle from the coins the NIST AES-256 CTR_DRBG gives for each count's seed.

The seeds and coins come from the library's DRBG (host code: no GPU needed);
the DRBG is pinned separately by the first seeds of NIST's published KAT files
and by an independent AES-256 / CTR_DRBG restatement in tests/test_kat_cpu.py.
Run from the repository root:  python tests/golden/make_kat_rsp.py
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from pqc_amd import kat  # noqa: E402
from tests import oracle_lib as O  # noqa: E402

COUNTS = {"kyber512": 4, "kyber768": 4, "kyber1024": 4, "frodo976": 1, "newhope1024": 3, "newhope512": 3}
K = {"kyber512": 2, "kyber768": 3, "kyber1024": 4}


def oracle_entries(alg, count):
    out = []
    for i, seed in enumerate(kat.kat_seeds(count)):
        kc, ec = kat.coins(alg, seed)
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
            pk, sk = O.kyber_keypair(K[alg], kc)
            ct, ss = O.kyber_enc(K[alg], pk, ec)
            assert O.kyber_dec(K[alg], ct, sk) == ss
        out.append({"count": i, "seed": seed, "pk": pk, "sk": sk, "ct": ct, "ss": ss})
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    for alg, n in COUNTS.items():
        path = os.path.join(here, "kat_%s.rsp" % alg)
        kat.write_rsp(path, alg, oracle_entries(alg, n))
        print("wrote", path)


if __name__ == "__main__":
    main()
