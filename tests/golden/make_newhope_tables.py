"""Writes tests/golden/newhope1024_tables.json and newhope512_tables.json: the
numeric NTT tables the reference tabulates in NewHope_NIST_GPU/precomp.cu
(omegas_inv_bitrev_montgomery, gammas_bitrev_montgomery,
gammas_inv_montgomery; for n = 1024 the __device__ arrays at
precomp.cu:191-371, for n = 512 the host arrays at precomp.cu:75-190) and the
bit-reversal tables of NewHope_NIST_GPU/ntt.cu:46 (10-bit) and :19 (9-bit).

Data only: it pins the oracle's closed-form table generator against the
reference's own tables.  Run from the repository root where the reference is
mounted:  python tests/golden/make_newhope_tables.py /root/reference
"""
import json
import re
import sys
from pathlib import Path


def extract(src, name, n, device=True):
    pat = r"%s\s+%s\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}" % ("__device__\\s+uint16_t" if device else "uint16_t", re.escape(name))
    m = re.search(pat, src)
    if not m:
        raise SystemExit("table %s not found" % name)
    vals = [int(v) for v in re.findall(r"\d+", m.group(1))]
    assert len(vals) == n, (name, len(vals))
    return vals


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "NewHope_NIST_GPU"
    pre = (ref / "precomp.cu").read_text()
    ntt = (ref / "ntt.cu").read_text()
    out = {
        "source": "NewHope_NIST_GPU/precomp.cu:191-371, ntt.cu:46",
        "omegas_inv_bitrev_montgomery": extract(pre, "omegas_inv_bitrev_montgomery", 512),
        "gammas_bitrev_montgomery": extract(pre, "gammas_bitrev_montgomery", 1024),
        "gammas_inv_montgomery": extract(pre, "gammas_inv_montgomery", 1024),
        "bitrev_table": extract(ntt, "bitrev_table", 1024),
    }
    dst = Path(__file__).with_name("newhope1024_tables.json")
    dst.write_text(json.dumps(out) + "\n")
    print("wrote", dst)
    # n = 512: the host arrays of the NEWHOPE_N == 512 branches (precomp.cu:75-190,
    # ntt.cu:19), which come first in each file
    out512 = {
        "source": "NewHope_NIST_GPU/precomp.cu:75-190, ntt.cu:19",
        "omegas_inv_bitrev_montgomery": extract(pre, "omegas_inv_bitrev_montgomery", 256, device=False),
        "gammas_bitrev_montgomery": extract(pre, "gammas_bitrev_montgomery", 512, device=False),
        "gammas_inv_montgomery": extract(pre, "gammas_inv_montgomery", 512, device=False),
        "bitrev_table": extract(ntt, "bitrev_table", 512, device=False),
    }
    dst = Path(__file__).with_name("newhope512_tables.json")
    dst.write_text(json.dumps(out512) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main()
