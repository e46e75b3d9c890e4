"""Writes tests/golden/kyber_zetas.json: the numeric twiddle tables that the
reference tabulates in Kyber_GPU_Batched/ntt.cu:50-68 (zetas, zetas_inv).

The fixture holds data only (two lists of 128 integers).  It pins the
oracle's closed-form twiddle generator against the reference's own table.
Run from the repository root in a container where the reference is mounted:
    python tests/golden/make_kyber_tables.py /root/reference
"""
import json
import re
import sys
from pathlib import Path


def extract(src: str, name: str):
    m = re.search(r"\b%s\s*\[\s*128\s*\]\s*=\s*\{([^}]*)\}" % re.escape(name), src)
    if not m:
        raise SystemExit("table %s not found" % name)
    vals = [int(v) for v in re.findall(r"-?\d+", m.group(1))]
    assert len(vals) == 128, len(vals)
    return vals


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = (ref / "Kyber_GPU_Batched" / "ntt.cu").read_text()
    out = {
        "source": "Kyber_GPU_Batched/ntt.cu:50-68",
        "zetas": extract(src, "zetas"),
        "zetas_inv": extract(src, "zetas_inv"),
    }
    dst = Path(__file__).with_name("kyber_zetas.json")
    dst.write_text(json.dumps(out) + "\n")
    print("wrote", dst)


if __name__ == "__main__":
    main()
