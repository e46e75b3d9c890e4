"""Exact FP32 arithmetic mod q used by the decapsulation tiles (CPU, no GPU).

`make -C tests/c f32-run` builds tests/c/f32_arith_test.cpp with the host
compiler (no floating-point contraction) and runs it: the products, reductions
and compressions of pqc_amd/csrc/kyber_f32.h exhaustively over the ranges the
kernels feed them, and the forward / inverse transform schedules by simulation
against exact 64-bit modular arithmetic.  IEEE binary32 fma, multiply and add in
round-to-nearest-even give the same bits on the host as on the GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_f32_arith_mod_q():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "c"), "f32-run"], capture_output=True, text=True,
                       timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "f32 arithmetic mod q: all checks hold" in out
