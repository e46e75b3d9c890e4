"""MI355X-native batched lattice-KEM engine (round-2 Kyber, NewHope-1024 CPA,
FrodoKEM-976) behind the C ABI of include/pqc_api.h.  This package is the
Python host binding of libpqc_amd.so; all arithmetic runs in its HIP kernels."""
from ._lib import PQCError, lib  # noqa: F401
from .kem import (ALGS, Context, DeviceBuffer, Event, Stream, crypto_kem_dec,  # noqa: F401
                  crypto_kem_enc, crypto_kem_keypair, keccak_batch, kyber_poly_transform,
                  randombytes_seed, randombytes_system, sizes)

__all__ = ["PQCError", "lib", "ALGS", "Context", "DeviceBuffer", "Event", "Stream", "crypto_kem_dec",
           "crypto_kem_enc", "crypto_kem_keypair", "keccak_batch", "kyber_poly_transform",
           "randombytes_seed", "randombytes_system", "sizes"]
