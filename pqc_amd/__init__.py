"""MI355X-native batched lattice-KEM engine (round-2 Kyber, NewHope-1024 CPA,
FrodoKEM-976) behind the C ABI of include/pqc_api.h.  This package is the
Python host binding of libpqc_amd.so; all arithmetic runs in its HIP kernels."""
from ._lib import PQCError, lib  # noqa: F401
from .kem import (ALGS, KEM_REQ_DTYPE, Context, DeviceBuffer, Event, Graph, PinnedBuffer, Stream,  # noqa: F401
                  crypto_kem_dec, crypto_kem_enc, crypto_kem_keypair, keccak_batch, kyber_poly_transform, newhope_poly_sample,
                  newhope_poly_transform,
                  randombytes_seed, randombytes_system, sizes)
from .mixed import DeviceArenas, MixedLayout  # noqa: F401

__all__ = ["PQCError", "lib", "ALGS", "KEM_REQ_DTYPE", "Context", "DeviceArenas", "DeviceBuffer", "Event",
           "Graph", "MixedLayout", "PinnedBuffer", "Stream", "crypto_kem_dec",
           "crypto_kem_enc", "crypto_kem_keypair", "keccak_batch", "kyber_poly_transform", "newhope_poly_sample",
           "newhope_poly_transform",
           "randombytes_seed", "randombytes_system", "sizes"]
