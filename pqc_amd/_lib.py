"""ctypes binding of libpqc_amd.so (the C ABI declared in include/pqc_api.h).

The library is the product: every KEM call goes to its HIP kernels.  There is
no CPU fallback; loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpqc_amd.so")

PQC_OK = 0
PQC_ERR_INVALID_ARG = -1
PQC_ERR_UNSUPPORTED = -2
PQC_ERR_BATCH_TOO_LARGE = -3
PQC_ERR_NO_MEMORY = -4
PQC_ERR_DEVICE = -5
PQC_ERR_NO_DEVICE = -6

ALG_KYBER512 = 1
ALG_KYBER768 = 2
ALG_KYBER1024 = 3
ALG_NEWHOPE1024_CPA = 4
ALG_FRODO976 = 5
ALG_NEWHOPE1024 = 6
ALG_NEWHOPE512 = 7
ALG_NEWHOPE512_CPA = 8

F_HOST_PTRS = 0x1
F_PLAIN = 0x2

_P = C.c_void_p
_SZ = C.c_size_t
_I = C.c_int
_U = C.c_uint

# name -> (restype, argtypes)
SIGNATURES = {
    "pqc_ctx_create": (_I, [C.POINTER(_P), _U, _SZ, _SZ]),
    "pqc_ctx_destroy": (None, [_P]),
    "pqc_ctx_device_bytes": (_SZ, [_P]),
    "pqc_ctx_set_tuning": (_I, [_P, _I, _I]),
    "pqc_alg_sizes": (_I, [_I, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ)]),
    "pqc_host_alloc": (_I, [_SZ, C.POINTER(_P)]),
    "randombytes_init": (None, [_P, _P, _I]),
    "randombytes": (_I, [_P, C.c_ulonglong]),
    "pqc_host_free": (_I, [_P]),
    "crypto_kem_keypair_derand_batch": (_I, [_P, _I, _SZ, _P, _P, _P, _U, _P]),
    "crypto_kem_enc_derand_batch": (_I, [_P, _I, _SZ, _P, _P, _P, _P, _U, _P]),
    "crypto_kem_dec_batch": (_I, [_P, _I, _SZ, _P, _P, _P, _U, _P]),
    "pqc_kyber_dec_batch_noalloc": (_I, [_I, _SZ, _P, _P, _P, _P, _P]),
    "crypto_kem_keypair_batch": (_I, [_P, _I, _SZ, _P, _P, _U, _P]),
    "crypto_kem_enc_batch": (_I, [_P, _I, _SZ, _P, _P, _P, _U, _P]),
    "crypto_kem_keypair": (_I, [_P, _P]),
    "crypto_kem_enc": (_I, [_P, _P, _P]),
    "crypto_kem_dec": (_I, [_P, _P, _P]),
    "pqc_randombytes_seed": (None, [C.c_uint64]),
    "pqc_randombytes_system": (None, []),
    "pqc_randombytes": (_I, [_P, _SZ]),
    "pqc_device_count": (_I, [C.POINTER(_I)]),
    "pqc_set_device": (_I, [_I]),
    "pqc_device_info": (_I, [C.c_char_p, _SZ, C.POINTER(_I), C.POINTER(_SZ)]),
    "pqc_device_malloc": (_I, [C.POINTER(_P), _SZ]),
    "pqc_device_free": (_I, [_P]),
    "pqc_memcpy_htod": (_I, [_P, _P, _SZ, _P]),
    "pqc_memcpy_dtoh": (_I, [_P, _P, _SZ, _P]),
    "pqc_memcpy_dtod": (_I, [_P, _P, _SZ, _P]),
    "pqc_stream_create": (_I, [C.POINTER(_P)]),
    "pqc_stream_destroy": (_I, [_P]),
    "pqc_stream_synchronize": (_I, [_P]),
    "pqc_event_create": (_I, [C.POINTER(_P)]),
    "pqc_event_destroy": (_I, [_P]),
    "pqc_event_record": (_I, [_P, _P]),
    "pqc_event_synchronize": (_I, [_P]),
    "pqc_event_elapsed_ms": (_I, [C.POINTER(C.c_float), _P, _P]),
    "pqc_keccak_batch": (_I, [_I, _SZ, _P, _SZ, _P, _SZ, _P]),
    "pqc_kyber_poly_transform": (_I, [_I, _SZ, _P, _P]),
    "pqc_newhope1024_poly_transform": (_I, [_I, _SZ, _P, _P]),
    "pqc_newhope1024_poly_sample": (_I, [_SZ, _P, _P, _SZ, _I, _P]),
    "pqc_frodo976_gemm": (_I, [_P, _SZ, _P, _P]),
    "pqc_valu_issue_rate": (_I, [_I, _I, _P, _P]),
    "pqc_keccak_rate": (_I, [_I, _I, C.POINTER(C.c_double), _P]),
    "pqc_newhope1024_cpa_keypair_batch": (_I, [_P, _SZ, _P, _P, _P, _U, _P]),
    "pqc_newhope1024_cpa_enc_batch": (_I, [_P, _SZ, _P, _P, _P, _P, _U, _P]),
    "pqc_newhope1024_cpa_dec_batch": (_I, [_P, _SZ, _P, _P, _P, _U, _P]),
    "pqc_newhope512_cpa_keypair_batch": (_I, [_P, _SZ, _P, _P, _P, _U, _P]),
    "pqc_newhope512_cpa_enc_batch": (_I, [_P, _SZ, _P, _P, _P, _P, _U, _P]),
    "pqc_newhope512_cpa_dec_batch": (_I, [_P, _SZ, _P, _P, _P, _U, _P]),
    "pqc_frodo976_keypair_batch": (_I, [_P, _SZ, _P, _P, _P, _U, _P]),
    "pqc_frodo976_load_key": (_I, [_P, _P, _I, _U, _P]),
    "pqc_frodo976_enc_batch": (_I, [_P, _SZ, _P, _P, _P, _U, _P]),
    "pqc_frodo976_dec_batch": (_I, [_P, _SZ, _P, _P, _U, _P]),
    "pqc_frodo976_products": (_I, [_P, _SZ, _P, _P, _U, _P]),
    "pqc_kem_enc_mixed": (_I, [_P, _SZ, _P, _P, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _P]),
    "pqc_kem_dec_mixed": (_I, [_P, _SZ, _P, _P, _SZ, _P, _SZ, _P, _SZ, _P]),
    "pqc_ctx_take_status": (_I, [_P, _P, C.POINTER(C.c_uint32)]),
    "pqc_graph_capture_begin": (_I, [_P]),
    "pqc_graph_capture_end": (_I, [_P, C.POINTER(_P)]),
    "pqc_graph_launch": (_I, [_P, _P]),
    "pqc_graph_destroy": (None, [_P]),
    "pqc_ctx_get_tuning": (_I, [_P, _P]),
    "pqc_ctx_put_tuning": (_I, [_P, _P]),
    "pqc_autotune": (_I, [_P, _P, _P, _P, _P, _P]),
    "pqc_tune_features": (_I, [_P, _P, _P, _I]),
    "pqc_tune_samples": (_I, [_P, _P, _P, _I]),
    "pqc_costmodel_create": (_I, [C.POINTER(_P), _I, _I, C.c_uint64]),
    "pqc_costmodel_destroy": (None, [_P]),
    "pqc_costmodel_num_params": (_I, [_P]),
    "pqc_costmodel_trained_steps": (_I, [_P]),
    "pqc_costmodel_get_params": (_I, [_P, _P, _SZ]),
    "pqc_costmodel_set_params": (_I, [_P, _P, _SZ]),
    "pqc_costmodel_predict": (_I, [_P, _I, _P, _P]),
    "pqc_costmodel_grad": (_I, [_P, _I, _P, _P, C.POINTER(C.c_float), _P, _SZ]),
    "pqc_costmodel_fit": (_I, [_P, _I, _P, _P, _I, C.c_float, C.POINTER(C.c_float)]),
    "pqc_ctx_costmodel": (_P, [_P]),
    "pqc_server_create": (_I, [C.POINTER(_P), _P]),
    "pqc_server_destroy": (None, [_P]),
    "pqc_server_key": (_I, [_P, _I, C.c_uint32, _P, _P]),
    "pqc_server_submit": (_I, [_P, _P, _SZ]),
    "pqc_server_poll": (_I, [_P, _P, _SZ, C.POINTER(_SZ), C.c_uint32]),
    "pqc_server_drain": (_I, [_P]),
    "pqc_server_get_stats": (_I, [_P, _P]),
    "pqc_server_now_us": (C.c_double, [_P]),
    "pqc_server_run_trace": (_I, [_P, _SZ, _P, _P, _P, _P]),
    "pqc_version": (C.c_char_p, []),
    "pqc_strerror": (C.c_char_p, [_I]),
}
for _alg in ("kyber512", "kyber768", "kyber1024", "frodo976", "newhope1024", "newhope512"):
    SIGNATURES["pqc_%s_crypto_kem_keypair" % _alg] = (_I, [_P, _P])
    SIGNATURES["pqc_%s_crypto_kem_enc" % _alg] = (_I, [_P, _P, _P])
    SIGNATURES["pqc_%s_crypto_kem_dec" % _alg] = (_I, [_P, _P, _P])

_lib = None


class PQCError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = "%s failed: %d" % (what, code)
        if _lib is not None:
            msg += " (%s)" % _lib.pqc_strerror(code).decode()
        super().__init__(msg)


def lib():
    """Load the built library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpqc_amd.so is not built (run __graft_entry__.build() or "
                               "make -C pqc_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != PQC_OK:
        raise PQCError(rc, what)
    return rc
