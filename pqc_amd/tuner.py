"""Python mirror of the autotuner and the learned cost model (pqc_autotune,
pqc_costmodel_* in include/pqc_api.h).  The tuner and the model run in the
library; this module only marshals arguments."""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check

FAMILIES = {"plain_kyber": 0, "fused_kyber": 1, "mixed_kyber": 2, "frodo": 3, "newhope": 4}
N_FEATURES = 14


class Tuning(C.Structure):
    _fields_ = [("plain_block", C.c_int32), ("newhope_block", C.c_int32), ("fused_waves_per_cu", C.c_int32),
                ("mixed_order", C.c_int32), ("frodo_stream_block", C.c_int32), ("lowlat_max", C.c_int32),
                ("frodo_chunk", C.c_int32), ("fused_dec_tile", C.c_int32),
                ("fused_enc_tile", C.c_int32), ("fused_key_tile", C.c_int32), ("frodo_lane_min", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TuneRequest(C.Structure):
    _fields_ = [("family", C.c_int32), ("k", C.c_int32), ("op", C.c_int32), ("budget", C.c_int32),
                ("n", C.c_uint64), ("use_model", C.c_int32), ("reps", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("best_ms", C.c_float), ("predicted_ms", C.c_float), ("measured", C.c_int32),
                ("reserved", C.c_int32)]


WORKLOAD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)


def get_tuning(ctx):
    t = Tuning()
    check(L.lib().pqc_ctx_get_tuning(ctx.handle, C.byref(t)), "pqc_ctx_get_tuning")
    return t


def put_tuning(ctx, **kw):
    t = get_tuning(ctx)
    for k, v in kw.items():
        setattr(t, k, v)
    check(L.lib().pqc_ctx_put_tuning(ctx.handle, C.byref(t)), "pqc_ctx_put_tuning")
    return t


def request(family, n, k=0, op=2, budget=0, use_model=False, reps=3):
    return TuneRequest(FAMILIES[family] if isinstance(family, str) else family, k, op, budget, n,
                       1 if use_model else 0, reps)


def autotune(ctx, req, run, stream):
    """run(stream) launches the workload once on `stream` (device pointers)."""
    err = []

    def cb(_user, _stream):
        try:
            run(stream)
            return 0
        except Exception as e:  # reported after the call
            err.append(e)
            return L.PQC_ERR_INVALID_ARG

    fn = WORKLOAD_FN(cb)
    res = TuneResult()
    rc = L.lib().pqc_autotune(ctx.handle, C.byref(req), fn, None, stream.handle, C.byref(res))
    if err:
        raise err[0]
    check(rc, "pqc_autotune")
    return res


def features(req, tuning):
    out = (C.c_float * N_FEATURES)()
    n = L.lib().pqc_tune_features(C.byref(req), C.byref(tuning), out, N_FEATURES)
    check(n if n < 0 else 0, "pqc_tune_features")
    return np.frombuffer(out, dtype=np.float32).copy()


def samples(ctx):
    n = L.lib().pqc_tune_samples(ctx.handle, None, None, 0)
    check(n if n < 0 else 0, "pqc_tune_samples")
    x = np.zeros((n, N_FEATURES), np.float32)
    y = np.zeros(n, np.float32)
    if n:
        L.lib().pqc_tune_samples(ctx.handle, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), n)
    return x, y


class CostModel:
    """The MLP (in -> 32 -> 32 -> 1) whose forward, backward and Adam run on the GPU."""

    def __init__(self, in_dim=N_FEATURES, max_samples=8192, seed=1, handle=None):
        self.owned = handle is None
        if handle is None:
            h = C.c_void_p()
            check(L.lib().pqc_costmodel_create(C.byref(h), in_dim, max_samples, seed), "pqc_costmodel_create")
            handle = h.value
        self.handle = handle
        self.in_dim = in_dim

    @classmethod
    def of_context(cls, ctx):
        h = L.lib().pqc_ctx_costmodel(ctx.handle)
        if not h:
            raise L.PQCError(L.PQC_ERR_NO_MEMORY, "pqc_ctx_costmodel")
        return cls(handle=h)

    @property
    def num_params(self):
        return L.lib().pqc_costmodel_num_params(self.handle)

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        check(L.lib().pqc_costmodel_get_params(self.handle, p.ctypes.data_as(C.c_void_p), p.size), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(L.lib().pqc_costmodel_set_params(self.handle, p.ctypes.data_as(C.c_void_p), p.size), "set_params")

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(len(x), np.float32)
        check(L.lib().pqc_costmodel_predict(self.handle, len(x), x.ctypes.data_as(C.c_void_p),
                                            y.ctypes.data_as(C.c_void_p)), "predict")
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_float()
        check(L.lib().pqc_costmodel_grad(self.handle, len(x), x.ctypes.data_as(C.c_void_p),
                                         t.ctypes.data_as(C.c_void_p), C.byref(loss),
                                         g.ctypes.data_as(C.c_void_p), g.size), "grad")
        return loss.value, g

    def fit(self, x, t, steps, lr=3e-3):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        check(L.lib().pqc_costmodel_fit(self.handle, len(x), x.ctypes.data_as(C.c_void_p),
                                        t.ctypes.data_as(C.c_void_p), steps, lr, C.byref(loss)), "fit")
        return loss.value

    def close(self):
        if self.owned and self.handle:
            L.lib().pqc_costmodel_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
