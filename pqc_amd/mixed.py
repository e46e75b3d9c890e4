"""Mixed Kyber-512/768/1024 batches: arenas and request descriptors.

A mixed batch names every request by a descriptor (pqc_kem_req: algorithm id and
byte offsets into four device arenas).  MixedLayout places the buffers of each
parameter set in one contiguous region per arena (in a shuffled order inside the
region), so the same arenas also feed the single-parameter-set calls, which is
how the plain per-op kernels are compared with the fused mixed path.
"""
import numpy as np

from . import _lib as L
from .kem import KEM_REQ_DTYPE, DeviceBuffer, sizes

KYBER_ALGS = (L.ALG_KYBER512, L.ALG_KYBER768, L.ALG_KYBER1024)


class MixedLayout:
    """Offsets for requests with algorithm ids `algs` (one per request).

    region[kind][alg] = (byte offset, element count) of that parameter set's slice
    of the arena `kind` ("pk", "sk", "ct", "ss", "coins", "kcoins"); slot[i] = the
    index of request i inside its slice.
    """

    def __init__(self, algs, rng=None):
        algs = np.asarray(algs, dtype=np.int32)
        self.algs = algs
        self.n = len(algs)
        rng = rng if rng is not None else np.random.default_rng(0)
        per = {a: sizes(a) for a in KYBER_ALGS}  # (pk, sk, ct, ss)
        elem = {"pk": {a: per[a][0] for a in KYBER_ALGS}, "sk": {a: per[a][1] for a in KYBER_ALGS},
                "ct": {a: per[a][2] for a in KYBER_ALGS}, "ss": {a: 32 for a in KYBER_ALGS},
                "coins": {a: 32 for a in KYBER_ALGS}, "kcoins": {a: 64 for a in KYBER_ALGS}}
        self.elem = elem
        self.slot = np.zeros(self.n, dtype=np.int64)
        self.count = {}
        for a in KYBER_ALGS:
            idx = np.nonzero(algs == a)[0]
            self.count[a] = len(idx)
            self.slot[idx] = rng.permutation(len(idx))
        self.region = {}
        self.nbytes = {}
        for kind, sz in elem.items():
            off = 0
            self.region[kind] = {}
            for a in KYBER_ALGS:
                self.region[kind][a] = (off, self.count[a])
                off += self.count[a] * sz[a]
            self.nbytes[kind] = max(off, 8)

    def offsets(self, kind):
        """Byte offset of every request's buffer in arena `kind`."""
        out = np.zeros(self.n, dtype=np.uint64)
        for a in KYBER_ALGS:
            m = self.algs == a
            base = self.region[kind][a][0]
            out[m] = base + self.slot[m] * self.elem[kind][a]
        return out

    def requests(self, key_kind):
        """Descriptor array; key_kind "pk" for encapsulation, "sk" for decapsulation."""
        r = np.zeros(self.n, dtype=KEM_REQ_DTYPE)
        r["alg"] = self.algs
        r["key_off"] = self.offsets(key_kind)
        r["ct_off"] = self.offsets("ct")
        r["ss_off"] = self.offsets("ss")
        r["coins_off"] = self.offsets("coins")
        return r

    def view(self, arena, kind, alg):
        """The element-contiguous slice of a host arena for one parameter set."""
        off, cnt = self.region[kind][alg]
        sz = self.elem[kind][alg]
        return arena[off:off + cnt * sz].reshape(cnt, sz)

    def element(self, arena, kind, i):
        off = int(self.offsets(kind)[i])
        return arena[off:off + self.elem[kind][int(self.algs[i])]]


class DeviceArenas:
    """Device arenas for a MixedLayout, plus the two descriptor arrays."""

    def __init__(self, layout):
        self.layout = layout
        self.buf = {kind: DeviceBuffer(nb) for kind, nb in layout.nbytes.items()}
        self.enc_reqs = DeviceBuffer.from_array(layout.requests("pk"))
        self.dec_reqs = DeviceBuffer.from_array(layout.requests("sk"))

    def region_ptr(self, kind, alg):
        """A DeviceBuffer-like view of one parameter set's slice (for the dense calls)."""
        off, cnt = self.layout.region[kind][alg]
        return _Slice(self.buf[kind], off, cnt * self.layout.elem[kind][alg])

    def download(self, kind):
        return self.buf[kind].download((self.layout.nbytes[kind],))


class _Slice:
    """A byte range of a DeviceBuffer that the binding accepts as a device pointer."""

    def __init__(self, parent, off, nbytes):
        self.ptr = parent.ptr + off
        self.nbytes = nbytes
