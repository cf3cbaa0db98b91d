"""Python view of the MatrixMul pipeline C ABI (include/nexus_mm.h)."""
import ctypes

import numpy as np

from . import Ciphertext, _check, _lib, _release, f64p, sz, vp

_lib.nx_mm_compress.restype = ctypes.c_int
_lib.nx_mm_compress.argtypes = [vp, vp, f64p, sz, ctypes.c_double, ctypes.c_uint64, vp]
_lib.nx_mm_expand.restype = ctypes.c_int
_lib.nx_mm_expand.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), vp, ctypes.POINTER(vp)]
_lib.nx_ct_batch_count.restype = ctypes.c_int
_lib.nx_ct_batch_count.argtypes = [vp, ctypes.POINTER(sz)]
_lib.nx_ct_batch_extract.restype = ctypes.c_int
_lib.nx_ct_batch_extract.argtypes = [vp, sz, vp]
_lib.nx_ct_batch_destroy.restype = ctypes.c_int
_lib.nx_ct_batch_destroy.argtypes = [vp]
_lib.nx_mm_accumulate.restype = ctypes.c_int
_lib.nx_mm_accumulate.argtypes = [ctypes.POINTER(vp), sz, ctypes.POINTER(vp), sz, sz, ctypes.POINTER(vp)]


def reference_elts(n, logm):
    """Galois elements of the expansion rounds: n / 2^i + 1 (src/ckks_evaluator.h:79-81)."""
    return [(n + (1 << i)) // (1 << i) for i in range(logm)]


class CtBatch:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def __del__(self, _fn=_lib.nx_ct_batch_destroy, _rel=_release):
        _rel(self, _fn)

    def __len__(self):
        c = sz()
        _check(_lib.nx_ct_batch_count(self.h, ctypes.byref(c)))
        return c.value

    def __getitem__(self, i):
        out = Ciphertext(self.ctx)
        _check(_lib.nx_ct_batch_extract(self.h, i, out.h))
        return out


def compress(ctx, sk, values, scale=1e10, seed=0):
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = Ciphertext(ctx)
    _check(_lib.nx_mm_compress(ctx.h, sk.h, v.ctypes.data_as(f64p), len(v), scale, seed, out.h))
    return out


def expand(ct, m, galois_keys, elts=None):
    logm = m.bit_length() - 1
    elts = elts or reference_elts(ct.ctx.n, logm)
    arr = (ctypes.c_uint32 * len(elts))(*elts)
    h = vp()
    _check(_lib.nx_mm_expand(ct.h, m, arr, galois_keys.h, ctypes.byref(h)))
    return CtBatch(ct.ctx, h)


def accumulate(a_pts, batches, cols):
    ctx = batches[0].ctx
    outs = [Ciphertext(ctx) for _ in range(cols)]
    a = (vp * len(a_pts))(*[p.h for p in a_pts])
    b = (vp * len(batches))(*[x.h for x in batches])
    o = (vp * cols)(*[c.h for c in outs])
    _check(_lib.nx_mm_accumulate(a, len(a_pts), b, len(batches), cols, o))
    return outs
