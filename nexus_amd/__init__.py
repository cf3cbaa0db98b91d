"""MI355X-native CKKS evaluation engine (Python view of the C ABI).

Thin ctypes layer over ``libnexus_ckks.so`` (include/nexus_ckks.h).  There is no
Python or CPU fallback: importing this package loads the HIP library, and any
call that needs the GPU fails with :class:`NexusError` when none is present.
"""
import atexit
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnexus_ckks.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libnexus_ckks.so is not built; run `python -m nexus_amd.build`")

_lib = ctypes.CDLL(LIB_PATH)

u64p = ctypes.POINTER(ctypes.c_uint64)
f64p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p
sz = ctypes.c_size_t


class NexusError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


_ERR_TYPES = {-1: ValueError, -2: RuntimeError, -3: IndexError}


def _check(rc):
    if rc != 0:
        msg = _lib.nx_last_error().decode()
        raise NexusError(rc, msg)


def _sig(name, args):
    f = getattr(_lib, name)
    f.restype = ctypes.c_int
    f.argtypes = args
    return f


_lib.nx_last_error.restype = ctypes.c_char_p
_lib.nx_last_error.argtypes = []
_sig("nx_device_count", [ctypes.POINTER(ctypes.c_int)])
_sig("nx_device_synchronize", [])
_sig("nx_context_create", [ctypes.c_int, ctypes.POINTER(ctypes.c_int), sz, ctypes.c_uint32, sz, ctypes.POINTER(vp)])
_sig("nx_context_create_hybrid", [ctypes.c_int, ctypes.POINTER(ctypes.c_int), sz, sz, ctypes.c_uint32, sz,
                                  ctypes.POINTER(vp)])
_sig("nx_context_key_info", [vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)])
_sig("nx_context_destroy", [vp])
_sig("nx_context_bind_lane", [vp, sz])
_sig("nx_context_info", [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(sz), ctypes.POINTER(sz)])
_sig("nx_context_moduli", [vp, u64p])
_sig("nx_context_synchronize", [vp])
_sig("nx_galois_elt_from_step", [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)])
_sig("nx_context_memory", [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)])
_sig("nx_context_trim", [vp])
for _n in ("ciphertext", "plaintext"):
    _sig("nx_%s_create" % _n, [vp, ctypes.POINTER(vp)])
    _sig("nx_%s_destroy" % _n, [vp])
    _sig("nx_%s_copy" % _n, [vp, vp])
    _sig("nx_%s_download" % _n, [vp, u64p])
    _sig("nx_%s_set_scale" % _n, [vp, ctypes.c_double])
_sig("nx_ciphertext_upload", [vp, u64p, sz, sz, ctypes.c_double, ctypes.c_int])
_sig("nx_ciphertext_device_data", [vp, ctypes.POINTER(u64p), ctypes.POINTER(sz)])
_sig("nx_ciphertext_info", [vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz),
                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)])
_sig("nx_plaintext_upload", [vp, u64p, sz, ctypes.c_double, ctypes.c_int])
_sig("nx_plaintext_info", [vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_double)])
_sig("nx_secret_key_generate", [vp, ctypes.c_uint64, ctypes.POINTER(vp)])
_sig("nx_secret_key_from_coeffs", [vp, ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(vp)])
_sig("nx_secret_key_coeffs", [vp, ctypes.POINTER(ctypes.c_int8)])
_sig("nx_secret_key_destroy", [vp])
_sig("nx_kswitch_keys_create", [vp, ctypes.POINTER(vp)])
_sig("nx_kswitch_keys_destroy", [vp])
_sig("nx_relin_key_generate", [vp, vp, ctypes.c_uint64, vp])
_sig("nx_galois_keys_generate", [vp, vp, ctypes.POINTER(ctypes.c_uint32), sz, ctypes.c_uint64, vp])
_sig("nx_kswitch_key_upload", [vp, ctypes.c_uint32, u64p])
_sig("nx_kswitch_key_download", [vp, ctypes.c_uint32, u64p])
_sig("nx_kswitch_keys_has", [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)])
_sig("nx_encrypt_symmetric", [vp, vp, vp, ctypes.c_uint64, vp])
_sig("nx_decrypt", [vp, vp, vp, vp])
_sig("nx_encode", [vp, f64p, f64p, sz, sz, ctypes.c_double, vp])
_sig("nx_encode_constant", [vp, ctypes.c_double, sz, ctypes.c_double, vp])
_sig("nx_decode", [vp, vp, f64p, f64p, sz])
for _n in ("add", "sub", "multiply"):
    _sig("nx_" + _n, [vp, vp, vp])
for _n in ("add_plain", "sub_plain", "multiply_plain"):
    _sig("nx_" + _n, [vp, vp, vp])
for _n in ("negate", "square", "rescale_to_next", "mod_switch_to_next", "transform_to_ntt", "transform_from_ntt"):
    _sig("nx_" + _n, [vp, vp])
_sig("nx_multiply_sum", [ctypes.POINTER(vp), ctypes.POINTER(vp), sz, vp])
_sig("nx_relinearize", [vp, vp, vp])
_sig("nx_mod_switch_to", [vp, sz, vp])
_sig("nx_plain_mod_switch_to", [vp, sz, vp])
_sig("nx_apply_galois", [vp, ctypes.c_uint32, vp, vp])
_sig("nx_rotate_vector", [vp, ctypes.c_int, vp, vp])
_sig("nx_complex_conjugate", [vp, vp, vp])
_sig("nx_add_many", [ctypes.POINTER(vp), sz, vp])
_sig("nx_multiply_const", [vp, ctypes.c_double, vp])
_sig("nx_add_const", [vp, ctypes.c_double, vp])
_sig("nx_multiply_power_of_x", [vp, ctypes.c_int64, vp])


_sig("nx_add_reduced_error", [vp, vp, vp])
_sig("nx_sub_reduced_error", [vp, vp, vp])
_sig("nx_multiply_reduced_error", [vp, vp, vp, vp])
_sig("nx_multiply_vector_reduced_error", [vp, f64p, f64p, sz, vp])
_sig("nx_double", [vp, vp])
_sig("nx_coeff_modulus_create", [sz, ctypes.POINTER(ctypes.c_int), sz, u64p])


def lib():
    return _lib


# Handles are released by their destructors while the interpreter runs; once
# it starts shutting down, nothing calls into the library any more (the
# process exit reclaims device memory), so teardown order cannot matter.
_state = {"alive": True}
atexit.register(lambda: _state.__setitem__("alive", False))


def _release(obj, fn, _st=_state):
    h = getattr(obj, "h", None)
    if h and _st["alive"]:
        fn(h)
    obj.h = None


def device_count():
    c = ctypes.c_int()
    _check(_lib.nx_device_count(ctypes.byref(c)))
    return c.value


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(u64p)


class Context:
    """CKKS parameters + device tables (SEALContext / PhantomContext)."""

    def __init__(self, log_n, bit_sizes, hamming_weight=0, sparse_slots=0, special_primes=1):
        """bit_sizes: data primes then special primes; special_primes > 1 selects
        hybrid key switching (digits of special_primes data primes)."""
        bits = (ctypes.c_int * len(bit_sizes))(*bit_sizes)
        h = vp()
        _check(_lib.nx_context_create_hybrid(log_n, bits, len(bit_sizes), special_primes, hamming_weight,
                                             sparse_slots, ctypes.byref(h)))
        self.h = h
        self.log_n = log_n
        self.n = 1 << log_n
        mods = np.zeros(len(bit_sizes), dtype=np.uint64)
        _check(_lib.nx_context_moduli(self.h, mods.ctypes.data_as(u64p)))
        self.moduli = [int(x) for x in mods]
        self.special_primes = special_primes
        self.L = len(bit_sizes) - special_primes
        d, kw = sz(), sz()
        _check(_lib.nx_context_key_info(self.h, None, ctypes.byref(d), ctypes.byref(kw)))
        self.digits, self.key_words = d.value, kw.value
        self.slot_count = self.n // 2

    def __del__(self, _fn=_lib.nx_context_destroy, _rel=_release):
        _rel(self, _fn)

    def limbs_of(self, chain_index):
        return self.L + 1 - chain_index

    def synchronize(self):
        _check(_lib.nx_context_synchronize(self.h))

    def bind_lane(self, lane):
        """Run this thread's operations on lane `lane` (1..64: a HIP stream with its
        own memory cache) of this context; 0 returns to the default stream."""
        _check(_lib.nx_context_bind_lane(self.h, lane))

    def elt_from_step(self, step):
        e = ctypes.c_uint32()
        _check(_lib.nx_galois_elt_from_step(self.h, step, ctypes.byref(e)))
        return e.value

    def memory(self):
        a, b = sz(), sz()
        _check(_lib.nx_context_memory(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def trim(self):
        """release the cached device blocks of every lane (after a device synchronisation)"""
        _check(_lib.nx_context_trim(self.h))


class Ciphertext:
    def __init__(self, ctx):
        self.ctx = ctx
        h = vp()
        _check(_lib.nx_ciphertext_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def __del__(self, _fn=_lib.nx_ciphertext_destroy, _rel=_release):
        _rel(self, _fn)

    def info(self):
        s, ci, cm, sc, nt = sz(), sz(), sz(), ctypes.c_double(), ctypes.c_int()
        _check(_lib.nx_ciphertext_info(self.h, ctypes.byref(s), ctypes.byref(ci), ctypes.byref(cm),
                                       ctypes.byref(sc), ctypes.byref(nt)))
        return dict(size=s.value, chain_index=ci.value, limbs=cm.value, scale=sc.value, ntt=bool(nt.value))

    @property
    def scale(self):
        return self.info()["scale"]

    @scale.setter
    def scale(self, v):
        _check(_lib.nx_ciphertext_set_scale(self.h, v))

    @property
    def chain_index(self):
        return self.info()["chain_index"]

    @property
    def coeff_modulus_size(self):
        return self.info()["limbs"]

    @classmethod
    def from_array(cls, ctx, data, chain_index, scale, ntt_form=True):
        """data: (size, limbs, n) uint64 residues."""
        ct = cls(ctx)
        arr, p = _u64(data)
        _check(_lib.nx_ciphertext_upload(ct.h, p, arr.shape[0], chain_index, scale, 1 if ntt_form else 0))
        return ct

    def to_array(self):
        i = self.info()
        out = np.zeros((i["size"], i["limbs"], self.ctx.n), dtype=np.uint64)
        _check(_lib.nx_ciphertext_download(self.h, out.ctypes.data_as(u64p)))
        return out

    def copy(self):
        c = Ciphertext(self.ctx)
        _check(_lib.nx_ciphertext_copy(c.h, self.h))
        return c

    def device_data(self):
        """(device address, words) of the residues (PhantomCiphertext::data())"""
        p, w = ctypes.POINTER(ctypes.c_uint64)(), sz()
        _check(_lib.nx_ciphertext_device_data(self.h, ctypes.byref(p), ctypes.byref(w)))
        return ctypes.cast(p, ctypes.c_void_p).value or 0, w.value


class Plaintext:
    def __init__(self, ctx):
        self.ctx = ctx
        h = vp()
        _check(_lib.nx_plaintext_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def __del__(self, _fn=_lib.nx_plaintext_destroy, _rel=_release):
        _rel(self, _fn)

    def info(self):
        ci, cm, sc = sz(), sz(), ctypes.c_double()
        _check(_lib.nx_plaintext_info(self.h, ctypes.byref(ci), ctypes.byref(cm), ctypes.byref(sc)))
        return dict(chain_index=ci.value, limbs=cm.value, scale=sc.value)

    @property
    def scale(self):
        return self.info()["scale"]

    @scale.setter
    def scale(self, v):
        _check(_lib.nx_plaintext_set_scale(self.h, v))

    @classmethod
    def from_array(cls, ctx, data, chain_index, scale, ntt_form=True):
        pt = cls(ctx)
        arr, p = _u64(data)
        _check(_lib.nx_plaintext_upload(pt.h, p, chain_index, scale, 1 if ntt_form else 0))
        return pt

    def to_array(self):
        i = self.info()
        out = np.zeros((i["limbs"], self.ctx.n), dtype=np.uint64)
        _check(_lib.nx_plaintext_download(self.h, out.ctypes.data_as(u64p)))
        return out


class SecretKey:
    def __init__(self, ctx, seed=None, coeffs=None):
        self.ctx = ctx
        h = vp()
        if coeffs is not None:
            c = np.ascontiguousarray(coeffs, dtype=np.int8)
            _check(_lib.nx_secret_key_from_coeffs(ctx.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                                  ctypes.byref(h)))
        else:
            _check(_lib.nx_secret_key_generate(ctx.h, seed or 0, ctypes.byref(h)))
        self.h = h

    def __del__(self, _fn=_lib.nx_secret_key_destroy, _rel=_release):
        _rel(self, _fn)

    def coeffs(self):
        out = np.zeros(self.ctx.n, dtype=np.int8)
        _check(_lib.nx_secret_key_coeffs(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))))
        return out


class KSwitchKeys:
    """A set of device-resident key-switching keys (relin = element 0)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = vp()
        _check(_lib.nx_kswitch_keys_create(ctx.h, ctypes.byref(h)))
        self.h = h

    def __del__(self, _fn=_lib.nx_kswitch_keys_destroy, _rel=_release):
        _rel(self, _fn)

    def generate_relin(self, sk, seed=1):
        _check(_lib.nx_relin_key_generate(self.ctx.h, sk.h, seed, self.h))
        return self

    def generate_galois(self, sk, elts, seed=2):
        arr = (ctypes.c_uint32 * len(elts))(*elts)
        _check(_lib.nx_galois_keys_generate(self.ctx.h, sk.h, arr, len(elts), seed, self.h))
        return self

    def key_array(self, elt):
        c = self.ctx
        out = np.zeros((c.digits, 2, c.L + c.special_primes, c.n), dtype=np.uint64)
        _check(_lib.nx_kswitch_key_download(self.h, elt, out.ctypes.data_as(u64p)))
        return out

    def upload(self, elt, data):
        arr, p = _u64(data)
        _check(_lib.nx_kswitch_key_upload(self.h, elt, p))

    def has(self, elt):
        v = ctypes.c_int()
        _check(_lib.nx_kswitch_keys_has(self.h, elt, ctypes.byref(v)))
        return bool(v.value)


# ---------------------------------------------------------------- encoder / crypto
def encode(ctx, values, scale, chain_index=1):
    v = np.asarray(values)
    re = np.ascontiguousarray(np.real(v), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(v), dtype=np.float64) if np.iscomplexobj(v) else None
    pt = Plaintext(ctx)
    _check(_lib.nx_encode(ctx.h, re.ctypes.data_as(f64p), im.ctypes.data_as(f64p) if im is not None else None,
                          len(re), chain_index, scale, pt.h))
    return pt


def encode_constant(ctx, value, scale, chain_index=1):
    pt = Plaintext(ctx)
    _check(_lib.nx_encode_constant(ctx.h, value, chain_index, scale, pt.h))
    return pt


def decode(ctx, pt, count=None, complex_out=False):
    count = count or ctx.slot_count
    re = np.zeros(count)
    im = np.zeros(count)
    _check(_lib.nx_decode(ctx.h, pt.h, re.ctypes.data_as(f64p), im.ctypes.data_as(f64p), count))
    return re + 1j * im if complex_out else re


def encrypt(ctx, sk, pt, seed=0):
    ct = Ciphertext(ctx)
    _check(_lib.nx_encrypt_symmetric(ctx.h, sk.h, pt.h, seed, ct.h))
    return ct


def decrypt(ctx, sk, ct):
    pt = Plaintext(ctx)
    _check(_lib.nx_decrypt(ctx.h, sk.h, ct.h, pt.h))
    return pt


# ---------------------------------------------------------------- evaluator
def _out(ctx, dest):
    return dest if dest is not None else Ciphertext(ctx)


def _binop(name):
    f = getattr(_lib, "nx_" + name)

    def op(a, b, dest=None):
        d = _out(a.ctx, dest)
        _check(f(a.h, b.h, d.h))
        return d

    op.__name__ = name
    return op


def _unop(name):
    f = getattr(_lib, "nx_" + name)

    def op(a, dest=None):
        d = _out(a.ctx, dest)
        _check(f(a.h, d.h))
        return d

    op.__name__ = name
    return op


add = _binop("add")
sub = _binop("sub")
multiply = _binop("multiply")
add_plain = _binop("add_plain")
sub_plain = _binop("sub_plain")
multiply_plain = _binop("multiply_plain")
negate = _unop("negate")
square = _unop("square")
rescale_to_next = _unop("rescale_to_next")
mod_switch_to_next = _unop("mod_switch_to_next")
transform_to_ntt = _unop("transform_to_ntt")
transform_from_ntt = _unop("transform_from_ntt")


def relinearize(a, relin_keys, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_relinearize(a.h, relin_keys.h, d.h))
    return d


def mod_switch_to(a, chain_index, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_mod_switch_to(a.h, chain_index, d.h))
    return d


def plain_mod_switch_to(p, chain_index):
    d = Plaintext(p.ctx)
    _check(_lib.nx_plain_mod_switch_to(p.h, chain_index, d.h))
    return d


def apply_galois(a, elt, galois_keys, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_apply_galois(a.h, elt, galois_keys.h, d.h))
    return d


def rotate_vector(a, steps, galois_keys, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_rotate_vector(a.h, steps, galois_keys.h, d.h))
    return d


def complex_conjugate(a, galois_keys, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_complex_conjugate(a.h, galois_keys.h, d.h))
    return d


def add_many(cts, dest=None):
    d = _out(cts[0].ctx, dest)
    arr = (vp * len(cts))(*[c.h for c in cts])
    _check(_lib.nx_add_many(arr, len(cts), d.h))
    return d


def multiply_const(a, value, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_multiply_const(a.h, value, d.h))
    return d


def add_const(a, value, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_add_const(a.h, value, d.h))
    return d


def add_reduced_error(a, b, dest=None):
    """SEAL-4.1-bs add_reduced_error (evaluator.cpp:420-466)"""
    d = _out(a.ctx, dest)
    _check(_lib.nx_add_reduced_error(a.h, b.h, d.h))
    return d


def sub_reduced_error(a, b, dest=None):
    """SEAL-4.1-bs sub_reduced_error (evaluator.cpp:468-514)"""
    d = _out(a.ctx, dest)
    _check(_lib.nx_sub_reduced_error(a.h, b.h, d.h))
    return d


def multiply_reduced_error(a, b, relin_keys, dest=None):
    """SEAL-4.1-bs multiply_reduced_error (evaluator.cpp:516-594), relinearised"""
    d = _out(a.ctx, dest)
    _check(_lib.nx_multiply_reduced_error(a.h, b.h, relin_keys.h, d.h))
    return d


def multiply_vector_reduced_error(a, values, dest=None):
    """values encoded at a's level and scale, then multiply_plain
    (cuda/src/ckks_evaluator.cuh:371-393)"""
    v = np.asarray(values)
    re = np.ascontiguousarray(np.real(v), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(v), dtype=np.float64) if np.iscomplexobj(v) else None
    d = _out(a.ctx, dest)
    _check(_lib.nx_multiply_vector_reduced_error(a.h, re.ctypes.data_as(f64p),
                                                 im.ctypes.data_as(f64p) if im is not None else None, len(re), d.h))
    return d


def double(a, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_double(a.h, d.h))
    return d


def coeff_modulus_create(n, bit_sizes):
    """SEAL CoeffModulus::Create on the host (no GPU needed)"""
    bits = (ctypes.c_int * len(bit_sizes))(*bit_sizes)
    out = np.zeros(len(bit_sizes), dtype=np.uint64)
    _check(_lib.nx_coeff_modulus_create(n, bits, len(bit_sizes), out.ctypes.data_as(u64p)))
    return [int(x) for x in out]


def multiply_power_of_x(a, exponent, dest=None):
    d = _out(a.ctx, dest)
    _check(_lib.nx_multiply_power_of_x(a.h, exponent, d.h))
    return d


_sig("nx_mod_raise", [vp, vp])
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sig("nx_hash_blake2b", [_u8p, sz, _u8p, sz])
_sig("nx_parms_id", [vp, sz, u64p])
for _n in ("ciphertext", "plaintext"):
    _sig("nx_%s_save" % _n, [vp, _u8p, sz, ctypes.POINTER(sz)])
    _sig("nx_%s_load" % _n, [vp, _u8p, sz, vp])
_sig("nx_kswitch_keys_save", [vp, ctypes.c_int, _u8p, sz, ctypes.POINTER(sz)])
_sig("nx_kswitch_keys_load", [vp, _u8p, sz, ctypes.c_int, vp])


def blake2b(data, digest_size=32):
    """BLAKE2b (RFC 7693) of bytes, as SEAL's HashFunction computes it."""
    b = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(bytes(data) or b"\0")
    out = (ctypes.c_uint8 * digest_size)()
    _check(_lib.nx_hash_blake2b(b, len(data), out, digest_size))
    return bytes(out)


def parms_id(ctx, chain_index):
    out = np.zeros(4, dtype=np.uint64)
    _check(_lib.nx_parms_id(ctx.h, chain_index, out.ctypes.data_as(u64p)))
    return [int(x) for x in out]


def _save(fn, *args):
    n = sz()
    _check(fn(*args, None, 0, ctypes.byref(n)))
    buf = (ctypes.c_uint8 * n.value)()
    _check(fn(*args, buf, n.value, ctypes.byref(n)))
    return bytes(buf)


def save(obj, galois=False):
    """SEAL 4.1 bytes of a Ciphertext, Plaintext or KSwitchKeys (compression none)."""
    if isinstance(obj, Ciphertext):
        return _save(_lib.nx_ciphertext_save, obj.h)
    if isinstance(obj, Plaintext):
        return _save(_lib.nx_plaintext_save, obj.h)
    return _save(_lib.nx_kswitch_keys_save, obj.h, 1 if galois else 0)


def load(ctx, data, kind, galois=False):
    """kind: "ciphertext", "plaintext" or "keys"."""
    buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
    if kind == "ciphertext":
        out = Ciphertext(ctx)
        _check(_lib.nx_ciphertext_load(ctx.h, buf, len(data), out.h))
    elif kind == "plaintext":
        out = Plaintext(ctx)
        _check(_lib.nx_plaintext_load(ctx.h, buf, len(data), out.h))
    else:
        out = KSwitchKeys(ctx)
        _check(_lib.nx_kswitch_keys_load(ctx.h, buf, len(data), 1 if galois else 0, out.h))
    return out
_sig("nx_rotate_hoisted", [vp, ctypes.POINTER(ctypes.c_int), sz, vp, ctypes.POINTER(vp)])
_sig("nx_rotate_hoisted_many", [ctypes.POINTER(vp), sz, ctypes.POINTER(ctypes.c_int), sz, vp, ctypes.POINTER(vp)])
_sig("nx_rotate_sum_many", [ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(sz), sz, vp,
                            ctypes.POINTER(vp)])
_sig("nx_hoisting", [vp, ctypes.POINTER(ctypes.c_int), sz, vp, vp])


_sig("nx_multiply_plain_accumulate", [ctypes.POINTER(vp), ctypes.POINTER(vp), sz, vp])


_sig("nx_rotate_sum", [ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int), sz, vp, vp])


def rotate_sum(cts, steps, gk):
    """sum_i rot(cts[i], steps[i]) with a single ModDown"""
    out = Ciphertext(cts[0].ctx)
    st = (ctypes.c_int * len(steps))(*steps)
    _check(_lib.nx_rotate_sum((vp * len(cts))(*[c.h for c in cts]), st, len(cts), gk.h, out.h))
    return out


def multiply_plain_accumulate(cts, pts, dest=None):
    """sum_r cts[r] * pts[r] in one pass."""
    d = _out(cts[0].ctx, dest)
    c = (vp * len(cts))(*[x.h for x in cts])
    p = (vp * len(pts))(*[x.h for x in pts])
    _check(_lib.nx_multiply_plain_accumulate(c, p, len(cts), d.h))
    return d


_sig("nx_linear_combination", [ctypes.POINTER(vp), f64p, f64p, sz, sz, vp])


def linear_combination(cts, values, scales, chain_index):
    """sum_k cts[k] * encode_constant(values[k], scales[k]) at chain_index, each
    input read at its first limbs (word for word mod_switch + multiply_plain + add)"""
    K = len(cts)
    out = Ciphertext(cts[0].ctx)
    v = (ctypes.c_double * K)(*values)
    sc = (ctypes.c_double * K)(*scales)
    _check(_lib.nx_linear_combination((vp * K)(*[c.h for c in cts]), v, sc, K, chain_index, out.h))
    return out


_sig("nx_multiply_plain_accumulate_multi", [ctypes.POINTER(vp), sz, ctypes.POINTER(vp), sz, ctypes.POINTER(vp)])


def multiply_plain_accumulate_multi(cts, rows):
    """[sum_r rows[g][r] * cts[r] for g] with each ciphertext read once per 4 rows
    (rows[g][r] may be None: no term)."""
    K, G = len(cts), len(rows)
    flat = []
    for row in rows:
        if len(row) != K:
            raise ValueError("every row needs one entry per ciphertext")
        flat += [p.h if p is not None else None for p in row]
    outs = [Ciphertext(cts[0].ctx) for _ in range(G)]
    _check(_lib.nx_multiply_plain_accumulate_multi((vp * K)(*[c.h for c in cts]), K, (vp * (G * K))(*flat), G,
                                                   (vp * G)(*[o.h for o in outs])))
    return outs


_sig("nx_rescale_to_next_many", [ctypes.POINTER(vp), sz, ctypes.POINTER(vp)])
_sig("nx_multiply_relin_rescale_many", [ctypes.POINTER(vp), ctypes.POINTER(vp), sz, vp, ctypes.POINTER(vp)])
_sig("nx_multiply_relin_rescale_fused_many", [ctypes.POINTER(vp), ctypes.POINTER(vp), sz, vp, ctypes.POINTER(vp)])


def rescale_to_next_many(cts):
    """[rescale_to_next(c) for c in cts] as one batched pass"""
    outs = [Ciphertext(cts[0].ctx) for _ in cts]
    _check(_lib.nx_rescale_to_next_many((vp * len(cts))(*[c.h for c in cts]), len(cts),
                                        (vp * len(outs))(*[o.h for o in outs])))
    return outs


def multiply_sum(a, b, dest=None):
    """sum_i a[i] * b[i] as one size-3 ciphertext (word for word multiply + add)"""
    d = dest or Ciphertext(a[0].ctx)
    _check(_lib.nx_multiply_sum((vp * len(a))(*[c.h for c in a]), (vp * len(b))(*[c.h for c in b]), len(a), d.h))
    return d


def multiply_relin_rescale_many(a, b, relin_keys, fused=False):
    """[rescale_to_next(relinearize(x * y)) for x, y in zip(a, b)] as one batched pass.

    fused=True relinearises and rescales in one rounding (hybrid key switching):
    every coefficient is within 1 of the unfused result."""
    outs = [Ciphertext(a[0].ctx) for _ in a]
    fn = _lib.nx_multiply_relin_rescale_fused_many if fused else _lib.nx_multiply_relin_rescale_many
    _check(fn((vp * len(a))(*[c.h for c in a]), (vp * len(b))(*[c.h for c in b]), len(a), relin_keys.h,
              (vp * len(outs))(*[o.h for o in outs])))
    return outs


_sig("nx_periodic_plain_create", [vp, f64p, sz, sz, sz, ctypes.c_double, ctypes.POINTER(vp)])
_sig("nx_periodic_plain_destroy", [vp])
_sig("nx_periodic_plain_expand", [vp, sz, vp])
_sig("nx_multiply_periodic_mix", [ctypes.POINTER(vp), sz, vp, ctypes.POINTER(ctypes.c_int32), sz,
                                  ctypes.POINTER(vp)])


class PeriodicPlaintexts:
    """Real slot vectors repeating every `period` slots (values: [count][period]),
    encoded exactly as polynomials in X^(N / (2 period)) and kept as 2 period
    NTT words per limb."""

    def __init__(self, ctx, values, scale, chain_index=1):
        v = np.ascontiguousarray(np.atleast_2d(np.asarray(values, dtype=np.float64)))
        self.ctx = ctx
        self.count, self.period = v.shape
        h = vp()
        _check(_lib.nx_periodic_plain_create(ctx.h, v.ctypes.data_as(f64p), self.period, self.count, chain_index,
                                             scale, ctypes.byref(h)))
        self.h = h

    def __del__(self, _fn=_lib.nx_periodic_plain_destroy, _rel=_release):
        _rel(self, _fn)

    def expand(self, index):
        """the full NTT-form plaintext of entry `index`"""
        pt = Plaintext(self.ctx)
        _check(_lib.nx_periodic_plain_expand(self.h, index, pt.h))
        return pt


def multiply_periodic_mix(cts, w, index):
    """[sum_k w[index[m, k]] (.) cts[k] for each row m of index] (-1: no term)"""
    idx = np.ascontiguousarray(np.asarray(index, dtype=np.int32))
    M, K = idx.shape
    if K != len(cts):
        raise ValueError("index needs one column per input")
    outs = [Ciphertext(cts[0].ctx) for _ in range(M)]
    c = (vp * K)(*[x.h for x in cts])
    o = (vp * M)(*[x.h for x in outs])
    _check(_lib.nx_multiply_periodic_mix(c, K, w.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), M, o))
    return outs


def rotate_hoisted(a, steps, galois_keys):
    """[rot(a, s) for s in steps] sharing one ModUp (Phantom hoisting)."""
    outs = [Ciphertext(a.ctx) for _ in steps]
    st = (ctypes.c_int * len(steps))(*steps)
    hs = (vp * len(steps))(*[o.h for o in outs])
    _check(_lib.nx_rotate_hoisted(a.h, st, len(steps), galois_keys.h, hs))
    return outs


def rotate_hoisted_many(cts, steps, galois_keys):
    """[[rot(c, s) for s in steps] for c in cts], hoisted per input, batched across inputs"""
    B, S = len(cts), len(steps)
    outs = [[Ciphertext(cts[0].ctx) for _ in steps] for _ in cts]
    st = (ctypes.c_int * S)(*steps)
    hs = (vp * (B * S))(*[o.h for row in outs for o in row])
    _check(_lib.nx_rotate_hoisted_many((vp * B)(*[c.h for c in cts]), B, st, S, galois_keys.h, hs))
    return outs


def rotate_sum_many(groups, steps, galois_keys):
    """[sum_i rot(g[i], s[i]) for g, s in zip(groups, steps)], the ModDowns batched"""
    flat = [c for g in groups for c in g]
    st = [s for ss in steps for s in ss]
    sizes = (sz * len(groups))(*[len(g) for g in groups])
    outs = [Ciphertext(flat[0].ctx) for _ in groups]
    _check(_lib.nx_rotate_sum_many((vp * len(flat))(*[c.h for c in flat]), (ctypes.c_int * len(st))(*st), sizes,
                                   len(groups), galois_keys.h, (vp * len(outs))(*[o.h for o in outs])))
    return outs


def hoisting(a, steps, galois_keys, dest=None):
    """sum_k rot(a, steps[k]) with one ModUp and one ModDown (Phantom hoisting_inplace)."""
    d = _out(a.ctx, dest)
    st = (ctypes.c_int * len(steps))(*steps)
    _check(_lib.nx_hoisting(a.h, st, len(steps), galois_keys.h, d.h))
    return d


def mod_raise(a, dest=None):
    """Lowest-level ciphertext -> top data level, scale unchanged (Bootstrapper.cpp:2937-2991)."""
    d = _out(a.ctx, dest)
    _check(_lib.nx_mod_raise(a.h, d.h))
    return d


# ---------------------------------------------------------------- evaluator programs (include/nexus_eval.h)
_sig("nx_evaluator_create", [vp, vp, vp, vp, ctypes.c_double, ctypes.POINTER(vp)])
_sig("nx_evaluator_destroy", [vp])
_sig("nx_evaluator_set_per_op", [vp, ctypes.c_int])
_sig("nx_eval_sgn", [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp])
_sig("nx_eval_invert_sqrt", [vp, vp, ctypes.c_int, ctypes.c_int, vp])
_sig("nx_eval_exp", [vp, vp, vp])
_sig("nx_eval_inverse", [vp, vp, ctypes.c_int, vp])
_sig("nx_eval_gelu", [vp, vp, vp])
_sig("nx_eval_gelu_many", [vp, ctypes.POINTER(vp), sz, ctypes.POINTER(vp)])
_sig("nx_eval_layer_norm", [vp, vp, ctypes.c_int, vp])
_sig("nx_eval_softmax", [vp, vp, ctypes.c_int, vp])


class CKKSEvaluator:
    """NEXUS CKKSEvaluator programs (GELU, LayerNorm, Softmax, sign, inverse...)."""

    def __init__(self, ctx, relin_keys, galois_keys=None, scale=2.0 ** 40, sk=None):
        self.ctx = ctx
        self._keep = (relin_keys, galois_keys, sk)
        h = vp()
        _check(_lib.nx_evaluator_create(ctx.h, sk.h if sk else None, relin_keys.h,
                                        galois_keys.h if galois_keys else None, scale, ctypes.byref(h)))
        self.h = h
        self.scale = scale

    def __del__(self, _fn=_lib.nx_evaluator_destroy, _rel=_release):
        _rel(self, _fn)

    def set_per_op(self, on=True):
        """per-op mode: the programs over this evaluator issue single-ciphertext
        operations only (the plain per-call drop-in); see include/nexus_eval.h"""
        _check(_lib.nx_evaluator_set_per_op(self.h, 1 if on else 0))

    def _run(self, fn, *args):
        out = Ciphertext(self.ctx)
        _check(fn(self.h, *args, out.h))
        return out

    def sgn_eval(self, x, d_g, d_f, sgn_factor=0.5):
        return self._run(_lib.nx_eval_sgn, x.h, d_g, d_f, sgn_factor)

    def invert_sqrt(self, x, d_newt=5, d_gold=2):
        return self._run(_lib.nx_eval_invert_sqrt, x.h, d_newt, d_gold)

    def exp(self, x):
        return self._run(_lib.nx_eval_exp, x.h)

    def inverse(self, x, iter=4):
        return self._run(_lib.nx_eval_inverse, x.h, iter)

    def gelu(self, x):
        return self._run(_lib.nx_eval_gelu, x.h)

    def gelu_many(self, xs):
        """[gelu(x) for x in xs] in lockstep"""
        outs = [Ciphertext(xs[0].ctx) for _ in xs]
        _check(_lib.nx_eval_gelu_many(self.h, (vp * len(xs))(*[c.h for c in xs]), len(xs),
                                      (vp * len(outs))(*[o.h for o in outs])))
        return outs

    def layer_norm(self, x, length):
        return self._run(_lib.nx_eval_layer_norm, x.h, length)

    def softmax(self, x, length):
        return self._run(_lib.nx_eval_softmax, x.h, length)


# ---------------------------------------------------------------- bootstrapping (include/nexus_eval.h)
_ip = ctypes.POINTER(ctypes.c_int)
_sig("nx_bootstrapper_create", [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.POINTER(vp)])
_sig("nx_bootstrapper_destroy", [vp])
_sig("nx_bootstrapper_rotation_steps", [vp, _ip, sz, ctypes.POINTER(sz)])
_sig("nx_bootstrapper_depth", [vp, _ip])
_sig("nx_autotune", [vp, ctypes.c_int, ctypes.c_char_p, sz, ctypes.POINTER(sz)])
_sig("nx_bootstrap", [vp, vp, vp])
_sig("nx_bootstrap_scaled", [vp, vp, ctypes.c_double, vp])
_sig("nx_bootstrap_many", [vp, ctypes.POINTER(vp), sz, f64p, ctypes.POINTER(vp)])
_sig("nx_boot_mod_raise", [vp, vp, vp])
_sig("nx_bootstrap_many_mixed", [ctypes.POINTER(vp), ctypes.POINTER(vp), sz, f64p, ctypes.POINTER(vp)])
_sig("nx_boot_sub_sum", [vp, vp, vp])
_sig("nx_boot_coeff_to_slot", [vp, vp, vp, vp])
_sig("nx_boot_eval_mod", [vp, vp, vp])
_sig("nx_boot_slot_to_coeff", [vp, vp, vp, ctypes.c_double, vp])
_sig("nx_eval_argmax", [vp, vp, ctypes.c_int, vp])
_sig("nx_boot_plan_info", [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip,
                           ctypes.POINTER(sz)])
_sig("nx_boot_plan_diagonals", [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _ip,
                                f64p, f64p])
_sig("nx_boot_eval_mod_coeffs", [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p, _ip])
_sig("nx_boot_plan_steps", [ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip, sz, ctypes.POINTER(sz)])


def boot_plan_steps(log_half_n, log_slots, K=25):
    """Rotation steps of the bootstrapping linear transforms (0 = conjugation); host only."""
    n = sz()
    _check(_lib.nx_boot_plan_steps(log_half_n, log_slots, float(K), None, 0, ctypes.byref(n)))
    buf = (ctypes.c_int * n.value)()
    _check(_lib.nx_boot_plan_steps(log_half_n, log_slots, float(K), buf, n.value, ctypes.byref(n)))
    return list(buf)


def boot_eval_mod_coeffs(K=25, degree=59, double_angles=2, log_width=10):
    """(Chebyshev coefficients of the EvalMod polynomial, levels it consumes); host only."""
    c = np.zeros(degree + 1)
    d = ctypes.c_int()
    _check(_lib.nx_boot_eval_mod_coeffs(K, degree, double_angles, log_width, c.ctypes.data_as(f64p), ctypes.byref(d)))
    return c, d.value


def boot_plan_level(log_half_n, log_slots, which, level, K=25):
    """One BSGS level of the plan: dict(stride, g, levels, diags={(j, i): complex[N/2]}); host only."""
    levels, stride, g, cnt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), sz()
    _check(_lib.nx_boot_plan_info(log_half_n, log_slots, float(K), which, level, ctypes.byref(levels),
                                  ctypes.byref(stride), ctypes.byref(g), ctypes.byref(cnt)))
    nh = 1 << log_half_n
    c = cnt.value
    js = (ctypes.c_int * c)()
    is_ = (ctypes.c_int * c)()
    re = np.zeros(c * nh)
    im = np.zeros(c * nh)
    _check(_lib.nx_boot_plan_diagonals(log_half_n, log_slots, float(K), which, level, js, is_,
                                       re.ctypes.data_as(f64p), im.ctypes.data_as(f64p)))
    d = (re + 1j * im).reshape(c, nh)
    return {"levels": levels.value, "stride": stride.value, "g": g.value,
            "diags": {(js[k], is_[k]): d[k] for k in range(c)}}


def autotune(ctx, reps=10):
    """Choose NTT tile widths and the multiply-accumulate layout for ctx's ring
    size by timing them (each candidate checked bit-exact against the default);
    returns the report as a dict."""
    import json
    need = sz()
    buf = ctypes.create_string_buffer(1 << 16)
    _check(_lib.nx_autotune(ctx.h, reps, buf, 1 << 16, ctypes.byref(need)))
    return json.loads(buf.value.decode())


class Bootstrapper:
    """CKKS bootstrapping (reference Bootstrapper::bootstrap_3): ModRaise, SubSum,
    CoeffToSlot, EvalMod, SlotToCoeff.  `ev` must hold relinearisation keys and
    Galois keys for every step of `rotation_steps()`."""

    def __init__(self, ev, log_slots=0, K=25, degree=59, double_angles=2, log_width=10, final_scale=0.0):
        self.ev = ev
        self.ctx = ev.ctx
        h = vp()
        _check(_lib.nx_bootstrapper_create(ev.h, log_slots, K, degree, double_angles, log_width, final_scale,
                                           ctypes.byref(h)))
        self.h = h
        self.log_slots = log_slots or self.ctx.log_n - 1
        self.dense = self.log_slots == self.ctx.log_n - 1

    def __del__(self, _fn=_lib.nx_bootstrapper_destroy, _rel=_release):
        _rel(self, _fn)

    def rotation_steps(self):
        n = sz()
        _check(_lib.nx_bootstrapper_rotation_steps(self.h, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_int * n.value)()
        _check(_lib.nx_bootstrapper_rotation_steps(self.h, buf, n.value, ctypes.byref(n)))
        return list(buf)

    def depth(self):
        d = ctypes.c_int()
        _check(_lib.nx_bootstrapper_depth(self.h, ctypes.byref(d)))
        return d.value

    def _run(self, fn, *args):
        out = Ciphertext(self.ctx)
        _check(fn(self.h, *args, out.h))
        return out

    def bootstrap(self, ct, factor=None):
        """fresh ciphertext of the message (times `factor`, folded into the last
        SlotToCoeff level, when given)"""
        if factor is None:
            return self._run(_lib.nx_bootstrap, ct.h)
        return self._run(_lib.nx_bootstrap_scaled, ct.h, float(factor))

    def bootstrap_many(self, cts, factors=None):
        """[bootstrap(c, f)] for all inputs, with one EvalMod over the whole set"""
        outs = [Ciphertext(self.ctx) for _ in cts]
        f = (ctypes.c_double * len(cts))(*(factors if factors is not None else [1.0] * len(cts)))
        _check(_lib.nx_bootstrap_many(self.h, (vp * len(cts))(*[c.h for c in cts]), len(cts), f,
                                      (vp * len(outs))(*[o.h for o in outs])))
        return outs

    @staticmethod
    def bootstrap_mixed(bts, cts, factors=None):
        """[bts[i].bootstrap(cts[i], factors[i])] as one batch: bootstrappers of one
        context with different slot counts, inputs at any level"""
        n = len(cts)
        outs = [Ciphertext(cts[0].ctx) for _ in cts]
        f = (ctypes.c_double * n)(*(factors if factors is not None else [1.0] * n))
        _check(_lib.nx_bootstrap_many_mixed((vp * n)(*[b.h for b in bts]), (vp * n)(*[c.h for c in cts]), n, f,
                                            (vp * n)(*[o.h for o in outs])))
        return outs

    def mod_raise(self, ct):
        return self._run(_lib.nx_boot_mod_raise, ct.h)

    def sub_sum(self, ct):
        return self._run(_lib.nx_boot_sub_sum, ct.h)

    def coeff_to_slot(self, ct):
        re = Ciphertext(self.ctx)
        im = Ciphertext(self.ctx) if self.dense else None
        _check(_lib.nx_boot_coeff_to_slot(self.h, ct.h, re.h, im.h if im is not None else None))
        return re, im

    def eval_mod(self, ct):
        return self._run(_lib.nx_boot_eval_mod, ct.h)

    def argmax(self, x, length):
        """ArgmaxEvaluator::argmax (src/argmax.cpp:9-64) with this bootstrapper."""
        out = Ciphertext(self.ctx)
        _check(_lib.nx_eval_argmax(self.h, x.h, length, out.h))
        return out

    def slot_to_coeff(self, f_re, f_im, input_scale):
        out = Ciphertext(self.ctx)
        _check(_lib.nx_boot_slot_to_coeff(self.h, f_re.h, f_im.h if f_im is not None else None, input_scale, out.h))
        return out


# ---------------------------------------------------------------- BERT layer (include/nexus_eval.h)
_f64p = ctypes.POINTER(ctypes.c_double)
_sig("nx_bert_layer_create", [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_f64p] * 6
     + [ctypes.c_int, ctypes.POINTER(vp)])
_sig("nx_bert_layer_destroy", [vp])
_sig("nx_bert_layer_rotation_steps", [sz, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), sz, ctypes.POINTER(sz)])
_sig("nx_bert_layer_forward", [vp, ctypes.POINTER(vp), sz, ctypes.c_int, ctypes.POINTER(vp)])
_sig("nx_bert_layer_forward_packed", [vp, ctypes.POINTER(vp), sz, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), sz, ctypes.POINTER(vp)])
_sig("nx_bert_layer_forward_sets", [vp, ctypes.POINTER(vp), sz, sz, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(vp)])
_sig("nx_bert_layer_linear", [vp, ctypes.c_int, ctypes.POINTER(vp), sz, ctypes.c_double, ctypes.POINTER(vp), sz])
_sig("nx_bert_layer_attention", [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), sz, ctypes.c_int,
                                 ctypes.POINTER(vp)])
_sig("nx_bert_layer_layer_norm", [vp, ctypes.POINTER(vp), sz, ctypes.POINTER(vp)])
_sig("nx_bert_layer_set_stage_timing", [vp, ctypes.c_int])
_sig("nx_bert_layer_stage_ms", [vp, ctypes.c_char_p, sz, _f64p, sz, ctypes.POINTER(sz)])

BERT_GROUP = 256
BERT_MATRICES = ("wq", "wk", "wv", "wo", "w1", "w2")


def _handles(cts):
    return (vp * len(cts))(*[c.h for c in cts])


def bert_rotation_steps(slots, hidden=768, heads=12, ffn=3072, baby=32):
    n = sz()
    _check(_lib.nx_bert_layer_rotation_steps(slots, hidden, heads, ffn, baby, None, 0, ctypes.byref(n)))
    buf = (ctypes.c_int * n.value)()
    _check(_lib.nx_bert_layer_rotation_steps(slots, hidden, heads, ffn, baby, buf, n.value, ctypes.byref(n)))
    return list(buf)


def bert_pack(x, slots):
    """token matrix (tokens x features) -> features / 256 slot vectors: token i
    in slots [256 i, 256 i + 256) of ciphertext c = feature // 256"""
    x = np.asarray(x, dtype=np.float64)
    s, d = x.shape
    if d % BERT_GROUP or s * BERT_GROUP > slots:
        raise ValueError("features must be a multiple of 256 and tokens fit the slots")
    out = []
    for c in range(d // BERT_GROUP):
        v = np.zeros(slots)
        v[: s * BERT_GROUP] = x[:, c * BERT_GROUP:(c + 1) * BERT_GROUP].reshape(-1)
        out.append(v)
    return out


def bert_unpack(vecs, tokens):
    """inverse of bert_pack for the first `tokens` tokens"""
    return np.concatenate([np.asarray(v)[: tokens * BERT_GROUP].reshape(tokens, BERT_GROUP) for v in vecs], axis=1)


class BertLayer:
    """Encrypted BERT-base encoder layer (include/nexus/bert.hpp) over a
    Bootstrapper; weights: dict of float64 matrices wq, wk, wv, wo (hidden x
    hidden), w1 (hidden x ffn), w2 (ffn x hidden), applied as x @ W."""

    def __init__(self, bt, weights, heads=None, baby=32):
        self.bt = bt
        self.ctx = bt.ctx
        ws = [np.ascontiguousarray(weights[k], dtype=np.float64) for k in BERT_MATRICES]
        self.hidden = ws[0].shape[0]
        self.ffn = ws[4].shape[1]
        self.heads = heads or self.hidden // 64
        h = vp()
        _check(_lib.nx_bert_layer_create(bt.h, self.hidden, self.heads, self.ffn,
                                         *[w.ctypes.data_as(_f64p) for w in ws], baby, ctypes.byref(h)))
        self.h = h

    def __del__(self, _fn=_lib.nx_bert_layer_destroy, _rel=_release):
        _rel(self, _fn)

    def _outs(self, n):
        return [Ciphertext(self.ctx) for _ in range(n)]

    def forward(self, x, seq_len=None, segments=None):
        """one sequence of seq_len rows, or several packed sequences given as
        segments [(start, len), ...] (attention stays within each)"""
        out = self._outs(len(x))
        if segments is None:
            _check(_lib.nx_bert_layer_forward(self.h, _handles(x), len(x), seq_len, _handles(out)))
        else:
            st = (ctypes.c_int * len(segments))(*[int(a) for a, _ in segments])
            ln = (ctypes.c_int * len(segments))(*[int(b) for _, b in segments])
            _check(_lib.nx_bert_layer_forward_packed(self.h, _handles(x), len(x), st, ln, len(segments),
                                                     _handles(out)))
        return out

    def forward_sets(self, xs, segments):
        """several packed sets in lockstep: xs[s] the set's ciphertexts,
        segments[s] its [(start, len), ...]; returns the outputs per set"""
        n = len(xs[0])
        flat = [c for x in xs for c in x]
        out = self._outs(len(flat))
        cnt = (ctypes.c_int * len(xs))(*[len(g) for g in segments])
        st = [int(a) for g in segments for a, _ in g]
        ln = [int(b) for g in segments for _, b in g]
        _check(_lib.nx_bert_layer_forward_sets(self.h, _handles(flat), len(xs), n, cnt, (ctypes.c_int * len(st))(*st),
                                               (ctypes.c_int * len(ln))(*ln), _handles(out)))
        return [out[s * n:(s + 1) * n] for s in range(len(xs))]

    def linear(self, matrix, x, factor=1.0):
        m = BERT_MATRICES.index(matrix)
        n_out = (self.ffn if matrix == "w1" else self.hidden) // BERT_GROUP
        out = self._outs(n_out)
        _check(_lib.nx_bert_layer_linear(self.h, m, _handles(x), len(x), factor, _handles(out), n_out))
        return out

    def attention(self, q, k, v, seq_len):
        out = self._outs(len(q))
        _check(_lib.nx_bert_layer_attention(self.h, _handles(q), _handles(k), _handles(v), len(q), seq_len,
                                            _handles(out)))
        return out

    def layer_norm(self, x):
        out = self._outs(len(x))
        _check(_lib.nx_bert_layer_layer_norm(self.h, _handles(x), len(x), _handles(out)))
        return out

    def set_stage_timing(self, on):
        _check(_lib.nx_bert_layer_set_stage_timing(self.h, 1 if on else 0))

    def stage_ms(self):
        """[(stage, ms)] of the last forward() with stage timing on"""
        n = sz()
        names = ctypes.create_string_buffer(1024)
        ms = (ctypes.c_double * 32)()
        _check(_lib.nx_bert_layer_stage_ms(self.h, names, 1024, ms, 32, ctypes.byref(n)))
        keys = names.value.decode().split(",") if n.value else []
        return list(zip(keys, list(ms)[: n.value]))


# ---------------------------------------------------------------- serving (include/nexus_eval.h)
_sig("nx_server_create", [vp, ctypes.c_int, sz, vp, ctypes.POINTER(vp)])
_sig("nx_server_destroy", [vp])
_sig("nx_server_submit", [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(vp), sz,
                          ctypes.POINTER(ctypes.c_uint64)])
_sig("nx_server_wait", [vp, ctypes.c_uint64, ctypes.POINTER(vp), sz])
_sig("nx_server_drain", [vp])
_sig("nx_server_stats", [vp, ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(sz)])

_sig("nx_server_create_bert", [vp, ctypes.c_int, sz, vp, vp, ctypes.c_int, sz, ctypes.POINTER(vp)])
_sig("nx_server_submit_bert", [vp, ctypes.POINTER(vp), sz, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)])
_sig("nx_server_submit_bert_set", [vp, ctypes.POINTER(vp), sz, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int), sz, ctypes.POINTER(ctypes.c_uint64)])
_sig("nx_server_pause", [vp, ctypes.c_int])
_sig("nx_server_set_bert_concurrency", [vp, ctypes.c_int])
_sig("nx_server_batching_stats", [vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)])

JOB_GELU, JOB_SOFTMAX, JOB_LAYER_NORM, JOB_BOOTSTRAP, JOB_BERT = 0, 1, 2, 3, 4


class Server:
    """Request loop: requests (lists of evaluator calls) run on `lanes` concurrent
    streams, admitted in arrival order within `memory_budget` bytes of footprint."""

    def __init__(self, ev, lanes=2, memory_budget=0, bootstrapper=None, bert=None, pack=True, bert_footprint=0):
        """bert: a BertLayer whose weights every lane shares (requests via
        submit_bert); pack: continuous batching of BERT requests"""
        self.ev = ev
        self.ctx = ev.ctx
        self._keep = (ev, bootstrapper, bert)
        h = vp()
        if bert is not None:
            _check(_lib.nx_server_create_bert(ev.h, lanes, memory_budget, bert.bt.h, bert.h, 1 if pack else 0,
                                              bert_footprint, ctypes.byref(h)))
        else:
            _check(_lib.nx_server_create(ev.h, lanes, memory_budget, bootstrapper.h if bootstrapper else None,
                                         ctypes.byref(h)))
        self.h = h
        self._counts = {}

    def __del__(self, _fn=_lib.nx_server_destroy, _rel=_release):
        _rel(self, _fn)

    def submit(self, jobs):
        """jobs: [(kind, row_length, ciphertext), ...] -> request id"""
        k = (ctypes.c_int * len(jobs))(*[j[0] for j in jobs])
        ln = (ctypes.c_int * len(jobs))(*[j[1] for j in jobs])
        ins = (vp * len(jobs))(*[j[2].h for j in jobs])
        rid = ctypes.c_uint64()
        _check(_lib.nx_server_submit(self.h, k, ln, ins, len(jobs), ctypes.byref(rid)))
        self._counts[rid.value] = len(jobs)
        return rid.value

    def submit_bert(self, cts, seq_len):
        """one BERT-layer request: a sequence at rows [0, seq_len) of hidden/256 ciphertexts"""
        rid = ctypes.c_uint64()
        _check(_lib.nx_server_submit_bert(self.h, _handles(cts), len(cts), seq_len, ctypes.byref(rid)))
        self._counts[rid.value] = len(cts)
        return rid.value

    def submit_bert_set(self, cts, segments):
        """one BERT-layer request over a packed set: segments = [(start, length), ...]"""
        st = (ctypes.c_int * len(segments))(*[int(a) for a, _ in segments])
        ln = (ctypes.c_int * len(segments))(*[int(b) for _, b in segments])
        rid = ctypes.c_uint64()
        _check(_lib.nx_server_submit_bert_set(self.h, _handles(cts), len(cts), st, ln, len(segments),
                                              ctypes.byref(rid)))
        self._counts[rid.value] = len(cts)
        return rid.value

    def set_bert_concurrency(self, n):
        """BERT layer calls in flight at once (1..lanes)"""
        _check(_lib.nx_server_set_bert_concurrency(self.h, n))

    def pause(self, paused=True):
        _check(_lib.nx_server_pause(self.h, 1 if paused else 0))

    def batching_stats(self):
        a, b, c = sz(), sz(), sz()
        _check(_lib.nx_server_batching_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"packed_calls": a.value, "packed_requests": b.value, "peak_footprint": c.value}

    def wait(self, rid):
        cnt = self._counts.pop(rid)
        outs = [Ciphertext(self.ctx) for _ in range(cnt)]
        arr = (vp * cnt)(*[o.h for o in outs])
        _check(_lib.nx_server_wait(self.h, rid, arr, cnt))
        return outs

    def drain(self):
        _check(_lib.nx_server_drain(self.h))

    def stats(self):
        c, pk = sz(), sz()
        p50, p99, mean = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(_lib.nx_server_stats(self.h, ctypes.byref(c), ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(mean),
                                    ctypes.byref(pk)))
        return {"completed": c.value, "p50_ms": p50.value, "p99_ms": p99.value, "mean_ms": mean.value,
                "peak_concurrent": pk.value}


# ---------------------------------------------------------------- live kernel timing (include/nexus_prof.h)
_sig("nx_profile_select", [ctypes.c_char_p])
_sig("nx_profile_read", [ctypes.c_char_p, ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double)])
_sig("nx_profile_read_flops", [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)])
_sig("nx_profile_classes", [ctypes.c_char_p, sz])


def profile_select(classes):
    """Time every launch of the given kernel classes ("*" = all) with HIP events."""
    _check(_lib.nx_profile_select(classes.encode()))


def profile_read(kernel_class):
    n, ms, b = sz(), ctypes.c_double(), ctypes.c_double()
    _check(_lib.nx_profile_read(kernel_class.encode(), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
    return n.value, ms.value, b.value


def profile_read_flops(kernel_class):
    """summed algorithmic double-precision operations of a class (0 unless it records them)"""
    f = ctypes.c_double()
    _check(_lib.nx_profile_read_flops(kernel_class.encode(), ctypes.byref(f)))
    return f.value


def profile_classes():
    buf = ctypes.create_string_buffer(4096)
    _check(_lib.nx_profile_classes(buf, 4096))
    return [c for c in buf.value.decode().split(",") if c]


_sig("nx_profile_valu_rate", [ctypes.c_int, ctypes.POINTER(ctypes.c_double)])


def valu_rate(kind):
    """measured VALU rate of the device: kind 0 = 64x64->128-bit products/s,
    kind 1 = double-precision FMAs/s"""
    v = ctypes.c_double()
    _check(_lib.nx_profile_valu_rate(kind, ctypes.byref(v)))
    return v.value


_sig("nx_op_tally", [ctypes.c_int, ctypes.c_char_p, sz, ctypes.POINTER(sz)])


def op_tally_start():
    """clear and start the operation tally (include/nexus_prof.h)"""
    _check(_lib.nx_op_tally(1, None, 0, None))


def op_tally_stop():
    """stop the tally: {kind: {limbs: count}}"""
    import json
    need = sz()
    _check(_lib.nx_op_tally(0, None, 0, ctypes.byref(need)))
    # the first call stopped the tally; read the kept counts with a buffer
    buf = ctypes.create_string_buffer(max(need.value, 2))
    _check(_lib.nx_op_tally(0, buf, len(buf), ctypes.byref(need)))
    return {k: {int(l): c for l, c in v.items()} for k, v in json.loads(buf.value.decode()).items()}


_sig("nx_rescale_to_next_noalloc", [vp, vp, sz, sz, vp])


def rescale_to_next_noalloc(ctx, in_ptr, polys, limbs, out_ptr):
    """allocation-free rescale of [polys][limbs][N] residues at device address
    in_ptr into [polys][limbs-1][N] at out_ptr (include/nexus_ckks.h)"""
    _check(_lib.nx_rescale_to_next_noalloc(ctx.h, vp(in_ptr), polys, limbs, vp(out_ptr)))
