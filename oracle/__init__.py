"""CPU oracle for the CKKS hot path (test infrastructure only).

ctypes bindings over ``oracle/liboracle.so`` (built from ``ckks_oracle.c`` by
``oracle/Makefile``), a plain-C restatement of SEAL-4.1-bs's arithmetic.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this package; the product library never does.
"""
import ctypes
import fractions
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_mulmod.restype = ctypes.c_uint64
        L.or_mulmod.argtypes = [ctypes.c_uint64] * 3
        L.or_powmod.restype = ctypes.c_uint64
        L.or_powmod.argtypes = [ctypes.c_uint64] * 3
        L.or_invmod.restype = ctypes.c_uint64
        L.or_invmod.argtypes = [ctypes.c_uint64] * 2
        L.or_is_prime.restype = ctypes.c_int
        L.or_is_prime.argtypes = [ctypes.c_uint64]
        L.or_get_primes.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t, u64p]
        L.or_coeff_modulus_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), ctypes.c_size_t, u64p]
        L.or_minimal_primitive_root.restype = ctypes.c_uint64
        L.or_minimal_primitive_root.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.or_ntt_tables.argtypes = [ctypes.c_int, ctypes.c_uint64, u64p, u64p]
        L.or_ntt_forward.argtypes = [u64p, ctypes.c_int, ctypes.c_uint64, u64p]
        L.or_ntt_inverse.argtypes = [u64p, ctypes.c_int, ctypes.c_uint64, u64p]
        L.or_elt_from_step.restype = ctypes.c_uint32
        L.or_elt_from_step.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_galois_table_ntt.argtypes = [ctypes.c_int, ctypes.c_uint32, u32p]
        L.or_apply_galois_ntt.argtypes = [u64p, ctypes.c_int, ctypes.c_uint32, u64p]
        L.or_apply_galois_coeff.argtypes = [u64p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, u64p]
        L.or_rescale_poly_ntt.argtypes = [u64p, ctypes.c_size_t, ctypes.c_int, u64p, ctypes.POINTER(u64p),
                                          ctypes.POINTER(u64p)]
        L.or_switch_key.argtypes = [u64p, u64p, ctypes.c_size_t, ctypes.c_int, u64p, ctypes.c_size_t, u64p,
                                    ctypes.POINTER(u64p), ctypes.POINTER(u64p)]
        L.or_rotate_sum.argtypes = [u64p, ctypes.POINTER(u64p), u32p, ctypes.POINTER(u64p), ctypes.c_size_t,
                                    ctypes.c_size_t, ctypes.c_int, u64p, ctypes.c_size_t, ctypes.POINTER(u64p),
                                    ctypes.POINTER(u64p)]
        L.or_tensor.argtypes = [u64p, u64p, ctypes.c_size_t, ctypes.c_int, u64p, u64p]
        L.or_dyadic_mul.argtypes = [u64p, u64p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, u64p, u64p]
        L.or_encoder_create.restype = ctypes.c_void_p
        L.or_encoder_create.argtypes = [ctypes.c_int]
        L.or_encoder_destroy.argtypes = [ctypes.c_void_p]
        L.or_ckks_encode.restype = ctypes.c_int
        L.or_ckks_encode.argtypes = [ctypes.c_void_p, f64p, f64p, ctypes.c_size_t, ctypes.c_double, u64p,
                                     ctypes.c_size_t, ctypes.POINTER(u64p), u64p]
        L.or_ckks_encode_coeffs.restype = None
        L.or_ckks_encode_coeffs.argtypes = [ctypes.c_void_p, f64p, f64p, ctypes.c_size_t, ctypes.c_double, f64p]
        L.or_ckks_encode_const.restype = ctypes.c_int
        L.or_ckks_encode_const.argtypes = [ctypes.c_double, ctypes.c_double, u64p, ctypes.c_size_t, ctypes.c_int,
                                           u64p]
        L.or_crt_compose.argtypes = [u64p, ctypes.c_size_t, ctypes.c_size_t, u64p, u64p]
        L.or_ckks_decode.argtypes = [ctypes.c_void_p, u64p, ctypes.c_size_t, u64p, ctypes.POINTER(u64p),
                                     ctypes.c_double, ctypes.c_size_t, f64p, f64p]
        vpp = ctypes.c_void_p
        L.or_port_create.restype = vpp
        L.or_port_create.argtypes = [ctypes.c_int, u64p, ctypes.c_size_t]
        L.or_port_destroy.argtypes = [vpp]
        L.or_port_ntt_forward.argtypes = [vpp, ctypes.c_size_t, u64p]
        L.or_port_ntt_inverse.argtypes = [vpp, ctypes.c_size_t, u64p]
        L.or_port_switch_key.argtypes = [vpp, u64p, u64p, ctypes.c_size_t, u64p]
        L.or_port_rescale.argtypes = [vpp, u64p, ctypes.c_size_t]
        L.or_port_multiply_plain.argtypes = [vpp, u64p, u64p, ctypes.c_size_t, ctypes.c_size_t, u64p]
        L.or_port_tensor.argtypes = [vpp, u64p, u64p, ctypes.c_size_t, u64p]
        L.or_port_add.argtypes = [vpp, u64p, u64p, ctypes.c_size_t, ctypes.c_size_t, u64p]
        L.or_port_encode_residues.argtypes = [vpp, f64p, ctypes.c_size_t, u64p]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def _arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


# ---------------------------------------------------------------- scalars
def mulmod(a, b, q):
    return int(lib().or_mulmod(a, b, q))


def invmod(a, q):
    return int(lib().or_invmod(a, q))


def is_prime(q):
    return bool(lib().or_is_prime(q))


def coeff_modulus_create(n, bits):
    out = np.zeros(len(bits), dtype=np.uint64)
    b = (ctypes.c_int * len(bits))(*bits)
    if lib().or_coeff_modulus_create(n, b, len(bits), _p(out)):
        raise ValueError("failed to find enough qualifying primes")
    return [int(v) for v in out]


def minimal_primitive_root(degree, q):
    return int(lib().or_minimal_primitive_root(degree, q))


# ---------------------------------------------------------------- NTT
_tables = {}


def ntt_tables(logn, q):
    key = (logn, int(q))
    if key not in _tables:
        n = 1 << logn
        rp = np.zeros(n, dtype=np.uint64)
        irp = np.zeros(n, dtype=np.uint64)
        if lib().or_ntt_tables(logn, q, _p(rp), _p(irp)):
            raise ValueError("invalid modulus for NTT")
        _tables[key] = (rp, irp)
    return _tables[key]


def ntt_forward(a, logn, q):
    a = _arr(a).copy()
    rp, _ = ntt_tables(logn, q)
    lib().or_ntt_forward(_p(a), logn, q, _p(rp))
    return a


def ntt_inverse(a, logn, q):
    a = _arr(a).copy()
    _, irp = ntt_tables(logn, q)
    lib().or_ntt_inverse(_p(a), logn, q, _p(irp))
    return a


def _table_ptrs(logn, moduli, which):
    arrs = [ntt_tables(logn, q)[which] for q in moduli]
    ptrs = (u64p * len(arrs))(*[_p(a) for a in arrs])
    return ptrs, arrs


# ---------------------------------------------------------------- galois
def elt_from_step(logn, step):
    return int(lib().or_elt_from_step(logn, step))


def galois_table_ntt(logn, elt):
    t = np.zeros(1 << logn, dtype=np.uint32)
    lib().or_galois_table_ntt(logn, elt, t.ctypes.data_as(u32p))
    return t


def apply_galois_ntt(a, logn, elt):
    a = _arr(a)
    out = np.zeros_like(a)
    lib().or_apply_galois_ntt(_p(a), logn, elt, _p(out))
    return out


def apply_galois_coeff(a, logn, elt, q):
    a = _arr(a)
    out = np.zeros_like(a)
    lib().or_apply_galois_coeff(_p(a), logn, elt, q, _p(out))
    return out


# ---------------------------------------------------------------- RNS ops
def rescale_poly_ntt(poly, moduli, logn):
    """poly: (l, n) NTT-form limbs for moduli[0..l-1]; returns (l-1, n)."""
    poly = _arr(poly).copy()
    l = poly.shape[0]
    mods = _arr(moduli[:l])
    rps, keep1 = _table_ptrs(logn, mods, 0)
    irps, keep2 = _table_ptrs(logn, mods, 1)
    lib().or_rescale_poly_ntt(_p(poly), l, logn, _p(mods), rps, irps)
    return poly[: l - 1].copy()


def switch_key(ct, target, key_moduli, key, logn):
    """ct: (2, l, n) destination; target: (l, n); key: (L, 2, L+1, n)."""
    ct = _arr(ct).copy()
    target = _arr(target)
    key = _arr(key)
    l = target.shape[0]
    kmods = _arr(key_moduli)
    rps, keep1 = _table_ptrs(logn, kmods, 0)
    irps, keep2 = _table_ptrs(logn, kmods, 1)
    lib().or_switch_key(_p(ct), _p(target), l, logn, _p(kmods), len(kmods), _p(key), rps, irps)
    return ct


def rotate_sum(cts, elts, keys, key_moduli, logn):
    """sum_r rot_r(cts[r]) with one ModDown, each rotation hoisted (its digits
    are sigma(lift(c1 mod q_j))); elts[r] = 0 adds cts[r] itself.
    cts: list of (2, l, n); keys: list of (L, 2, L+1, n) (None for identity terms)."""
    cs = [_arr(c) for c in cts]
    l = cs[0].shape[1]
    n = 1 << logn
    kmods = _arr(key_moduli)
    ks = [_arr(k) if k is not None else cs[0] for k in keys]
    out = np.zeros((2, l, n), dtype=np.uint64)
    rps, keep1 = _table_ptrs(logn, kmods, 0)
    irps, keep2 = _table_ptrs(logn, kmods, 1)
    cp = (u64p * len(cs))(*[_p(c) for c in cs])
    kp = (u64p * len(ks))(*[_p(k) for k in ks])
    el = np.ascontiguousarray(np.asarray(elts, dtype=np.uint32))
    lib().or_rotate_sum(_p(out), cp, el.ctypes.data_as(u32p), kp, len(cs), l, logn, _p(kmods), len(kmods), rps, irps)
    return out


def rotate_hoisted(ct, elt, key, key_moduli, logn):
    """one hoisted rotation: sigma(c0) + ModDown(sum_j sigma(lift(c1 mod q_j)) key[j])"""
    return rotate_sum([ct], [elt], [key], key_moduli, logn)


def tensor(a, b, moduli, logn):
    a = _arr(a)
    b = _arr(b)
    l = a.shape[1]
    out = np.zeros((3, l, 1 << logn), dtype=np.uint64)
    lib().or_tensor(_p(a), _p(b), l, logn, _p(_arr(moduli[:l])), _p(out))
    return out


def dyadic_mul(a, pt, moduli, logn):
    """a: (polys, l, n) times plaintext pt: (l, n)."""
    a = _arr(a)
    pt = _arr(pt)
    out = np.zeros_like(a)
    lib().or_dyadic_mul(_p(a), _p(pt), a.shape[0], a.shape[1], logn, _p(_arr(moduli[: a.shape[1]])), _p(out))
    return out


def crt_compose(limbs, moduli):
    limbs = _arr(limbs)
    l, n = limbs.shape
    out = np.zeros((n, l), dtype=np.uint64)
    lib().or_crt_compose(_p(limbs), l, n, _p(_arr(moduli[:l])), _p(out))
    return out


# ---------------------------------------------------------------- encoder
class Encoder:
    def __init__(self, logn):
        self.logn = logn
        self.h = lib().or_encoder_create(logn)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_encoder_destroy(self.h)
            self.h = None

    def encode(self, values, scale, moduli):
        v = np.asarray(values)
        re = np.ascontiguousarray(np.real(v), dtype=np.float64)
        im = np.ascontiguousarray(np.imag(v), dtype=np.float64) if np.iscomplexobj(v) else None
        l = len(moduli)
        mods = _arr(moduli)
        out = np.zeros((l, 1 << self.logn), dtype=np.uint64)
        rps, keep = _table_ptrs(self.logn, mods, 0)
        r = lib().or_ckks_encode(self.h, re.ctypes.data_as(f64p), im.ctypes.data_as(f64p) if im is not None else None,
                                 len(re), scale, _p(mods), l, rps, _p(out))
        if r:
            raise ValueError("encoded values are too large")
        return out

    def encode_coeffs(self, values, scale):
        """Unrounded plaintext coefficients (float64, length n) in SEAL's order."""
        v = np.asarray(values)
        re = np.ascontiguousarray(np.real(v), dtype=np.float64)
        im = np.ascontiguousarray(np.imag(v), dtype=np.float64) if np.iscomplexobj(v) else None
        out = np.zeros(1 << self.logn)
        lib().or_ckks_encode_coeffs(self.h, re.ctypes.data_as(f64p), im.ctypes.data_as(f64p) if im is not None else None,
                                    len(re), scale, out.ctypes.data_as(f64p))
        return out

    @staticmethod
    def residues_exact(coeffs, moduli):
        """std::round(c) (halves away from zero) mod q for every coefficient, with
        exact Python integers (any size)."""
        ints = []
        for c in coeffs:
            x = fractions.Fraction(float(c))
            m = math.floor(abs(x) + fractions.Fraction(1, 2))
            ints.append(-m if x < 0 else m)
        return np.array([[x % int(q) for x in ints] for q in moduli], dtype=np.uint64)

    def encode_const(self, value, scale, moduli):
        l = len(moduli)
        mods = _arr(moduli)
        out = np.zeros((l, 1 << self.logn), dtype=np.uint64)
        if lib().or_ckks_encode_const(value, scale, _p(mods), l, self.logn, _p(out)):
            raise ValueError("encoded value is too large")
        return out

    def decode(self, plain, moduli, scale, sparse_slots=0, complex_out=False):
        plain = _arr(plain)
        l = plain.shape[0]
        mods = _arr(moduli[:l])
        slots = sparse_slots or (1 << (self.logn - 1))
        re = np.zeros(slots)
        im = np.zeros(slots)
        irps, keep = _table_ptrs(self.logn, mods, 1)
        lib().or_ckks_decode(self.h, _p(plain), l, _p(mods), irps, scale, sparse_slots, re.ctypes.data_as(f64p),
                             im.ctypes.data_as(f64p))
        return re + 1j * im if complex_out else re


# ---------------------------------------------------------------- performance-faithful CPU port
class Port:
    """SEAL's CPU methods (lazy Harvey NTT, Shoup and Barrett products, SEAL-form
    key switching) over `moduli` (L data primes then one special prime), for the
    CPU baseline's timings; results equal the oracle's word for word.  ctypes
    releases the GIL in every call, so threads run calls concurrently."""

    def __init__(self, logn, moduli):
        self.logn = logn
        self.n = 1 << logn
        self.moduli = [int(q) for q in moduli]
        m = _arr(self.moduli)
        self.h = lib().or_port_create(logn, _p(m), len(self.moduli))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_port_destroy(self.h)
            self.h = None

    def ntt_forward(self, a, prime):
        a = _arr(a).copy()
        lib().or_port_ntt_forward(self.h, prime, _p(a))
        return a

    def ntt_inverse(self, a, prime):
        a = _arr(a).copy()
        lib().or_port_ntt_inverse(self.h, prime, _p(a))
        return a

    def switch_key(self, ct, target, key):
        ct = _arr(ct).copy()
        target = _arr(target)
        lib().or_port_switch_key(self.h, _p(ct), _p(target), target.shape[0], _p(_arr(key)))
        return ct

    def rescale(self, poly):
        poly = _arr(poly).copy()
        lib().or_port_rescale(self.h, _p(poly), poly.shape[0])
        return poly[:-1].copy()

    def multiply_plain(self, a, pt):
        a = _arr(a)
        out = np.zeros_like(a)
        lib().or_port_multiply_plain(self.h, _p(a), _p(_arr(pt)), a.shape[0], a.shape[1], _p(out))
        return out

    def tensor(self, a, b):
        a = _arr(a)
        out = np.zeros((3,) + a.shape[1:], dtype=np.uint64)
        lib().or_port_tensor(self.h, _p(a), _p(_arr(b)), a.shape[1], _p(out))
        return out

    def add(self, a, b):
        a = _arr(a)
        out = np.zeros_like(a)
        lib().or_port_add(self.h, _p(a), _p(_arr(b)), a.shape[0], a.shape[1], _p(out))
        return out

    def encode_residues(self, coeffs, limbs):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        out = np.zeros((limbs, self.n), dtype=np.uint64)
        lib().or_port_encode_residues(self.h, c.ctypes.data_as(f64p), limbs, _p(out))
        return out
