"""Oracle restatements of NEXUS programs over the CPU primitives (test
infrastructure only): ciphertexts are numpy arrays [size][limbs][n] of NTT-form
residues, evaluated with the plain C oracle in SEAL's order of operations.

Sources: src/matrix_mul.cpp:101-170 (expand_ciphertext, multiply_power_of_x),
SEAL evaluator.cpp:2563-2665 (apply_galois_inplace), SEAL evaluator.cpp:395-594
(multiply_const, *_reduced_error),
util/polyarithsmallmod.h negacyclic_shift_poly_coeffmod.
"""
import numpy as np

import oracle as orc


def negacyclic_shift(a, k, q):
    """a * x^k mod (x^n + 1) in the coefficient domain (index_raw = i + k; the
    coefficient lands at index_raw mod n, negated when bit n of index_raw is set)."""
    n = len(a)
    idx = np.arange(n, dtype=np.int64) + k
    pos = idx % n
    neg = ((idx // n) % 2) == 1
    a = np.asarray(a, dtype=np.uint64)
    vals = np.where(neg & (a != 0), np.uint64(q) - a, a)
    out = np.zeros(n, dtype=np.uint64)
    out[pos] = vals
    return out


def multiply_power_of_x(ct, k, moduli, logn):
    """src/matrix_mul.cpp:131-170: INTT, negacyclic shift by k, NTT (per poly, limb)."""
    out = np.zeros_like(ct)
    for p in range(ct.shape[0]):
        for i in range(ct.shape[1]):
            q = moduli[i]
            coef = orc.ntt_inverse(ct[p, i], logn, q)
            out[p, i] = orc.ntt_forward(negacyclic_shift(coef, k, q), logn, q)
    return out


def add(a, b, moduli):
    q = np.array(moduli[: a.shape[1]], dtype=object)[None, :, None]
    return ((a.astype(object) + b.astype(object)) % q).astype(np.uint64)


def apply_galois(ct, elt, key, key_moduli, logn):
    """SEAL evaluator.cpp:2563-2665: permute c0, c1 in the NTT domain, clear c1,
    key-switch the permuted c1."""
    l = ct.shape[1]
    c0 = np.stack([orc.apply_galois_ntt(ct[0, i], logn, elt) for i in range(l)])
    c1 = np.stack([orc.apply_galois_ntt(ct[1, i], logn, elt) for i in range(l)])
    base = np.stack([c0, np.zeros_like(c0)])
    return orc.switch_key(base, c1, key_moduli, key, logn)


def expand_ciphertext(ct, m, elts, keys, key_moduli, logn):
    """src/matrix_mul.cpp:101-129 (keys: dict elt -> key array)."""
    n = 1 << logn
    logm = m.bit_length() - 1
    temp = [ct]
    for i in range(logm):
        newtemp = [None] * (2 * len(temp))
        index_raw = (n << 1) - (1 << i)
        index = (index_raw * elts[i]) % (n << 1)
        for a in range(len(temp)):
            rot = apply_galois(temp[a], elts[i], keys[elts[i]], key_moduli, logn)
            newtemp[a] = add(temp[a], rot, key_moduli)
            sh = multiply_power_of_x(temp[a], index_raw, key_moduli, logn)
            rsh = multiply_power_of_x(rot, index, key_moduli, logn)
            newtemp[a + len(temp)] = add(sh, rsh, key_moduli)
        temp = newtemp
    return temp


# ---------------------------------------------------------------- reduced-error ops
def sub(a, b, moduli):
    q = np.array(moduli[: a.shape[1]], dtype=object)[None, :, None]
    return ((a.astype(object) - b.astype(object)) % q).astype(np.uint64)


def multiply_const(ct, value, scale, moduli, logn):
    """SEAL evaluator.cpp:403-409: the constant encoded at `scale` (SEAL's
    constant encoder, ckks.cpp:73-200), then multiply_plain"""
    l = ct.shape[1]
    pt = orc.Encoder(logn).encode_const(value, scale, moduli[:l])
    return orc.dyadic_mul(ct, pt, moduli, logn)


def rescale(ct, moduli, logn):
    """evaluator.cpp:1402-1481: every poly divided and rounded by its last prime"""
    return np.stack([orc.rescale_poly_ntt(ct[p], moduli, logn) for p in range(ct.shape[0])])


def reduced_error(op, a, sa, b, sb, moduli, logn, key=None, key_moduli=None):
    """SEAL-4.1-bs add/sub/multiply_inplace_reduced_error (evaluator.cpp:420-594)
    on (a, scale sa) and (b, scale sb): returns (result, scale).  op is "add",
    "sub" or "mul" (relinearised with key over key_moduli).  The scale formulas
    are SEAL's, in its order of double operations."""
    n1, n2 = a.shape[1], b.shape[1]

    def apply(x, y):
        if op == "add":
            return add(x, y, moduli)
        if op == "sub":
            return sub(x, y, moduli)
        return orc.tensor(x, y, moduli, logn)

    if n1 == n2:
        x, y, sx, sy = a, b, sb, sb  # encrypted1 takes encrypted2's scale
    elif n1 < n2:
        qlast = float(moduli[n2 - 1])
        adjust = sa * qlast / (sb * sb)
        e2 = multiply_const(b, adjust, sb, moduli, logn)
        s2 = sa * qlast  # scale override, then the rescale divides by q_last
        e2 = rescale(e2, moduli, logn)[:, :n1].copy()
        s2 = s2 / qlast
        x, y, sx, sy = a, e2, s2, s2
    else:
        qlast = float(moduli[n1 - 1])
        adjust = sb * qlast / (sa * sa)
        e1 = multiply_const(a, adjust, sa, moduli, logn)
        e1 = rescale(e1, moduli, logn)[:, :n2].copy()
        x, y, sx, sy = e1, b, sb, sb  # override to encrypted2's scale
    if op == "mul":
        t = orc.tensor(x, y, moduli, logn)
        return orc.switch_key(t[:2], t[2], key_moduli, key, logn), sx * sy
    return apply(x, y), sx
