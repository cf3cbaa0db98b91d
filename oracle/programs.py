"""Oracle restatements of NEXUS programs over the CPU primitives (test
infrastructure only): ciphertexts are numpy arrays [size][limbs][n] of NTT-form
residues, evaluated with the plain C oracle in SEAL's order of operations.

Sources: src/matrix_mul.cpp:101-170 (expand_ciphertext, multiply_power_of_x),
SEAL evaluator.cpp:2563-2665 (apply_galois_inplace),
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
