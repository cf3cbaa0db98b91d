"""Pins the hoisted-rotation oracles (orc.rotate_sum / orc.rotate_hoisted and
hybrid.rotate_sum_hybrid), which restate Phantom's hoisting_inplace
(cuda/thirdparty/phantom-fhe/src/evaluate.cu:1727-1920): one ModUp of c1 per
input, the automorphism applied to the lifted digits, one ModDown per sum.

* Closed form: sigma applied in the coefficient domain to the exact integer
  lift of each digit ([0, q_j) representative, signs flipped where x^(i elt)
  wraps), schoolbook negacyclic products with the key, the sum over terms,
  and round(X / p) of the CRT value over q_i p -- an independent formulation.
  A permutation applied to the wrong operand, a dropped term or a wrong sign
  fails it.
* The decryption invariant with real Galois keys: rot(ct) decrypts to
  sigma(m) up to key-switching noise, for one rotation and for sums.
* Hybrid: the same sums decrypt to the sum of the rotated messages.
"""
import random

import numpy as np
import pytest

import oracle as orc
from oracle import hybrid

from test_oracle_pins import crt, negacyclic_mul


def auto_coeff_int(a, elt, n):
    """x -> x^elt on integer coefficients mod x^n + 1 (signs flipped on wrap)"""
    out = [0] * n
    for i, v in enumerate(a):
        k = (i * elt) % (2 * n)
        if k < n:
            out[k] += v
        else:
            out[k - n] -= v
    return out


def test_rotate_sum_closed_form():
    logn = 4
    n = 1 << logn
    kmods = orc.coeff_modulus_create(n, [40, 45, 50, 60])  # q0, q1, q2 | p
    L = len(kmods) - 1
    l = 2
    p = kmods[-1]
    rng = random.Random(11)
    elts = [orc.elt_from_step(logn, 1), orc.elt_from_step(logn, 3), 0, orc.elt_from_step(logn, -2)]
    keys_c, keys, cts_c, cts = [], [], [], []
    for _ in elts:
        kc = [[[[rng.randrange(kmods[i]) for _ in range(n)] for i in range(L + 1)] for _ in range(2)]
              for _ in range(L)]
        keys_c.append(kc)
        keys.append(np.array([[[orc.ntt_forward(kc[j][k][i], logn, kmods[i]) for i in range(L + 1)]
                               for k in range(2)] for j in range(L)], dtype=np.uint64))
        cc = [[[rng.randrange(kmods[i]) for _ in range(n)] for i in range(l)] for _ in range(2)]
        cts_c.append(cc)
        cts.append(np.array([[orc.ntt_forward(cc[k][i], logn, kmods[i]) for i in range(l)] for k in range(2)]))
    got = orc.rotate_sum(cts, elts, [k if e else None for k, e in zip(keys, elts)], kmods, logn)

    outs = list(range(l)) + [L]
    for k in range(2):
        acc = {I: [0] * n for I in outs}
        base = [[0] * n for _ in range(l)]
        for r, elt in enumerate(elts):
            if elt == 0:
                for i in range(l):
                    base[i] = [x + y for x, y in zip(base[i], cts_c[r][k][i])]
                continue
            if k == 0:
                for i in range(l):
                    base[i] = [x + y for x, y in zip(base[i], auto_coeff_int(cts_c[r][0][i], elt, n))]
            for I in outs:
                qI = kmods[I]
                for J in range(l):
                    lifted = auto_coeff_int(cts_c[r][1][J], elt, n)  # sigma of the [0, q_J) integers
                    prod = negacyclic_mul([v % qI for v in lifted], keys_c[r][J][k][I], qI)
                    acc[I] = [(x + y) % qI for x, y in zip(acc[I], prod)]
        for i in range(l):
            qi = kmods[i]
            res = orc.ntt_inverse(got[k][i], logn, qi)
            for c in range(n):
                X = crt([acc[i][c], acc[L][c]], [qi, p])
                expect = (base[i][c] + (X + (p - 1) // 2) // p) % qi
                assert int(res[c]) == expect, (k, i, c)


def _ntt(coeffs, q, logn):
    return orc.ntt_forward(np.array([int(c) % q for c in coeffs], dtype=np.uint64), logn, q)


def _galois_key(rng, s, elt, kmods, L, logn, P, alpha=1):
    """key_j = (-a s + e + [i in digit j] P sigma(s), a): switches sigma(s) to s"""
    n = 1 << logn
    s_rot = auto_coeff_int(list(s), elt, n)
    digits = (L + alpha - 1) // alpha
    K = len(kmods)
    key = np.zeros((digits, 2, K, n), dtype=np.uint64)
    for j in range(digits):
        e = rng.integers(-3, 4, n)
        for i, q in enumerate(kmods):
            a = rng.integers(0, q, n, dtype=np.uint64)
            b = (-(a.astype(object) * _ntt(s, q, logn).astype(object)) + _ntt(e, q, logn).astype(object)) % q
            if i < L and j * alpha <= i < (j + 1) * alpha:
                b = (b + (P % q) * _ntt(s_rot, q, logn).astype(object)) % q
            key[j, 0, i] = np.array(b, dtype=np.uint64)
            key[j, 1, i] = a
    return key


def _phase(ct, s, mods, logn):
    """coefficients of c0 + c1 s, centred"""
    out = []
    for i, q in enumerate(mods):
        v = (ct[0][i].astype(object) + ct[1][i].astype(object) * _ntt(s, q, logn).astype(object)) % q
        c = orc.ntt_inverse(np.array([int(x) for x in v], dtype=np.uint64), logn, q)
        out.append([int(x) - q if int(x) > q // 2 else int(x) for x in c])
    return out


@pytest.mark.parametrize("steps", [[1], [5, -3], [2, 0, 7, 2]])
def test_rotate_sum_decrypts(steps):
    logn = 6
    n = 1 << logn
    kmods = orc.coeff_modulus_create(n, [50, 40, 40, 58])
    L, P = 3, kmods[-1]
    l = 3
    rng = np.random.default_rng(len(steps) * 7 + steps[0])
    s = rng.integers(-1, 2, n)
    elts = [orc.elt_from_step(logn, st) if st else 0 for st in steps]
    keys = [_galois_key(rng, s, e, kmods, L, logn, P) if e else None for e in elts]
    # ciphertexts of small messages m_r: c0 = m - c1 s + e
    cts, msgs = [], []
    for _ in steps:
        m = rng.integers(-1000, 1001, n)
        c1 = np.array([rng.integers(0, q, n, dtype=np.uint64) for q in kmods[:l]])
        c0 = np.array([((_ntt(m, q, logn).astype(object) - c1[i].astype(object) * _ntt(s, q, logn).astype(object))
                        % q).astype(np.uint64) for i, q in enumerate(kmods[:l])])
        cts.append(np.stack([c0, c1]))
        msgs.append(m)
    got = orc.rotate_sum(cts, elts, keys, kmods, logn)
    expect = [0] * n
    for m, e in zip(msgs, elts):
        r = auto_coeff_int(list(m), e, n) if e else list(m)
        expect = [x + y for x, y in zip(expect, r)]
    ph = _phase(got, s, kmods[:l], logn)
    for i in range(l):
        err = max(abs(a - b) for a, b in zip(ph[i], expect))
        assert err < 8 * n, (i, err)
    if len(steps) == 1:
        one = orc.rotate_hoisted(cts[0], elts[0], keys[0], kmods, logn)
        assert np.array_equal(one, got)


def test_rotate_sum_identity_only():
    logn = 4
    n = 1 << logn
    kmods = orc.coeff_modulus_create(n, [40, 45, 60])
    rng = np.random.default_rng(2)
    a = np.array([[rng.integers(0, q, n, dtype=np.uint64) for q in kmods[:2]] for _ in range(2)])
    b = np.array([[rng.integers(0, q, n, dtype=np.uint64) for q in kmods[:2]] for _ in range(2)])
    got = orc.rotate_sum([a, b], [0, 0], [None, None], kmods, logn)
    q = np.array(kmods[:2], dtype=object)[None, :, None]
    assert np.array_equal(got, ((a.astype(object) + b.astype(object)) % q).astype(np.uint64))


@pytest.mark.parametrize("L,alpha,l,steps", [(4, 2, 4, [1, -2]), (5, 2, 3, [3]), (3, 3, 3, [1, 0, 2])])
def test_rotate_sum_hybrid_decrypts(L, alpha, l, steps):
    logn = 5
    n = 1 << logn
    kmods = [int(q) for q in orc.coeff_modulus_create(n, [40] * L + [41] * alpha)]
    P = 1
    for p in kmods[L:]:
        P *= p
    rng = np.random.default_rng(L * 31 + alpha)
    s = rng.integers(-1, 2, n)
    elts = [orc.elt_from_step(logn, st) if st else 0 for st in steps]
    keys = [_galois_key(rng, s, e, kmods, L, logn, P, alpha) if e else None for e in elts]
    cts, msgs = [], []
    for _ in steps:
        m = rng.integers(-1000, 1001, n)
        c1 = np.array([rng.integers(0, q, n, dtype=np.uint64) for q in kmods[:l]])
        c0 = np.array([((_ntt(m, q, logn).astype(object) - c1[i].astype(object) * _ntt(s, q, logn).astype(object))
                        % q).astype(np.uint64) for i, q in enumerate(kmods[:l])])
        cts.append(np.stack([c0, c1]))
        msgs.append(m)
    got = hybrid.rotate_sum_hybrid(cts, elts, keys, kmods, L, alpha, logn)
    expect = [0] * n
    for m, e in zip(msgs, elts):
        r = auto_coeff_int(list(m), e, n) if e else list(m)
        expect = [x + y for x, y in zip(expect, r)]
    ph = _phase(got, s, kmods[:l], logn)
    for i in range(l):
        err = max(abs(a - b) for a, b in zip(ph[i], expect))
        assert err < 2 ** 10, (i, err)
    # one term at alpha = 1 is the SEAL-form hoisted rotation
    if alpha == 1 and len(steps) == 1:
        assert np.array_equal(got, orc.rotate_hoisted(cts[0], elts[0], keys[0], kmods, logn))


def test_rotate_sum_hybrid_alpha1_equals_c_oracle():
    """with one special prime the hybrid restatement is the C restatement word for word"""
    logn = 5
    n = 1 << logn
    kmods = [int(q) for q in orc.coeff_modulus_create(n, [40, 40, 40, 45])]
    L = 3
    rng = np.random.default_rng(9)
    elts = [orc.elt_from_step(logn, 1), orc.elt_from_step(logn, 6)]
    keys = [np.array([[[rng.integers(0, q, n, dtype=np.uint64) for q in kmods] for _ in range(2)]
                      for _ in range(L)]) for _ in elts]
    cts = [np.array([[rng.integers(0, q, n, dtype=np.uint64) for q in kmods[:3]] for _ in range(2)]) for _ in elts]
    a = hybrid.rotate_sum_hybrid(cts, elts, keys, kmods, L, 1, logn)
    b = orc.rotate_sum(cts, elts, keys, kmods, logn)
    assert np.array_equal(a, b)
