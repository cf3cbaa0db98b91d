"""Pins the hybrid key-switching oracle (oracle/hybrid.py) by the defining
identity of key switching: with key_j = (-a_j s + e_j + [i in digit j] P s', a_j),
the switched pair satisfies ks0 + ks1 s = c2 s' + (small) mod every q_i.  A
dropped term, a wrong conversion constant or index leaves an error of the
size of the moduli instead."""
import numpy as np
import pytest

import oracle as orc
from oracle import hybrid


def _ntt(coeffs, q, logn):
    return orc.ntt_forward(np.array([int(c) % q for c in coeffs], dtype=np.uint64), logn, q)


def _centred(a, q):
    return np.array([int(x) - q if int(x) > q // 2 else int(x) for x in a], dtype=object)


@pytest.mark.parametrize("L,alpha,l", [(5, 2, 5), (5, 2, 4), (6, 3, 5), (4, 1, 4), (3, 3, 2)])
def test_hybrid_switch_identity(L, alpha, l):
    logn = 6
    n = 1 << logn
    moduli = [int(q) for q in orc.coeff_modulus_create(n, [40] * L + [41] * alpha)]
    rng = np.random.default_rng(L * 100 + alpha * 10 + l)
    s = rng.integers(-1, 2, n)
    s2 = rng.integers(-1, 2, n)
    K = L + alpha
    P = 1
    for p in moduli[L:]:
        P *= p
    digits = (L + alpha - 1) // alpha
    key = np.zeros((digits, 2, K, n), dtype=np.uint64)
    for j in range(digits):
        e = rng.integers(-3, 4, n)
        for i, q in enumerate(moduli):
            a = rng.integers(0, q, n, dtype=np.uint64)
            sn = _ntt(s, q, logn).astype(object)
            b = (-(a.astype(object) * sn) + _ntt(e, q, logn).astype(object)) % q
            if i < L and j * alpha <= i < (j + 1) * alpha:
                b = (b + (P % q) * _ntt(s2, q, logn).astype(object)) % q
            key[j, 0, i] = np.array(b, dtype=np.uint64)
            key[j, 1, i] = a
    c2 = np.array([rng.integers(0, moduli[i], n, dtype=np.uint64) for i in range(l)])
    ks0, ks1 = hybrid.switch_key_hybrid(c2, key, moduli, L, alpha, logn)
    for i in range(l):
        q = moduli[i]
        lhs = (ks0[i].astype(object) + ks1[i].astype(object) * _ntt(s, q, logn).astype(object)) % q
        rhs = c2[i].astype(object) * _ntt(s2, q, logn).astype(object) % q
        diff = np.array([int(v) for v in (lhs - rhs) % q], dtype=np.uint64)
        err = _centred(orc.ntt_inverse(diff, logn, q), q)
        # noise: rounding times s plus sum_j d_j e_j / P; the moduli are 2^40
        assert max(abs(int(v)) for v in err) < 2 ** 8, (i, max(abs(int(v)) for v in err))


def test_centred_conversion_against_exact_integers():
    """the ModDown conversion's float64 quotient r = round(sum_k v_k / p_k)
    gives the representative of y in (-P/2, P/2] -- checked against exact CRT
    composition with Python integers on random residues; within about 2^-40 P
    of P/2 it may give the other representative (off by P)"""
    import math
    ps = [int(p) for p in orc.coeff_modulus_create(1 << 6, [51] * 8)]
    P = math.prod(ps)
    t = int(orc.coeff_modulus_create(1 << 6, [46])[0])
    rng = np.random.default_rng(8)
    ys = [int(rng.integers(0, 1 << 62)) * int(rng.integers(0, 1 << 62)) % P for _ in range(3000)]
    ys += [0, 1, P - 1, P // 3, 2 * P // 3]
    vals = [np.array([y % p for y in ys], dtype=np.uint64) for p in ps]
    got = hybrid._conv_centred(vals, ps, t)
    want = [(y - P if y > P // 2 else y) % t for y in ys]
    assert [int(g) for g in got] == want
    # within about 2^-40 P of P/2 the float64 quotient may round to the other
    # representative, y - P instead of y: the stated window, and nothing else
    edge = [P // 2 + d for d in (-(P >> 45), -2, -1, 0, 1, 2, P >> 45)]
    vals = [np.array([y % p for y in edge], dtype=np.uint64) for p in ps]
    got = hybrid._conv_centred(vals, ps, t)
    for y, g in zip(edge, got):
        assert int(g) in ((y - P) % t, y % t)
        if abs(y - P // 2) >= P >> 45:
            assert int(g) == ((y - P if y > P // 2 else y) % t)
