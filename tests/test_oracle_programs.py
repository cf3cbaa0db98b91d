"""CPU pins for the oracle program restatements (oracle/programs.py)."""
import numpy as np

import oracle as orc
from oracle import programs


def test_negacyclic_shift_matches_monomial_product():
    q, n = 97, 16
    rng = np.random.default_rng(1)
    a = rng.integers(0, q, n, dtype=np.uint64)
    for k in [0, 1, 5, 15, 16, 17, 31]:
        mono = [0] * n
        mono[k % n] = 1 if (k // n) % 2 == 0 else q - 1
        expect = [0] * n
        for i in range(n):
            for j in range(n):
                t = i + j
                v = int(a[i]) * mono[j]
                if t < n:
                    expect[t] = (expect[t] + v) % q
                else:
                    expect[t - n] = (expect[t - n] - v) % q
        assert [int(x) for x in programs.negacyclic_shift(a, k, q)] == expect


def test_expand_round_is_linear_in_coefficients():
    """With identity-free algebra: expanding by one round maps m(x) to
    (m(x) + m(-x), x^-1 (m(x) - m(-x))) when the key switch is replaced by the
    exact automorphism on a noiseless 'ciphertext' (c0 = m, c1 = 0)."""
    logn, n = 5, 32
    mods = orc.coeff_modulus_create(n, [40, 50])
    q = mods[0]
    rng = np.random.default_rng(2)
    coef = rng.integers(0, q, n, dtype=np.uint64)
    ntt = orc.ntt_forward(coef, logn, q)
    elt = n + 1
    rot = orc.apply_galois_ntt(ntt, logn, elt)  # m(x^(n+1)) = m(-x)
    back = orc.ntt_inverse(rot, logn, q)
    expect = [(int(c) * (-1) ** i) % q for i, c in enumerate(coef)]
    assert [int(x) for x in back] == expect
