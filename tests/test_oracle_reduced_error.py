"""Pins oracle/programs.reduced_error (SEAL-4.1-bs evaluator.cpp:420-594) by what
the operations mean: ciphertexts of known messages at different levels and
scales, encrypted on the CPU under a ternary secret, combined by the
restatement, decrypt (at the returned scale) to m1 + m2, m1 - m2 or m1 m2.  A
wrong adjust constant, a scale override on the wrong operand, a missing
rescale or the wrong prime fails by far more than the bound; the returned
scale is checked against SEAL's formulas, and the equal-level case against
plain addition."""
import numpy as np
import pytest

import oracle as orc
from oracle import programs

LOGN = 6
N = 1 << LOGN
BITS = [50, 40, 40, 40, 58]


def _ntt(coeffs, q):
    return orc.ntt_forward(np.array([int(c) % q for c in coeffs], dtype=np.uint64), LOGN, q)


@pytest.fixture(scope="module")
def env():
    mods = orc.coeff_modulus_create(N, BITS)
    rng = np.random.default_rng(21)
    s = rng.integers(-1, 2, N)
    s_ntt = [_ntt(s, q) for q in mods]
    L = len(mods) - 1
    P = mods[-1]
    key = np.zeros((L, 2, L + 1, N), dtype=np.uint64)
    for j in range(L):
        e = rng.integers(-3, 4, N)
        for i, q in enumerate(mods):
            a = rng.integers(0, q, N, dtype=np.uint64)
            b = (-(a.astype(object) * s_ntt[i].astype(object)) + _ntt(e, q).astype(object)) % q
            if i == j:
                b = (b + (P % q) * (s_ntt[i].astype(object) ** 2 % q)) % q
            key[j, 0, i] = np.array(b, dtype=np.uint64)
            key[j, 1, i] = a
    return mods, s_ntt, key, rng


def _encrypt(env, m, scale, limbs):
    mods, s_ntt, _, rng = env
    pt = orc.Encoder(LOGN).encode(m, scale, mods[:limbs])
    c1 = np.array([rng.integers(0, q, N, dtype=np.uint64) for q in mods[:limbs]])
    e = rng.integers(-3, 4, N)
    c0 = np.array([((pt[i].astype(object) - c1[i].astype(object) * s_ntt[i].astype(object)
                     + _ntt(e, q).astype(object)) % q).astype(np.uint64) for i, q in enumerate(mods[:limbs])])
    return np.stack([c0, c1])


def _decrypt(env, ct, scale):
    mods, s_ntt, _, _ = env
    l = ct.shape[1]
    plain = np.array([((ct[0][i].astype(object) + ct[1][i].astype(object) * s_ntt[i].astype(object)) % q)
                      .astype(np.uint64) for i, q in enumerate(mods[:l])])
    return orc.Encoder(LOGN).decode(plain, mods, scale)


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
@pytest.mark.parametrize("l1,l2", [(4, 4), (2, 4), (4, 3), (3, 1 + 1)])
def test_reduced_error_decrypts(env, op, l1, l2):
    mods, _, key, rng = env
    slots = N // 2
    m1 = rng.uniform(-1, 1, slots)
    m2 = rng.uniform(-1, 1, slots)
    s1, s2 = 2.0 ** 40 * 1.03, 2.0 ** 40 * 0.97  # different scales on purpose
    a = _encrypt(env, m1, s1, l1)
    b = _encrypt(env, m2, s2, l2)
    res, scale = programs.reduced_error(op, a, s1, b, s2, mods, LOGN, key, mods)
    # level: the lower of the two; scale: SEAL's formulas
    assert res.shape[1] == min(l1, l2)
    if l1 < l2:
        expect_scale = s1 * float(mods[l2 - 1]) / float(mods[l2 - 1])
    else:
        expect_scale = s2
    if op == "mul":
        expect_scale = expect_scale * expect_scale
    assert scale == expect_scale
    # at equal levels SEAL only overwrites encrypted1's scale: its message reads as m1 s1 / s2
    x1 = m1 * (s1 / s2) if l1 == l2 else m1
    want = {"add": x1 + m2, "sub": x1 - m2, "mul": x1 * m2}[op]
    got = _decrypt(env, res, scale)
    assert np.max(np.abs(got - want)) < 1e-5, np.max(np.abs(got - want))


def test_equal_levels_is_plain_add(env):
    mods, _, _, rng = env
    a = _encrypt(env, rng.uniform(-1, 1, N // 2), 2.0 ** 40, 3)
    b = _encrypt(env, rng.uniform(-1, 1, N // 2), 2.0 ** 40 * 1.5, 3)
    res, scale = programs.reduced_error("add", a, 2.0 ** 40, b, 2.0 ** 40 * 1.5, mods, LOGN)
    assert scale == 2.0 ** 40 * 1.5
    assert np.array_equal(res, programs.add(a, b, mods))
