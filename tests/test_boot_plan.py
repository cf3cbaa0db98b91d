"""Bootstrapping linear-transform plan (host only, no GPU): the three
CoeffToSlot and three SlotToCoeff BSGS levels built by the library reproduce
the closed-form slot maps, and the slot convention they assume is the oracle
encoder's."""
import numpy as np
import pytest

import nexus_amd as nx
import oracle as orc
from oracle import boot as ob


@pytest.mark.parametrize("logN,logn", [(8, 4), (8, 7), (9, 6)])
def test_slot_matrix_is_the_encoder_slot_map(logN, logn):
    N = 1 << logN
    n = 1 << logn
    gap = N // (2 * n)
    mods = orc.coeff_modulus_create(N, [50, 50])
    enc = orc.Encoder(logN)
    rng = np.random.default_rng(logN * 100 + logn)
    a = rng.integers(-1000, 1000, 2 * n)
    coef = np.zeros(N, dtype=np.int64)
    coef[::gap] = a
    pt = np.array([orc.ntt_forward((coef % q).astype(np.uint64), logN, q) for q in mods])
    slots = enc.decode(pt, mods, 1.0, complex_out=True)
    model = ob.decode_sparse(coef, logN, logn)
    assert np.max(np.abs(slots[:n] - model)) < 1e-8
    assert np.max(np.abs(ob.slot_matrix(logn) @ (a[:n] + 1j * a[n:]) - model)) < 1e-8
    # the slot vector of a sparse polynomial is n-periodic
    assert np.max(np.abs(slots - np.tile(slots[:n], N // 2 // n))) < 1e-8


def _levels(logNh, logn, which, K=25):
    first = nx.boot_plan_level(logNh, logn, which, 0, K)
    return [first] + [nx.boot_plan_level(logNh, logn, which, k, K) for k in range(1, first["levels"])]


@pytest.mark.parametrize("logNh,logn", [(7, 5), (7, 7), (9, 6), (6, 3), (5, 1)])
def test_coeff_to_slot_plan(logNh, logn):
    K = 25
    Nh, n = 1 << logNh, 1 << logn
    c = Nh // n
    rng = np.random.default_rng(logNh * 10 + logn)
    alpha = rng.uniform(-K, K, 2 * n)
    z = alpha[:n] + 1j * alpha[n:]
    w = np.tile(c * (ob.slot_matrix(logn) @ z), Nh // n)  # slots after SubSum
    x = w
    for lv in _levels(logNh, logn, 0, K):
        x = ob.apply_level(lv, x)
    br = ob.brv_perm(logn)
    if logn < logNh:
        u = x + np.conj(x)
        assert np.max(np.abs(u.imag)) < 1e-9
        s = np.arange(Nh)
        expect = alpha[br[s % n] + n * ((s % (2 * n)) >= n)] / K
        assert np.max(np.abs(u.real - expect)) < 1e-9
    else:
        expect = z[br] / (2 * K)
        assert np.max(np.abs(x - expect)) < 1e-9


@pytest.mark.parametrize("logNh,logn", [(7, 5), (7, 7), (9, 6), (6, 3), (5, 1)])
def test_slot_to_coeff_plan(logNh, logn):
    Nh, n = 1 << logNh, 1 << logn
    rng = np.random.default_rng(logNh * 7 + logn)
    a = rng.uniform(-0.5, 0.5, n)
    b = rng.uniform(-0.5, 0.5, n)
    br = ob.brv_perm(logn)
    if logn < logNh:
        f = np.concatenate([a[br], b[br]])
        x = np.tile(f, Nh // (2 * n)).astype(complex)
    else:
        x = a[br] + 1j * b[br]
    for lv in _levels(logNh, logn, 1):
        x = ob.apply_level(lv, x)
    if logn < logNh:
        x = x + np.roll(x, -n)
    expect = np.tile(ob.slot_matrix(logn) @ (a + 1j * b), Nh // n) / (2 * np.pi)
    assert np.max(np.abs(x - expect)) < 1e-9


@pytest.mark.parametrize("logNh,logn", [(14, 13), (15, 15), (15, 12), (7, 5)])
def test_plan_rotation_steps(logNh, logn):
    steps = set(nx.boot_plan_steps(logNh, logn))
    Nh = 1 << logNh
    assert 0 in steps  # conjugation
    used = set()
    for which in (0, 1):
        for lv in _levels(logNh, logn, which):
            for (j, i) in lv["diags"]:
                if i:
                    used.add(i * lv["stride"])
                if j:
                    used.add(j * lv["g"] * lv["stride"])
    if logn < logNh:
        used |= {1 << i for i in range(logn, logNh)}
    assert used | {0} == steps
    assert all(abs(s) < Nh for s in steps)
    # three levels each way; the reference's key lists are 42 / 47 / 38 elements
    assert len(steps) <= 48


def test_eval_mod_polynomial():
    """the library's EvalMod polynomial: same fit as the oracle's, error
    against sin(2 pi K u) on the windows, and depth 6 at degree 59 (the
    reference's EvalMod budget: 9 levels with two double angles)"""
    K = 25
    c, depth = nx.boot_eval_mod_coeffs(K, 59, 2, 10)
    assert depth == 6
    assert nx.boot_eval_mod_coeffs(K, 63, 2, 10)[1] == 7
    xs = np.concatenate([(i + np.linspace(-2.0 ** -10, 2.0 ** -10, 33)) / K for i in range(-K + 1, K)])
    mine = np.polynomial.chebyshev.chebval(xs, c)
    ref = np.polynomial.chebyshev.chebval(xs, ob.eval_mod_fit(K, 59, 2, 10))
    assert np.max(np.abs(mine - ref)) < 1e-9
    y = mine
    for _ in range(2):
        y = 2 * y * y - 1
    assert np.max(np.abs(y - np.sin(2 * np.pi * K * xs))) < 1e-8
    # the fit degrades gracefully outside the windows (inputs with |m / q0| up to 2^-7)
    xw = np.concatenate([(i + np.linspace(-2.0 ** -7, 2.0 ** -7, 65)) / K for i in range(-K + 1, K)])
    assert np.max(np.abs(ob.eval_mod_model(xw) - np.sin(2 * np.pi * K * xw))) < 1e-6


def test_mod_raise_oracle_is_centred_lift():
    N = 64
    mods = [int(q) for q in orc.coeff_modulus_create(N, [51, 46, 46, 51])]
    q0 = mods[0]
    rng = np.random.default_rng(5)
    x = np.array([int(v) for v in rng.integers(0, q0, N, dtype=np.uint64)] + [0, q0 >> 1, (q0 >> 1) + 1, q0 - 1],
                 dtype=object)
    out = ob.mod_raise(x, mods)
    Q = 1
    for q in mods:
        Q *= q
    for k, xv in enumerate(x):
        centred = int(xv) - q0 if int(xv) > (q0 >> 1) else int(xv)
        # CRT composition of the lifted residues is the centred value
        acc = 0
        for j, q in enumerate(mods):
            Mj = Q // q
            acc += int(out[j][k]) * Mj * pow(Mj, -1, q)
        acc %= Q
        if acc > Q // 2:
            acc -= Q
        assert acc == centred
