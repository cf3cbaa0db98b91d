"""Pins each function of oracle/approx.py (the FP64 restatement of NEXUS's
polynomial approximations) by something other than its own formula:

* the reference's own data: the input/calibration vectors of data/ (softmax,
  LayerNorm, argmax; GELU regenerated from data/data_generation.py), within the
  approximation error the reference's method leaves;
* closed forms: the Newton product of `inverse` telescopes to
  (1 - (1 - x)^(2^(k+1))) / x; exp(x) ~ (1 + x/128)^128 equals
  exp(128 log1p(x / 128)) and differs from exp(x) by exp(x) x^2 / 256 to
  leading order; sgn_eval is odd and within 1e-10 of sign(x) on [0.03, 1];
  invert_sqrt is within 1e-4 of 1/sqrt on the range its initial guess is
  tuned for;
* a planted mistake per function (wrong rotation direction or sign
  composition in Argmax, a dropped Newton step in the softmax inverse, the
  wrong length in LayerNorm, a flipped GELU coefficient, a wrong exponent in
  exp) fails the same check.
"""
import numpy as np

import fixtures
from oracle import approx as A


def test_softmax_matches_reference_calibration():
    x, y = fixtures.softmax_vectors()  # data/input/softmax_input_128_128.txt, calibration
    err = np.max(np.abs(A.softmax_approx(x) - y))
    assert err < 5e-4, err
    # planted: one Newton step fewer in the inverse of the sum
    e = A.exp_approx(x)
    bad = e * (A.inverse(np.sum(e) * 0.01, 3) * 0.01)
    assert np.max(np.abs(bad - y)) > 2e-3


def test_layer_norm_matches_reference_calibration():
    x, y = fixtures.layernorm_vectors()  # 768 values on [-10, 10] and sqrt(768) x / |x|
    err = np.max(np.abs(A.layer_norm_approx(x) - y))
    assert err < 2e-6, err  # the calibration file's 6-decimal rounding
    # planted: the mean over the padded length 1024 instead of 768
    bad = x * A.invert_sqrt(np.sum(x * x) / 1024.0, 4, 2)
    assert np.max(np.abs(bad - y)) > 1e-2


def test_argmax_matches_reference_calibration():
    x, y = fixtures.argmax_vectors()  # 8 inputs and the one-hot calibration
    v = np.zeros(64)
    v[:8] = x
    got = A.argmax_approx(v, 8)[:8]
    assert int(np.argmax(got)) == int(np.argmax(y))
    assert np.max(np.abs(got - y)) < 2e-2

    # planted: the rotation in the wrong direction, or single sign compositions
    def planted(x, length, shift, dg):
        x = x + np.roll(x, length)
        a = x.copy()
        for i in range(int(np.log2(length))):
            b = np.roll(a, shift * (1 << i))
            a = (a + b) * 0.5 + (a - b) * A.sgn_eval(a - b, dg, dg)
        return A.sgn_eval(x - a, dg, dg, 1.0) + 1.0

    assert np.max(np.abs(planted(v, 8, -1, 2)[:8] - got)) < 1e-12  # the same program, restated
    assert np.max(np.abs(planted(v, 8, 1, 2)[:8] - y)) > 0.5
    assert np.max(np.abs(planted(v, 8, -1, 1)[:8] - y)) > 0.5


def test_gelu_matches_reference_calibration():
    x, y = fixtures.gelu_vectors()  # data_generation.py's GELU vectors on [-5, 5]
    err = np.max(np.abs(A.gelu_approx(x) - y))
    assert err < 2e-2, err
    inner = np.abs(x) <= 4.5  # the degree-12 even polynomial A(x) drifts near the ends of [-5, 5]
    assert np.max(np.abs(A.gelu_approx(x[inner]) - y[inner])) < 8e-3
    old = A.GELU_A
    try:
        A.GELU_A = old[:7] + [-old[7]]  # planted: a flipped coefficient
        assert np.max(np.abs(A.gelu_approx(x) - y)) > 1.0
    finally:
        A.GELU_A = old


def test_inverse_closed_form():
    x = np.linspace(0.01, 1.99, 2001)
    for k in (1, 3, 4, 6):
        want = (1.0 - (1.0 - x) ** (2 ** (k + 1))) / x
        assert np.max(np.abs(A.inverse(x, k) - want)) < 1e-12
    # converges to 1/x on (0, 2)
    assert np.max(np.abs(A.inverse(x[(x > 0.3) & (x < 1.7)], 6) * x[(x > 0.3) & (x < 1.7)] - 1.0)) < 1e-12


def test_exp_closed_form():
    x = np.linspace(-10.0, 0.0, 1001)
    got = A.exp_approx(x)
    assert np.max(np.abs(got - np.exp(128.0 * np.log1p(x / 128.0))) / np.exp(x)) < 1e-12
    # leading term of exp(x) - (1 + x/n)^n is exp(x) x^2 / (2n): n = 128
    rel = (np.exp(x) - got) / np.exp(x)
    sel = np.abs(x) > 1.0
    ratio = rel[sel] / (x[sel] ** 2 / 256.0)
    assert np.all((ratio > 0.85) & (ratio < 1.01)), (ratio.min(), ratio.max())
    bad = (1.0 + x / 64.0) ** 64  # planted: the wrong exponent
    assert np.max(np.abs((np.exp(x[sel]) - bad[sel]) / np.exp(x[sel]) / (x[sel] ** 2 / 256.0))) > 1.5


def test_sgn_eval_closed_form():
    t = np.linspace(0.03, 1.0, 4001)
    # the reference scales the last g4 step as well as the last f4 step by the
    # factor (src/ckks_evaluator.cpp:181-212), so a factor 0.5 leaves 1.7e-3
    for f, tol in ((1.0, 1e-10), (0.5, 2e-3)):
        s = A.sgn_eval(t, 2, 2, f)
        assert np.max(np.abs(s - f)) < tol
        assert np.max(np.abs(A.sgn_eval(-t, 2, 2, f) + s)) < 1e-14  # odd
    # f4 and g4 fix +-1 up to their published rounding: f4(1) = 1, g4(1) = 1022/1024
    assert abs(A.odd9(A.F4, 1.0) - 1.0) < 1e-15
    assert abs(A.odd9(A.G4, 1.0) - 1022.0 / 1024.0) < 1e-15


def test_invert_sqrt_on_its_tuned_range():
    v = np.linspace(5.0, 700.0, 5001)  # the range of the initial guess -1.29e-4 v + 0.129
    assert np.max(np.abs(A.invert_sqrt(v, 4, 2) * np.sqrt(v) - 1.0)) < 1e-4
    assert np.max(np.abs(A.invert_sqrt(v, 5, 2) * np.sqrt(v) - 1.0)) < 1e-8
