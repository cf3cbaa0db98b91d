"""NEXUS evaluator programs on the GPU at the reference's nonlinear chain
(N = 2^16, {58, 40 x 18, 58}, scale 2^40; src/main.cpp:27-30), on the
reference's own input vectors.

Checks: decrypted output vs the FP64 plaintext model of the same approximation
(oracle/approx.py) within 2^-10 (observed ~1e-6), and mean absolute error vs
the reference's calibration files no worse than the FP64 model's own error
plus 1e-4.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
from oracle import approx  # noqa: E402
from tests import fixtures  # noqa: E402

CHAIN = [58] + [40] * 18 + [58]
SCALE = 2.0 ** 40
TOL = 2.0 ** -10


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("keygen N=2^16 L=19")
    ctx = nx.Context(16, CHAIN)
    sk = nx.SecretKey(ctx, seed=2024)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=1)
    steps = [-1024, -128] + [1 << i for i in range(10)]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=2)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    return ctx, sk, ev


def enc(ctx, sk, v):
    return nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=7)


def dec(ctx, sk, ct, count):
    return nx.decode(ctx, nx.decrypt(ctx, sk, ct))[:count]


def test_gelu(env):
    ctx, sk, ev = env
    _note("gelu 32768")
    x, calib = fixtures.gelu_vectors()
    out = dec(ctx, sk, ev.gelu(enc(ctx, sk, x)), len(x))
    model = approx.gelu_approx(x)
    assert np.max(np.abs(out - model)) < TOL
    mae = np.mean(np.abs(out - calib))
    assert mae <= np.mean(np.abs(model - calib)) + 1e-4
    print("gelu: max|dec-fp64 model| %.3g  MAE vs calibration %.3g" % (np.max(np.abs(out - model)), mae))


def test_gelu_many_matches_single(env):
    """three inputs in lockstep: each output equals its own gelu() word for word"""
    ctx, sk, ev = env
    _note("gelu_many x3")
    rng = np.random.default_rng(21)
    xs = [enc(ctx, sk, rng.uniform(-6, 6, ctx.slot_count)) for _ in range(3)]
    many = ev.gelu_many(xs)
    for x, m in zip(xs, many):
        assert np.array_equal(m.to_array(), ev.gelu(x).to_array())


def test_layer_norm(env):
    ctx, sk, ev = env
    _note("layernorm 768 len 1024")
    x, calib = fixtures.layernorm_vectors()
    out = dec(ctx, sk, ev.layer_norm(enc(ctx, sk, x), 1024), len(x))
    model = approx.layer_norm_approx(x)
    assert np.max(np.abs(out - model)) < TOL
    mae = np.mean(np.abs(out - calib))
    assert mae <= np.mean(np.abs(model - calib)) + 1e-4
    print("layernorm: max|dec-fp64 model| %.3g  MAE %.3g" % (np.max(np.abs(out - model)), mae))


def test_softmax(env):
    ctx, sk, ev = env
    _note("softmax 128")
    x, calib = fixtures.softmax_vectors()
    out = dec(ctx, sk, ev.softmax(enc(ctx, sk, x), 128), len(x))
    model = approx.softmax_approx(x)
    assert np.max(np.abs(out - model)) < TOL
    mae = np.mean(np.abs(out - calib))
    assert mae <= np.mean(np.abs(model - calib)) + 1e-4
    print("softmax: max|dec-fp64 model| %.3g  MAE %.3g" % (np.max(np.abs(out - model)), mae))


def test_helpers(env):
    ctx, sk, ev = env
    _note("sgn / inverse / exp / invert_sqrt")
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, ctx.slot_count)
    out = dec(ctx, sk, ev.sgn_eval(enc(ctx, sk, x), 2, 2), ctx.slot_count)
    assert np.max(np.abs(out - approx.sgn_eval(x, 2, 2))) < TOL
    v = rng.uniform(0.1, 1.8, ctx.slot_count)
    out = dec(ctx, sk, ev.inverse(enc(ctx, sk, v)), ctx.slot_count)
    assert np.max(np.abs(out - approx.inverse(v))) < TOL
    e = rng.uniform(-4, 0, ctx.slot_count)
    out = dec(ctx, sk, ev.exp(enc(ctx, sk, e)), ctx.slot_count)
    assert np.max(np.abs(out - approx.exp_approx(e))) < TOL
    s = rng.uniform(1, 100, ctx.slot_count)
    out = dec(ctx, sk, ev.invert_sqrt(enc(ctx, sk, s), 5, 2), ctx.slot_count)
    assert np.max(np.abs(out - approx.invert_sqrt(s, 5, 2))) < TOL
