"""Batched rescale and multiply -> relinearize -> rescale on the GPU: every
result equals the per-ciphertext operations word for word (one-prime and
hybrid key switching, batches crossing the internal chunk size, size-3
ciphertexts); the fused form (relinearisation and rescale in one rounding)
is within one of it on every coefficient; invalid batches are rejected on
the host."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=[(12, [58, 40, 40, 40, 58], 1), (13, [51, 46, 46, 46, 51, 58, 58], 2),
                                        (16, [51, 46, 46, 58, 58], 2), (16, [51, 46, 46, 46, 51, 51], 2)],
                ids=["seal", "hybrid", "n16", "n16-fp"])
def env(request):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn, chain, special = request.param
    ctx = nx.Context(logn, chain, special_primes=special)
    sk = nx.SecretKey(ctx, seed=5)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=6)
    return ctx, sk, rk


def _cts(ctx, sk, count, seed, scale=2.0 ** 40):
    rng = np.random.default_rng(seed)
    return [nx.encrypt(ctx, sk, nx.encode(ctx, rng.uniform(-1, 1, ctx.slot_count), scale), seed=seed + i)
            for i in range(count)]


@pytest.mark.parametrize("count", [1, 5, 17])
def test_multiply_relin_rescale_many_matches(env, count):
    ctx, sk, rk = env
    _note("mul-relin-rescale many logn=%d count=%d" % (ctx.log_n, count))
    a = _cts(ctx, sk, count, 10)
    b = _cts(ctx, sk, count, 50)
    b[0] = a[0]  # a square in the batch
    got = nx.multiply_relin_rescale_many(a, b, rk)
    for i in range(count):
        want = nx.rescale_to_next(nx.relinearize(nx.multiply(a[i], b[i]), rk))
        assert np.array_equal(got[i].to_array(), want.to_array()), i
        assert got[i].info()["scale"] == want.info()["scale"]


def _centred_diff(f, u, q):
    """(f - u) mod q as -1, 0, 1 (2 marks any other difference)"""
    d = np.where(f >= u, f - u, f + np.uint64(q) - u)
    out = np.full(d.shape, 2, dtype=np.int8)
    out[d == 0] = 0
    out[d == 1] = 1
    out[d == np.uint64(q - 1)] = -1
    return out


@pytest.mark.parametrize("count", [1, 17])
def test_fused_relin_rescale_within_one(env, count):
    """relinearisation and rescale in one rounding: every coefficient of the
    fused result is the unfused one plus d, d in {-1, 0, 1} and the same on
    every limb.  Rounding X / P and then by q differs from rounding X / (P q)
    only where X / (P q) lies within 1 / (2 q) of a half-integer, so d != 0 on
    at most a handful of coefficients (a one-prime key switch has nothing to
    fuse: word for word)"""
    ctx, sk, rk = env
    _note("fused mul-relin-rescale logn=%d count=%d" % (ctx.log_n, count))
    a = _cts(ctx, sk, count, 210)
    b = _cts(ctx, sk, count, 250)
    b[0] = a[0]
    fused = nx.multiply_relin_rescale_many(a, b, rk, fused=True)
    plain = nx.multiply_relin_rescale_many(a, b, rk)
    changed = 0
    for i in range(count):
        assert fused[i].info()["scale"] == plain[i].info()["scale"]
        assert fused[i].chain_index == plain[i].chain_index
        if ctx.special_primes == 1:
            assert np.array_equal(fused[i].to_array(), plain[i].to_array())
            continue
        f = nx.transform_from_ntt(fused[i]).to_array()
        u = nx.transform_from_ntt(plain[i]).to_array()
        d0 = None
        for j in range(f.shape[1]):
            d = _centred_diff(f[:, j], u[:, j], ctx.moduli[j])
            assert not np.any(d == 2), (i, j)
            if d0 is None:
                d0 = d
            assert np.array_equal(d, d0), (i, j)
        changed += int(np.count_nonzero(d0))
        x = nx.decode(ctx, nx.decrypt(ctx, sk, fused[i]))
        y = nx.decode(ctx, nx.decrypt(ctx, sk, plain[i]))
        assert np.max(np.abs(x - y)) < 1e-4
    assert changed <= 8


def test_linear_combination_matches(env):
    """sum_k x_k * encode_constant(v_k, s_k) at a lower level with inputs at
    three levels (6 terms: two launches), negative and tiny constants: word for
    word mod_switch + multiply_plain + add per term; every term's scale must
    equal the sum's, as SEAL's add requires"""
    ctx, sk, rk = env
    _note("linear combination logn=%d" % ctx.log_n)
    L = ctx.L
    target = min(L, 3)
    cts = _cts(ctx, sk, 6, 300)
    cts = [nx.mod_switch_to(c, 1 + (i % target)) for i, c in enumerate(cts)]
    # products stay below the target level's modulus (one 51-bit prime at n16)
    values = [0.75, -1.25, 3e-3, -0.5, 1.0, 2.0 ** -6]
    scales = [2.0 ** 8] * 6
    got = nx.linear_combination(cts, values, scales, target)
    want = None
    for c, v, sc in zip(cts, values, scales):
        t = nx.multiply_plain(nx.mod_switch_to(c, target), nx.encode_constant(ctx, v, sc, target))
        want = t if want is None else nx.add(want, t)
    assert got.chain_index == target
    assert np.array_equal(got.to_array(), want.to_array())
    assert got.info()["scale"] == want.info()["scale"]
    # a term whose scale differs from the sum's is rejected, as by add
    with pytest.raises(nx.NexusError, match="scale mismatch"):
        nx.linear_combination(cts[:2], values[:2], [2.0 ** 8, 2.0 ** 8 * 1.5], target)
    # a term at another scale brought to the sum's scale by its constant's scale
    c2 = cts[1].copy()
    c2.scale = cts[1].scale * 2.0
    got = nx.linear_combination([cts[0], c2], values[:2], [2.0 ** 8, 2.0 ** 7], target)
    t0 = nx.multiply_plain(nx.mod_switch_to(cts[0], target), nx.encode_constant(ctx, values[0], 2.0 ** 8, target))
    t1 = nx.multiply_plain(nx.mod_switch_to(c2, target), nx.encode_constant(ctx, values[1], 2.0 ** 7, target))
    assert np.array_equal(got.to_array(), nx.add(t0, t1).to_array())


@pytest.mark.parametrize("count", [1, 20])
def test_rescale_many_matches(env, count):
    ctx, sk, rk = env
    _note("rescale many logn=%d count=%d" % (ctx.log_n, count))
    a = _cts(ctx, sk, count, 90)
    prods = [nx.multiply(x, x) for x in a]  # size 3
    for batch in (a, prods):
        got = nx.rescale_to_next_many(batch)
        for i in range(count):
            assert np.array_equal(got[i].to_array(), nx.rescale_to_next(batch[i]).to_array()), i


def test_batched_rejects_mixed_levels(env):
    ctx, sk, rk = env
    a = _cts(ctx, sk, 2, 130)
    low = nx.rescale_to_next(a[1])
    with pytest.raises(nx.NexusError):
        nx.rescale_to_next_many([a[0], low])
    with pytest.raises(nx.NexusError):
        nx.multiply_relin_rescale_many([a[0], low], [a[0], low], rk)
    with pytest.raises(nx.NexusError):
        nx.multiply_relin_rescale_many([a[0], low], [a[0], low], rk, fused=True)


def test_multiply_sum_matches(env):
    """sum_i a_i * b_i in one pass equals multiply and add per pair word for
    word (more pairs than one launch takes)"""
    ctx, sk, rk = env
    _note("multiply_sum logn=%d" % ctx.log_n)
    a = _cts(ctx, sk, 37, 410)
    b = _cts(ctx, sk, 37, 460)
    got = nx.multiply_sum(a, b)
    want = nx.multiply(a[0], b[0])
    for x, y in zip(a[1:], b[1:]):
        want = nx.add(want, nx.multiply(x, y))
    assert got.info()["size"] == 3 and got.info()["scale"] == want.info()["scale"]
    assert np.array_equal(got.to_array(), want.to_array())
