"""Batched rescale and multiply -> relinearize -> rescale on the GPU: every
result equals the per-ciphertext operations word for word (one-prime and
hybrid key switching, batches crossing the internal chunk size, size-3
ciphertexts), and invalid batches are rejected on the host."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=[(12, [58, 40, 40, 40, 58], 1), (13, [51, 46, 46, 46, 51, 58, 58], 2),
                                        (16, [51, 46, 46, 58, 58], 2)], ids=["seal", "hybrid", "n16"])
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
