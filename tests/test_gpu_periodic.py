"""Periodic plaintexts and the block mix on the GPU.

* The exact periodic encoding decodes to the periodic slot vector, and its
  coefficients (inverse NTT by the oracle) differ from the oracle encoder's
  rounded coefficients of the tiled vector by at most a couple of units (the
  two round different floating-point evaluations of the same polynomial).
* Every block mix output equals nx_multiply_plain_accumulate over the expanded
  plaintexts word for word, for the fast (primes < 2^58) and the generic
  (60-bit primes) kernels, for ragged index patterns, outputs without terms,
  several periods and the largest LDS tile.
* Invalid shapes are rejected on the host before any launch.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
import oracle as orc  # noqa: E402


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=[[50, 40, 40, 40, 50], [60, 40, 40, 60]], ids=["fast", "generic"])
def env(request):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    chain = request.param
    ctx = nx.Context(12, chain)
    sk = nx.SecretKey(ctx, seed=3)
    return ctx, sk, chain


def _encrypt(ctx, sk, rng, count, scale, chain_index, seed0):
    return [nx.encrypt(ctx, sk, nx.encode(ctx, rng.uniform(-1, 1, ctx.slot_count), scale, chain_index),
                       seed=seed0 + i) for i in range(count)]


@pytest.mark.parametrize("period", [1, 4, 64, 2048])
def test_periodic_encoding_decodes_and_matches_oracle(env, period):
    ctx, sk, chain = env
    _note("periodic encode period=%d chain=%s" % (period, chain))
    rng = np.random.default_rng(period)
    vals = rng.uniform(-1, 1, (3, period))
    scale = 2.0 ** 40
    W = nx.PeriodicPlaintexts(ctx, vals, scale, chain_index=2)
    limbs = ctx.limbs_of(2)
    enc = orc.Encoder(ctx.log_n)
    for e in range(3):
        pt = W.expand(e)
        tiled = np.tile(vals[e], ctx.slot_count // period)
        got = nx.decode(ctx, pt)
        assert np.max(np.abs(got - tiled)) < 1e-9
        # coefficients against the oracle encoder's rounding of the same vector
        ref = enc.encode(tiled, scale, ctx.moduli[:limbs])
        arr = np.zeros((limbs, ctx.n), dtype=np.uint64)
        nx.lib().nx_plaintext_download(pt.h, arr.ctypes.data_as(nx.u64p))
        q = ctx.moduli[0]
        a = orc.ntt_inverse(arr[0], ctx.log_n, q).astype(object)
        b = orc.ntt_inverse(ref[0], ctx.log_n, q).astype(object)
        d = [(int(x) - int(y)) % q for x, y in zip(a, b)]
        d = [x - q if x > q // 2 else x for x in d]
        assert max(abs(x) for x in d) <= 2
        # non-zero coefficients only at multiples of n / (2 period)
        stride = ctx.n // (2 * period)
        nz = [j for j in range(ctx.n) if int(a[j]) != 0]
        assert all(j % stride == 0 for j in nz)


def _check_mix(ctx, cts, W, index, pts):
    outs = nx.multiply_periodic_mix(cts, W, index)
    for m, o in enumerate(outs):
        terms = [(k, int(e)) for k, e in enumerate(index[m]) if e >= 0]
        if not terms:
            assert not o.to_array().any()
            continue
        ref = nx.multiply_plain_accumulate([cts[k] for k, _ in terms], [pts[e] for _, e in terms])
        assert np.array_equal(o.to_array(), ref.to_array()), "output %d" % m
    return outs


@pytest.mark.parametrize("K,M,period", [(7, 5, 64), (33, 40, 64), (120, 64, 64), (9, 128, 4), (5, 3, 2048)])
def test_block_mix_equals_plain_accumulate(env, K, M, period):
    ctx, sk, chain = env
    _note("block mix K=%d M=%d period=%d chain=%s" % (K, M, period, chain))
    rng = np.random.default_rng(K * 1000 + M)
    count = 16
    W = nx.PeriodicPlaintexts(ctx, rng.uniform(-1, 1, (count, period)), 2.0 ** 30, chain_index=2)
    pts = [W.expand(e) for e in range(count)]
    cts = _encrypt(ctx, sk, rng, K, 2.0 ** 30, 2, 100)
    index = rng.integers(-1, count, (M, K)).astype(np.int32)
    index[rng.uniform(size=(M, K)) < 0.3] = -1  # ragged
    index[M // 2] = -1  # an output without terms
    if K >= 20:
        index[0] = rng.integers(0, count, K)  # a full row (more than one fold of terms)
    outs = _check_mix(ctx, cts, W, index, pts)
    # decrypted: sum_k w[e] x_k slot-wise
    xs = [nx.decode(ctx, nx.decrypt(ctx, sk, c)) for c in cts]
    wv = [nx.decode(ctx, p) for p in pts]
    m = 0
    want = sum(wv[e] * xs[k] for k, e in enumerate(index[m]) if e >= 0)
    got = nx.decode(ctx, nx.decrypt(ctx, sk, outs[m]))
    assert np.max(np.abs(got - want)) < 1e-4


def test_block_mix_hankel_pattern(env):
    """the attention compaction pattern: out[d] = sum_t A_(t + d) (.) P_t over
    offsets d in (-64, 64), with one-hot masks A_p of period 64"""
    ctx, sk, chain = env
    _note("block mix hankel chain=%s" % chain)
    rng = np.random.default_rng(7)
    K = 16
    W = nx.PeriodicPlaintexts(ctx, np.eye(64), 2.0 ** 30, chain_index=2)
    pts = [W.expand(e) for e in range(64)]
    cts = _encrypt(ctx, sk, rng, K, 2.0 ** 30, 2, 200)
    ds = list(range(-(K - 1), 64))
    index = np.full((len(ds), K), -1, dtype=np.int32)
    for m, d in enumerate(ds):
        for t in range(K):
            if 0 <= t + d < 64:
                index[m, t] = t + d
    _check_mix(ctx, cts, W, index, pts)


def test_block_mix_rejects_bad_shapes(env):
    ctx, sk, chain = env
    rng = np.random.default_rng(9)
    W = nx.PeriodicPlaintexts(ctx, np.eye(64), 2.0 ** 30, chain_index=2)
    cts = _encrypt(ctx, sk, rng, 2, 2.0 ** 30, 2, 300)
    with pytest.raises(nx.NexusError):
        nx.multiply_periodic_mix(cts, W, np.full((1, 2), 64, dtype=np.int32))  # entry out of range
    with pytest.raises(nx.NexusError):
        nx.multiply_periodic_mix(cts, W, np.zeros((129, 2), dtype=np.int32))  # too many outputs
    low = _encrypt(ctx, sk, rng, 2, 2.0 ** 30, 3, 400)
    with pytest.raises(nx.NexusError):
        nx.multiply_periodic_mix(low, W, np.zeros((1, 2), dtype=np.int32))  # level mismatch
    with pytest.raises(nx.NexusError):
        nx.PeriodicPlaintexts(ctx, np.ones((1, 3)), 2.0 ** 30)  # period not a power of two
    # 127 inputs at N = 2^12 with period 64 need more than 64 KiB of LDS
    many = [cts[0]] * 127
    with pytest.raises(nx.NexusError):
        nx.multiply_periodic_mix(many, W, np.zeros((1, 127), dtype=np.int32))
