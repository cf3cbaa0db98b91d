"""MatrixMul (src/matrix_mul.cpp) on the GPU.

* Expansion residues are bit-exact against the oracle restatement of
  expand_ciphertext (SEAL's INTT -> negacyclic shift -> NTT for x^k, SEAL's
  automorphism + key switch) at the reference's MatMul parameters
  (N = 2^13, {60, 40, 60}) for the first rounds.
* The full 4096x768 (plaintext) x 768x64 (encrypted) product of the
  reference's MM_test (src/main.cpp:322-382) on numpy-seed-42 data
  (data/data_generation.py:52-75) decrypts to the numpy product.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
from nexus_amd import mm  # noqa: E402
from oracle import programs  # noqa: E402

CHAIN = [60, 40, 60]
LOGN = 13


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    ctx = nx.Context(LOGN, CHAIN)
    sk = nx.SecretKey(ctx, seed=42)
    elts = mm.reference_elts(ctx.n, LOGN)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, elts, seed=43)
    return ctx, sk, gk, elts


def test_expand_parity(env):
    ctx, sk, gk, elts = env
    _note("expand m=8 parity")
    rng = np.random.default_rng(8)
    vals = rng.uniform(-1, 1, ctx.n)
    ct = mm.compress(ctx, sk, vals, 1e10, seed=5)
    m = 8
    batch = mm.expand(ct, m, gk)
    assert len(batch) == m
    keys = {e: gk.key_array(e) for e in elts[:3]}
    ref = programs.expand_ciphertext(ct.to_array(), m, elts, keys, ctx.moduli, LOGN)
    for k in range(m):
        assert np.array_equal(batch[k].to_array(), ref[k]), k


def test_expand_decrypts_to_coefficients(env):
    ctx, sk, gk, elts = env
    _note("expand m=8192")
    rng = np.random.default_rng(9)
    vals = rng.uniform(-1, 1, ctx.n)
    ct = mm.compress(ctx, sk, vals, 1e10, seed=6)
    batch = mm.expand(ct, ctx.n, gk)
    assert len(batch) == ctx.n
    for k in [0, 1, 2, 77, 4095, 8191]:
        d = nx.decode(ctx, nx.decrypt(ctx, sk, batch[k]))
        # the k-th expanded ciphertext is the constant n * vals[k]
        assert np.max(np.abs(d / ctx.n - vals[k])) < 1e-6, k


def test_matrix_mul_reference_shape(env):
    ctx, sk, gk, elts = env
    _note("matmul 4096x768 x 768x64")
    np.random.seed(42)
    x = np.random.uniform(-1, 1, (128 * 32, 768))
    y = np.random.uniform(-1, 1, (768, 64))
    x, y = np.round(x, 4), np.round(y, 4)  # written with %.4f (data_generation.py:59-60)
    xT, yT = x.T, y.T
    scale = 2.0 ** 40
    a_pts = [nx.encode(ctx, xT[i], scale) for i in range(768)]
    flat = yT.reshape(-1)
    packs = [flat[i * ctx.n:(i + 1) * ctx.n] for i in range(768 * 64 // ctx.n)]
    batches = [mm.expand(mm.compress(ctx, sk, p, 1e10, seed=100 + i), ctx.n, gk) for i, p in enumerate(packs)]
    res = mm.accumulate(a_pts, batches, 64)
    ref = x @ y
    for col in [0, 1, 63]:
        r = res[col]
        r.scale = r.scale * 4096
        r = nx.rescale_to_next(r)
        d = nx.decode(ctx, nx.decrypt(ctx, sk, r))[:4096]
        err = np.mean(np.abs(d / 2.0 - ref[:, col]))
        print("column %d: mean |err| %.3g" % (col, err))
        assert err < 1e-3


def test_matrix_mul_config1():
    """The stated ct-pt configuration: N = 2^13, {60, 40, 40, 60}, an encrypted
    128x768 input ~U[-1, 1] times a 768x768 plaintext ~N(0, 0.02), seed 42.
    Through the MatrixMul pipeline as (X W) row i = sum_j X[i, j] W[j, :]: X is
    coefficient-packed (12 ciphertexts), expanded to 128 x 768 scalars and
    accumulated against the 768 weight rows, 64 rows of X at a time."""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("matmul config 1: 128x768 (encrypted) x 768x768 (plaintext), {60, 40, 40, 60}")
    ctx = nx.Context(LOGN, [60, 40, 40, 60])
    sk = nx.SecretKey(ctx, seed=42)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, mm.reference_elts(ctx.n, LOGN), seed=44)
    rng = np.random.default_rng(42)
    X = rng.uniform(-1, 1, (128, 768))
    W = rng.normal(0, 0.02, (768, 768))
    ref = X @ W
    w_pts = [nx.encode(ctx, W[j], 2.0 ** 40) for j in range(768)]
    rows = 64  # 64 x 768 = 6 packed ciphertexts of 8192 coefficients
    errs = []
    for r0 in range(0, 128, rows):
        flat = X[r0:r0 + rows].reshape(-1)
        packs = [flat[i * ctx.n:(i + 1) * ctx.n] for i in range(len(flat) // ctx.n)]
        batches = [mm.expand(mm.compress(ctx, sk, p, 1e10, seed=200 + r0 + i), ctx.n, gk) for i, p in enumerate(packs)]
        res = mm.accumulate(w_pts, batches, rows)
        del batches
        for i in range(rows):
            r = res[i]
            r.scale = r.scale * (ctx.n // 2)
            d = nx.decode(ctx, nx.decrypt(ctx, sk, nx.rescale_to_next(r)))[:768] / 2.0
            errs.append(np.abs(d - ref[r0 + i]))
    errs = np.array(errs)
    print("config 1: max |err| %.3g, mean %.3g" % (errs.max(), errs.mean()))
    assert errs.max() < 1e-4
