"""Hybrid key switching (alpha special primes, digits of alpha data primes) on
the GPU.

* Residues: relinearisation and Galois key switching are bit-exact against the
  oracle's restatement of Phantom's ModUp / inner product / ModDown
  (oracle/hybrid.py), with the device-generated keys downloaded.
* Decryption: multiply+relinearize, rotations, hoisted rotations and a full
  bootstrap in hybrid mode decode to the expected values within noise.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
import oracle as orc  # noqa: E402
from oracle import hybrid  # noqa: E402


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


def _rand_ct(rng, ctx, size, limbs):
    return np.stack([np.stack([rng.integers(0, ctx.moduli[i], ctx.n, dtype=np.uint64) for i in range(limbs)])
                     for _ in range(size)])


@pytest.mark.parametrize("logn,L,alpha,chain_index", [(8, 6, 2, 1), (8, 6, 2, 2), (8, 7, 3, 1), (8, 7, 3, 3),
                                                      (10, 5, 5, 1)])
def test_hybrid_relin_bit_exact(logn, L, alpha, chain_index):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("hybrid relin logn=%d L=%d alpha=%d chain=%d" % (logn, L, alpha, chain_index))
    ctx = nx.Context(logn, [45] * L + [50] * alpha, special_primes=alpha)
    assert ctx.digits == (L + alpha - 1) // alpha
    sk = nx.SecretKey(ctx, seed=5)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=9)
    key = rk.key_array(0)
    rng = np.random.default_rng(logn + L + alpha + chain_index)
    limbs = ctx.limbs_of(chain_index)
    data = _rand_ct(rng, ctx, 3, limbs)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
    got = nx.relinearize(ct, rk).to_array()
    ks0, ks1 = hybrid.switch_key_hybrid(data[2], key, ctx.moduli, L, alpha, logn)
    mods = np.array(ctx.moduli[:limbs], dtype=object)[:, None]
    exp0 = ((data[0].astype(object) + ks0.astype(object)) % mods).astype(np.uint64)
    exp1 = ((data[1].astype(object) + ks1.astype(object)) % mods).astype(np.uint64)
    assert np.array_equal(got[0], exp0)
    assert np.array_equal(got[1], exp1)


@pytest.mark.parametrize("chain,alpha,chain_index", [([46] * 6 + [58] * 2, 2, 1), ([51] * 7 + [58] * 3, 3, 2),
                                                     ([58] + [46] * 5 + [58] * 2, 2, 1),
                                                     ([58] + [46] * 5 + [51] * 2, 2, 1),
                                                     ([51] * 9 + [51] * 8, 8, 1)],
                         ids=["46-58", "51-58", "mixed", "mixed-fp-inputs", "51-alpha8"])
def test_hybrid_relin_bit_exact_wide(chain, alpha, chain_index):
    """58-bit special primes (ModDown inputs beyond 2^51: integer products), a
    58-bit data prime (the integer path for that output) next to 51-bit special
    primes (double-precision products for the other outputs), and 8 special
    primes of 51 bits as in the BERT chain"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn, L = 8, len(chain) - alpha
    _note("hybrid relin wide chain=%s alpha=%d chain_index=%d" % (chain, alpha, chain_index))
    ctx = nx.Context(logn, chain, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=15)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=19)
    key = rk.key_array(0)
    rng = np.random.default_rng(L + alpha + chain_index)
    limbs = ctx.limbs_of(chain_index)
    data = _rand_ct(rng, ctx, 3, limbs)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
    got = nx.relinearize(ct, rk).to_array()
    ks0, ks1 = hybrid.switch_key_hybrid(data[2], key, ctx.moduli, L, alpha, logn)
    mods = np.array(ctx.moduli[:limbs], dtype=object)[:, None]
    assert np.array_equal(got[0], ((data[0].astype(object) + ks0.astype(object)) % mods).astype(np.uint64))
    assert np.array_equal(got[1], ((data[1].astype(object) + ks1.astype(object)) % mods).astype(np.uint64))


@pytest.mark.parametrize("logn,chain,alpha,chain_index",
                         [(14, [45] * 6 + [50] * 2, 2, 1), (14, [51] * 9 + [51] * 8, 8, 1),
                          (15, [46] * 7 + [51] * 3, 3, 2), (16, [51] + [46] * 11 + [51] * 8, 8, 1),
                          (16, [51] + [46] * 11 + [51] * 8, 8, 4)],
                         ids=["n14-a2", "n14-a8", "n15-a3-partial", "n16-a8-partial", "n16-a8-l9"])
def test_hybrid_relin_bit_exact_two_pass(logn, chain, alpha, chain_index):
    """ring sizes of the two-pass NTT with every prime below 2^51: the ModUp
    conversion runs inside the forward column pass (full and partial last
    digits, below the top level)"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    L = len(chain) - alpha
    _note("hybrid relin two-pass logn=%d chain=%s alpha=%d chain_index=%d" % (logn, chain, alpha, chain_index))
    ctx = nx.Context(logn, chain, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=25)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=29)
    key = rk.key_array(0)
    rng = np.random.default_rng(logn + L + alpha + chain_index)
    limbs = ctx.limbs_of(chain_index)
    data = _rand_ct(rng, ctx, 3, limbs)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
    got = nx.relinearize(ct, rk).to_array()
    ks0, ks1 = hybrid.switch_key_hybrid(data[2], key, ctx.moduli, L, alpha, logn)
    mods = np.array(ctx.moduli[:limbs], dtype=object)[:, None]
    assert np.array_equal(got[0], ((data[0].astype(object) + ks0.astype(object)) % mods).astype(np.uint64))
    assert np.array_equal(got[1], ((data[1].astype(object) + ks1.astype(object)) % mods).astype(np.uint64))


def test_hybrid_rotations_two_pass():
    """hoisted rotations (one shared ModUp) and a sum of rotations of four
    inputs (one batched ModUp over a source list) at N = 2^14, bit-exact"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn, L, alpha = 14, 7, 3
    _note("hybrid rotations two-pass logn=%d L=%d alpha=%d" % (logn, L, alpha))
    ctx = nx.Context(logn, [46] * L + [51] * alpha, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=31)
    steps = [1, -2, 0, 77]
    el = [ctx.elt_from_step(s) if s else 0 for s in steps]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [e for e in el if e], seed=32)
    kk = [gk.key_array(e) if e else None for e in el]
    rng = np.random.default_rng(33)
    data = _rand_ct(rng, ctx, 2, L)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40)
    outs = nx.rotate_hoisted(ct, steps, gk)
    for k, s in enumerate(steps):
        ref = data if s == 0 else hybrid.rotate_sum_hybrid([data], [el[k]], [kk[k]], ctx.moduli, L, alpha, logn)
        assert np.array_equal(outs[k].to_array(), ref), s
    others = [_rand_ct(rng, ctx, 2, L) for _ in steps]
    cts = [nx.Ciphertext.from_array(ctx, d, 1, 2.0 ** 40) for d in others]
    got = nx.rotate_sum(cts, steps, gk)
    assert np.array_equal(got.to_array(), hybrid.rotate_sum_hybrid(others, el, kk, ctx.moduli, L, alpha, logn))


def test_hybrid_galois_bit_exact():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn, L, alpha = 8, 6, 2
    ctx = nx.Context(logn, [45] * L + [50] * alpha, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=6)
    elt = ctx.elt_from_step(3)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [elt], seed=3)
    key = gk.key_array(elt)
    rng = np.random.default_rng(13)
    data = _rand_ct(rng, ctx, 2, L)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40)
    got = nx.apply_galois(ct, elt, gk).to_array()
    c0 = np.stack([orc.apply_galois_ntt(data[0, i], logn, elt) for i in range(L)])
    c1 = np.stack([orc.apply_galois_ntt(data[1, i], logn, elt) for i in range(L)])
    ks0, ks1 = hybrid.switch_key_hybrid(c1, key, ctx.moduli, L, alpha, logn)
    mods = np.array(ctx.moduli[:L], dtype=object)[:, None]
    assert np.array_equal(got[0], ((c0.astype(object) + ks0.astype(object)) % mods).astype(np.uint64))
    assert np.array_equal(got[1], ks1)


@pytest.fixture(scope="module")
def env16():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("hybrid N=2^16 keys: 19 x 40-bit data primes, 4 x 58-bit special primes")
    ctx = nx.Context(16, [58] + [40] * 18 + [58] * 4, special_primes=4)
    sk = nx.SecretKey(ctx, seed=12)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=1)
    steps = [1, 2, 5, -7, 1024]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=2)
    return ctx, sk, rk, gk, steps


def test_hybrid_n16_multiply_rotate(env16):
    ctx, sk, rk, gk, steps = env16
    _note("hybrid N=2^16 multiply / rotate / hoisted")
    scale = 2.0 ** 40
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, ctx.slot_count)
    b = rng.uniform(-1, 1, ctx.slot_count)
    ca = nx.encrypt(ctx, sk, nx.encode(ctx, a, scale), seed=1)
    cb = nx.encrypt(ctx, sk, nx.encode(ctx, b, scale), seed=2)
    prod = nx.rescale_to_next(nx.relinearize(nx.multiply(ca, cb), rk))
    dec = lambda c: nx.decode(ctx, nx.decrypt(ctx, sk, c))
    assert np.max(np.abs(dec(prod) - a * b)) < 1e-5
    for s in steps:
        assert np.max(np.abs(dec(nx.rotate_vector(ca, s, gk)) - np.roll(a, -s))) < 1e-5, s
    outs = nx.rotate_hoisted(prod, steps, gk)
    for s, o in zip(steps, outs):
        assert np.max(np.abs(dec(o) - np.roll(a * b, -s))) < 1e-5, s
    hs = nx.hoisting(prod, steps, gk)
    assert np.max(np.abs(dec(hs) - sum(np.roll(a * b, -s) for s in steps))) < 1e-4


def test_hybrid_bootstrap():
    """bootstrap at N=2^15 with 2^13 slots over the reference's data chain and
    three special primes (key-switching digits of ten or eleven primes)"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    from oracle import boot as ob
    logN, logn, K = 15, 13, 25
    chain = [51] + [46] * 16 + [51] * 14 + [58] * 3
    _note("hybrid bootstrap keys logN=15 alpha=3")
    ctx = nx.Context(logN, chain, hamming_weight=192, special_primes=3)
    sk = nx.SecretKey(ctx, seed=44)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=3)
    steps = nx.boot_plan_steps(logN - 1, logn, K)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=4)
    ev = nx.CKKSEvaluator(ctx, rk, gk, 2.0 ** 46, sk)
    bt = nx.Bootstrapper(ev, log_slots=logn, K=K)
    n = 1 << logn
    v = np.random.default_rng(9).uniform(-1, 1, n)
    ct = nx.encrypt(ctx, sk, nx.encode(ctx, np.tile(v, (1 << (logN - 1)) // n), 2.0 ** 46), seed=5)
    ct = nx.mod_switch_to(ct, ctx.L)
    _note("hybrid bootstrap")
    out = bt.bootstrap(ct)
    after = nx.decode(ctx, nx.decrypt(ctx, sk, out))[:n]
    err = np.abs(after - v)
    print("hybrid bootstrap: max %.3g mean %.3g" % (err.max(), err.mean()))
    assert err.max() < 1e-3 and err.mean() < 1e-4
    assert out.chain_index == 1 + 14
    # lockstep EvalMod of two inputs: products relinearised and rescaled in one
    # rounding, so the outputs match the single bootstrap to rounding noise
    _note("hybrid bootstrap_many x2")
    w = np.random.default_rng(10).uniform(-1, 1, n)
    ct2 = nx.mod_switch_to(nx.encrypt(ctx, sk, nx.encode(ctx, np.tile(w, (1 << (logN - 1)) // n), 2.0 ** 46),
                                      seed=6), ctx.L)
    many = bt.bootstrap_many([ct, ct2])
    got = [nx.decode(ctx, nx.decrypt(ctx, sk, m))[:n] for m in many]
    assert np.max(np.abs(got[0] - after)) < 1e-5
    err2 = np.abs(got[1] - w)
    assert err2.max() < 1e-3 and err2.mean() < 1e-4
    assert all(m.chain_index == out.chain_index and m.scale == out.scale for m in many)
