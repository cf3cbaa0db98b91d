"""Conversions fused into the forward NTT's column pass (ring sizes of the
two-pass transform, every prime below 2^51), bit-exact against the oracle:

* rescale (util/rns.cpp:830-905): the centred last limb reduced to each
  remaining prime inside the column pass, single and batched;
* SEAL-exact key switching with one special prime (evaluator.cpp:2724-3021):
  the ModDown's centring of the special limb inside the column pass;
* hybrid multiply-relinearise-rescale in one rounding: the conversion by
  P q_(l-1) inside the column pass equals the per-ciphertext fused products
  word for word (the batched and single calls share no launch);
* key-switching keys generated here, whose uniform half the inner products
  regenerate from its seed, against the same words uploaded as a plain key.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
import oracle as orc  # noqa: E402

CHAINS = {14: [51, 46, 46, 46, 46, 51], 15: [51] + [46] * 6 + [51], 16: [51, 46, 46, 46, 51]}


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


def _rand(rng, mods, size, limbs, n):
    return np.stack([np.stack([rng.integers(0, mods[i], n, dtype=np.uint64) for i in range(limbs)])
                     for _ in range(size)])


@pytest.mark.parametrize("logn,chain_index", [(14, 1), (14, 3), (15, 1), (16, 2)])
def test_rescale_column_fusion(logn, chain_index):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("rescale (centring in the column pass) logn=%d chain_index=%d" % (logn, chain_index))
    ctx = nx.Context(logn, CHAINS[logn])
    rng = np.random.default_rng(100 + logn + chain_index)
    limbs = ctx.limbs_of(chain_index)
    datas = [_rand(rng, ctx.moduli, 2, limbs, ctx.n) for _ in range(3)]
    cts = [nx.Ciphertext.from_array(ctx, d, chain_index, 2.0 ** 80) for d in datas]
    refs = [np.stack([orc.rescale_poly_ntt(d[p], ctx.moduli[:limbs], logn) for p in range(2)]) for d in datas]
    assert np.array_equal(nx.rescale_to_next(cts[0]).to_array(), refs[0])
    many = nx.rescale_to_next_many(cts)
    for got, ref in zip(many, refs):
        assert np.array_equal(got.to_array(), ref)


@pytest.mark.parametrize("logn", [14, 16])
def test_seal_key_switch_column_fusion(logn):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("one-special-prime relinearize (centring in the column pass) logn=%d" % logn)
    ctx = nx.Context(logn, CHAINS[logn])
    sk = nx.SecretKey(ctx, seed=81)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=82)
    rng = np.random.default_rng(83 + logn)
    data = _rand(rng, ctx.moduli, 3, ctx.L, ctx.n)
    got = nx.relinearize(nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 80), rk).to_array()
    assert np.array_equal(got, orc.switch_key(data[:2], data[2], ctx.moduli, rk.key_array(0), logn))


def test_fused_relin_rescale_batch_matches_single():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn, chain, special = 16, [51, 46, 46, 46, 46, 51, 51, 51], 3
    _note("fused relin+rescale batched vs single logn=%d alpha=%d" % (logn, special))
    ctx = nx.Context(logn, chain, special_primes=special)
    sk = nx.SecretKey(ctx, seed=91)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=92)
    rng = np.random.default_rng(93)
    enc = lambda s: nx.encrypt(ctx, sk, nx.encode(ctx, rng.uniform(-1, 1, ctx.slot_count), 2.0 ** 40), seed=s)
    a = [enc(10 + i) for i in range(9)]
    b = [enc(30 + i) for i in range(9)]
    many = nx.multiply_relin_rescale_many(a, b, rk, fused=True)
    for i in range(9):
        one = nx.multiply_relin_rescale_many([a[i]], [b[i]], rk, fused=True)[0]
        assert np.array_equal(many[i].to_array(), one.to_array()), i
        dec = nx.decode(ctx, nx.decrypt(ctx, sk, many[i]))
        want = nx.decode(ctx, nx.decrypt(ctx, sk, a[i])) * nx.decode(ctx, nx.decrypt(ctx, sk, b[i]))
        assert np.max(np.abs(dec - want)) < 1e-5


@pytest.mark.parametrize("special", [1, 3])
def test_generated_keys_regenerate_uniform_half(special):
    """keys generated here are used with their uniform half regenerated from
    its seed inside the inner products; the same words uploaded as a plain key
    (read in full) give identical relinearisation, rotation and rotation-sum
    residues"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn = 14
    chain = [51, 46, 46, 46, 46, 46, 51] + ([51, 51] if special == 3 else [])
    _note("seeded keys vs uploaded keys logn=%d special=%d" % (logn, special))
    ctx = nx.Context(logn, chain, special_primes=special)
    sk = nx.SecretKey(ctx, seed=61)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=62)
    steps = [1, -3, 17]
    elts = [ctx.elt_from_step(s) for s in steps]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, elts, seed=63)
    rk2 = nx.KSwitchKeys(ctx)
    rk2.upload(0, rk.key_array(0))
    gk2 = nx.KSwitchKeys(ctx)
    for e in elts:
        gk2.upload(e, gk.key_array(e))
    rng = np.random.default_rng(64)
    L = ctx.L
    d3 = _rand(rng, ctx.moduli, 3, L, ctx.n)
    ct3 = nx.Ciphertext.from_array(ctx, d3, 1, 2.0 ** 40)
    assert np.array_equal(nx.relinearize(ct3, rk).to_array(), nx.relinearize(ct3, rk2).to_array())
    cts = [nx.Ciphertext.from_array(ctx, _rand(rng, ctx.moduli, 2, L, ctx.n), 1, 2.0 ** 40) for _ in steps]
    for s in steps:
        assert np.array_equal(nx.rotate_vector(cts[0], s, gk).to_array(), nx.rotate_vector(cts[0], s, gk2).to_array())
    for a, b in zip(nx.rotate_hoisted(cts[0], steps, gk), nx.rotate_hoisted(cts[0], steps, gk2)):
        assert np.array_equal(a.to_array(), b.to_array())
    assert np.array_equal(nx.rotate_sum(cts, steps, gk).to_array(), nx.rotate_sum(cts, steps, gk2).to_array())
