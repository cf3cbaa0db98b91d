"""Hoisted rotations (one ModUp shared by all steps; reference Phantom
hoisting_inplace, cuda/thirdparty/phantom-fhe/src/evaluate.cu:1727-1920).

Hoisting decomposes sigma(c1) differently from the one-at-a-time rotation
(sigma of the lifted digits instead of the lift of sigma(c1)), so residues are
not compared; decryptions are: each hoisted rotation and the hoisted sum must
decode to the rotated message within key-switching noise, like the plain
rotation does.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")

SCALE = 2.0 ** 40


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=[(13, [60, 40, 40, 60]), (14, [50] * 9 + [60])], ids=["l3", "l9"])
def env(request):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logN, chain = request.param
    _note("hoist keys logN=%d L=%d" % (logN, len(chain) - 1))
    ctx = nx.Context(logN, chain)
    sk = nx.SecretKey(ctx, seed=31)
    steps = sorted(set([1, 2, 3, 5, 7, 8, 16, 100, -1, -3, -64] + list(range(9, 45))))
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=5)
    return ctx, sk, gk, steps


def _enc(ctx, sk, v):
    return nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=3)


def _dec(ctx, sk, ct):
    return nx.decode(ctx, nx.decrypt(ctx, sk, ct), complex_out=True)


def test_rotate_hoisted_matches_rotation(env):
    ctx, sk, gk, _ = env
    _note("rotate_hoisted logN=%d" % ctx.log_n)
    rng = np.random.default_rng(1)
    v = rng.uniform(-1, 1, ctx.slot_count) + 1j * rng.uniform(-1, 1, ctx.slot_count)
    ct = _enc(ctx, sk, v)
    steps = [0, 1, 3, -1, 8, 100, -64, 7]
    outs = nx.rotate_hoisted(ct, steps, gk)
    for s, o in zip(steps, outs):
        assert o.chain_index == ct.chain_index and o.scale == ct.scale
        got = _dec(ctx, sk, o)
        plain = _dec(ctx, sk, nx.rotate_vector(ct, s, gk)) if s else _dec(ctx, sk, ct)
        exp = np.roll(v, -s)
        # same noise as the plain rotation: the 60-bit q0 over a 60-bit special
        # prime gives both a heavy tail near 1e-6 at scale 2^40
        e_h, e_p = np.abs(got - exp), np.abs(plain - exp)
        assert e_h.max() < 1e-5 and e_h.mean() < 2 * e_p.mean() + 1e-9, s


def test_rotate_hoisted_many_steps(env):
    """more steps than one launch takes (chunks of 32)"""
    ctx, sk, gk, steps = env
    _note("rotate_hoisted %d steps" % len(steps))
    rng = np.random.default_rng(2)
    v = rng.uniform(-1, 1, ctx.slot_count)
    ct = nx.mod_switch_to_next(_enc(ctx, sk, v))
    outs = nx.rotate_hoisted(ct, steps, gk)
    for s, o in zip(steps, outs):
        e = np.abs(_dec(ctx, sk, o).real - np.roll(v, -s))
        assert e.max() < 1e-5 and e.mean() < 1e-7, s


@pytest.mark.parametrize("steps", [[1, 2, 3], [0, 5, -3, 16], "all"])
def test_hoisting_sum(env, steps):
    ctx, sk, gk, all_steps = env
    if steps == "all":
        steps = all_steps
    _note("hoisting sum of %d steps" % len(steps))
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, ctx.slot_count)
    ct = _enc(ctx, sk, v)
    out = nx.hoisting(ct, steps, gk)
    exp = sum(np.roll(v, -s) for s in steps)
    got = _dec(ctx, sk, out).real
    assert np.max(np.abs(got - exp)) < 1e-5 * len(steps)


@pytest.mark.parametrize("count", [1, 7, 70])
def test_multiply_plain_accumulate_bit_exact(env, count):
    """one-pass sum_r ct_r * pt_r equals the chain of multiply_plain + add residue for residue"""
    ctx, sk, _, _ = env
    _note("multiply_plain_accumulate count=%d" % count)
    rng = np.random.default_rng(count)
    cts = [_enc(ctx, sk, rng.uniform(-1, 1, ctx.slot_count)) for _ in range(count)]
    pts = [nx.encode(ctx, rng.uniform(-1, 1, ctx.slot_count), SCALE) for _ in range(count)]
    got = nx.multiply_plain_accumulate(cts, pts)
    ref = nx.multiply_plain(cts[0], pts[0])
    for c, p in zip(cts[1:], pts[1:]):
        ref = nx.add(ref, nx.multiply_plain(c, p))
    assert got.scale == ref.scale
    assert np.array_equal(got.to_array(), ref.to_array())


def test_multiply_plain_accumulate_multi_bit_exact(env):
    """every giant step of a BSGS output in one pass (37 ciphertexts, 6 rows:
    both chunk sizes crossed, holes in every row, one row with a single term)
    equals multiply_plain_accumulate over each row's terms word for word"""
    ctx, sk, _, _ = env
    _note("multiply_plain_accumulate_multi 37 x 6")
    rng = np.random.default_rng(77)
    K, G = 37, 6
    cts = [_enc(ctx, sk, rng.uniform(-1, 1, ctx.slot_count)) for _ in range(K)]
    rows = []
    for g in range(G):
        keep = rng.random(K) < 0.7 if g != 3 else np.arange(K) == 33
        rows.append([nx.encode(ctx, rng.uniform(-1, 1, ctx.slot_count), SCALE) if keep[r] else None
                     for r in range(K)])
    got = nx.multiply_plain_accumulate_multi(cts, rows)
    assert len(got) == G
    for g in range(G):
        terms = [(c, p) for c, p in zip(cts, rows[g]) if p is not None]
        ref = nx.multiply_plain_accumulate([c for c, _ in terms], [p for _, p in terms])
        assert np.array_equal(got[g].to_array(), ref.to_array()), g
        assert got[g].info()["scale"] == ref.info()["scale"]
    with pytest.raises(nx.NexusError):  # a row without terms
        nx.multiply_plain_accumulate_multi(cts[:2], [[None, None]])


def test_rotate_sum_matches_separate_rotations():
    """sum_i rot(ct_i, s_i) with one ModDown decrypts like the sum of separate
    rotations (identity steps and repeated steps included)"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    logn = 13
    ctx = nx.Context(logn, [60, 40, 40, 60], special_primes=1)
    sk = nx.SecretKey(ctx, seed=61)
    steps = [0, 1, 3, -5, 1, 64]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in set(steps) if s], seed=62)
    rng = np.random.default_rng(63)
    vals = [rng.uniform(-1, 1, ctx.slot_count) for _ in steps]
    cts = [nx.encrypt(ctx, sk, nx.encode(ctx, v, 2.0 ** 40), seed=70 + i) for i, v in enumerate(vals)]
    got = nx.decode(ctx, nx.decrypt(ctx, sk, nx.rotate_sum(cts, steps, gk)))
    want = sum(np.roll(v, -s) for v, s in zip(vals, steps))
    sep = cts[0] if steps[0] == 0 else nx.rotate_vector(cts[0], steps[0], gk)
    for c, s in zip(cts[1:], steps[1:]):
        sep = nx.add(sep, c if s == 0 else nx.rotate_vector(c, s, gk))
    ref = nx.decode(ctx, nx.decrypt(ctx, sk, sep))
    print("rotate_sum: max err %.3g, separate %.3g" % (np.max(np.abs(got - want)), np.max(np.abs(ref - want))))
    # six fresh encryptions at scale 2^40 carry ~1e-6 of noise; one ModDown
    # instead of five must not add to it
    assert np.max(np.abs(got - want)) < 1e-5
    assert np.max(np.abs(got - want)) <= 2 * np.max(np.abs(ref - want)) + 1e-9
