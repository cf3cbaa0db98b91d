"""GPU parity, through the C ABI, of the operations whose residues the other
primitive tests reach only through programs:

* mod-switch drop of ciphertexts and plaintexts (SEAL evaluator.cpp:1483-1580,
  :1630-1680) against array slices, at N = 2^12 and 2^16;
* the reduced-error family (SEAL-4.1-bs evaluator.cpp:420-594) against its
  oracle restatement (oracle/programs.reduced_error), bit-exact, with the
  result scale equal as a double;
* hoisted rotations, hoisting sums and sums of rotations with one ModDown
  (Phantom hoisting_inplace, evaluate.cu:1727-1920) against orc.rotate_sum
  (one special prime) and hybrid.rotate_sum_hybrid (alpha special primes),
  bit-exact;
* key switching at full size: relinearisation and a Galois key switch at
  N = 2^16 on the reference's nonlinear chain {58, 40 x 18, 58}, l = 19,
  bit-exact against orc.switch_key.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
import oracle as orc  # noqa: E402
from oracle import hybrid, programs  # noqa: E402


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")


def _rand(rng, moduli, size, limbs, n):
    return np.stack([np.stack([rng.integers(0, moduli[i], n, dtype=np.uint64) for i in range(limbs)])
                     for _ in range(size)])


# ---------------------------------------------------------------- mod switch
@pytest.mark.parametrize("logn,chain", [(12, [58, 40, 40, 40, 58]), (16, [58] + [40] * 18 + [58])])
def test_mod_switch_drop_parity(logn, chain):
    _note("mod switch drop logn=%d" % logn)
    ctx = nx.Context(logn, chain)
    rng = np.random.default_rng(logn)
    L = ctx.L
    for size in (2, 3):
        data = _rand(rng, ctx.moduli, size, L, ctx.n)
        ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40)
        nxt = nx.mod_switch_to_next(ct)
        i = nxt.info()
        assert i["chain_index"] == 2 and i["limbs"] == L - 1 and i["size"] == size and i["scale"] == 2.0 ** 40
        assert np.array_equal(nxt.to_array(), data[:, :L - 1])
        for target in (1, 2, L // 2, L):
            got = nx.mod_switch_to(ct, target)
            assert got.chain_index == target
            assert np.array_equal(got.to_array(), data[:, :ctx.limbs_of(target)]), (size, target)
        # in place (dest aliases the input)
        c2 = ct.copy()
        nx.mod_switch_to(c2, L, dest=c2)
        assert np.array_equal(c2.to_array(), data[:, :1])
        with pytest.raises(nx.NexusError):
            nx.mod_switch_to(nxt, 1)  # cannot switch up
        with pytest.raises(nx.NexusError):
            nx.mod_switch_to_next(c2)  # end of the chain
    pdata = _rand(rng, ctx.moduli, 1, L, ctx.n)[0]
    pt = nx.Plaintext.from_array(ctx, pdata, 1, 2.0 ** 40)
    for target in (1, 3, L):
        got = nx.plain_mod_switch_to(pt, target)
        assert got.info()["chain_index"] == target and got.scale == 2.0 ** 40
        assert np.array_equal(got.to_array(), pdata[:ctx.limbs_of(target)])
    with pytest.raises(nx.NexusError):
        nx.plain_mod_switch_to(nx.plain_mod_switch_to(pt, 3), 2)


# ---------------------------------------------------------------- reduced error
RE_CHAIN = [58, 40, 40, 40, 40, 58]


@pytest.fixture(scope="module")
def re_env():
    ctx = nx.Context(12, RE_CHAIN)
    sk = nx.SecretKey(ctx, seed=3)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=4)
    return ctx, rk, rk.key_array(0)


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
@pytest.mark.parametrize("c1,c2", [(1, 1), (2, 2), (3, 1), (1, 3), (2, 4), (4, 2)])
def test_reduced_error_parity(re_env, op, c1, c2):
    ctx, rk, key = re_env
    _note("reduced_error %s chain %d/%d" % (op, c1, c2))
    rng = np.random.default_rng(100 * c1 + 10 * c2 + len(op))
    l1, l2 = ctx.limbs_of(c1), ctx.limbs_of(c2)
    s1, s2 = 2.0 ** 40 * 1.0009765625, 2.0 ** 40 * 0.99951171875
    a = _rand(rng, ctx.moduli, 2, l1, ctx.n)
    b = _rand(rng, ctx.moduli, 2, l2, ctx.n)
    A = nx.Ciphertext.from_array(ctx, a, c1, s1)
    B = nx.Ciphertext.from_array(ctx, b, c2, s2)
    if op == "add":
        got = nx.add_reduced_error(A, B)
    elif op == "sub":
        got = nx.sub_reduced_error(A, B)
    else:
        got = nx.multiply_reduced_error(A, B, rk)
    ref, scale = programs.reduced_error(op, a, s1, b, s2, ctx.moduli, ctx.log_n, key, ctx.moduli)
    assert got.scale == scale
    assert np.array_equal(got.to_array(), ref)
    # dest aliasing the first operand (the _inplace form) gives the same
    fn = {"add": nx.add_reduced_error, "sub": nx.sub_reduced_error}.get(op)
    if fn is not None:
        fn(A, B, dest=A)
        assert np.array_equal(A.to_array(), ref) and A.scale == scale


def test_multiply_vector_and_double_parity(re_env):
    ctx, _, _ = re_env
    _note("multiply_vector_reduced_error / double")
    rng = np.random.default_rng(7)
    a = _rand(rng, ctx.moduli, 2, 3, ctx.n)
    A = nx.Ciphertext.from_array(ctx, a, ctx.L - 2, 2.0 ** 40)
    for vals in (rng.uniform(-1, 1, ctx.slot_count), rng.uniform(-1, 1, 100) + 1j * rng.uniform(-1, 1, 100)):
        got = nx.multiply_vector_reduced_error(A, vals)
        pt = orc.Encoder(ctx.log_n).encode(vals, 2.0 ** 40, ctx.moduli[:3])
        assert np.array_equal(got.to_array(), orc.dyadic_mul(a, pt, ctx.moduli, ctx.log_n))
        assert got.scale == 2.0 ** 80
    assert np.array_equal(nx.double(A).to_array(), programs.add(a, a, ctx.moduli))


# ---------------------------------------------------------------- hoisting
@pytest.mark.parametrize("logn,chain,chain_index", [(12, [58, 40, 40, 40, 58], 1), (12, [58, 40, 40, 40, 58], 3),
                                                    (16, [58, 40, 40, 40, 58], 2)])
def test_hoisting_parity(logn, chain, chain_index):
    _note("hoisting logn=%d chain_index=%d" % (logn, chain_index))
    ctx = nx.Context(logn, chain)
    sk = nx.SecretKey(ctx, seed=31)
    steps = [1, -3, 17, 0, 256]
    elts = [ctx.elt_from_step(s) for s in steps if s]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, elts, seed=32)
    keys = {e: gk.key_array(e) for e in elts}
    rng = np.random.default_rng(logn + chain_index)
    limbs = ctx.limbs_of(chain_index)
    data = _rand(rng, ctx.moduli, 2, limbs, ctx.n)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
    el = [ctx.elt_from_step(s) if s else 0 for s in steps]
    kk = [keys[e] if e else None for e in el]
    # one hoisted rotation per step
    outs = nx.rotate_hoisted(ct, steps, gk)
    for k, (s, e) in enumerate(zip(steps, el)):
        ref = data if s == 0 else orc.rotate_hoisted(data, e, kk[k], ctx.moduli, logn)
        assert np.array_equal(outs[k].to_array(), ref), s
    # Phantom's hoisting sum: one ModUp, one ModDown
    got = nx.hoisting(ct, steps, gk)
    assert np.array_equal(got.to_array(), orc.rotate_sum([data] * len(steps), el, kk, ctx.moduli, logn))
    # sum of rotations of distinct inputs with one ModDown
    others = [_rand(rng, ctx.moduli, 2, limbs, ctx.n) for _ in steps]
    cts = [nx.Ciphertext.from_array(ctx, d, chain_index, 2.0 ** 40) for d in others]
    got = nx.rotate_sum(cts, steps, gk)
    assert np.array_equal(got.to_array(), orc.rotate_sum(others, el, kk, ctx.moduli, logn))


@pytest.mark.parametrize("L,alpha,chain_index", [(6, 2, 1), (6, 2, 2), (7, 3, 1)])
def test_hoisting_hybrid_parity(L, alpha, chain_index):
    logn = 8
    _note("hoisting hybrid L=%d alpha=%d chain_index=%d" % (L, alpha, chain_index))
    ctx = nx.Context(logn, [45] * L + [50] * alpha, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=41)
    steps = [1, -2, 0, 5]
    el = [ctx.elt_from_step(s) if s else 0 for s in steps]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [e for e in el if e], seed=42)
    kk = [gk.key_array(e) if e else None for e in el]
    rng = np.random.default_rng(L + alpha + chain_index)
    limbs = ctx.limbs_of(chain_index)
    data = _rand(rng, ctx.moduli, 2, limbs, ctx.n)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
    outs = nx.rotate_hoisted(ct, steps, gk)
    for k, s in enumerate(steps):
        ref = data if s == 0 else hybrid.rotate_sum_hybrid([data], [el[k]], [kk[k]], ctx.moduli, L, alpha, logn)
        assert np.array_equal(outs[k].to_array(), ref), s
    others = [_rand(rng, ctx.moduli, 2, limbs, ctx.n) for _ in steps]
    cts = [nx.Ciphertext.from_array(ctx, d, chain_index, 2.0 ** 40) for d in others]
    got = nx.rotate_sum(cts, steps, gk)
    assert np.array_equal(got.to_array(), hybrid.rotate_sum_hybrid(others, el, kk, ctx.moduli, L, alpha, logn))
    got = nx.hoisting(ct, steps, gk)
    assert np.array_equal(got.to_array(),
                          hybrid.rotate_sum_hybrid([data] * len(steps), el, kk, ctx.moduli, L, alpha, logn))


# ---------------------------------------------------------------- full-size key switch
NL_CHAIN = [58] + [40] * 18 + [58]


@pytest.fixture(scope="module")
def nl_env():
    ctx = nx.Context(16, NL_CHAIN)
    sk = nx.SecretKey(ctx, seed=51)
    return ctx, sk


def test_relinearize_full_size(nl_env):
    ctx, sk = nl_env
    _note("relinearize N=2^16 l=19 (reference nonlinear chain)")
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=52)
    key = rk.key_array(0)
    rng = np.random.default_rng(53)
    data = _rand(rng, ctx.moduli, 3, ctx.L, ctx.n)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 80)
    got = nx.relinearize(ct, rk).to_array()
    assert np.array_equal(got, orc.switch_key(data[:2], data[2], ctx.moduli, key, 16))


def test_apply_galois_full_size(nl_env):
    ctx, sk = nl_env
    _note("apply_galois N=2^16 l=19 step 3")
    elt = ctx.elt_from_step(3)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [elt], seed=54)
    key = gk.key_array(elt)
    rng = np.random.default_rng(55)
    data = _rand(rng, ctx.moduli, 2, ctx.L, ctx.n)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40)
    got = nx.apply_galois(ct, elt, gk).to_array()
    assert np.array_equal(got, programs.apply_galois(data, elt, key, ctx.moduli, 16))


@pytest.mark.parametrize("alpha", [1, 2])
def test_rotations_many_match_single(alpha):
    """the batched hoisted rotations and sums of rotations equal the per-input
    calls word for word"""
    logn = 12
    _note("rotations many alpha=%d" % alpha)
    ctx = nx.Context(logn, [58, 40, 40, 40, 40] + [58] * alpha, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=61)
    steps = [0, 1, -2, 5, 64]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps if s], seed=62)
    rng = np.random.default_rng(63 + alpha)
    # groups at different scales, as the giant sums of differently scaled inputs
    cts = [nx.Ciphertext.from_array(ctx, _rand(rng, ctx.moduli, 2, 4, ctx.n), 2, 2.0 ** 40 * (1 + (k >= 3)))
           for k in range(9)]
    many = nx.rotate_hoisted_many(cts, steps, gk)
    for c, row in zip(cts, many):
        single = nx.rotate_hoisted(c, steps, gk)
        for a, b in zip(row, single):
            assert np.array_equal(a.to_array(), b.to_array())
    groups = [cts[:3], cts[3:8], [cts[8]]]
    gsteps = [[1, 0, 5], [0, 0, 64, 1, -2], [0]]
    sums = nx.rotate_sum_many(groups, gsteps, gk)
    for g, s, got in zip(groups, gsteps, sums):
        want = nx.rotate_sum(g, s, gk)
        assert np.array_equal(got.to_array(), want.to_array()) and got.scale == want.scale
