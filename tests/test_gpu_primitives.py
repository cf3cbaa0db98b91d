"""GPU parity of the CKKS primitives against the CPU oracle (bit-exact residues).

Each test feeds the same seeded residues (and, for key switching, the same key
material, downloaded from the device) to the HIP path through the C ABI and to
the oracle, and compares every output word.
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


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")


def rand_poly(rng, moduli, n):
    return np.stack([rng.integers(0, q, size=n, dtype=np.uint64) for q in moduli])


def rand_ct(rng, ctx, size, chain_index):
    limbs = ctx.limbs_of(chain_index)
    return np.stack([rand_poly(rng, ctx.moduli[:limbs], ctx.n) for _ in range(size)])


CHAINS = {
    6: [40, 30, 30, 50],
    12: [58, 40, 40, 40, 58],
    13: [60, 40, 60],
    14: [58, 40, 40, 58],
    15: [51, 46, 46, 51],
    16: [58, 40, 40, 40, 58],
}


@pytest.mark.parametrize("logn", sorted(CHAINS))
def test_ntt_roundtrip_parity(logn):
    _note("ntt logn=%d" % logn)
    ctx = nx.Context(logn, CHAINS[logn])
    rng = np.random.default_rng(logn)
    data = rand_ct(rng, ctx, 2, 1)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 1.0, ntt_form=False)
    ntt = nx.transform_to_ntt(ct).to_array()
    for p in range(2):
        for i in range(data.shape[1]):
            ref = orc.ntt_forward(data[p, i], logn, ctx.moduli[i])
            assert np.array_equal(ntt[p, i], ref), (p, i)
    back = nx.transform_from_ntt(nx.transform_to_ntt(ct)).to_array()
    assert np.array_equal(back, data)


@pytest.mark.parametrize("logn", [12, 14, 16])
def test_ntt_extreme_values(logn):
    """all q - 1, alternating 0 / q - 1, and a single non-zero word, on primes
    of the double-precision path (30..51 bits) and of the 64-bit path (58, 60)"""
    _note("ntt extremes logn=%d" % logn)
    ctx = nx.Context(logn, [51, 50, 46, 30, 58, 60])
    n = ctx.n
    limbs = ctx.limbs_of(1)
    for name in ("max", "alt", "one"):
        data = np.zeros((2, limbs, n), dtype=np.uint64)
        for i in range(limbs):
            q = ctx.moduli[i]
            if name == "max":
                data[:, i] = q - 1
            elif name == "alt":
                data[:, i, ::2] = q - 1
            else:
                data[:, i, n // 3] = q - 1
        ct = nx.Ciphertext.from_array(ctx, data, 1, 1.0, ntt_form=False)
        fwd = nx.transform_to_ntt(ct).to_array()
        inv = nx.transform_from_ntt(nx.Ciphertext.from_array(ctx, data, 1, 1.0, ntt_form=True)).to_array()
        for i in range(limbs):
            q = ctx.moduli[i]
            assert np.array_equal(fwd[0, i], orc.ntt_forward(data[0, i], logn, q)), (name, i)
            assert np.array_equal(inv[0, i], orc.ntt_inverse(data[0, i], logn, q)), (name, i)


@pytest.mark.parametrize("logn", [12, 16])
def test_inverse_ntt_parity(logn):
    _note("intt logn=%d" % logn)
    ctx = nx.Context(logn, CHAINS[logn])
    rng = np.random.default_rng(100 + logn)
    data = rand_ct(rng, ctx, 2, 1)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 1.0, ntt_form=True)
    out = nx.transform_from_ntt(ct).to_array()
    for p in range(2):
        for i in range(data.shape[1]):
            assert np.array_equal(out[p, i], orc.ntt_inverse(data[p, i], logn, ctx.moduli[i])), (p, i)


@pytest.mark.parametrize("logn,chain_index", [(12, 1), (12, 2), (16, 1), (13, 1)])
def test_rescale_parity(logn, chain_index):
    _note("rescale logn=%d chain=%d" % (logn, chain_index))
    ctx = nx.Context(logn, CHAINS[logn])
    rng = np.random.default_rng(7 + logn)
    data = rand_ct(rng, ctx, 2, chain_index)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 80)
    out = nx.rescale_to_next(ct)
    got = out.to_array()
    limbs = data.shape[1]
    for p in range(2):
        ref = orc.rescale_poly_ntt(data[p], ctx.moduli[:limbs], logn)
        assert np.array_equal(got[p], ref), p
    assert out.scale == 2.0 ** 80 / ctx.moduli[limbs - 1]
    assert out.chain_index == chain_index + 1


@pytest.mark.parametrize("logn", [12, 16])
def test_dyadic_parity(logn):
    _note("dyadic logn=%d" % logn)
    ctx = nx.Context(logn, CHAINS[logn])
    rng = np.random.default_rng(3)
    a = rand_ct(rng, ctx, 2, 1)
    b = rand_ct(rng, ctx, 2, 1)
    pt = rand_poly(rng, ctx.moduli[: a.shape[1]], ctx.n)
    A = nx.Ciphertext.from_array(ctx, a, 1, 2.0 ** 20)
    B = nx.Ciphertext.from_array(ctx, b, 1, 2.0 ** 20)
    P = nx.Plaintext.from_array(ctx, pt, 1, 2.0 ** 20)
    mods = np.array(ctx.moduli[: a.shape[1]], dtype=object)[:, None]
    ao, bo, po = a.astype(object), b.astype(object), pt.astype(object)
    assert np.array_equal(nx.add(A, B).to_array().astype(object), (ao + bo) % mods)
    assert np.array_equal(nx.sub(A, B).to_array().astype(object), (ao - bo) % mods)
    assert np.array_equal(nx.negate(A).to_array().astype(object), (-ao) % mods)
    assert np.array_equal(nx.multiply_plain(A, P).to_array(), orc.dyadic_mul(a, pt, ctx.moduli, logn))
    ap = nx.add_plain(A, P).to_array().astype(object)
    assert np.array_equal(ap[0], (ao[0] + po) % mods) and np.array_equal(ap[1], ao[1])
    t = nx.multiply(A, B)
    assert np.array_equal(t.to_array(), orc.tensor(a, b, ctx.moduli, logn))
    assert t.scale == 2.0 ** 40
    assert np.array_equal(nx.square(A).to_array(), orc.tensor(a, a, ctx.moduli, logn))


@pytest.mark.parametrize("logn,chain_index", [(12, 1), (12, 3), (16, 2)])
def test_relinearize_parity(logn, chain_index):
    _note("relin logn=%d chain=%d" % (logn, chain_index))
    ctx = nx.Context(logn, CHAINS[logn])
    sk = nx.SecretKey(ctx, seed=5)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=9)
    key = rk.key_array(0)
    rng = np.random.default_rng(11)
    data = rand_ct(rng, ctx, 3, chain_index)
    ct = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
    got = nx.relinearize(ct, rk).to_array()
    ref = orc.switch_key(data[:2], data[2], ctx.moduli, key, logn)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("logn,step", [(12, 1), (12, -3), (12, 0), (16, 5)])
def test_apply_galois_parity(logn, step):
    _note("galois logn=%d step=%d" % (logn, step))
    ctx = nx.Context(logn, CHAINS[logn])
    sk = nx.SecretKey(ctx, seed=6)
    elt = ctx.elt_from_step(step)
    assert elt == orc.elt_from_step(logn, step)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [elt], seed=3)
    key = gk.key_array(elt)
    rng = np.random.default_rng(13)
    data = rand_ct(rng, ctx, 2, 1)
    ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40)
    got = nx.apply_galois(ct, elt, gk).to_array()
    limbs = data.shape[1]
    c0 = np.stack([orc.apply_galois_ntt(data[0, i], logn, elt) for i in range(limbs)])
    c1 = np.stack([orc.apply_galois_ntt(data[1, i], logn, elt) for i in range(limbs)])
    base = np.stack([c0, np.zeros_like(c0)])
    ref = orc.switch_key(base, c1, ctx.moduli, key, logn)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("logn", [10, 13, 16])
def test_encode_parity(logn):
    _note("encode logn=%d" % logn)
    chain = CHAINS.get(logn, [60, 40, 40, 60])
    ctx = nx.Context(logn, chain)
    rng = np.random.default_rng(21)
    vals = rng.uniform(-1, 1, ctx.slot_count)
    scale = 2.0 ** 40
    pt = nx.encode(ctx, vals, scale)
    enc = orc.Encoder(logn)
    limbs = ctx.L
    ref = enc.encode(vals, scale, ctx.moduli[:limbs])
    assert np.array_equal(pt.to_array(), ref)
    cv = rng.uniform(-1, 1, 37) + 1j * rng.uniform(-1, 1, 37)
    ptc = nx.encode(ctx, cv, scale, chain_index=2)
    assert np.array_equal(ptc.to_array(), enc.encode(cv, scale, ctx.moduli[: limbs - 1]))
    # decode returns doubles: same algorithm, compared to 1e-12 (far below 2^-20)
    dec = nx.decode(ctx, pt)
    dref = enc.decode(ref, ctx.moduli[:limbs], scale)
    assert np.max(np.abs(dec - dref)) < 1e-12
    assert np.max(np.abs(dec - vals)) < 1e-7
    c = nx.encode_constant(ctx, -1.25, scale)
    assert np.array_equal(c.to_array(), enc.encode_const(-1.25, scale, ctx.moduli[:limbs]))


def test_encode_wide_coefficients():
    """Coefficients past 2^64 (SEAL's 128-bit decomposition) and past 2^128
    (multi-word), a ragged count, an empty vector and the size check."""
    logn = 12
    _note("encode wide logn=%d" % logn)
    ctx = nx.Context(logn, [60, 60, 60, 60, 60])
    enc = orc.Encoder(logn)
    rng = np.random.default_rng(31)
    mods = ctx.moduli[: ctx.L]
    v = rng.uniform(-(2.0 ** 20), 2.0 ** 20, ctx.slot_count) + 1j * rng.uniform(-1, 1, ctx.slot_count)
    pt = nx.encode(ctx, v, 2.0 ** 60)
    assert np.array_equal(pt.to_array(), enc.encode(v, 2.0 ** 60, mods))
    # |coefficients| ~ 2^134: exact big-integer residues of SEAL's unrounded coefficients
    w = rng.uniform(-(2.0 ** 30), 2.0 ** 30, 301)
    c = enc.encode_coeffs(w, 2.0 ** 110)
    assert np.max(np.abs(c)) > 2.0 ** 128
    r = orc.Encoder.residues_exact(c, mods)
    ref = np.stack([orc.ntt_forward(r[j], logn, mods[j]) for j in range(len(mods))])
    assert np.array_equal(nx.encode(ctx, w, 2.0 ** 110).to_array(), ref)
    # empty: the zero polynomial
    assert not nx.encode(ctx, np.zeros(0), 2.0 ** 40).to_array().any()
    # coefficients past the modulus: SEAL's "encoded values are too large"
    with pytest.raises(Exception, match="too large"):
        nx.encode(ctx, np.full(4, 2.0 ** 170), 2.0 ** 100)


def test_encrypt_multiply_rotate_decrypt():
    _note("end-to-end logn=13")
    logn = 13
    ctx = nx.Context(logn, [60, 40, 40, 60])
    sk = nx.SecretKey(ctx, seed=1)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in (1, 2, 4, -1, 0)])
    rng = np.random.default_rng(0)
    scale = 2.0 ** 40
    x = rng.uniform(-1, 1, ctx.slot_count)
    y = rng.uniform(-1, 1, ctx.slot_count)
    cx = nx.encrypt(ctx, sk, nx.encode(ctx, x, scale), seed=1)
    cy = nx.encrypt(ctx, sk, nx.encode(ctx, y, scale), seed=2)
    prod = nx.rescale_to_next(nx.relinearize(nx.multiply(cx, cy), rk))
    out = nx.decode(ctx, nx.decrypt(ctx, sk, prod))
    assert np.max(np.abs(out - x * y)) < 1e-6
    for step in (1, 3, -1):
        r = nx.rotate_vector(cx, step, gk)
        o = nx.decode(ctx, nx.decrypt(ctx, sk, r))
        assert np.max(np.abs(o - np.roll(x, -step))) < 1e-6, step
    cc = nx.complex_conjugate(nx.encrypt(ctx, sk, nx.encode(ctx, x + 0.5j * y, scale)), gk)
    oc = nx.decode(ctx, nx.decrypt(ctx, sk, cc), complex_out=True)
    assert np.max(np.abs(oc - (x - 0.5j * y))) < 1e-6
    mc = nx.rescale_to_next(nx.multiply_const(cx, 0.25))
    assert np.max(np.abs(nx.decode(ctx, nx.decrypt(ctx, sk, mc)) - 0.25 * x)) < 1e-6
    ac = nx.add_const(cx, -0.75)
    assert np.max(np.abs(nx.decode(ctx, nx.decrypt(ctx, sk, ac)) - (x - 0.75))) < 1e-6
