"""The performance-faithful CPU port (oracle/cpu_port.c: lazy Harvey NTT,
Shoup and Barrett products, SEAL-form key switching), whose timings are the
CPU baseline, equals the plain oracle (oracle/ckks_oracle.c, pinned by SEAL's
KATs and closed forms) word for word: NTTs, key switch, rescale, ct x pt and
ct x ct products, additions and the encoder's residues."""
import numpy as np
import pytest

import oracle as orc

LOGN = 10
N = 1 << LOGN


@pytest.fixture(scope="module", params=[[58, 40, 40, 40, 58], [51, 46, 46, 51, 51], [60, 50, 60]],
                ids=["58-40", "51-46", "60-50"])
def env(request):
    mods = orc.coeff_modulus_create(N, request.param)
    return mods, orc.Port(LOGN, mods)


def _rand(rng, mods, shape_polys, l, bound=1):
    return np.stack([np.stack([rng.integers(0, bound * mods[i], N, dtype=np.uint64) for i in range(l)])
                     for _ in range(shape_polys)])


def test_ntt(env):
    mods, port = env
    rng = np.random.default_rng(1)
    for i, q in enumerate(mods):
        a = rng.integers(0, q, N, dtype=np.uint64)
        a[:3] = [0, q - 1, 1]
        f = port.ntt_forward(a, i)
        assert np.array_equal(f, orc.ntt_forward(a, LOGN, q))
        assert np.array_equal(port.ntt_inverse(f, i), a)
        assert np.array_equal(port.ntt_inverse(a, i), orc.ntt_inverse(a, LOGN, q))


def test_switch_key(env):
    mods, port = env
    rng = np.random.default_rng(2)
    L = len(mods) - 1
    key = np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in mods]) for _ in range(2)])
                    for _ in range(L)])
    for l in range(1, L + 1):
        ct = _rand(rng, mods, 2, l)
        target = _rand(rng, mods, 1, l)[0]
        assert np.array_equal(port.switch_key(ct, target, key), orc.switch_key(ct, target, mods, key, LOGN)), l


def test_rescale_products_add(env):
    mods, port = env
    rng = np.random.default_rng(3)
    L = len(mods) - 1
    for l in range(2, L + 1):
        p = _rand(rng, mods, 1, l)[0]
        assert np.array_equal(port.rescale(p), orc.rescale_poly_ntt(p, mods, LOGN)), l
    a = _rand(rng, mods, 2, L)
    b = _rand(rng, mods, 2, L)
    pt = _rand(rng, mods, 1, L)[0]
    assert np.array_equal(port.multiply_plain(a, pt), orc.dyadic_mul(a, pt, mods, LOGN))
    assert np.array_equal(port.tensor(a, b), orc.tensor(a, b, mods, LOGN))
    qs = np.array(mods[:L], dtype=object)[None, :, None]
    assert np.array_equal(port.add(a, b), ((a.astype(object) + b.astype(object)) % qs).astype(np.uint64))


def test_encode_residues(env):
    mods, port = env
    rng = np.random.default_rng(4)
    L = len(mods) - 1
    enc = orc.Encoder(LOGN)
    v = rng.uniform(-1, 1, N // 2)
    coeffs = enc.encode_coeffs(v, 2.0 ** 40)
    assert np.array_equal(port.encode_residues(coeffs, L), enc.encode(v, 2.0 ** 40, mods[:L]))
