"""SEAL 4.1 serialization on device objects: container layout, parms_id,
round trips of ciphertexts, plaintexts and key sets, and rejection of
malformed or foreign data.  No SEAL build exists here, so compatibility with
SEAL's own files is pinned by the format restatement (header, member order,
nested data array) and by parms_id = BLAKE2b-256 of the parameter words, not
by SEAL-written fixtures."""
import hashlib
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")

SCALE = 2.0 ** 40


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("serialize keys logN=12")
    ctx = nx.Context(12, [60, 40, 40, 60])
    sk = nx.SecretKey(ctx, seed=3)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=4)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in (1, -1, 5)], seed=5)
    return ctx, sk, rk, gk


def _header(b, off=0):
    magic, hs, major, minor, compr, res, size = struct.unpack_from("<HBBBBHQ", b, off)
    return magic, hs, major, minor, compr, res, size


def test_parms_id_is_blake2b_of_parameters(env):
    ctx = env[0]
    for ci in (0, 1, 2, 3):
        mods = ctx.moduli if ci == 0 else ctx.moduli[: ctx.limbs_of(ci)]
        words = struct.pack("<%dQ" % (len(mods) + 3), 2, ctx.n, *mods, 0)
        d = hashlib.blake2b(words, digest_size=32).digest()
        assert nx.parms_id(ctx, ci) == list(struct.unpack("<4Q", d))


def test_ciphertext_format_and_round_trip(env):
    ctx, sk, _, _ = env
    _note("ciphertext save/load")
    rng = np.random.default_rng(1)
    ct = nx.encrypt(ctx, sk, nx.encode(ctx, rng.uniform(-1, 1, ctx.slot_count), SCALE, 2), seed=9)
    b = nx.save(ct)
    magic, hs, major, minor, compr, res, size = _header(b)
    assert (magic, hs, major, minor, compr, res, size) == (0xA15E, 16, 4, 1, 0, 0, len(b))
    pid = struct.unpack_from("<4Q", b, 16)
    assert list(pid) == nx.parms_id(ctx, 2)
    is_ntt, csize, n, limbs, scale, corr = struct.unpack_from("<BQQQdQ", b, 48)
    assert (is_ntt, csize, n, limbs, scale, corr) == (1, 2, ctx.n, ctx.limbs_of(2), SCALE, 1)
    m2 = _header(b, 89)
    assert m2[0] == 0xA15E and m2[6] == 16 + 8 + 8 * 2 * limbs * ctx.n and 89 + m2[6] == len(b)
    (count,) = struct.unpack_from("<Q", b, 105)
    assert count == 2 * limbs * ctx.n
    data = np.frombuffer(b, dtype=np.uint64, count=count, offset=113).reshape(2, limbs, ctx.n)
    assert np.array_equal(data, ct.to_array())
    back = nx.load(ctx, b, "ciphertext")
    assert back.chain_index == 2 and back.scale == SCALE
    assert np.array_equal(back.to_array(), ct.to_array())


def test_plaintext_round_trip(env):
    ctx = env[0]
    pt = nx.encode(ctx, np.linspace(-1, 1, ctx.slot_count), SCALE)
    b = nx.save(pt)
    assert list(struct.unpack_from("<4Q", b, 16)) == nx.parms_id(ctx, 1)
    back = nx.load(ctx, b, "plaintext")
    assert np.array_equal(back.to_array(), pt.to_array())


def test_key_sets_round_trip_and_work(env):
    ctx, sk, rk, gk = env
    _note("key sets save/load")
    rb = nx.save(rk)
    assert list(struct.unpack_from("<4Q", rb, 16)) == nx.parms_id(ctx, 0)
    assert struct.unpack_from("<Q", rb, 48)[0] == 1
    rk2 = nx.load(ctx, rb, "keys")
    assert np.array_equal(rk2.key_array(0), rk.key_array(0))
    gb = nx.save(gk, galois=True)
    assert struct.unpack_from("<Q", gb, 48)[0] == ctx.n
    gk2 = nx.load(ctx, gb, "keys", galois=True)
    for s in (1, -1, 5):
        e = ctx.elt_from_step(s)
        assert np.array_equal(gk2.key_array(e), gk.key_array(e))
    v = np.random.default_rng(2).uniform(-1, 1, ctx.slot_count)
    ct = nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=1)
    got = nx.decode(ctx, nx.decrypt(ctx, sk, nx.rotate_vector(ct, 5, gk2)))
    assert np.max(np.abs(got - np.roll(v, -5))) < 1e-5


def test_rejects_malformed(env):
    ctx, sk, _, _ = env
    ct = nx.encrypt(ctx, sk, nx.encode(ctx, np.ones(8), SCALE), seed=2)
    b = bytearray(nx.save(ct))
    bad = bytearray(b)
    bad[0] ^= 1
    with pytest.raises(nx.NexusError):
        nx.load(ctx, bytes(bad), "ciphertext")
    with pytest.raises(nx.NexusError):
        nx.load(ctx, bytes(b[:-8]), "ciphertext")
    other = nx.Context(12, [60, 40, 60])
    with pytest.raises(nx.NexusError):
        nx.load(other, bytes(b), "ciphertext")
    bad = bytearray(b)
    struct.pack_into("<Q", bad, 113, (1 << 64) - 1)  # first residue out of range
    with pytest.raises(nx.NexusError):
        nx.load(ctx, bytes(bad), "ciphertext")
