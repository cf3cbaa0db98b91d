"""Launch-shape autotuner: after tuning, the NTT and the fused
multiply-accumulate still agree bit-exactly with the oracle / with the plain
multiply-and-add path (the tuner only keeps shapes whose output equals the
default's on its own data)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
import oracle as orc  # noqa: E402


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("logn,chain", [(14, [58, 40, 40, 58]), (16, [51] + [46] * 19 + [58] * 4)])
def test_autotune_then_parity(logn, chain):
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("autotune logn=%d primes=%d" % (logn, len(chain)))
    ctx = nx.Context(logn, chain)
    rep = nx.autotune(ctx, reps=3)
    print(rep)
    assert rep["logn"] == logn
    assert len(rep["ntt"]) == 20  # 2 directions x batches 1..512
    for e in rep["ntt"]:
        assert e["col"] in (0, 4, 8, 16) and e["row"] in (0, 4, 8)
        assert e["radix16"][0] in (0, 1) and e["radix16"][1] in (0, 1)
        assert e["radix16_matched"] == [1, 1]  # radix-16 passes give the default's words
    assert rep["dot_plain"] and all(d["layout"] in (0, 1) for d in rep["dot_plain"])
    # NTT round trip and forward parity with the oracle under the tuned shapes
    rng = np.random.default_rng(logn)
    L = ctx.L
    data = np.stack([np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in ctx.moduli[:L]])
                     for _ in range(2)])
    ct = nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40, ntt_form=False)
    fwd = nx.transform_to_ntt(ct).to_array()
    for i in range(0, L, max(1, L // 4)):
        assert np.array_equal(fwd[1, i], orc.ntt_forward(data[1, i], logn, ctx.moduli[i]))
    back = nx.transform_from_ntt(nx.transform_to_ntt(ct)).to_array()
    assert np.array_equal(back, data)
    # fused multiply-accumulate vs multiply_plain + add at several levels
    for ci in (1, L - 1):
        limbs = L + 1 - ci
        mods = ctx.moduli[:limbs]
        cts = [nx.Ciphertext.from_array(ctx, np.stack([np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64)
                                                                for q in mods]) for _ in range(2)]), ci, 2.0 ** 40)
               for _ in range(5)]
        pts = [nx.Plaintext.from_array(ctx, np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in mods]), ci,
                                       2.0 ** 40) for _ in range(5)]
        got = nx.multiply_plain_accumulate(cts, pts).to_array()
        want = nx.multiply_plain(cts[0], pts[0])
        for c, p in zip(cts[1:], pts[1:]):
            want = nx.add(want, nx.multiply_plain(c, p))
        assert np.array_equal(got, want.to_array())


def test_autotune_modup_then_parity():
    """hybrid context with every prime below 2^51: the tuner times the ModUp and
    ModDown conversions fused into the column pass against the separate
    launches (each candidate checked word for word), and relinearisation stays
    bit-exact under its choices"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    from oracle import hybrid
    logn, chain, alpha = 14, [51] + [46] * 9 + [51] * 4, 4
    _note("autotune hybrid ModUp logn=%d alpha=%d" % (logn, alpha))
    ctx = nx.Context(logn, chain, special_primes=alpha)
    rep = nx.autotune(ctx, reps=3)
    print(rep["modup"])
    assert [e["batch"] for e in rep["modup"]] == [1, 2, 4, 8, 16, 32]
    for e in rep["modup"]:
        assert e["mode"] in (-1, 0, 4, 8, 12, 64, 256, 264, 268, 512, 520, 524)
        assert all(t > 0 for t in e["fused_us"])  # every fused candidate equalled the unfused words
    assert [e["batch"] for e in rep["moddown"]] == [1, 2, 4, 8, 16]
    for e in rep["moddown"]:
        assert e["mode"] in (-1, 0, 4, 8, 12, 64, 256, 264, 268, 512, 520, 524)
        assert all(t > 0 for t in e["fused_us"])
    L = ctx.L
    sk = nx.SecretKey(ctx, seed=71)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=72)
    rng = np.random.default_rng(73)
    data = np.stack([np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in ctx.moduli[:L]])
                     for _ in range(3)])
    got = nx.relinearize(nx.Ciphertext.from_array(ctx, data, 1, 2.0 ** 40), rk).to_array()
    ks0, ks1 = hybrid.switch_key_hybrid(data[2], rk.key_array(0), ctx.moduli, L, alpha, logn)
    mods = np.array(ctx.moduli[:L], dtype=object)[:, None]
    assert np.array_equal(got[0], ((data[0].astype(object) + ks0.astype(object)) % mods).astype(np.uint64))
    assert np.array_equal(got[1], ((data[1].astype(object) + ks1.astype(object)) % mods).astype(np.uint64))
