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
