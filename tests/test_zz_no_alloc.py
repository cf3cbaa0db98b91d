"""The allocation-free rescale (nx_rescale_to_next_noalloc) against the
allocating one (nx_rescale_to_next), word for word, through the C ABI, on two
calls in a row (the second on a smaller input, so the shared scratch array
holds the first call's contents), with caller-owned device memory; inputs
beyond its 32 MiB scratch are rejected."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


def _rand(rng, moduli, size, limbs, n):
    return np.stack([np.stack([rng.integers(0, moduli[i], n, dtype=np.uint64) for i in range(limbs)])
                     for _ in range(size)])


def test_rescale_noalloc_matches():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    ctx = nx.Context(16, [58] + [40] * 18 + [58])
    rng = np.random.default_rng(91)
    for polys, chain_index in ((3, 1), (2, 14)):  # 3 x 19 limbs (3.7 Mi words), then 2 x 6 limbs
        limbs = ctx.limbs_of(chain_index)
        _note("rescale noalloc polys=%d limbs=%d" % (polys, limbs))
        data = _rand(rng, ctx.moduli, polys, limbs, ctx.n)
        src = nx.Ciphertext.from_array(ctx, data, chain_index, 2.0 ** 40)
        want = nx.rescale_to_next(src).to_array()
        # caller-owned output memory: a ciphertext of the result's shape
        dst = nx.Ciphertext.from_array(ctx, np.zeros((polys, limbs - 1, ctx.n), dtype=np.uint64), chain_index + 1,
                                       2.0 ** 40)
        (pin, win), (pout, wout) = src.device_data(), dst.device_data()
        assert win == polys * limbs * ctx.n and wout == polys * (limbs - 1) * ctx.n
        nx.rescale_to_next_noalloc(ctx, pin, polys, limbs, pout)
        assert np.array_equal(dst.to_array(), want)
    # beyond the 32 MiB scratch: 3 x 22 limbs at N = 2^16 is rejected before any launch
    big = nx.Context(16, [58] + [40] * 21 + [58])
    a = nx.Ciphertext.from_array(big, np.zeros((3, 22, big.n), dtype=np.uint64), 1, 2.0 ** 40)
    b = nx.Ciphertext.from_array(big, np.zeros((3, 21, big.n), dtype=np.uint64), 2, 2.0 ** 40)
    with pytest.raises(nx.NexusError, match="scratch"):
        nx.rescale_to_next_noalloc(big, a.device_data()[0], 3, 22, b.device_data()[0])
