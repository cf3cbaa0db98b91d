"""Host-side pieces of the encrypted BERT layer: token-group packing, the
layer's rotation steps, and the FP64 model (oracle/bert.py) against the exact
layer -- the approximations' own error at BERT-base size."""
import numpy as np

import nexus_amd as nx
from oracle import approx
from oracle import bert as ob


def test_pack_round_trip():
    rng = np.random.default_rng(0)
    x = rng.normal(0, 1, (5, 768))
    vecs = nx.bert_pack(x, 2048)
    assert len(vecs) == 3
    # token i, feature f at slot 256 i + (f mod 256) of ciphertext f // 256
    assert vecs[1][256 * 3 + 7] == x[3, 256 + 7]
    assert np.array_equal(nx.bert_unpack(vecs, 5), x)
    assert not np.any(vecs[0][5 * 256:])


def test_rotation_steps_cover_the_program():
    for slots in (2048, 32768):
        steps = set(nx.bert_rotation_steps(slots))
        T = slots // 256
        tb = 1
        while tb * tb < T:
            tb *= 2
        want = set(range(1, 32)) | {32 * a for a in range(-8, 8) if a} | {256 * b for b in range(1, min(tb, T))}
        want |= {256 * tb * a for a in range(1, (T + tb - 1) // tb)}
        want |= {8 * a for a in range(-8, 8) if a} | {1 << k for k in range(8)} | {-(1 << k) for k in range(8)}
        norm = set()
        for w in want:
            w %= slots
            norm.add(w - slots if w > slots // 2 else w)
        assert norm - {0} == steps


def test_model_within_reference_approximation_error():
    W = ob.random_weights(768, 3072, seed=1)
    x = np.random.default_rng(2).normal(0, 1, (48, 768))
    a = ob.layer_approx(x, W)
    e = ob.layer_exact(x, W)
    assert np.max(np.abs(a - e)) < 2.0 ** -10
    # a planted mistake (softmax without the 1/s scaling of the inverse) is caught
    bad = ob._softmax_rows_approx
    try:
        ob._softmax_rows_approx = lambda S, s: approx.exp_approx(S) * approx.inverse(
            np.sum(approx.exp_approx(S), axis=-1, keepdims=True) / s)
        assert np.max(np.abs(ob.layer_approx(x, W) - e)) > 1e-2
    finally:
        ob._softmax_rows_approx = bad
