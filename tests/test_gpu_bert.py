"""Encrypted BERT layer on the GPU against its FP64 model (oracle/bert.py).

Small ring: N = 2^12 (8 tokens per ciphertext in 256-slot token groups),
hidden 256 (4 heads of 64), FFN 512, chain {51, 46 x 19, 51 x 14, 58 x 3}
with three special primes, scale 2^46, dense bootstrapping.  Each stage
(plaintext-weight linear map, attention with its internal bootstrap,
LayerNorm) and the whole layer are compared with the float64 evaluation of
the same approximations; the tolerance is 2^-10 absolute for the layer, far
tighter for the linear maps (one rescale of CKKS noise).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")
from oracle import approx  # noqa: E402
from oracle import bert as ob  # noqa: E402

LOGN = 12
CHAIN = [51] + [46] * 19 + [51] * 14 + [58] * 3
SCALE = 2.0 ** 46
HIDDEN, HEADS, FFN = 256, 4, 512


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


class Env:
    def __init__(self):
        _note("bert keys logN=%d L=%d" % (LOGN, len(CHAIN) - 3))
        self.ctx = nx.Context(LOGN, CHAIN, hamming_weight=64, special_primes=3)
        self.slots = self.ctx.slot_count
        self.T = self.slots // nx.BERT_GROUP
        self.sk = nx.SecretKey(self.ctx, seed=41)
        self.rk = nx.KSwitchKeys(self.ctx).generate_relin(self.sk, seed=42)
        steps = sorted(set(nx.bert_rotation_steps(self.slots, HIDDEN, HEADS, FFN))
                       | set(nx.boot_plan_steps(LOGN - 1, LOGN - 1, 25)))
        self.gk = nx.KSwitchKeys(self.ctx).generate_galois(self.sk, [self.ctx.elt_from_step(s) for s in steps],
                                                            seed=43)
        self.ev = nx.CKKSEvaluator(self.ctx, self.rk, self.gk, SCALE, self.sk)
        self.bt = nx.Bootstrapper(self.ev)
        self.W = ob.random_weights(HIDDEN, FFN, seed=44)
        self.layer = nx.BertLayer(self.bt, self.W)

    def enc(self, x, chain_index=1):
        return [nx.encrypt(self.ctx, self.sk, nx.encode(self.ctx, v, SCALE, chain_index), seed=100 + i)
                for i, v in enumerate(nx.bert_pack(x, self.slots))]

    def dec(self, cts, tokens):
        return nx.bert_unpack([nx.decode(self.ctx, nx.decrypt(self.ctx, self.sk, c)) for c in cts], tokens)


@pytest.fixture(scope="module")
def env():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    return Env()


def test_linear_maps(env):
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (env.T, HIDDEN))
    for name, factor in (("wq", 1.0), ("wo", 0.125), ("w1", 0.25)):
        _note("bert linear %s" % name)
        y = env.layer.linear(name, env.enc(x, chain_index=len(CHAIN) - 3 - 1), factor)
        assert all(c.chain_index == len(CHAIN) - 3 for c in y)  # one limb left
        got = env.dec(y, env.T)
        err = np.max(np.abs(got - factor * x @ env.W[name]))
        print("linear %s: max err %.3g" % (name, err))
        assert err < 1e-5
    _note("bert linear w2")
    h = rng.normal(0, 1, (env.T, FFN))
    got = env.dec(env.layer.linear("w2", env.enc(h), 1.0), env.T)
    assert np.max(np.abs(got - h @ env.W["w2"])) < 1e-5


def _attention_model(q, k, v, s):
    o = np.zeros_like(q)
    for h in range(HEADS):
        sl = slice(64 * h, 64 * (h + 1))
        o[:, sl] = ob._softmax_rows_approx(q[:, sl] @ k[:, sl].T, s) @ v[:, sl]
    return o


@pytest.mark.parametrize("s", [8, 5])
def test_attention(env, s):
    _note("bert attention s=%d" % s)
    rng = np.random.default_rng(2 + s)
    q = np.zeros((env.T, HIDDEN))
    k = np.zeros((env.T, HIDDEN))
    v = np.zeros((env.T, HIDDEN))
    q[:s] = rng.normal(0, 0.2, (s, HIDDEN))
    k[:s] = rng.normal(0, 0.2, (s, HIDDEN))
    v[:s] = rng.normal(0, 1, (s, HIDDEN))
    top = 1 + 14  # the bootstrapper's output level
    o = env.layer.attention(env.enc(q, top), env.enc(k, top), env.enc(v, top), s)
    got = env.dec(o, env.T)
    want = _attention_model(q[:s], k[:s], v[:s], s)
    err = np.max(np.abs(got[:s] - want))
    print("attention s=%d: max err %.3g, padded rows %.3g" % (s, err, np.max(np.abs(got[s:])) if s < env.T else 0))
    assert err < 2.0 ** -10
    if s < env.T:
        assert np.max(np.abs(got[s:])) < 2.0 ** -10


def test_layer_norm(env):
    _note("bert layer_norm")
    rng = np.random.default_rng(3)
    r = rng.normal(0, 1, (env.T, HIDDEN)) * np.sqrt(30.0)
    got = env.dec(env.layer.layer_norm(env.enc(r, 1 + 14)), env.T)
    m = np.sum(r * r, axis=1, keepdims=True) / HIDDEN
    want = r * approx.invert_sqrt(m, 4, 2)
    err = np.max(np.abs(got - want))
    print("layer_norm: max err %.3g" % err)
    assert err < 2.0 ** -10


@pytest.mark.parametrize("s", [8, 6])
def test_layer_forward(env, s):
    _note("bert layer forward s=%d" % s)
    rng = np.random.default_rng(10 + s)
    x = np.zeros((env.T, HIDDEN))
    x[:s] = rng.normal(0, 1, (s, HIDDEN))
    # the layer's input arrives at two limbs, as a previous layer leaves it
    out = env.layer.forward(env.enc(x, chain_index=len(CHAIN) - 3 - 1), s)
    assert all(c.chain_index == len(CHAIN) - 3 - 1 for c in out)
    got = env.dec(out, env.T)[:s]
    want = ob.layer_approx(x[:s], env.W)
    exact = ob.layer_exact(x[:s], env.W)
    err = np.max(np.abs(got - want))
    print("layer s=%d: max |enc - fp64 model| %.3g mean %.3g; max |enc - exact| %.3g"
          % (s, err, np.mean(np.abs(got - want)), np.max(np.abs(got - exact))))
    assert err < 2.0 ** -10


def test_layer_forward_per_op(env):
    """the plain per-call drop-in (per-op mode: single-ciphertext operations
    only) computes the same layer: within 2^-10 of the FP64 model and within
    noise of the batched program's output"""
    s = 8
    _note("bert layer forward per-op s=%d" % s)
    rng = np.random.default_rng(10 + s)
    x = np.zeros((env.T, HIDDEN))
    x[:s] = rng.normal(0, 1, (s, HIDDEN))
    enc = env.enc(x, chain_index=len(CHAIN) - 3 - 1)
    batched = env.dec(env.layer.forward(enc, s), env.T)[:s]
    env.ev.set_per_op(True)
    try:
        got = env.dec(env.layer.forward(enc, s), env.T)[:s]
    finally:
        env.ev.set_per_op(False)
    want = ob.layer_approx(x[:s], env.W)
    err = np.max(np.abs(got - want))
    print("layer per-op: max |enc - fp64 model| %.3g, max |per-op - batched| %.3g"
          % (err, np.max(np.abs(got - batched))))
    assert err < 2.0 ** -10
    assert np.max(np.abs(got - batched)) < 2.0 ** -12


def test_layer_packed_sequences(env):
    """three sequences packed in one set: each equals its own layer"""
    segs = [(0, 3), (3, 2), (5, 3)]
    _note("bert layer packed %s" % segs)
    rng = np.random.default_rng(30)
    x = np.zeros((env.T, HIDDEN))
    x[:8] = rng.normal(0, 1, (8, HIDDEN))
    out = env.layer.forward(env.enc(x, chain_index=len(CHAIN) - 3 - 1), segments=segs)
    got = env.dec(out, env.T)
    for a, n in segs:
        want = ob.layer_approx(x[a:a + n], env.W)
        err = np.max(np.abs(got[a:a + n] - want))
        print("packed segment (%d, %d): max err %.3g" % (a, n, err))
        assert err < 2.0 ** -10


def test_server_continuous_batching(env):
    """BERT requests queued together are packed into shared layer calls; each
    result matches its own solo layer call and the FP64 model"""
    lens = [3, 2, 3, 4, 4]
    _note("bert server lens=%s" % lens)
    rng = np.random.default_rng(40)
    xs = [rng.normal(0, 1, (n, HIDDEN)) for n in lens]
    two = len(CHAIN) - 3 - 1
    encs = []
    for k, n in enumerate(lens):
        rows = np.zeros((env.T, HIDDEN))
        rows[:n] = xs[k]
        encs.append(env.enc(rows, chain_index=two))
    srv = nx.Server(env.ev, lanes=2, bert=env.layer, pack=True)
    srv.pause(True)
    ids = [srv.submit_bert(c, n) for c, n in zip(encs, lens)]
    srv.pause(False)
    outs = [srv.wait(i) for i in ids]
    st = srv.batching_stats()
    del srv
    print("server: %s" % st)
    assert st["packed_calls"] >= 1 and st["packed_requests"] >= 2
    for k, n in enumerate(lens):
        got = env.dec(outs[k], env.T)
        solo = env.dec(env.layer.forward(encs[k], n), env.T)[:n]
        want = ob.layer_approx(xs[k], env.W)
        e_model = np.max(np.abs(got[:n] - want))
        e_solo = np.max(np.abs(got[:n] - solo))
        print("request %d (len %d): max err vs model %.3g, vs solo call %.3g, other rows %.3g"
              % (k, n, e_model, e_solo, np.max(np.abs(got[n:]))))
        assert e_model < 2.0 ** -10 and e_solo < 2.0 ** -10
        assert np.max(np.abs(got[n:])) < 2.0 ** -10


def test_forward_sets_lockstep(env):
    """three packed sets in lockstep (every bootstrapping step over all sets,
    real messages paired across sets): each sequence within 2^-10 of the FP64
    model and of its set's own layer call"""
    segs = [[(0, 3), (3, 4)], [(0, 8)], [(1, 2), (4, 3)]]
    _note("bert forward_sets %s" % segs)
    rng = np.random.default_rng(60)
    two = len(CHAIN) - 3 - 1
    xs, sets = [], []
    for g in segs:
        x = np.zeros((env.T, HIDDEN))
        for a, n in g:
            x[a:a + n] = rng.normal(0, 1, (n, HIDDEN))
        xs.append(x)
        sets.append(env.enc(x, chain_index=two))
    outs = env.layer.forward_sets(sets, segs)
    assert len(outs) == 3 and all(len(o) == len(s) for o, s in zip(outs, sets))
    for x, g, s, o in zip(xs, segs, sets, outs):
        got = env.dec(o, env.T)
        alone = env.dec(env.layer.forward(s, segments=g), env.T)
        for a, n in g:
            want = ob.layer_approx(x[a:a + n], env.W)
            e_model = np.max(np.abs(got[a:a + n] - want))
            e_alone = np.max(np.abs(got[a:a + n] - alone[a:a + n]))
            print("set segment (%d, %d): max err vs model %.3g, vs its own call %.3g" % (a, n, e_model, e_alone))
            assert e_model < 2.0 ** -10 and e_alone < 2.0 ** -10


def test_server_packed_set(env):
    """a set the caller packed, served as one request on a lane: word for word
    the direct layer call over the same segments; two such sets run at once"""
    segs = [(0, 3), (3, 4)]
    _note("bert server packed set %s" % segs)
    rng = np.random.default_rng(50)
    two = len(CHAIN) - 3 - 1
    sets = []
    for _ in range(2):
        x = np.zeros((env.T, HIDDEN))
        x[:7] = rng.normal(0, 1, (7, HIDDEN))
        sets.append(env.enc(x, chain_index=two))
    srv = nx.Server(env.ev, lanes=2, bert=env.layer, pack=False)
    srv.set_bert_concurrency(2)
    ids = [srv.submit_bert_set(s, segs) for s in sets]
    outs = [srv.wait(i) for i in ids]
    del srv
    for s, o in zip(sets, outs):
        direct = env.layer.forward(s, segments=segs)
        assert len(o) == len(direct)
        for a, b in zip(o, direct):
            assert np.array_equal(a.to_array(), b.to_array())
    with pytest.raises(nx.NexusError):
        srv = nx.Server(env.ev, lanes=1, bert=env.layer, pack=False)
        srv.submit_bert_set(sets[0], [(0, 3), (2, 3)])  # overlapping rows


def test_bert_base_full_size():
    """BERT-base (hidden 768, 12 heads, FFN 3072) at N = 2^16 with 128-token
    sets, chain {51, 46 x 19, 51 x 14, 51 x 8}: a 40-token sequence, two packed
    sequences, and four uneven requests served with continuous batching (two
    calls in flight), each against the FP64 model (and its solo call) within
    2^-10"""
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("bert-base full size keys")
    chain = [51] + [46] * 19 + [51] * 14 + [51] * 8
    ctx = nx.Context(16, chain, hamming_weight=192, special_primes=8)
    sk = nx.SecretKey(ctx, seed=81)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=82)
    steps = sorted(set(nx.bert_rotation_steps(ctx.slot_count)) | set(nx.boot_plan_steps(15, 15, 25)))
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=83)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    bt = nx.Bootstrapper(ev)
    W = ob.random_weights(768, 3072, seed=84)
    layer = nx.BertLayer(bt, W)
    rng = np.random.default_rng(85)
    x = np.zeros((128, 768))
    x[:40] = rng.normal(0, 1, (40, 768))
    x[40:100] = rng.normal(0, 1, (60, 768))
    enc = [nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE, ctx.L - 1), seed=90 + i)
           for i, v in enumerate(nx.bert_pack(x, ctx.slot_count))]
    dec = lambda cts: nx.bert_unpack([nx.decode(ctx, nx.decrypt(ctx, sk, c)) for c in cts], 128)
    _note("bert-base full size: one sequence of 40")
    single = [nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE, ctx.L - 1), seed=95 + i)
              for i, v in enumerate(nx.bert_pack(np.vstack([x[:40], np.zeros((88, 768))]), ctx.slot_count))]
    got = dec(layer.forward(single, 40))[:40]
    err = np.max(np.abs(got - ob.layer_approx(x[:40], W)))
    print("bert-base s=40: max err vs model %.3g" % err)
    assert err < 2.0 ** -10
    _note("bert-base full size: packed 40 + 60")
    got = dec(layer.forward(enc, segments=[(0, 40), (40, 60)]))
    for a, n in ((0, 40), (40, 60)):
        e = np.max(np.abs(got[a:a + n] - ob.layer_approx(x[a:a + n], W)))
        print("bert-base packed (%d, %d): max err vs model %.3g" % (a, n, e))
        assert e < 2.0 ** -10

    # served (continuous batching, two BERT calls in flight on two lanes): four
    # requests of uneven length, each within 2^-10 of its solo call and the model
    lens = [17, 45, 30, 64]
    _note("bert-base served: requests of lengths %s" % lens)
    reqs, xs = [], []
    for k, n in enumerate(lens):
        rows = np.zeros((128, 768))
        rows[:n] = rng.normal(0, 1, (n, 768))
        xs.append(rows[:n])
        reqs.append([nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE, ctx.L - 1), seed=200 + 10 * k + i)
                     for i, v in enumerate(nx.bert_pack(rows, ctx.slot_count))])
    srv = nx.Server(ev, lanes=2, bert=layer, pack=True)
    srv.set_bert_concurrency(2)
    srv.pause(True)
    ids = [srv.submit_bert(c, n) for c, n in zip(reqs, lens)]
    srv.pause(False)
    outs = [srv.wait(i) for i in ids]
    st = srv.batching_stats()
    del srv
    print("bert-base served: %s" % st)
    for k, n in enumerate(lens):
        got = dec(outs[k])
        solo = dec(layer.forward(reqs[k], n))[:n]
        e_model = np.max(np.abs(got[:n] - ob.layer_approx(xs[k], W)))
        e_solo = np.max(np.abs(got[:n] - solo))
        print("served request %d (len %d): max err vs model %.3g, vs solo %.3g" % (k, n, e_model, e_solo))
        assert e_model < 2.0 ** -10 and e_solo < 2.0 ** -10
        if st["packed_requests"] == len(lens):  # rows past a packed request are cleared
            assert np.max(np.abs(got[n:])) < 2.0 ** -10
