"""Lanes and the request loop.

* Two threads on two lanes of one context compute the same results, residue for
  residue, as one thread on the default stream.
* Requests served concurrently (several lanes, memory-budget admission) give,
  for every request, the residues the plain per-call evaluator gives alone.
"""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")

CHAIN = [58] + [40] * 18 + [58]
SCALE = 2.0 ** 40


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("server keys N=2^14 L=19")
    ctx = nx.Context(14, CHAIN)
    sk = nx.SecretKey(ctx, seed=21)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=22)
    steps = [-1024, -128] + [1 << i for i in range(10)]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=23)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    return ctx, sk, ev


def _inputs(ctx, sk, seed):
    rng = np.random.default_rng(seed)
    g = np.clip(rng.normal(0, 1, ctx.slot_count), -5, 5)
    s = np.zeros(ctx.slot_count)
    for r in range(ctx.slot_count // 256):
        row = np.clip(rng.normal(0, 3, 128), -10, 10)
        s[256 * r: 256 * r + 128] = np.clip(row - row.max(), -10, 0)
    ln = np.zeros(ctx.slot_count)
    for r in range(ctx.slot_count // 2048):
        ln[2048 * r: 2048 * r + 768] = rng.normal(0, 1, 768) * np.sqrt(30.0)
    enc = lambda v, k: nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=seed * 10 + k)
    return enc(g, 1), enc(s, 2), enc(ln, 3)


def test_two_lanes_match_default_stream(env):
    ctx, sk, ev = env
    _note("two lanes vs default stream")
    a, b, _ = _inputs(ctx, sk, 1)
    ref_a = ev.gelu(a).to_array()
    ref_b = ev.softmax(b, 128).to_array()
    out = {}

    def run(lane, fn, key):
        ctx.bind_lane(lane)
        try:
            r = fn()
            ctx.synchronize()
            out[key] = r.to_array()
        finally:
            ctx.bind_lane(0)

    t1 = threading.Thread(target=run, args=(1, lambda: ev.gelu(a), "a"))
    t2 = threading.Thread(target=run, args=(2, lambda: ev.softmax(b, 128), "b"))
    t1.start()
    t2.start()
    t1.join()
    t2.join()
    assert np.array_equal(out["a"], ref_a)
    assert np.array_equal(out["b"], ref_b)


@pytest.mark.parametrize("lanes,budget", [(1, 0), (3, 0), (3, 1 << 28)])
def test_served_requests_match_alone(env, lanes, budget):
    ctx, sk, ev = env
    _note("server lanes=%d budget=%d" % (lanes, budget))
    reqs = []
    for i in range(6):
        g, s, ln = _inputs(ctx, sk, 100 + i)
        jobs = [(nx.JOB_GELU, 0, g), (nx.JOB_SOFTMAX, 128, s), (nx.JOB_LAYER_NORM, 1024, ln)][: 1 + i % 3]
        reqs.append(jobs)
    alone = []
    for jobs in reqs:
        res = []
        for kind, ln_, ct in jobs:
            if kind == nx.JOB_GELU:
                res.append(ev.gelu(ct).to_array())
            elif kind == nx.JOB_SOFTMAX:
                res.append(ev.softmax(ct, ln_).to_array())
            else:
                res.append(ev.layer_norm(ct, ln_).to_array())
        alone.append(res)
    srv = nx.Server(ev, lanes=lanes, memory_budget=budget)
    ids = [srv.submit(jobs) for jobs in reqs]
    for rid, ref in zip(ids, alone):
        outs = srv.wait(rid)
        assert len(outs) == len(ref)
        for o, r in zip(outs, ref):
            assert np.array_equal(o.to_array(), r)
    st = srv.stats()
    print("stats", st)
    assert st["completed"] == len(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["peak_concurrent"] <= lanes
    if lanes > 1 and budget == 0:
        assert st["peak_concurrent"] > 1
