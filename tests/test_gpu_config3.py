"""Batched bootstrapping at the stated configuration: N = 2^16, 64 ciphertexts
with 2^15 slots (dense) and 16 with 2^12 slots (sparse) in one batch, input
levels 0-2, values ~U[-1, 1], key switching with 3 digits (31 data primes, 11
special primes), K = 25, EvalMod degree 59 with two double angles.  Every
CoeffToSlot / SlotToCoeff level runs once per chunk of inputs of one
bootstrapper and one EvalMod covers all 144 halves
(cuda/src/bootstrapping/Bootstrapper.cu:2987-3113 bootstraps one at a time).
Each output decrypts to its message, and sampled outputs match their own
single bootstrap within the key-switching rounding of the fused products."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

nx = pytest.importorskip("nexus_amd")

CHAIN = [51] + [46] * 16 + [51] * 14 + [51] * 11
SPECIAL = 11
SCALE = 2.0 ** 46


def _note(msg):
    print("[case]", msg, flush=True)
    sys.stdout.flush()


def test_batched_bootstrap_config3():
    if nx.device_count() < 1:
        pytest.skip("no GPU visible")
    _note("config 3: keys N=2^16, 31 data primes, 11 special primes")
    ctx = nx.Context(16, CHAIN, hamming_weight=192, special_primes=SPECIAL)
    assert ctx.digits == 3
    sk = nx.SecretKey(ctx, seed=301)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=302)
    steps = sorted(set(nx.boot_plan_steps(15, 15, 25)) | set(nx.boot_plan_steps(15, 12, 25)))
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=303)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    dense = nx.Bootstrapper(ev, log_slots=15)
    sparse = nx.Bootstrapper(ev, log_slots=12)
    rng = np.random.default_rng(304)
    Nh = ctx.slot_count
    cts, bts, msgs = [], [], []
    for i in range(80):
        n = Nh if i < 64 else 1 << 12
        v = rng.uniform(-1, 1, n)
        c = nx.encrypt(ctx, sk, nx.encode(ctx, np.tile(v, Nh // n), SCALE), seed=400 + i)
        cts.append(nx.mod_switch_to(c, ctx.L - (i % 3)))  # input levels 0, 1, 2
        bts.append(dense if i < 64 else sparse)
        msgs.append(v)
    order = rng.permutation(80)  # dense and sparse interleaved in the batch
    _note("config 3: 64 dense + 16 sparse bootstraps in one batch")
    nx.Bootstrapper.bootstrap_mixed([bts[i] for i in order[:4]], [cts[i] for i in order[:4]])  # diagonals encoded
    ctx.synchronize()
    t0 = time.perf_counter()
    outs = nx.Bootstrapper.bootstrap_mixed([bts[i] for i in order], [cts[i] for i in order])
    ctx.synchronize()
    dt = time.perf_counter() - t0
    print("config 3: %d bootstraps in %.2f s (%.1f ms each)" % (len(outs), dt, dt / len(outs) * 1e3))
    worst = 0.0
    for k, i in enumerate(order):
        d = nx.decode(ctx, nx.decrypt(ctx, sk, outs[k]))[:len(msgs[i])]
        err = float(np.max(np.abs(d - msgs[i])))
        worst = max(worst, err)
        assert outs[k].chain_index == outs[0].chain_index
    print("config 3: max |error| %.3g" % worst)
    assert worst < 1e-3
    for k in (0, int(np.where(order >= 64)[0][0])):  # a dense and a sparse one
        i = order[k]
        single = bts[i].bootstrap(cts[i])
        a = nx.decode(ctx, nx.decrypt(ctx, sk, single))[:len(msgs[i])]
        b = nx.decode(ctx, nx.decrypt(ctx, sk, outs[k]))[:len(msgs[i])]
        assert np.max(np.abs(a - b)) < 1e-5
