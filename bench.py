#!/usr/bin/env python3
"""Benchmark of the CKKS evaluation engine on one MI355X.

Workload (one "step"): the nonlinear evaluator block of one encrypted
BERT-base layer at the reference's full nonlinear chain -- N = 2^16, moduli
{58, 40 x 18, 58} (src/main.cpp:27-30), scale 2^40 -- through the plain
per-call drop-in (one evaluator call per ciphertext):
  * GELU on a 128 x 3072 tensor            -> 12 ciphertexts of 32768 slots
  * softmax over 12 x 128 x 128 scores     -> 12 ciphertexts (128 rows, stride 256)
  * 2 x LayerNorm on 128 x 768             -> 16 ciphertexts (16 rows, stride 2048)
Inputs are synthetic: GELU ~N(0,1) clipped to [-5,5]; softmax scores ~N(0,3)
shifted to max 0 and clipped to [-10,0]; LayerNorm ~N(0,1)*sqrt(30) (the range
the reference's inverse-sqrt initial guess is tuned for).  Matrix products and
bootstrapping of the layer are not part of this step.

Prints one JSON line.  The roofline entry is the dominant kernel class of the
step, timed live with HIP events on its launch stream during the timed steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "enc. BERT-base layers/sec + p50/p99 latency, batched vs plain GPU, ≤2^-10 err"
CHAIN = [58] + [40] * 18 + [58]
SCALE = 2.0 ** 40
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_inputs(rng, slots):
    gelu = [np.clip(rng.normal(0, 1, slots), -5, 5) for _ in range(12)]
    sm = []
    for _ in range(12):
        v = np.zeros(slots)
        rows = rng.normal(0, 3, (128, 128))
        rows = np.clip(rows - rows.max(axis=1, keepdims=True), -10, 0)
        for r in range(128):
            v[256 * r: 256 * r + 128] = rows[r]
        sm.append(v)
    ln = []
    for _ in range(16):
        v = np.zeros(slots)
        rows = rng.normal(0, 1, (16, 768)) * np.sqrt(30.0)
        for r in range(16):
            v[2048 * r: 2048 * r + 768] = rows[r]
        ln.append(v)
    return gelu, sm, ln


def cpu_baseline(gpu_ks_bytes, n, budget_s):
    """Time the CPU oracle's SEAL-restated key switch on the host (1 thread) and
    scale it to one step by the step's key-switch work measured on the GPU."""
    import oracle as orc

    logn = 16
    mods = orc.coeff_modulus_create(n, CHAIN)
    L = len(CHAIN) - 1
    rng = np.random.default_rng(1)
    key = np.stack([np.stack([np.stack([rng.integers(0, q, n, dtype=np.uint64) for q in mods])
                              for _ in range(2)]) for _ in range(L)])
    for q in mods:
        orc.ntt_tables(logn, q)
    l = L
    tgt = np.stack([rng.integers(0, q, n, dtype=np.uint64) for q in mods[:l]])
    ct = np.zeros((2, l, n), dtype=np.uint64)
    reps, t0 = 0, time.perf_counter()
    while True:
        orc.switch_key(ct, tgt, mods, key, logn)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 3:
            break
    t_ks = (time.perf_counter() - t0) / reps
    # ks_inner algorithmic bytes of one top-level key switch (l = 19, 20 output primes)
    top_bytes = n * (l + 1) * (24.0 * l + 16.0)
    ks_equiv = gpu_ks_bytes / top_bytes
    step_s = t_ks * ks_equiv
    return {
        "value": 1.0 / step_s,
        "unit": "layer nonlinear blocks/s",
        "cores": 1,
        "kind": "port",
        "sample": ("%d key switch(es) at N=2^16, l=19 on the CPU oracle (oracle/ckks_oracle.c, 1 thread): "
                   "%.2f s each; scaled by the step's key-switch work (%.1f top-level key-switch "
                   "equivalents, from ks_inner bytes); rescale/NTT-only and elementwise work not counted"
                   % (reps, t_ks, ks_equiv)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU, so --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import nexus_amd as nx
    from oracle import approx

    if nx.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    log("setup: context + keys (N=2^16, L=19)")
    ctx = nx.Context(16, CHAIN)
    sk = nx.SecretKey(ctx, seed=11)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=12)
    steps_needed = [-1024, -128] + [1 << i for i in range(10)]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps_needed], seed=13)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    rng = np.random.default_rng(2025)
    g_in, s_in, l_in = make_inputs(rng, ctx.slot_count)
    enc = lambda v: nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=int(rng.integers(1 << 62)))
    g_ct = [enc(v) for v in g_in]
    s_ct = [enc(v) for v in s_in]
    l_ct = [enc(v) for v in l_in]
    ctx.synchronize()

    def step():
        outs = []
        for c in g_ct:
            outs.append(ev.gelu(c))
        for c in s_ct:
            outs.append(ev.softmax(c, 128))
        for c in l_ct:
            outs.append(ev.layer_norm(c, 1024))
        return outs

    # warm-up; the first warm-up step ranks the kernel classes by event-timed duration
    nx.profile_select("*")
    for w in range(max(1, args.warmup)):
        log("warmup step %d" % w)
        step()
        ctx.synchronize()
        if w == 0:
            ranking = sorted(((nx.profile_read(c)[1], c) for c in nx.profile_classes()), reverse=True)
            nx.profile_select("")
    dominant = ranking[0][1]
    log("kernel time ranking (ms): " + ", ".join("%s=%.1f" % (c, t) for t, c in ranking))

    # timed region: only the dominant class carries events
    nx.profile_select(dominant + ",ks_inner")
    ctx.synchronize()
    t0 = time.perf_counter()
    lat = []
    for k in range(args.steps):
        ts = time.perf_counter()
        outs = step()
        ctx.synchronize()
        lat.append(time.perf_counter() - ts)
        log("step %d: %.1f ms" % (k, lat[-1] * 1e3))
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    launches, k_ms, k_bytes = nx.profile_read(dominant)
    _, _, ks_bytes = nx.profile_read("ks_inner")
    nx.profile_select("")
    ms_per_step = elapsed / args.steps * 1e3

    # accuracy of the last step vs the FP64 plaintext model of the same approximations
    def decv(ct, count):
        return nx.decode(ctx, nx.decrypt(ctx, sk, ct))[:count]
    err_g = np.max(np.abs(decv(outs[0], ctx.slot_count) - approx.gelu_approx(g_in[0])))
    sm_model = np.zeros(ctx.slot_count)
    for r in range(128):
        sm_model[256 * r: 256 * r + 128] = approx.softmax_approx(s_in[0][256 * r: 256 * r + 128])
    d = decv(outs[12], ctx.slot_count)
    err_s = max(np.max(np.abs(d[256 * r: 256 * r + 128] - sm_model[256 * r: 256 * r + 128])) for r in range(128))
    d = decv(outs[24], ctx.slot_count)
    err_l = max(np.max(np.abs(d[2048 * r: 2048 * r + 768] - approx.layer_norm_approx(l_in[0][2048 * r: 2048 * r + 768])))
                for r in range(16))
    max_err = float(max(err_g, err_s, err_l))

    avg_launch_ms = k_ms / max(launches, 1)
    achieved = (k_bytes / max(launches, 1)) / (avg_launch_ms * 1e-3) / 1e9 if launches else 0.0
    res = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_step,
        "unit": "layer nonlinear blocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (exact RNS modular arithmetic)",
        "data": "synthetic",
        "max_err_vs_fp64_model": max_err,
        "err_ok": bool(max_err <= 2.0 ** -10),
        "config": {
            "workload": "BERT-base layer nonlinear block, plain per-call drop-in: GELU 128x3072 (12 ct), "
                        "softmax 12x128x128 (12 ct), 2x LayerNorm 128x768 (16 ct); matmuls and bootstrapping "
                        "not included",
            "poly_degree": 1 << 16,
            "coeff_modulus_bits": CHAIN,
            "scale": "2^40",
            "ciphertexts_per_step": len(g_ct) + len(s_ct) + len(l_ct),
        },
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "launches": launches,
            "avg_launch_us": avg_launch_ms * 1e3,
            "kernel_share_of_step": k_ms / (elapsed * 1e3),
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle key switch)")
        res["cpu_baseline"] = cpu_baseline(ks_bytes / args.steps, ctx.n, args.cpu_budget)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
