#!/usr/bin/env python3
"""Benchmark of the CKKS evaluation engine on one MI355X.

Workload (one "step"): one encrypted BERT-base encoder layer (hidden 768, 12
heads, FFN 3072, weights ~N(0, 0.02) random-init, no biases, LayerNorm gamma 1
beta 0) over 8 sequences of uneven length ~U{32..128} with inputs ~N(0, 1), at
N = 2^16 with bootstrapping between blocks (chain {51, 46 x 19, 51 x 14, 51 x 8},
8 special primes (5 key-switching digits), Hamming weight 192, scale 2^46).  The batched path packs the
sequences into sets of 128 token rows (first-fit decreasing; attention stays
within each sequence) and runs one layer call per set; the plain path runs
one call per sequence.  `value` is the batched path's sequence-layers per
second; the plain path's rate, per-sequence p50/p99 latency and the decrypted
error against the FP64 model of the same approximations (and against the
exact FP64 layer) are reported beside it.

Extras (after the timed steps): the nonlinear evaluators of one layer through
the plain drop-in at the reference's chain {58, 40 x 18, 58}, their serving
measurement, bootstrapping alone, and the reference's MatrixMul.

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
BERT_CHAIN = [51] + [46] * 19 + [51] * 14 + [51] * 8  # special primes below 2^51: double-precision paths
BERT_SPECIAL = 8
BERT_SCALE = 2.0 ** 46
BERT_TOKENS = 128
SCALE = 2.0 ** 40
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X published double-precision peak
# HBM bytes moved per algorithmic byte, from FETCH_SIZE / WRITE_SIZE counter runs
# of `bench.py --only ntt` / `--only dot` calibrated against a device copy of
# known size (profiles/ntt_fwd_n65536_38limbs_traffic.md,
# profiles/dot_plain_n65536_64terms_traffic.md)
TRAFFIC_PER_ALGORITHMIC_BYTE = {"ntt_fwd_col": 40.38 / 39.93, "ntt_fwd_row": 50.57 / 60.73,
                                "dot_plain": 1930.5 / 1932.5, "ks_modup_col": 1528.2 / 1457.6}
TRAFFIC_PROFILE = {"ntt_fwd_col": "profiles/ntt_fwd_n65536_38limbs_traffic.md",
                   "ntt_fwd_row": "profiles/ntt_fwd_n65536_38limbs_traffic.md",
                   "dot_plain": "profiles/dot_plain_n65536_64terms_traffic.md",
                   "ks_modup_col": "profiles/ks_kernels_n65536_l20_hybrid8.md"}


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


def cpu_baseline(tally, units, unit):
    """The reference's CPU path (SEAL-4.1-bs: one special prime, one digit per
    data prime, single-threaded) timed on this host with oracle/cpu_port.c, the
    port of SEAL's methods (lazy Harvey NTT, Shoup / Barrett products), which
    equals the oracle word for word.  Each operation kind is timed at sampled
    levels of the BERT chain's data primes and fitted in the level (key switch:
    quadratic; the others: linear), then composed with the operation mix that
    the plain per-call path issued for the same sequences (the engine's tally).
    Throughput: independent key switches on every host thread at once, the
    parallel efficiency applied to the single-thread composition."""
    import concurrent.futures

    import oracle as orc

    logn = 16
    n = 1 << logn
    data_bits = BERT_CHAIN[:len(BERT_CHAIN) - BERT_SPECIAL]
    lmax = 28
    mods = orc.coeff_modulus_create(n, data_bits[:lmax] + [51])
    port = orc.Port(logn, mods)
    rng = np.random.default_rng(1)
    qmin = min(mods)
    key = rng.integers(0, qmin, size=(lmax, 2, lmax + 1, n), dtype=np.uint64)
    ct = {}

    def rnd(polys, l):
        k = (polys, l)
        if k not in ct:
            ct[k] = rng.integers(0, qmin, size=(polys, l, n), dtype=np.uint64)
        return ct[k]

    def timed(fn, reps=1):
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps

    enc = orc.Encoder(logn)
    vals = rng.uniform(-1, 1, n // 2)
    t_fft = timed(lambda: enc.encode_coeffs(vals, 2.0 ** 46), 3)
    coeffs = enc.encode_coeffs(vals, 2.0 ** 46)
    samples = {
        "ks": ([6, 16, 28], lambda l: port.switch_key(rnd(2, l), rnd(1, l)[0], key), 1, 2),
        "rescale": ([4, 20], lambda l: [port.rescale(rnd(2, l)[p]) for p in range(2)], 3, 1),
        "tensor": ([4, 20], lambda l: port.tensor(rnd(2, l), rnd(2, l)), 3, 1),
        "mul_plain": ([4, 20], lambda l: port.multiply_plain(rnd(2, l), rnd(1, l)[0]), 3, 1),
        "add": ([4, 20], lambda l: port.add(rnd(2, l), rnd(2, l)), 5, 1),
        "encode": ([4, 20], lambda l: port.encode_residues(coeffs, l), 3, 1),
    }
    fits, measured = {}, {}
    for kind, (ls, fn, reps, deg) in samples.items():
        ts = [timed(lambda: fn(l), reps) for l in ls]
        if kind == "encode":
            ts = [t + t_fft for t in ts]
        fits[kind] = np.polyfit(ls, ts, deg)
        measured[kind] = {str(l): round(t * 1e3, 3) for l, t in zip(ls, ts)}
    single_s = 0.0
    per_kind = {}
    for kind, levels in tally.items():
        base = "ks" if kind.startswith("ks") else kind
        if base not in fits:
            continue
        t = sum(c * max(float(np.polyval(fits[base], l)), 0.0) for l, c in levels.items())
        per_kind[kind] = round(t, 3)
        single_s += t
    threads = max(1, min(16, os.cpu_count() or 1))
    a, b = rnd(2, 16), rnd(1, 16)[0]
    t1 = timed(lambda: port.switch_key(a, b, key))
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda _: port.switch_key(a, b, key), range(threads)))
        tw = time.perf_counter() - t0
    efficiency = min(1.0, t1 / tw)
    return {
        "value": units / single_s * threads * efficiency,
        "unit": unit,
        "cores": threads,
        "kind": "port",
        "single_thread_value": units / single_s,
        "single_thread_seconds_per_step": single_s,
        "parallel_efficiency": efficiency,
        "seconds_by_kind": per_kind,
        "sampled_ms": measured,
        "sample": ("oracle/cpu_port.c (SEAL's CPU methods; word for word the oracle): key switch (one special "
                   "prime, one digit per data prime) at l = 6/16/28, rescale, tensor, ct x pt, add and encode at "
                   "l = 4/20 limbs, N = 2^16, one thread; fitted per level and composed with the operation mix "
                   "the plain per-call path issued for the step (%d sequences); %d host threads running "
                   "independent key switches give the parallel efficiency" % (units, threads)),
    }


def served_requests(nx, ctx, sk, rng, count):
    """Requests of the serving workload: the nonlinear block of one BERT-base
    layer for a sequence of length S ~ U{16..128}: GELU over S x 3072 values,
    softmax over 12 heads x S rows (rows of 128, stride 256), LayerNorm over 2 x
    S rows of 768 (16 rows per ciphertext).  Inputs are encrypted up front."""
    slots = ctx.slot_count
    reqs = []
    for i in range(count):
        S = int(rng.integers(16, 129))
        jobs = []
        for _ in range(-(-S * 3072 // slots)):
            v = np.clip(rng.normal(0, 1, slots), -5, 5)
            jobs.append((nx.JOB_GELU, 0, v))
        for _ in range(-(-12 * S // 128)):
            v = np.zeros(slots)
            rows = np.clip(rng.normal(0, 3, (128, 128)), -10, 10)
            rows = rows - rows.max(axis=1, keepdims=True)
            for r in range(128):
                v[256 * r: 256 * r + 128] = np.clip(rows[r], -10, 0)
            jobs.append((nx.JOB_SOFTMAX, 128, v))
        for _ in range(2 * (-(-S // 16))):
            v = np.zeros(slots)
            for r in range(16):
                v[2048 * r: 2048 * r + 768] = rng.normal(0, 1, 768) * np.sqrt(30.0)
            jobs.append((nx.JOB_LAYER_NORM, 1024, v))
        enc = [(k, ln, nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=int(rng.integers(1 << 62))))
               for k, ln, v in jobs]
        reqs.append((S, enc))
    return reqs


def run_served(nx, ev, reqs, arrivals, lanes):
    """Submit the requests at the given arrival times (s) and wait for all of them;
    returns (makespan s, server stats)."""
    import threading
    srv = nx.Server(ev, lanes=lanes, memory_budget=64 << 30)
    ids = [None] * len(reqs)
    t0 = time.perf_counter()

    def submitter():
        for i, (S, jobs) in enumerate(reqs):
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids[i] = srv.submit(jobs)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    srv.drain()
    makespan = time.perf_counter() - t0
    st = srv.stats()
    for rid in ids:
        srv.wait(rid)
    del srv
    return makespan, st


def bench_bootstrap(nx, logN, logn, chain, alpha, reps):
    """Seconds per bootstrap of one ciphertext (ModRaise .. SlotToCoeff) and its
    error against the input message; diagonals are encoded by an untimed first call."""
    ctx = nx.Context(logN, chain, hamming_weight=192, special_primes=alpha)
    sk = nx.SecretKey(ctx, seed=5)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=6)
    steps = nx.boot_plan_steps(logN - 1, logn, 25)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=7)
    ev = nx.CKKSEvaluator(ctx, rk, gk, 2.0 ** 46, sk)
    bt = nx.Bootstrapper(ev, log_slots=logn)
    n = 1 << logn
    v = np.random.default_rng(3).uniform(-1, 1, n)
    ct = nx.encrypt(ctx, sk, nx.encode(ctx, np.tile(v, (1 << (logN - 1)) // n), 2.0 ** 46), seed=8)
    ct = nx.mod_switch_to(ct, ctx.L)
    t0 = time.perf_counter()
    out = bt.bootstrap(ct)
    ctx.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        out = bt.bootstrap(ct)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    err = np.abs(nx.decode(ctx, nx.decrypt(ctx, sk, out))[:n] - v)
    return {"ms_per_bootstrap": ms, "first_call_ms": first_ms, "slots": n, "poly_degree": 1 << logN, "data_primes": ctx.L,
            "special_primes": alpha, "key_switch_digits": ctx.digits, "depth": bt.depth(),
            "rotation_keys": len(steps), "max_err": float(err.max()), "mean_err": float(err.mean())}


def bench_matmul(nx, reps):
    """The reference's MatrixMul (src/main.cpp:306-382): 4096x768 plaintext times
    768x64 encrypted (coefficient-packed in 6 ciphertexts, expanded, accumulated);
    N = 2^13, {60, 40, 60}; seconds per full product."""
    from nexus_amd import mm
    ctx = nx.Context(13, [60, 40, 60])
    sk = nx.SecretKey(ctx, seed=4)
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, mm.reference_elts(ctx.n, 13), seed=2)
    rng = np.random.default_rng(42)
    x = np.round(rng.uniform(-1, 1, (4096, 768)), 4)
    y = np.round(rng.uniform(-1, 1, (768, 64)), 4)
    a_pts = [nx.encode(ctx, x.T[i], 2.0 ** 40) for i in range(768)]
    flat = y.T.reshape(-1)
    cts = [mm.compress(ctx, sk, flat[i * ctx.n:(i + 1) * ctx.n], 1e10, seed=10 + i) for i in range(6)]

    def run():
        return mm.accumulate(a_pts, [mm.expand(c, ctx.n, gk) for c in cts], 64)

    res = run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        res = run()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    r = res[0]
    r.scale = r.scale * 4096
    d = nx.decode(ctx, nx.decrypt(ctx, sk, nx.rescale_to_next(r)))[:4096] / 2.0
    err = float(np.mean(np.abs(d - (x @ y)[:, 0])))
    return {"ms_per_matmul": ms, "shape": "4096x768 (plaintext) x 768x64 (encrypted)", "poly_degree": 1 << 13,
            "key_switches": 6 * (ctx.n - 1), "mean_err_col0": err}


def bench_ntt(nx, reps):
    """The dominant kernel class of the nonlinear step (forward NTT, column and
    row passes) at its shape: one ciphertext of 2 x 19 limbs at N = 2^16, next to
    a device-to-device copy of the same 19.9 MB (the copy calibrates HBM counters)."""
    ctx = nx.Context(16, CHAIN)
    rng = np.random.default_rng(1)
    data = np.stack([np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in ctx.moduli[:ctx.L]])
                     for _ in range(2)])
    ct = nx.Ciphertext.from_array(ctx, data, 1, SCALE, ntt_form=False)
    nx.transform_to_ntt(ct)  # first launch loads the code object: untimed
    ctx.synchronize()
    nx.profile_select("ntt_fwd_col,ntt_fwd_row")
    t0 = time.perf_counter()
    for _ in range(reps):
        out = nx.transform_to_ntt(ct)
        cp = ct.copy()
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / reps * 1e3
    res = {"ms_per_rep": wall, "limbs_per_launch": 2 * ctx.L, "copy_bytes": int(data.nbytes)}
    for cls in ("ntt_fwd_col", "ntt_fwd_row"):
        launches, ms, b = nx.profile_read(cls)
        res[cls] = {"launches": launches, "avg_us": ms / max(launches, 1) * 1e3,
                    "algorithmic_bytes_per_launch": b / max(launches, 1),
                    "achieved_gbs": (b / max(launches, 1)) / (ms / max(launches, 1) * 1e-3) / 1e9}
    nx.profile_select("")
    del out, cp
    # the BERT chain at its top compute level (one 51-bit and nineteen 46-bit
    # primes) and at the bootstrapping top (34 data primes), both directions
    bctx = nx.Context(16, BERT_CHAIN, special_primes=BERT_SPECIAL)
    for limbs in (20, 34):
        mods = bctx.moduli[:limbs]
        d = np.stack([np.stack([rng.integers(0, q, bctx.n, dtype=np.uint64) for q in mods]) for _ in range(2)])
        c = nx.Ciphertext.from_array(bctx, d, bctx.L + 1 - limbs, BERT_SCALE, ntt_form=False)
        f = nx.transform_to_ntt(c)
        g = nx.transform_from_ntt(f)  # first launches load the code objects: untimed
        bctx.synchronize()
        classes = ("ntt_fwd_col", "ntt_fwd_row", "ntt_inv_col", "ntt_inv_row")
        nx.profile_select(",".join(classes))
        for _ in range(reps):
            f = nx.transform_to_ntt(c)
            g = nx.transform_from_ntt(f)
        bctx.synchronize()
        sub = {"limbs_per_launch": 2 * limbs}
        for cls in classes:
            launches, ms, b = nx.profile_read(cls)
            sub[cls] = {"launches": launches, "avg_us": ms / max(launches, 1) * 1e3,
                        "algorithmic_bytes_per_launch": b / max(launches, 1),
                        "achieved_gbs": (b / max(launches, 1)) / (ms / max(launches, 1) * 1e-3) / 1e9}
        nx.profile_select("")
        res["bert_chain_%d_limbs" % limbs] = sub
        del c, f, g
    return res


def bench_dot(nx, reps):
    """The dominant kernel class of the BERT layer (dot_plain: sum of 64
    ciphertext x plaintext products) at the attention's shape: 64 ciphertexts of
    2 x 19 limbs and 64 plaintexts at N = 2^16, next to a device-to-device copy
    of one ciphertext (19.9 MB; the copy calibrates HBM counters)."""
    ctx = nx.Context(16, BERT_CHAIN, special_primes=BERT_SPECIAL)
    rng = np.random.default_rng(1)
    limbs = 19
    mods = ctx.moduli[:limbs]
    ci = ctx.L + 1 - limbs
    cts = [nx.Ciphertext.from_array(
        ctx, np.stack([np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in mods]) for _ in range(2)]),
        ci, BERT_SCALE) for _ in range(64)]
    pts = [nx.Plaintext.from_array(ctx, np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in mods]), ci,
                                   BERT_SCALE) for _ in range(64)]
    out = nx.multiply_plain_accumulate(cts, pts)  # first launch loads the code object: untimed
    ctx.synchronize()
    nx.profile_select("dot_plain")
    t0 = time.perf_counter()
    for _ in range(reps):
        out = nx.multiply_plain_accumulate(cts, pts)
        cp = cts[0].copy()
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / reps * 1e3
    launches, ms, b = nx.profile_read("dot_plain")
    nx.profile_select("")
    del out, cp
    return {"ms_per_rep": wall, "terms": 64, "limbs": limbs, "copy_bytes": 2 * limbs * ctx.n * 8,
            "dot_plain": {"launches": launches, "avg_us": ms / max(launches, 1) * 1e3,
                          "algorithmic_bytes_per_launch": b / max(launches, 1),
                          "achieved_gbs": (b / max(launches, 1)) / (ms / max(launches, 1) * 1e-3) / 1e9}}


KS_CLASSES = ("ks_modup_conv", "ks_inner_rot_sum", "ks_inner_rot", "ks_moddown_conv", "modup_copy", "ks_inner")


def bench_ks(nx, reps):
    """The key-switching kernel classes at the BERT layer's top compute level:
    16 ciphertexts of 2 x 20 limbs at N = 2^16 on the BERT chain (8 special
    primes: 3 key-switching digits at 20 limbs) -- one sum of 16 rotations with
    one ModDown (batched ModUp conversion, ks_inner_rot_sum, ModDown
    conversion), 16 hoisted rotations of one ciphertext (ks_inner_rot) and 16
    fused multiply-relinearise-rescale products; next to a device-to-device
    copy of one ciphertext (20.9 MB; the copy calibrates HBM counters)."""
    ctx = nx.Context(16, BERT_CHAIN, hamming_weight=192, special_primes=BERT_SPECIAL)
    sk = nx.SecretKey(ctx, seed=61)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=62)
    steps = list(range(1, 17))
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=63)
    rng = np.random.default_rng(64)
    limbs = 20
    ci = ctx.L + 1 - limbs
    cts = [nx.Ciphertext.from_array(
        ctx, np.stack([np.stack([rng.integers(0, q, ctx.n, dtype=np.uint64) for q in ctx.moduli[:limbs]])
                       for _ in range(2)]), ci, BERT_SCALE) for _ in range(16)]

    def run():
        a = nx.rotate_sum(cts, steps, gk)
        b = nx.rotate_hoisted(cts[0], steps, gk)
        c = nx.multiply_relin_rescale_many(cts, cts[::-1], rk, fused=True)
        return a, b, c

    run()  # first launches load the code objects: untimed
    ctx.synchronize()
    nx.profile_select(",".join(KS_CLASSES))
    t0 = time.perf_counter()
    for _ in range(reps):
        out = run()
        cp = cts[0].copy()
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / reps * 1e3
    res = {"ms_per_rep": wall, "limbs": limbs, "digits": ctx.digits, "copy_bytes": 2 * limbs * ctx.n * 8}
    for cls in KS_CLASSES:
        launches, ms, b = nx.profile_read(cls)
        if launches:
            res[cls] = {"launches": launches, "avg_us": ms / launches * 1e3,
                        "algorithmic_bytes_per_launch": b / launches,
                        "achieved_gbs": (b / launches) / (ms / launches * 1e-3) / 1e9}
    nx.profile_select("")
    res["valu_int64_products_per_s"] = nx.valu_rate(0)
    res["valu_fp64_fma_per_s"] = nx.valu_rate(1)
    del out, cp
    return res


def bert_setup(nx):
    """context, keys, bootstrapper and BERT-base layer at N = 2^16"""
    from nexus_amd import bert_rotation_steps, boot_plan_steps
    ctx = nx.Context(16, BERT_CHAIN, hamming_weight=192, special_primes=BERT_SPECIAL)
    sk = nx.SecretKey(ctx, seed=21)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=22)
    steps = sorted(set(bert_rotation_steps(ctx.slot_count)) | set(boot_plan_steps(15, 15, 25)))
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps], seed=23)
    tune = nx.autotune(ctx, reps=5)  # launch shapes for this ring size (bit-exact candidates only)
    ev = nx.CKKSEvaluator(ctx, rk, gk, BERT_SCALE, sk)
    bt = nx.Bootstrapper(ev)
    rng = np.random.default_rng(42)
    W = {k: rng.normal(0, 0.02, shape) for k, shape in
         (("wq", (768, 768)), ("wk", (768, 768)), ("wv", (768, 768)), ("wo", (768, 768)),
          ("w1", (768, 3072)), ("w2", (3072, 768)))}
    layer = nx.BertLayer(bt, W)
    return {"ctx": ctx, "sk": sk, "rk": rk, "gk": gk, "ev": ev, "bt": bt, "layer": layer, "W": W,
            "keys": len(steps) + 1, "autotune": tune}


def bench_bert_call(nx, reps):
    """one BERT-base layer call on a 128-token sequence (after a warm-up call
    that encodes the weight diagonals): seconds per call and its stages"""
    B = bert_setup(nx)
    rows = np.random.default_rng(5).normal(0, 1, (BERT_TOKENS, 768))
    cts = bert_encrypt(nx, B, rows)
    B["layer"].forward(cts, BERT_TOKENS)
    B["ctx"].synchronize()
    w0 = time.clock_gettime_ns(time.CLOCK_MONOTONIC)
    t0 = time.perf_counter()
    for _ in range(reps):
        B["layer"].forward(cts, BERT_TOKENS)
    B["ctx"].synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    # the timed calls' window, for nexus_amd/trace_busy.py on a kernel trace of this run
    log("window monotonic_ns %d %d" % (w0, time.clock_gettime_ns(time.CLOCK_MONOTONIC)))
    B["layer"].set_stage_timing(True)
    B["layer"].forward(cts, BERT_TOKENS)
    return {"ms_per_layer_call": ms, "tokens": BERT_TOKENS, "stages_ms": dict(B["layer"].stage_ms()),
            "autotune_ntt": B["autotune"]["ntt"],
            "autotune_modup": B["autotune"].get("modup", []), "autotune_moddown": B["autotune"].get("moddown", []),
            "autotune_epilogue": B["autotune"].get("epilogue")}


def fused_mode(mode):
    """the autotuner's choice for a conversion fused into the NTT column pass"""
    if mode < 0:
        return "unfused"
    return "conversion in the column pass, %s outputs per workgroup, %d-column tiles, %d threads" % (
        mode & 255 or "default", 4 if mode & 256 else 8, 512 if mode & 512 else 256)


def bench_bert_sets(nx, reps, count=8):
    """`count` full 128-token sets: one layer call per set against every set
    in lockstep (forward_sets: the bootstrapping steps of all sets together)"""
    B = bert_setup(nx)
    rng = np.random.default_rng(6)
    sets = [bert_encrypt(nx, B, rng.normal(0, 1, (BERT_TOKENS, 768))) for _ in range(count)]
    segs = [[(0, BERT_TOKENS)]] * count
    layer, ctx = B["layer"], B["ctx"]
    layer.forward(sets[0], BERT_TOKENS)
    layer.forward_sets(sets[:2], segs[:2])
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for x in sets:
            layer.forward(x, BERT_TOKENS)
    ctx.synchronize()
    t_calls = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        layer.forward_sets(sets, segs)
    ctx.synchronize()
    t_lock = (time.perf_counter() - t0) / reps
    return {"sets": count, "per_set_calls_s": t_calls, "lockstep_s": t_lock, "lockstep_speedup": t_calls / t_lock}


def pack_sets(lens, capacity):
    """first-fit decreasing: [[(sequence, start), ...], ...]"""
    sets, fill = [], []
    for k in sorted(range(len(lens)), key=lambda k: -lens[k]):
        for j in range(len(sets)):
            if fill[j] + lens[k] <= capacity:
                sets[j].append((k, fill[j]))
                fill[j] += lens[k]
                break
        else:
            sets.append([(k, 0)])
            fill.append(lens[k])
    return sets


def bert_encrypt(nx, B, rows):
    ctx = B["ctx"]
    return [nx.encrypt(ctx, B["sk"], nx.encode(ctx, v, BERT_SCALE, ctx.L - 1), seed=1000 + i)
            for i, v in enumerate(nx.bert_pack(rows, ctx.slot_count))]


def bert_decrypt(nx, B, cts):
    ctx = B["ctx"]
    return nx.bert_unpack([nx.decode(ctx, nx.decrypt(ctx, B["sk"], c)) for c in cts], BERT_TOKENS)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sequences", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the nonlinear, bootstrapping and MatMul extras")
    ap.add_argument("--no-served", action="store_true", help="skip the serving measurement")
    ap.add_argument("--lockstep-steps", type=int, default=2,
                    help="steps of the cross-set lockstep mode (every set of a step in one forward_sets call)")
    ap.add_argument("--requests", type=int, default=100, help="served BERT requests of the latency pass "
                                                                 "(p50/p99 samples)")
    ap.add_argument("--capacity-requests", type=int, default=48,
                    help="served BERT requests queued at once for the capacity pass")
    ap.add_argument("--load", type=float, default=0.8,
                    help="arrival rate of the served latency pass as a fraction of the measured capacity")
    ap.add_argument("--served-plain-requests", type=int, default=0,
                    help="also serve the first k requests through the plain per-call drop-in")
    ap.add_argument("--nonlinear-requests", type=int, default=12, help="requests of the served nonlinear extra")
    ap.add_argument("--lanes", type=int, default=3, help="lanes of the served nonlinear-block measurement")
    ap.add_argument("--bert-lanes", type=int, default=1,
                    help="lanes of the served BERT measurement, each with one BERT call in flight (each lane "
                         "caches its own call working set)")
    ap.add_argument("--only", choices=["bert", "bert_served", "bert_sets", "dot", "bootstrap_dense",
                                       "bootstrap_sparse", "matmul", "ntt", "ks", "valu"],
                    help="run one of the extra workloads alone (for profiling) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU, so --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import nexus_amd as nx

    if nx.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.only:
        if args.only == "bert":
            r = bench_bert_call(nx, args.steps)
        elif args.only == "bert_sets":
            r = bench_bert_sets(nx, args.steps)
        elif args.only == "bert_served":
            B = bert_setup(nx)
            rows = np.random.default_rng(5).normal(0, 1, (BERT_TOKENS, 768))
            cts = bert_encrypt(nx, B, rows)
            B["layer"].forward(cts, BERT_TOKENS)
            B["ctx"].synchronize()
            t0 = time.perf_counter()
            B["layer"].forward(cts, BERT_TOKENS)
            B["ctx"].synchronize()
            call_s = time.perf_counter() - t0
            # one call per sequence: the call rate stands for both paths' rates here
            r = served_bert(nx, B, args, 1.0 / call_s, 1.0 / call_s)
        elif args.only == "bootstrap_dense":
            r = bench_bootstrap(nx, 16, 15, [51] + [46] * 16 + [51] * 14 + [58] * 3, 3, args.steps)
        elif args.only == "bootstrap_sparse":
            r = bench_bootstrap(nx, 15, 13, [51] + [46] * 16 + [51] * 14 + [51], 1, args.steps)
        elif args.only == "ntt":
            r = bench_ntt(nx, args.steps)
        elif args.only == "dot":
            r = bench_dot(nx, args.steps)
        elif args.only == "ks":
            r = bench_ks(nx, args.steps)
        elif args.only == "valu":
            r = {"int64_products_per_s": nx.valu_rate(0), "fp64_fma_per_s": nx.valu_rate(1)}
        else:
            r = bench_matmul(nx, args.steps)
        print(json.dumps({"workload": args.only, **r}), flush=True)
        return

    log("setup: BERT-base layer, N=2^16, L=%d, %d special primes" % (len(BERT_CHAIN) - BERT_SPECIAL, BERT_SPECIAL))
    t0 = time.perf_counter()
    B = bert_setup(nx)
    ctx, layer = B["ctx"], B["layer"]
    rng = np.random.default_rng(2025)
    lens = [int(x) for x in rng.integers(32, BERT_TOKENS + 1, args.sequences)]
    xs = [rng.normal(0, 1, (n, 768)) for n in lens]
    sets = pack_sets(lens, BERT_TOKENS)
    set_rows, set_segs = [], []
    for members in sets:
        rows = np.zeros((BERT_TOKENS, 768))
        segs = []
        for k, a in members:
            rows[a:a + lens[k]] = xs[k]
            segs.append((a, lens[k]))
        set_rows.append(rows)
        set_segs.append(segs)
    single_rows = []
    for k in range(len(lens)):
        rows = np.zeros((BERT_TOKENS, 768))
        rows[:lens[k]] = xs[k]
        single_rows.append(rows)
    enc_sets = [bert_encrypt(nx, B, r) for r in set_rows]
    enc_single = [bert_encrypt(nx, B, r) for r in single_rows]
    ctx.synchronize()
    log("setup %.1f s; %d sequences (lengths %s) -> %d packed sets" % (time.perf_counter() - t0, len(lens), lens,
                                                                      len(sets)))

    # warm-up: the first call encodes the weight diagonals; one more call ranks the kernel classes
    for w in range(max(1, args.warmup)):
        tw = time.perf_counter()
        layer.forward(enc_sets[0], segments=set_segs[0])
        ctx.synchronize()
        log("warmup %d: %.1f s" % (w, time.perf_counter() - tw))
    nx.profile_select("*")
    layer.forward(enc_sets[0], segments=set_segs[0])
    ctx.synchronize()
    ranking = sorted(((nx.profile_read(c)[1], c) for c in nx.profile_classes()), reverse=True)
    # every top class's rates over that one call (events around each launch)
    class_rates = []
    for t, c in ranking[:10]:
        nl, ms_c, by = nx.profile_read(c)
        fl = nx.profile_read_flops(c)
        class_rates.append({"class": c, "ms_per_call": ms_c, "launches": nl,
                            "algorithmic_gbs": by / (ms_c * 1e-3) / 1e9 if ms_c else 0.0,
                            "hbm_frac": by / (ms_c * 1e-3) / 1e9 / HBM_PEAK_GBS if ms_c else 0.0,
                            "fp64_tflops": fl / (ms_c * 1e-3) / 1e12 if ms_c and fl else None})
    nx.profile_select("")
    dominant = ranking[0][1]
    log("kernel time ranking of one call (ms): " + ", ".join("%s=%.0f" % (c, t) for t, c in ranking[:10]))

    # plain per-call drop-in: one call per sequence, every operation a
    # single-ciphertext call (per-op mode: no hoisting, batching or lockstep)
    plain_lat, plain_out = [], []
    B["ev"].set_per_op(True)
    nx.op_tally_start()
    ctx.synchronize()
    tp = time.perf_counter()
    for k in range(len(lens)):
        ts = time.perf_counter()
        plain_out.append(layer.forward(enc_single[k], lens[k]))
        ctx.synchronize()
        plain_lat.append(time.perf_counter() - ts)
        log("plain per-call sequence %d (length %d): %.2f s" % (k, lens[k], plain_lat[-1]))
    plain_s = time.perf_counter() - tp
    tally = nx.op_tally_stop()
    B["ev"].set_per_op(False)
    log("plain per-call: %d sequences in %.2f s" % (len(lens), plain_s))

    # batched: packed sets (the timed steps)
    nx.profile_select(dominant + ",ks_inner")
    ctx.synchronize()
    tb = time.perf_counter()
    set_lat, outs = [], None
    for k in range(args.steps):
        outs = []
        for j in range(len(sets)):
            ts = time.perf_counter()
            outs.append(layer.forward(enc_sets[j], segments=set_segs[j]))
            ctx.synchronize()
            set_lat.append(time.perf_counter() - ts)
        log("step %d: %.2f s" % (k, sum(set_lat[-len(sets):])))
    ctx.synchronize()
    elapsed = time.perf_counter() - tb
    launches, k_ms, k_bytes = nx.profile_read(dominant)
    k_flops = nx.profile_read_flops(dominant)
    _, _, ks_bytes = nx.profile_read("ks_inner")
    nx.profile_select("")
    ms_per_step = elapsed / args.steps * 1e3
    # cross-set lockstep: every set of the step in one call (each bootstrapping
    # step over all sets, odd ciphertexts paired across sets); latency = the step
    lock = None
    if args.lockstep_steps > 0:
        ctx.trim()  # the lockstep working set differs from the per-set calls'
        layer.forward_sets(enc_sets[:2], set_segs[:2])  # warm the lockstep working set
        ctx.synchronize()
        tl = time.perf_counter()
        for _ in range(args.lockstep_steps):
            lock_out = layer.forward_sets(enc_sets, set_segs)
            ctx.synchronize()
        lock_s = (time.perf_counter() - tl) / args.lockstep_steps
        got_lock = {}
        for j, members in enumerate(sets):
            dec = bert_decrypt(nx, B, lock_out[j])
            for k, a in members:
                got_lock[k] = dec[a:a + lens[k]]
        del lock_out
        lock = {"value": len(lens) / lock_s, "seconds_per_step": lock_s, "latency_ms": lock_s * 1e3,
                "speedup_over_per_set_calls": (ms_per_step / 1e3) / lock_s, "steps": args.lockstep_steps,
                "definition": "BertLayer::forward_sets over the step's packed sets: every stage over all sets, "
                              "each bootstrapping step refreshing every set's ciphertexts together (odd ones "
                              "paired across sets); every sequence's result arrives at the end of the step"}
        log("lockstep: %.2f s per step (%.3f sequence-layers/s)" % (lock_s, len(lens) / lock_s))
    # per-sequence latency of the batched path = its set's call
    seq_lat = []
    for j, members in enumerate(sets):
        seq_lat += [set_lat[-len(sets) + j]] * len(members)
    got, got_plain = {}, {}
    for j, members in enumerate(sets):
        dec = bert_decrypt(nx, B, outs[j])
        for k, a in members:
            got[k] = dec[a:a + lens[k]]
    for k in range(len(lens)):
        got_plain[k] = bert_decrypt(nx, B, plain_out[k])[:lens[k]]
    del plain_out
    # decrypted error: against the FP64 model of the same approximations (the
    # 2^-10 bound) and against the exact FP64 layer (the approximations' own error)
    from oracle import bert as obert
    W = B["W"]
    model = {k: obert.layer_approx(xs[k], W) for k in range(len(lens))}
    exact = {k: obert.layer_exact(xs[k], W) for k in range(len(lens))}
    err_b = max(float(np.max(np.abs(got[k] - model[k]))) for k in model)
    err_p = max(float(np.max(np.abs(got_plain[k] - model[k]))) for k in model)
    err_x = max(float(np.max(np.abs(got[k] - exact[k]))) for k in model)
    if lock is not None:
        lock["max_err_vs_fp64_model"] = max(float(np.max(np.abs(got_lock[k] - model[k]))) for k in model)
    model_x = max(float(np.max(np.abs(model[k] - exact[k]))) for k in model)

    avg_launch_ms = k_ms / max(launches, 1)
    achieved = (k_bytes / max(launches, 1)) / (avg_launch_ms * 1e-3) / 1e9 if launches else 0.0
    # a class that records FP64 operations (the conversions fused into the NTT
    # column pass) is bound by the double-precision VALU rate, not by HBM
    achieved_tf = (k_flops / max(launches, 1)) / (avg_launch_ms * 1e-3) / 1e12 if launches else 0.0
    valu = k_flops > 0
    res = {
        "metric": METRIC,
        "value": len(lens) * 1000.0 / ms_per_step,
        "unit": "BERT-base sequence-layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "p50_ms": float(np.percentile(np.array(seq_lat) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(seq_lat) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (exact RNS modular arithmetic)",
        "data": "synthetic (random-init weights ~N(0,0.02), inputs ~N(0,1))",
        "plain_per_call": {
            "value": len(lens) / plain_s,
            "seconds": plain_s,
            "p50_ms": float(np.percentile(np.array(plain_lat) * 1e3, 50)),
            "p99_ms": float(np.percentile(np.array(plain_lat) * 1e3, 99)),
        },
        "batched_over_plain": plain_s / (ms_per_step / 1e3),
        "lockstep": lock,
        "plain_definition": "the same layer program in per-op mode: every operation one single-ciphertext "
                            "call (rotate_vector, multiply_plain, add, multiply, relinearize, rescale_to_next per "
                            "term; one bootstrap and one GELU per ciphertext), one layer call per sequence",
        "max_err_vs_fp64_model": err_b,
        "max_err_vs_fp64_model_plain": err_p,
        "err_bound": 2.0 ** -10,
        "err_ok": bool(err_b <= 2.0 ** -10 and err_p <= 2.0 ** -10 and
                       (lock is None or lock["max_err_vs_fp64_model"] <= 2.0 ** -10)),
        "err_reference": "FP64 model of the reference's own approximations (oracle/bert.py layer_approx); the "
                         "approximations themselves differ from the exact FP64 layer by model_vs_exact",
        "max_err_vs_exact_fp64_layer": err_x,
        "model_vs_exact_fp64_layer": model_x,
        "plain_op_mix": {k: int(sum(v.values())) for k, v in tally.items()},
        "config": {
            "workload": "BERT-base encoder layer (hidden 768, 12 heads, FFN 3072) over %d sequences of lengths %s, "
                        "packed into %d sets of %d token rows, bootstrapping between blocks"
                        % (len(lens), lens, len(sets), BERT_TOKENS),
            "poly_degree": 1 << 16,
            "coeff_modulus_bits": BERT_CHAIN,
            "special_primes": BERT_SPECIAL,
            "scale": "2^46",
            "galois_keys": B["keys"],
            "autotuned": {
                "ntt_tiles": ["%s batch %d: col %d row %d radix16 %s (%.1f vs %.1f us)"
                              % (e["dir"], e["batch"], e["col"], e["row"], e.get("radix16"), e["us"], e["default_us"])
                              for e in B["autotune"]["ntt"] if e["col"] or e["row"] or any(e.get("radix16", []))],
                "dot_plain_layouts": ["%d limbs: %d" % (d["limbs"], d["layout"]) for d in B["autotune"]["dot_plain"]],
                "modup": ["batch %d at %d limbs: %s (%.1f vs %.1f us unfused)"
                          % (e["batch"], e["limbs"], fused_mode(e["mode"]), e["us"], e["unfused_us"])
                          for e in B["autotune"].get("modup", [])],
                "moddown": ["batch %d at %d limbs: %s (%.1f vs %.1f us unfused)"
                            % (e["batch"], e["limbs"], fused_mode(e["mode"]), e["us"], e["unfused_us"])
                            for e in B["autotune"].get("moddown", [])],
                "moddown_epilogue": B["autotune"].get("epilogue"),
            },
        },
        "roofline": {
            "kernel": dominant,
            "bound": "valu-fp64" if valu else "hbm",
            "achieved": achieved_tf if valu else achieved,
            "peak": FP64_PEAK_TFLOPS if valu else HBM_PEAK_GBS,
            "unit": "TFLOP/s" if valu else "GB/s",
            "frac": achieved_tf / FP64_PEAK_TFLOPS if valu else achieved / HBM_PEAK_GBS,
            "achieved_hbm_gbs": achieved,
            "hbm_frac": achieved / HBM_PEAK_GBS,
            "algorithmic_flops_per_launch": k_flops / max(launches, 1),
            "algorithmic_bytes_per_launch": k_bytes / max(launches, 1),
            "traffic": (TRAFFIC_PER_ALGORITHMIC_BYTE[dominant] * k_bytes / max(launches, 1)
                        if dominant in TRAFFIC_PER_ALGORITHMIC_BYTE else None),
            "traffic_source": "calibrated FETCH_SIZE+WRITE_SIZE per algorithmic byte of this kernel class "
                              "(%s) times this step's average algorithmic bytes per launch"
                              % TRAFFIC_PROFILE.get(dominant, "no counter run for this class"),
            "launches": launches,
            "avg_launch_us": avg_launch_ms * 1e3,
            "kernel_share_of_step": k_ms / (elapsed * 1e3),
            "top_classes_one_call": class_rates,
        },
    }
    if not args.no_served:
        ctx.trim()  # the serving lane starts on a device without the lockstep phase's cached blocks
        res["served"] = served_bert(nx, B, args, len(lens) / plain_s, len(lens) * 1000.0 / ms_per_step)
    del enc_sets, enc_single, outs, layer, B, ctx
    if not args.no_extra:
        res["extra"] = extras(nx, args)
    if not args.no_cpu_baseline:
        log("cpu baseline: the SEAL-form port timed on the host, composed with the plain path's operation mix")
        res["cpu_baseline"] = cpu_baseline(tally, len(lens), "BERT-base sequence-layers/s")
    print(json.dumps(res), flush=True)


def served_bert(nx, B, args, plain_rate, batched_rate):
    """Single-sequence BERT-layer requests, lengths ~U{16..128}, served with
    continuous batching (queued requests packed into 128-token sets), in two
    passes: (1) capacity -- --capacity-requests requests all queued at once,
    the sustained sequence-layers/s; (2) latency -- --requests requests with
    Poisson arrivals at --load times that capacity, p50/p99 from submission to
    completion.  Optionally (--served-plain-requests) the plain per-call drop-in
    serves the first requests of the latency pass (one lane, per-op mode, one
    call per request) at the same arrival times."""
    rng = np.random.default_rng(77)
    n = args.requests
    n_cap = args.capacity_requests
    count_all = max(n, n_cap)
    lens = [int(x) for x in rng.integers(16, BERT_TOKENS + 1, count_all)]
    reqs = []
    for L_ in lens:
        rows = np.zeros((BERT_TOKENS, 768))
        rows[:L_] = rng.normal(0, 1, (L_, 768))
        reqs.append((L_, bert_encrypt(nx, B, rows)))
    B["ctx"].synchronize()
    import threading

    def run(count, lanes, pack, per_op, arrivals):
        B["ev"].set_per_op(per_op)
        srv = nx.Server(B["ev"], lanes=lanes, bert=B["layer"], pack=pack, memory_budget=96 << 30)
        srv.set_bert_concurrency(lanes)
        B["ev"].set_per_op(False)
        ids = [None] * count
        if arrivals is None:  # everything queued before the server admits anything
            srv.pause(True)
        t0 = time.perf_counter()

        def submitter():
            for i in range(count):
                L_, cts = reqs[i]
                if arrivals is not None:
                    delay = arrivals[i] - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                ids[i] = srv.submit_bert(cts, L_)

        th = threading.Thread(target=submitter)
        th.start()
        if arrivals is None:
            th.join()
            t0 = time.perf_counter()
            srv.pause(False)
        done = 0
        for i in range(count):  # progress from the work itself
            while ids[i] is None:
                time.sleep(0.05)
            srv.wait(ids[i])
            done += 1
            if done % 20 == 0:
                log("served: %d of %d requests done" % (done, count))
        th.join()
        srv.drain()
        span = time.perf_counter() - t0
        st = srv.stats()
        st.update(srv.batching_stats())
        del srv
        return span, st

    log("served BERT capacity: %d requests queued at once" % n_cap)
    span_c, st_c = run(n_cap, args.bert_lanes, True, False, None)
    capacity = n_cap / span_c
    log("served capacity (%d lanes, packing): %.1f s, %.3f sequence-layers/s, %d packed calls carrying %d requests"
        % (args.bert_lanes, span_c, capacity, st_c["packed_calls"], st_c["packed_requests"]))
    rate = args.load * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    log("served BERT latency: %d requests, Poisson arrivals at %.3f/s (%.2f of capacity)" % (n, rate, args.load))
    spanb, stb = run(n, args.bert_lanes, True, False, arrivals)
    log("served (%d lanes, packing): %.1f s, p50 %.0f ms, p99 %.0f ms, %d packed calls carrying %d requests"
        % (args.bert_lanes, spanb, stb["p50_ms"], stb["p99_ms"], stb["packed_calls"], stb["packed_requests"]))
    res = {
        "workload": "single-sequence BERT-base layer requests, lengths ~U{16..128}; served = %d lanes, continuous "
                    "batching of queued requests into 128-token sets; capacity: %d requests queued at once; "
                    "latency: %d requests with Poisson arrivals at %.3f requests/s (%.2f of capacity)"
                    % (args.bert_lanes, n_cap, n, rate, args.load),
        "requests": n, "lanes": args.bert_lanes, "mean_length": float(np.mean(lens[:n])),
        "capacity": {"layers_per_s": capacity, "requests": n_cap, "makespan_s": span_c,
                     "packed_calls": st_c["packed_calls"], "packed_requests": st_c["packed_requests"]},
        "served": {"layers_per_s": n / spanb, "arrival_rate": rate, "p50_ms": stb["p50_ms"], "p99_ms": stb["p99_ms"],
                   "mean_ms": stb["mean_ms"], "makespan_s": spanb, "packed_calls": stb["packed_calls"],
                   "packed_requests": stb["packed_requests"], "peak_concurrent": stb["peak_concurrent"],
                   "latency_samples": stb["completed"]},
        "plain_per_call_rate": plain_rate,
        "throughput_ratio": capacity / plain_rate,
        "throughput_ratio_definition": "served capacity (sequence-layers/s with continuous batching, requests "
                                       "queued) over the plain per-call drop-in's sequence-layers/s (one call at "
                                       "a time, per-op mode)",
    }
    k = min(args.served_plain_requests, n)
    if k:
        span1, st1 = run(k, 1, False, True, arrivals)
        log("plain per-call served (%d requests): %.1f s, p50 %.0f ms, p99 %.0f ms"
            % (k, span1, st1["p50_ms"], st1["p99_ms"]))
        res["plain_per_call_served"] = {"requests": k, "layers_per_s": k / span1, "p50_ms": st1["p50_ms"],
                                        "p99_ms": st1["p99_ms"], "makespan_s": span1}
    return res


def extras(nx, args):
    """the nonlinear block at the reference's chain (plain and served),
    bootstrapping alone, and the reference's MatrixMul"""
    extra = {}
    log("nonlinear block: context + keys (N=2^16, L=19)")
    ctx = nx.Context(16, CHAIN)
    sk = nx.SecretKey(ctx, seed=11)
    rk = nx.KSwitchKeys(ctx).generate_relin(sk, seed=12)
    steps_needed = [-1024, -128] + [1 << i for i in range(10)]
    gk = nx.KSwitchKeys(ctx).generate_galois(sk, [ctx.elt_from_step(s) for s in steps_needed], seed=13)
    ev = nx.CKKSEvaluator(ctx, rk, gk, SCALE, sk)
    rng = np.random.default_rng(2025)
    g_in, s_in, l_in = make_inputs(rng, ctx.slot_count)
    enc = lambda v: nx.encrypt(ctx, sk, nx.encode(ctx, v, SCALE), seed=int(rng.integers(1 << 62)))
    cts = [(ev.gelu, enc(v), None) for v in g_in] + [(ev.softmax, enc(v), 128) for v in s_in] + \
          [(ev.layer_norm, enc(v), 1024) for v in l_in]

    def step():
        return [f(c) if a is None else f(c, a) for f, c, a in cts]

    step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        step()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / 3 * 1e3
    extra["nonlinear_block_plain"] = {
        "workload": "GELU 128x3072 (12 ct), softmax 12x128x128 (12 ct), 2x LayerNorm 128x768 (16 ct) through "
                    "the plain per-call drop-in, N=2^16, {58, 40 x 18, 58}, scale 2^40",
        "ms_per_block": ms, "blocks_per_s": 1000.0 / ms}
    log("nonlinear block: %.0f ms" % ms)
    if not args.no_served:
        log("served nonlinear blocks: %d requests, lengths U{16..128}, Poisson arrivals" % args.nonlinear_requests)
        srng = np.random.default_rng(77)
        reqs = served_requests(nx, ctx, sk, srng, args.nonlinear_requests)
        n_ct = sum(len(j) for _, j in reqs)
        per_ct_s = ms / 1e3 / len(cts)
        rate = 1.2 / (per_ct_s * n_ct / len(reqs))
        arrivals = np.cumsum(srng.exponential(1.0 / rate, len(reqs)))
        arrivals -= arrivals[0]
        ctx.synchronize()
        span1, st1 = run_served(nx, ev, reqs, arrivals, 1)
        log("plain per-call (1 lane): %.2f s, p50 %.0f ms, p99 %.0f ms" % (span1, st1["p50_ms"], st1["p99_ms"]))
        span3, st3 = run_served(nx, ev, reqs, arrivals, args.lanes)
        log("served (%d lanes): %.2f s, p50 %.0f ms, p99 %.0f ms" % (args.lanes, span3, st3["p50_ms"], st3["p99_ms"]))
        extra["nonlinear_served"] = {
            "workload": "nonlinear block of one BERT-base layer per request, S ~ U{16..128}, Poisson arrivals at "
                        "%.2f requests/s (1.2x the plain path's capacity)" % rate,
            "requests": len(reqs), "ciphertexts": n_ct, "mean_seq_len": float(np.mean([S for S, _ in reqs])),
            "lanes": args.lanes,
            "served": {"blocks_per_s": len(reqs) / span3, "p50_ms": st3["p50_ms"], "p99_ms": st3["p99_ms"],
                       "mean_ms": st3["mean_ms"], "peak_concurrent": st3["peak_concurrent"], "makespan_s": span3},
            "plain_per_call": {"blocks_per_s": len(reqs) / span1, "p50_ms": st1["p50_ms"], "p99_ms": st1["p99_ms"],
                               "mean_ms": st1["mean_ms"], "makespan_s": span1},
            "throughput_ratio": span1 / span3,
        }
        del reqs
    del cts, ev, gk, rk, sk, ctx
    log("bootstrap N=2^16, 2^15 slots (dense), hybrid key switching (3 special primes)")
    extra["bootstrap_dense_n65536"] = bench_bootstrap(nx, 16, 15, [51] + [46] * 16 + [51] * 14 + [58] * 3, 3, 3)
    log("bootstrap N=2^15, 2^13 slots (sparse), the reference's chain (cuda/src/bootstrapping.cu:22-60)")
    extra["bootstrap_sparse_n32768_reference_chain"] = bench_bootstrap(
        nx, 15, 13, [51] + [46] * 16 + [51] * 14 + [51], 1, 3)
    log("bootstrap N=2^15, 2^13 slots, same data chain with 4 special primes")
    extra["bootstrap_sparse_n32768_hybrid"] = bench_bootstrap(
        nx, 15, 13, [51] + [46] * 16 + [51] * 14 + [58] * 4, 4, 3)
    log("matmul 4096x768 x 768x64 (N=2^13)")
    extra["matmul_reference_mm"] = bench_matmul(nx, 2)
    return extra


if __name__ == "__main__":
    main()
