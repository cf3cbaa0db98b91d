"""Oracle of hybrid key switching (test infrastructure only): alpha special
primes p_0..p_{alpha-1}, digits of alpha consecutive data primes, fast basis
extension for ModUp and ModDown.

Restates Phantom's CKKS path (reference cuda/thirdparty/phantom-fhe/src/
rns_bconv.cu:455-634 modup, :776-828 moddown_from_NTT; eval_key_switch.cu:12-66
inner product) with Python integers, so it pins every residue of the device
path at small N.  One departure from Phantom, whose ModDown uses the plain
fast conversion (bConv_BEHZ): the P part is converted exactly as its centred
representative, so ModDown rounds like SEAL's single-special-prime path
instead of carrying a bias of up to alpha per coefficient.  Inputs and outputs are NTT
form ([limbs][n] uint64 arrays).
"""
import numpy as np

import oracle as orc


def _conv(values, src_mods, dst_mod):
    """fast basis extension: sum_i [x_i * Qhat_i^-1]_{q_i} * (Qhat_i mod t) mod t"""
    t = int(dst_mod)
    acc = np.zeros(len(values[0]), dtype=object)
    for i, qi in enumerate(src_mods):
        qi = int(qi)
        qhat_qi = 1
        qhat_t = 1
        for i2, q2 in enumerate(src_mods):
            if i2 != i:
                qhat_qi = qhat_qi * int(q2) % qi
                qhat_t = qhat_t * int(q2) % t
        w = pow(qhat_qi, -1, qi)
        y = np.array([int(x) * w % qi for x in values[i]], dtype=object)
        acc = acc + y * qhat_t
    return np.array([int(v) % t for v in acc], dtype=np.uint64)


def _conv_centred(values, src_mods, dst_mod):
    """exact conversion of the representative in (-P/2, P/2]: sum_k v_k * Phat_k
    - r * P mod t with v_k = [y_k Phat_k^-1]_{p_k}, r = round(sum_k v_k / p_k)
    taken in float64 (summed in k order, as the device does)."""
    t = int(dst_mod)
    Pt = 1
    for p in src_mods:
        Pt = Pt * int(p) % t
    n = len(values[0])
    acc = np.zeros(n, dtype=object)
    frac = np.zeros(n, dtype=np.float64)
    for k, pk in enumerate(src_mods):
        pk = int(pk)
        ph_pk = 1
        ph_t = 1
        for k2, p2 in enumerate(src_mods):
            if k2 != k:
                ph_pk = ph_pk * int(p2) % pk
                ph_t = ph_t * int(p2) % t
        w = pow(ph_pk, -1, pk)
        v = [int(x) * w % pk for x in values[k]]
        frac = frac + np.array([float(x) for x in v]) / float(pk)
        acc = acc + np.array(v, dtype=object) * ph_t
    r = np.floor(frac + 0.5).astype(np.int64)
    return np.array([(int(a) - int(rr) * Pt) % t for a, rr in zip(acc, r)], dtype=np.uint64)


def _modup(target, moduli, L, alpha, logn):
    """ModUp of target ([l][n], NTT): T[o][j] = NTT_o(conv(digit j)) over the
    output primes q_0..q_{l-1}, p_0..p_{alpha-1} (rns_bconv.cu:530-634)."""
    l = target.shape[0]
    qs = [int(q) for q in moduli[:l]]
    ps = [int(p) for p in moduli[L:L + alpha]]
    outs = qs + ps
    coef = [orc.ntt_inverse(target[i].copy(), logn, qs[i]) for i in range(l)]
    digits = (l + alpha - 1) // alpha
    T = [[None] * digits for _ in outs]
    for j in range(digits):
        idx = list(range(j * alpha, min(l, (j + 1) * alpha)))
        src = [coef[i] for i in idx]
        src_mods = [qs[i] for i in idx]
        for o, t in enumerate(outs):
            T[o][j] = orc.ntt_forward(_conv(src, src_mods, t), logn, t)
    return T, outs


def _inner(T, key, outs, l, L, perm=None):
    """acc[k][o] = sum_j perm(T[o][j]) key[j][k][limb of o] (eval_key_switch.cu:12-66),
    as Python integers (not yet reduced)"""
    n = len(T[0][0])
    out_key_limb = list(range(l)) + [L + k for k in range(len(outs) - l)]
    acc = [[np.zeros(n, dtype=object) for _ in outs] for _ in range(2)]
    for k in range(2):
        for o in range(len(outs)):
            for j in range(len(T[o])):
                t = T[o][j] if perm is None else T[o][j][perm]
                acc[k][o] = acc[k][o] + t.astype(object) * key[j, k, out_key_limb[o]].astype(object)
    return acc


def _moddown(acc, moduli, l, L, alpha, logn):
    """(acc_q - centred(INTT acc_P)) * P^-1 mod q_i = round(acc / P) (rns_bconv.cu:776-828,
    with the exact centred conversion)"""
    qs = [int(q) for q in moduli[:l]]
    ps = [int(p) for p in moduli[L:L + alpha]]
    res = []
    for k in range(2):
        yP = [orc.ntt_inverse(np.array([int(v) % ps[m] for v in acc[k][l + m]], dtype=np.uint64), logn, ps[m])
              for m in range(alpha)]
        polys = []
        for i, q in enumerate(qs):
            z = orc.ntt_forward(_conv_centred(yP, ps, q), logn, q)
            P = 1
            for p in ps:
                P = P * p % q
            pinv = pow(P, -1, q)
            d = (acc[k][i] - z.astype(object)) % q
            polys.append(np.array([int(v) * pinv % q for v in d], dtype=np.uint64))
        res.append(np.array(polys))
    return res[0], res[1]


def switch_key_hybrid(target, key, moduli, L, alpha, logn):
    """Key switch of `target` ([l][n], NTT) with key [dnum][2][L+alpha][n]:
    returns (ks0, ks1), each [l][n] NTT, to be added to (c0, c1)."""
    l = target.shape[0]
    T, outs = _modup(target, moduli, L, alpha, logn)
    return _moddown(_inner(T, key, outs, l, L), moduli, l, L, alpha, logn)


def rotate_sum_hybrid(cts, elts, keys, moduli, L, alpha, logn):
    """sum_r rot_r(cts[r]) under hybrid key switching with one ModDown; each
    rotation is hoisted: its digits are sigma(ModUp(c1)) (Phantom
    hoisting_inplace, evaluate.cu:1727-1920).  elts[r] = 0 adds cts[r] itself.
    cts: (2, l, n) NTT arrays; returns (2, l, n)."""
    l = cts[0].shape[1]
    n = 1 << logn
    qs = [int(q) for q in moduli[:l]]
    base = np.zeros((2, l, n), dtype=object)
    acc = None
    for ct, elt, key in zip(cts, elts, keys):
        if elt == 0:
            base = base + ct.astype(object)
            continue
        perm = orc.galois_table_ntt(logn, elt).astype(np.int64)
        base[0] = base[0] + ct[0][:, perm].astype(object)
        T, outs = _modup(ct[1], moduli, L, alpha, logn)
        a = _inner(T, key, outs, l, L, perm)
        acc = a if acc is None else [[x + y for x, y in zip(ra, rb)] for ra, rb in zip(acc, a)]
    out = np.zeros((2, l, n), dtype=np.uint64)
    ks = _moddown(acc, moduli, l, L, alpha, logn) if acc is not None else None
    for k in range(2):
        for i, q in enumerate(qs):
            v = base[k][i] + (ks[k][i].astype(object) if ks is not None else 0)
            out[k][i] = np.array([int(x) % q for x in v], dtype=np.uint64)
    return out
