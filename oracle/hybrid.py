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


def switch_key_hybrid(target, key, moduli, L, alpha, logn):
    """Key switch of `target` ([l][n], NTT) with key [dnum][2][L+alpha][n]:
    returns (ks0, ks1), each [l][n] NTT, to be added to (c0, c1)."""
    l = target.shape[0]
    n = 1 << logn
    qs = [int(q) for q in moduli[:l]]
    ps = [int(p) for p in moduli[L:L + alpha]]
    outs = qs + ps
    out_key_limb = list(range(l)) + [L + k for k in range(alpha)]
    coef = [orc.ntt_inverse(target[i].copy(), logn, qs[i]) for i in range(l)]
    digits = (l + alpha - 1) // alpha
    # ModUp: T[o][j] in NTT form
    T = [[None] * digits for _ in outs]
    for j in range(digits):
        idx = list(range(j * alpha, min(l, (j + 1) * alpha)))
        src = [coef[i] for i in idx]
        src_mods = [qs[i] for i in idx]
        for o, t in enumerate(outs):
            T[o][j] = orc.ntt_forward(_conv(src, src_mods, t), logn, t)
    # inner product
    acc = [[None] * len(outs) for _ in range(2)]
    for k in range(2):
        for o, t in enumerate(outs):
            s = np.zeros(n, dtype=object)
            for j in range(digits):
                s = s + T[o][j].astype(object) * key[j, k, out_key_limb[o]].astype(object)
            acc[k][o] = np.array([int(v) % t for v in s], dtype=np.uint64)
    # ModDown: (acc_q - centred(INTT acc_P)) * P^-1 mod q_i = round(acc / P)
    res = []
    for k in range(2):
        yP = [orc.ntt_inverse(acc[k][l + m].copy(), logn, ps[m]) for m in range(alpha)]
        polys = []
        for i, q in enumerate(qs):
            z = orc.ntt_forward(_conv_centred(yP, ps, q), logn, q)
            P = 1
            for p in ps:
                P = P * p % q
            pinv = pow(P, -1, q)
            d = (acc[k][i].astype(object) - z.astype(object)) % q
            polys.append(np.array([int(v) * pinv % q for v in d], dtype=np.uint64))
        res.append(np.array(polys))
    return res[0], res[1]
