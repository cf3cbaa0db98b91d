/*
 * ckks_oracle.c -- CPU restatement of the SEAL-4.1-bs CKKS arithmetic that
 * NEXUS's CKKSEvaluator drives (the reference's bit-exact oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load it.
 * The product (libnexus_ckks.so) never links, loads or calls it.
 *
 * Arithmetic is written the plainest way possible (unsigned __int128 products
 * and `%`), never with the lazy Harvey/Shoup/Barrett reductions the GPU
 * kernels use: every SEAL routine restated here produces a value that is
 * fully determined modulo q, so a plain formulation is the cleanest oracle.
 * Where SEAL's floating-point order of operations matters (CKKS encode/decode)
 * the same order is kept and the file is compiled with -ffp-contract=off.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * thirdparty/SEAL-4.1-bs/native/src/seal/).
 *
 * Pins (tests/test_oracle_pins.py): SEAL's own KATs (tests/seal/util/ntt.cpp:53-97,
 * galois.cpp:28-113, rns.cpp:1013-1075), O(N^2) brute-force negacyclic evaluation,
 * exact big-integer CRT closed forms for rescale / ModUp / ModDown, coefficient-
 * domain automorphisms, decryption invariants, and mpmath high-precision DFTs
 * for encode/decode.  The encode rounding at exact .5 ties is the one place no
 * check pins beyond the high-precision DFT ("parity unpinned at ties").
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* scalar modular arithmetic (util/uintarithsmallmod.h:167-330)       */
/* ------------------------------------------------------------------ */
static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }
static inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) { u128 s = (u128)a + b; return (uint64_t)(s % q); }
static inline uint64_t submod(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : q - (b - a); }

uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q) { return mulmod(a % q, b % q, q); }

uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q)
{
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

/* modular inverse via extended Euclid (util/numth.h try_invert_uint_mod) */
uint64_t or_invmod(uint64_t a, uint64_t q)
{
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr) {
        __int128 quo = r / nr, tmp;
        tmp = t - quo * nt; t = nt; nt = tmp;
        tmp = r - quo * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += q;
    return (uint64_t)t;
}

static uint32_t reverse_bits32(uint32_t x, int bits)
{
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* ------------------------------------------------------------------ */
/* primes (util/numth.cpp:180-310, modulus.cpp:143-184)               */
/* ------------------------------------------------------------------ */
/* Deterministic Miller-Rabin: the first 12 prime bases decide every
 * 64-bit integer, so the answer equals SEAL's randomised test (numth.cpp:237). */
int or_is_prime(uint64_t n)
{
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n == small[i]) return 1;
        if (n % small[i] == 0) return 0;
    }
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) { d >>= 1; r++; }
    for (int i = 0; i < 12; i++) {
        uint64_t x = or_powmod(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int j = 1; j < r; j++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* numth.cpp:278-310: descending primes = 1 mod factor below 2^bit_size */
int or_get_primes(uint64_t factor, int bit_size, size_t count, uint64_t *out)
{
    uint64_t value = ((((uint64_t)1) << bit_size) - 1) / factor * factor + 1;
    uint64_t lower = ((uint64_t)1) << (bit_size - 1);
    size_t found = 0;
    while (found < count && value > lower) {
        if (or_is_prime(value)) out[found++] = value;
        value -= factor;
    }
    return found == count ? 0 : -1;
}

/* modulus.cpp:143-184: for each distinct bit size take `count` primes
 * (descending) and hand them out from the back, i.e. the first occurrence of a
 * bit size receives the smallest of the primes found. */
int or_coeff_modulus_create(size_t n, const int *bits, size_t k, uint64_t *out)
{
    uint64_t factor = 2 * (uint64_t)n;
    for (size_t i = 0; i < k; i++) out[i] = 0;
    for (size_t i = 0; i < k; i++) {
        if (out[i]) continue;
        size_t cnt = 0;
        for (size_t j = 0; j < k; j++) cnt += (bits[j] == bits[i]);
        uint64_t *pr = (uint64_t *)malloc(cnt * sizeof(uint64_t));
        if (or_get_primes(factor, bits[i], cnt, pr)) { free(pr); return -1; }
        size_t take = cnt;
        for (size_t j = 0; j < k; j++)
            if (bits[j] == bits[i]) out[j] = pr[--take];
        free(pr);
    }
    return 0;
}

/* numth.cpp:313-412: minimal primitive `degree`-th root of unity mod q.
 * A primitive root r satisfies r^(degree/2) = -1; the minimal one is the
 * smallest among the odd powers of any primitive root. */
uint64_t or_minimal_primitive_root(uint64_t degree, uint64_t q)
{
    if ((q - 1) % degree) return 0;
    uint64_t quot = (q - 1) / degree, root = 0;
    for (uint64_t g = 2; g < q; g++) {
        uint64_t c = or_powmod(g, quot, q);
        if (or_powmod(c, degree >> 1, q) == q - 1) { root = c; break; }
    }
    if (!root) return 0;
    uint64_t gsq = mulmod(root, root, q), cur = root, best = root;
    for (uint64_t i = 0; i < degree; i += 2) {
        if (cur < best) best = cur;
        cur = mulmod(cur, gsq, q);
    }
    return best;
}

/* ------------------------------------------------------------------ */
/* NTT tables and transforms (util/ntt.cpp:241-300, dwthandler.h)     */
/* ------------------------------------------------------------------ */
/* root_powers[brv(i)] = psi^i (i>=1), root_powers[0]=1;
 * inv_root_powers[brv(i-1)+1] = psi^-i, inv_root_powers[0]=1  (ntt.cpp:265-289) */
int or_ntt_tables(int logn, uint64_t q, uint64_t *root_powers, uint64_t *inv_root_powers)
{
    size_t n = (size_t)1 << logn;
    uint64_t psi = or_minimal_primitive_root(2 * n, q);
    if (!psi) return -1;
    uint64_t ipsi = or_invmod(psi, q);
    uint64_t p = psi;
    for (size_t i = 1; i < n; i++) {
        root_powers[reverse_bits32((uint32_t)i, logn)] = p;
        p = mulmod(p, psi, q);
    }
    root_powers[0] = 1;
    p = ipsi;
    for (size_t i = 1; i < n; i++) {
        inv_root_powers[reverse_bits32((uint32_t)(i - 1), logn) + 1] = p;
        p = mulmod(p, ipsi, q);
    }
    inv_root_powers[0] = 1;
    return 0;
}

/* dwthandler.h:94-200 transform_to_rev + ntt.cpp:394-420 final reduction:
 * natural-order input, bit-reversed output, values in [0,q). */
void or_ntt_forward(uint64_t *a, int logn, uint64_t q, const uint64_t *rp)
{
    size_t n = (size_t)1 << logn;
    size_t gap = n >> 1, idx = 0;
    for (size_t m = 1; m < n; m <<= 1) {
        for (size_t i = 0; i < m; i++) {
            uint64_t r = rp[++idx];
            uint64_t *x = a + 2 * i * gap, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                uint64_t u = x[j] % q, v = mulmod(y[j] % q, r, q);
                x[j] = addmod(u, v, q);
                y[j] = submod(u, v, q);
            }
        }
        gap >>= 1;
    }
}

/* dwthandler.h:202-360 transform_from_rev with n^-1 + ntt.cpp:439-477:
 * bit-reversed input, natural-order output, values in [0,q). */
void or_ntt_inverse(uint64_t *a, int logn, uint64_t q, const uint64_t *irp)
{
    size_t n = (size_t)1 << logn;
    size_t gap = 1, idx = 0;
    for (size_t m = n >> 1; m >= 1; m >>= 1) {
        for (size_t i = 0; i < m; i++) {
            uint64_t r = irp[++idx];
            uint64_t *x = a + 2 * i * gap, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                uint64_t u = x[j] % q, v = y[j] % q;
                x[j] = addmod(u, v, q);
                y[j] = mulmod(submod(u, v, q), r, q);
            }
        }
        gap <<= 1;
        if (m == 1) break;
    }
    uint64_t ninv = or_invmod((uint64_t)(n % q), q);
    for (size_t j = 0; j < n; j++) a[j] = mulmod(a[j], ninv, q);
}

/* ------------------------------------------------------------------ */
/* Galois automorphisms (util/galois.cpp:18-218)                      */
/* ------------------------------------------------------------------ */
/* galois.cpp:53-90: rotation step -> Galois element 5^step mod 2N.  The bs fork
 * sets GaloisTool::generator_ = 5 (util/galois.h:169), matching the encoder's
 * slot order (ckks.cpp:36). */
uint32_t or_elt_from_step(int logn, int step)
{
    uint32_t n = 1u << logn, m = 2 * n;
    if (step == 0) return m - 1;
    int sign = step < 0;
    uint32_t pos = (uint32_t)(sign ? -step : step);
    if (pos >= (n >> 1)) return 0;
    pos &= m - 1;
    int s = sign ? (int)(n >> 1) - (int)pos : (int)pos;
    uint64_t e = 1;
    while (s--) e = (e * 5) & (m - 1);
    return (uint32_t)e;
}

/* galois.cpp:18-51: NTT-domain permutation table */
void or_galois_table_ntt(int logn, uint32_t elt, uint32_t *table)
{
    uint32_t n = 1u << logn, mask = n - 1;
    for (uint32_t i = n; i < 2 * n; i++) {
        uint32_t rev = reverse_bits32(i, logn + 1);
        uint64_t raw = ((uint64_t)elt * rev) >> 1;
        table[i - n] = reverse_bits32((uint32_t)(raw & mask), logn);
    }
}

/* galois.cpp:192-218: out[i] = in[table[i]] */
void or_apply_galois_ntt(const uint64_t *in, int logn, uint32_t elt, uint64_t *out)
{
    size_t n = (size_t)1 << logn;
    uint32_t *t = (uint32_t *)malloc(n * sizeof(uint32_t));
    or_galois_table_ntt(logn, elt, t);
    for (size_t i = 0; i < n; i++) out[i] = in[t[i]];
    free(t);
}

/* galois.cpp:148-190: coefficient-domain x -> x^elt with negacyclic sign */
void or_apply_galois_coeff(const uint64_t *in, int logn, uint32_t elt, uint64_t q, uint64_t *out)
{
    uint64_t n = (uint64_t)1 << logn, raw = 0;
    for (uint64_t i = 0; i < n; i++, raw += elt) {
        uint64_t idx = raw & (n - 1), v = in[i];
        if ((raw >> logn) & 1) v = v ? q - v : 0;
        out[idx] = v;
    }
}

/* ------------------------------------------------------------------ */
/* Rescale (util/rns.cpp:830-905)                                     */
/* ------------------------------------------------------------------ */
/* In-place on one polynomial of `l` NTT-form limbs; limbs 0..l-2 receive
 * round(x / q_last) (q_last/2 added for rounding).  tables: rp[i], irp[i]
 * are the NTT tables of moduli[i] (each n words). */
void or_rescale_poly_ntt(uint64_t *poly, size_t l, int logn, const uint64_t *moduli, const uint64_t *const *rp,
                         const uint64_t *const *irp)
{
    size_t n = (size_t)1 << logn;
    uint64_t ql = moduli[l - 1], half = ql >> 1;
    uint64_t *last = poly + (l - 1) * n;
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    or_ntt_inverse(last, logn, ql, irp[l - 1]);
    for (size_t j = 0; j < n; j++) last[j] = addmod(last[j], half, ql);
    for (size_t i = 0; i + 1 < l; i++) {
        uint64_t qi = moduli[i];
        uint64_t neg_half = submod(0, half % qi, qi);
        for (size_t j = 0; j < n; j++) tmp[j] = addmod(last[j] % qi, neg_half, qi);
        or_ntt_forward(tmp, logn, qi, rp[i]);
        uint64_t inv = or_invmod(ql % qi, qi);
        uint64_t *ci = poly + i * n;
        for (size_t j = 0; j < n; j++) ci[j] = mulmod(submod(ci[j], tmp[j], qi), inv, qi);
    }
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* Key switching (evaluator.cpp:2724-3021, CKKS branch)               */
/* ------------------------------------------------------------------ */
/*
 * ct      : destination ciphertext, 2 polys x l limbs x n (NTT), updated in place
 * target  : l limbs x n (NTT form) -- the polynomial being switched
 * key_mod : L+1 key-level moduli (q_0..q_{L-1}, p)
 * key     : [L digits][2 polys][L+1 limbs][n]  (KSwitchKeys, keygenerator.cpp:303-336)
 * rp/irp  : NTT tables for the L+1 key moduli
 */
/* ModUp of `target` and its inner product with `key` (evaluator.cpp:2812-2910),
 * accumulated into acc [2][l+1][n] (output primes q_0..q_{l-1}, p).  With a
 * Galois table `perm` (NTT domain, out[i] = in[perm[i]]) every lifted digit is
 * permuted before the product: the digits of a hoisted rotation are
 * sigma(lift(target_j)) (Phantom hoisting_inplace, evaluate.cu:1727-1920). */
static void ks_accumulate(uint64_t *acc, const uint64_t *target, size_t l, int logn, const uint64_t *key_mod,
                          size_t L1, const uint64_t *key, const uint32_t *perm, const uint64_t *const *rp,
                          const uint64_t *const *irp)
{
    size_t n = (size_t)1 << logn, L = L1 - 1;
    /* t_target = INTT(target)   (evaluator.cpp:2812-2818) */
    uint64_t *tt = (uint64_t *)malloc(l * n * sizeof(uint64_t));
    memcpy(tt, target, l * n * sizeof(uint64_t));
    for (size_t j = 0; j < l; j++) or_ntt_inverse(tt + j * n, logn, key_mod[j], irp[j]);
    uint64_t *tn = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *tp = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (size_t OI = 0; OI <= l; OI++) {
        size_t ki = (OI == l) ? L : OI;
        uint64_t qI = key_mod[ki];
        for (size_t DJ = 0; DJ < l; DJ++) {
            const uint64_t *op;
            if (OI == DJ) {
                op = target + DJ * n; /* NTT form already present (:2836-2839) */
            } else {
                for (size_t c = 0; c < n; c++) tn[c] = tt[DJ * n + c] % qI; /* (:2843-2852) */
                or_ntt_forward(tn, logn, qI, rp[ki]);
                op = tn;
            }
            if (perm) {
                for (size_t c = 0; c < n; c++) tp[c] = op[perm[c]];
                op = tp;
            }
            for (size_t k = 0; k < 2; k++) {
                const uint64_t *kp = key + ((DJ * 2 + k) * L1 + ki) * n;
                uint64_t *ap = acc + (k * (l + 1) + OI) * n;
                for (size_t c = 0; c < n; c++) ap[c] = addmod(ap[c], mulmod(op[c], kp[c], qI), qI);
            }
        }
    }
    free(tp);
    free(tn);
    free(tt);
}

/* ModDown with rounding (evaluator.cpp:2953-3018): ct[k][i] += (acc_i -
 * [acc_p + (p-1)/2]_qi + [(p-1)/2]_qi) p^-1 for both polys. */
static void ks_moddown(uint64_t *ct, uint64_t *acc, size_t l, int logn, const uint64_t *key_mod, size_t L1,
                       const uint64_t *const *rp, const uint64_t *const *irp)
{
    size_t n = (size_t)1 << logn, L = L1 - 1;
    uint64_t p = key_mod[L];
    uint64_t half = p >> 1;
    uint64_t *tn = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (size_t k = 0; k < 2; k++) {
        uint64_t *last = acc + (k * (l + 1) + l) * n;
        or_ntt_inverse(last, logn, p, irp[L]);
        for (size_t c = 0; c < n; c++) last[c] = addmod(last[c], half, p);
        for (size_t i = 0; i < l; i++) {
            uint64_t qi = key_mod[i];
            uint64_t fix = submod(0, half % qi, qi);
            for (size_t c = 0; c < n; c++) tn[c] = addmod(last[c] % qi, fix, qi);
            or_ntt_forward(tn, logn, qi, rp[i]);
            uint64_t pinv = or_invmod(p % qi, qi);
            uint64_t *ai = acc + (k * (l + 1) + i) * n;
            uint64_t *ci = ct + (k * l + i) * n;
            for (size_t c = 0; c < n; c++) ci[c] = addmod(ci[c], mulmod(submod(ai[c], tn[c], qi), pinv, qi), qi);
        }
    }
    free(tn);
}

void or_switch_key(uint64_t *ct, const uint64_t *target, size_t l, int logn, const uint64_t *key_mod, size_t L1,
                   const uint64_t *key, const uint64_t *const *rp, const uint64_t *const *irp)
{
    size_t n = (size_t)1 << logn;
    uint64_t *acc = (uint64_t *)calloc(2 * (l + 1) * n, sizeof(uint64_t));
    ks_accumulate(acc, target, l, logn, key_mod, L1, key, NULL, rp, irp);
    ks_moddown(ct, acc, l, logn, key_mod, L1, rp, irp);
    free(acc);
}

/* Sum of rotations with one ModDown (hoisted key switching; Phantom
 * hoisting_inplace, evaluate.cu:1727-1920, and the BSGS giant-step sums of
 * Bootstrapper.cu:1855-1919): out = sum_r rot_r(cts[r]) where rot_r is the
 * automorphism elts[r] (0 = identity: the input itself) applied as
 *   sigma_r(c0_r) + ModDown(sum_r sum_j sigma_r(lift(c1_r mod q_j)) key_r[j]).
 * cts: count ciphertexts of 2 x l x n (NTT), keys: one key per term (unused
 * for identity terms); out: 2 x l x n, overwritten.  With count = 1 it is
 * one hoisted rotation; the same input repeated is Phantom's hoisting sum. */
void or_rotate_sum(uint64_t *out, const uint64_t *const *cts, const uint32_t *elts, const uint64_t *const *keys,
                   size_t count, size_t l, int logn, const uint64_t *key_mod, size_t L1, const uint64_t *const *rp,
                   const uint64_t *const *irp)
{
    size_t n = (size_t)1 << logn;
    uint64_t *acc = (uint64_t *)calloc(2 * (l + 1) * n, sizeof(uint64_t));
    uint32_t *perm = (uint32_t *)malloc(n * sizeof(uint32_t));
    memset(out, 0, 2 * l * n * sizeof(uint64_t));
    int switched = 0;
    for (size_t r = 0; r < count; r++) {
        const uint64_t *c0 = cts[r], *c1 = cts[r] + l * n;
        if (elts[r] == 0) {
            for (size_t i = 0; i < 2 * l; i++)
                for (size_t c = 0; c < n; c++)
                    out[i * n + c] = addmod(out[i * n + c], cts[r][i * n + c], key_mod[i % l]);
            continue;
        }
        or_galois_table_ntt(logn, elts[r], perm);
        for (size_t i = 0; i < l; i++)
            for (size_t c = 0; c < n; c++) out[i * n + c] = addmod(out[i * n + c], c0[i * n + perm[c]], key_mod[i]);
        ks_accumulate(acc, c1, l, logn, key_mod, L1, keys[r], perm, rp, irp);
        switched = 1;
    }
    if (switched) ks_moddown(out, acc, l, logn, key_mod, L1, rp, irp);
    free(perm);
    free(acc);
}

/* ------------------------------------------------------------------ */
/* Dyadic ops (evaluator.cpp:770-909 ckks_multiply, :2336-2373)       */
/* ------------------------------------------------------------------ */
/* a,b: 2 polys x l limbs; out: 3 polys x l limbs: (a0b0, a0b1+a1b0, a1b1) */
void or_tensor(const uint64_t *a, const uint64_t *b, size_t l, int logn, const uint64_t *moduli, uint64_t *out)
{
    size_t n = (size_t)1 << logn;
    for (size_t i = 0; i < l; i++) {
        uint64_t q = moduli[i];
        for (size_t c = 0; c < n; c++) {
            uint64_t a0 = a[i * n + c], a1 = a[(l + i) * n + c];
            uint64_t b0 = b[i * n + c], b1 = b[(l + i) * n + c];
            out[i * n + c] = mulmod(a0, b0, q);
            out[(l + i) * n + c] = addmod(mulmod(a0, b1, q), mulmod(a1, b0, q), q);
            out[(2 * l + i) * n + c] = mulmod(a1, b1, q);
        }
    }
}

/* pointwise mod ops over `polys` polys of l limbs each */
void or_dyadic_mul(const uint64_t *a, const uint64_t *b, size_t polys, size_t l, int logn, const uint64_t *moduli,
                   uint64_t *out)
{
    size_t n = (size_t)1 << logn;
    for (size_t p = 0; p < polys; p++)
        for (size_t i = 0; i < l; i++)
            for (size_t c = 0; c < n; c++) {
                size_t o = (p * l + i) * n + c;
                out[o] = mulmod(a[o], b[i * n + c], moduli[i]);
            }
}

/* ------------------------------------------------------------------ */
/* CKKS encoder (ckks.cpp:13-71, ckks.h:457-760, croots.cpp)          */
/* ------------------------------------------------------------------ */
static const double OR_PI = 3.1415926535897932384626433832795028842;

typedef struct {
    int logn;
    size_t n, slots;
    size_t *index_map;      /* matrix_reps_index_map_ (ckks.cpp:33-50) */
    double complex *roots;  /* root_powers_ */
    double complex *iroots; /* inv_root_powers_ */
} or_encoder;

static double complex croot_base(const double complex *tab, size_t m, size_t index)
{
    /* croots.cpp:44-73: 8-fold symmetry of the m-th roots */
    index &= m - 1;
    if (index <= m / 8) return tab[index];
    if (index <= m / 4) { double complex a = tab[m / 4 - index]; return cimag(a) + I * creal(a); }
    if (index <= m / 2) { double complex a = croot_base(tab, m, m / 2 - index); return -creal(a) + I * cimag(a); }
    if (index <= 3 * m / 4) { double complex a = croot_base(tab, m, index - m / 2); return -creal(a) - I * cimag(a); }
    { double complex a = croot_base(tab, m, m - index); return creal(a) - I * cimag(a); }
}

void *or_encoder_create(int logn)
{
    or_encoder *e = (or_encoder *)calloc(1, sizeof(or_encoder));
    e->logn = logn;
    e->n = (size_t)1 << logn;
    e->slots = e->n >> 1;
    size_t n = e->n, m = 2 * n;
    e->index_map = (size_t *)malloc(n * sizeof(size_t));
    uint64_t pos = 1;
    for (size_t i = 0; i < e->slots; i++) {
        uint64_t i1 = (pos - 1) >> 1, i2 = (m - pos - 1) >> 1;
        e->index_map[i] = reverse_bits32((uint32_t)i1, logn);
        e->index_map[e->slots | i] = reverse_bits32((uint32_t)i2, logn);
        pos = (pos * 5) & (m - 1);
    }
    double complex *tab = (double complex *)malloc((m / 8 + 1) * sizeof(double complex));
    for (size_t i = 0; i <= m / 8; i++) {
        double ang = 2 * OR_PI * (double)i / (double)m;
        tab[i] = cos(ang) + I * sin(ang);
    }
    e->roots = (double complex *)calloc(n, sizeof(double complex));
    e->iroots = (double complex *)calloc(n, sizeof(double complex));
    for (size_t i = 1; i < n; i++) {
        e->roots[i] = croot_base(tab, m, reverse_bits32((uint32_t)i, logn));
        e->iroots[i] = conj(croot_base(tab, m, reverse_bits32((uint32_t)(i - 1), logn) + 1));
    }
    free(tab);
    return e;
}

void or_encoder_destroy(void *h)
{
    or_encoder *e = (or_encoder *)h;
    free(e->index_map);
    free(e->roots);
    free(e->iroots);
    free(e);
}

/* complex product written out as (ac-bd) + (ad+bc)i, the order GCC emits for
 * std::complex<double> multiplication of finite values */
static inline double complex cmul(double complex a, double complex b)
{
    double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    return (ar * br - ai * bi) + I * (ar * bi + ai * br);
}

/* dwthandler.h transform_from_rev for std::complex<double> with a real scalar */
static void fft_from_rev(double complex *v, int logn, const double complex *roots, double scalar)
{
    size_t n = (size_t)1 << logn, gap = 1, ridx = 0;
    size_t m = n >> 1;
    for (; m > 1; m >>= 1) {
        size_t off = 0;
        for (size_t i = 0; i < m; i++) {
            double complex r = roots[++ridx];
            double complex *x = v + off, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                double complex u = x[j], w = y[j];
                x[j] = u + w;
                y[j] = cmul(u - w, r);
            }
            off += gap << 1;
        }
        gap <<= 1;
    }
    double complex r = roots[++ridx];
    double complex sr = creal(r) * scalar + I * (cimag(r) * scalar);
    double complex *x = v, *y = v + gap;
    for (size_t j = 0; j < gap; j++) {
        double complex u = x[j], w = y[j];
        double complex s = u + w;
        x[j] = creal(s) * scalar + I * (cimag(s) * scalar);
        y[j] = cmul(u - w, sr);
    }
}

/* dwthandler.h transform_to_rev for std::complex<double>, no scalar */
static void fft_to_rev(double complex *v, int logn, const double complex *roots)
{
    size_t n = (size_t)1 << logn, gap = n >> 1, ridx = 0, m = 1;
    for (; m < (n >> 1); m <<= 1) {
        size_t off = 0;
        for (size_t i = 0; i < m; i++) {
            double complex r = roots[++ridx];
            double complex *x = v + off, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                double complex u = x[j], w = cmul(y[j], r);
                x[j] = u + w;
                y[j] = u - w;
            }
            off += gap << 1;
        }
        gap >>= 1;
    }
    for (size_t i = 0; i < m; i++) {
        double complex r = roots[++ridx];
        double complex u = v[0], w = cmul(v[1], r);
        v[0] = u + w;
        v[1] = u - w;
        v += 2;
    }
}

/* ckks.h:457-620 encode_internal for complex values; coefficients whose
 * magnitude fits 64 or 128 bits (the only cases a 2^40..2^60 scale reaches
 * for |v| <= 2^20).  out: l limbs x n, NTT form.  Returns -1 if too large. */
int or_ckks_encode(void *h, const double *re, const double *im, size_t values_size, double scale,
                   const uint64_t *moduli, size_t l, const uint64_t *const *rp, uint64_t *out)
{
    or_encoder *e = (or_encoder *)h;
    size_t n = e->n, slots = e->slots;
    double complex *cv = (double complex *)calloc(n, sizeof(double complex));
    for (size_t i = 0; i < values_size; i++) {
        double complex v = re[i] + I * (im ? im[i] : 0.0);
        cv[e->index_map[i]] = v;
        cv[e->index_map[i + slots]] = conj(v);
    }
    double fix = scale / (double)n;
    fft_from_rev(cv, e->logn, e->iroots, fix);
    double max_coeff = 0;
    for (size_t i = 0; i < n; i++) max_coeff = fmax(max_coeff, fabs(creal(cv[i])));
    int bits = (int)ceil(log2(fmax(max_coeff, 1.0))) + 1;
    if (bits > 128) { free(cv); return -1; }
    double two64 = pow(2.0, 64);
    for (size_t i = 0; i < n; i++) {
        double cd = round(creal(cv[i]));
        int neg = signbit(cd);
        cd = fabs(cd);
        u128 cu;
        if (bits <= 64) cu = (u128)(uint64_t)cd;
        else cu = ((u128)(uint64_t)(cd / two64) << 64) | (uint64_t)fmod(cd, two64);
        for (size_t j = 0; j < l; j++) {
            uint64_t r = (uint64_t)(cu % moduli[j]);
            out[j * n + i] = neg ? (r ? moduli[j] - r : 0) : r;
        }
    }
    for (size_t j = 0; j < l; j++) or_ntt_forward(out + j * n, e->logn, moduli[j], rp[j]);
    free(cv);
    return 0;
}

/* ckks.h:457-540 encode_internal up to the rounding: the real parts of the
 * scaled inverse transform (the unrounded plaintext coefficients).  Checks of
 * coefficients beyond 128 bits (SEAL's multi-word decomposition, ckks.h:
 * 590-610) reduce these with exact big integers. */
void or_ckks_encode_coeffs(void *h, const double *re, const double *im, size_t values_size, double scale,
                           double *out)
{
    or_encoder *e = (or_encoder *)h;
    size_t n = e->n, slots = e->slots;
    double complex *cv = (double complex *)calloc(n, sizeof(double complex));
    for (size_t i = 0; i < values_size; i++) {
        double complex v = re[i] + I * (im ? im[i] : 0.0);
        cv[e->index_map[i]] = v;
        cv[e->index_map[i + slots]] = conj(v);
    }
    fft_from_rev(cv, e->logn, e->iroots, scale / (double)n);
    for (size_t i = 0; i < n; i++) out[i] = creal(cv[i]);
    free(cv);
}

/* ckks.cpp:73-...: constant encode; every coefficient-domain limb holds
 * round(value*scale) mod q_j in coefficient 0 only, i.e. the constant in every
 * NTT slot.  Returns -1 when |value*scale| needs more than 128 bits. */
int or_ckks_encode_const(double value, double scale, const uint64_t *moduli, size_t l, int logn, uint64_t *out)
{
    size_t n = (size_t)1 << logn;
    value *= scale;
    int bits = (int)log2(fabs(value)) + 2;
    if (bits > 128) return -1;
    double two64 = pow(2.0, 64);
    double cd = round(value);
    int neg = signbit(cd);
    cd = fabs(cd);
    u128 cu;
    if (bits <= 64) cu = (u128)(uint64_t)cd;
    else cu = ((u128)(uint64_t)(cd / two64) << 64) | (uint64_t)fmod(cd, two64);
    for (size_t j = 0; j < l; j++) {
        uint64_t r = (uint64_t)(cu % moduli[j]);
        uint64_t v = neg ? (r ? moduli[j] - r : 0) : r;
        for (size_t c = 0; c < n; c++) out[j * n + c] = v;
    }
    return 0;
}

/* ---- multi-precision helpers for CRT composition (rns.cpp:354-400) ---- */
/* Little-endian words.  Q = prod moduli.  x = sum_i [x_i * (Q/q_i)^-1]_{q_i} * (Q/q_i) mod Q */
static void mp_mul_word(const uint64_t *a, size_t w, uint64_t b, uint64_t *out /* w+1 words */)
{
    u128 carry = 0;
    for (size_t i = 0; i < w; i++) {
        u128 t = (u128)a[i] * b + carry;
        out[i] = (uint64_t)t;
        carry = t >> 64;
    }
    out[w] = (uint64_t)carry;
}
static int mp_geq(const uint64_t *a, const uint64_t *b, size_t w)
{
    for (size_t i = w; i-- > 0;) {
        if (a[i] != b[i]) return a[i] > b[i];
    }
    return 1;
}
static void mp_sub(uint64_t *a, const uint64_t *b, size_t w)
{
    uint64_t borrow = 0;
    for (size_t i = 0; i < w; i++) {
        uint64_t bi = b[i] + borrow;
        uint64_t nb = (bi < borrow) || (a[i] < bi);
        a[i] -= bi;
        borrow = nb;
    }
}
static void mp_add(uint64_t *a, const uint64_t *b, size_t w)
{
    uint64_t carry = 0;
    for (size_t i = 0; i < w; i++) {
        u128 s = (u128)a[i] + b[i] + carry;
        a[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
}

/* coeff-domain limbs (l x n) -> n multi-precision integers of l words each in [0,Q) */
void or_crt_compose(const uint64_t *limbs, size_t l, size_t n, const uint64_t *moduli, uint64_t *out)
{
    uint64_t *Q = (uint64_t *)calloc(l + 1, sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)calloc(l + 1, sizeof(uint64_t));
    Q[0] = 1;
    for (size_t i = 0; i < l; i++) {
        mp_mul_word(Q, l, moduli[i], tmp);
        memcpy(Q, tmp, l * sizeof(uint64_t));
    }
    uint64_t *punct = (uint64_t *)calloc(l * (l + 1), sizeof(uint64_t));
    uint64_t *invp = (uint64_t *)malloc(l * sizeof(uint64_t));
    for (size_t i = 0; i < l; i++) {
        uint64_t *pi = punct + i * (l + 1);
        pi[0] = 1;
        for (size_t j = 0; j < l; j++) {
            if (j == i) continue;
            mp_mul_word(pi, l, moduli[j], tmp);
            memcpy(pi, tmp, l * sizeof(uint64_t));
        }
        uint64_t r = 1;
        for (size_t j = 0; j < l; j++)
            if (j != i) r = mulmod(r, moduli[j] % moduli[i], moduli[i]);
        invp[i] = or_invmod(r, moduli[i]);
    }
    uint64_t *accw = (uint64_t *)calloc(l + 1, sizeof(uint64_t));
    for (size_t c = 0; c < n; c++) {
        memset(accw, 0, (l + 1) * sizeof(uint64_t));
        for (size_t i = 0; i < l; i++) {
            uint64_t t = mulmod(limbs[i * n + c], invp[i], moduli[i]);
            mp_mul_word(punct + i * (l + 1), l, t, tmp);
            mp_add(accw, tmp, l + 1);
            if (mp_geq(accw, Q, l + 1)) mp_sub(accw, Q, l + 1);
        }
        memcpy(out + c * l, accw, l * sizeof(uint64_t));
    }
    free(accw);
    free(invp);
    free(punct);
    free(tmp);
    free(Q);
}

/* ckks.h:624-735 decode_internal (full or sparse slots). `plain` is NTT form
 * l x n; out_re/out_im receive `sparse_slots` values. */
void or_ckks_decode(void *h, const uint64_t *plain, size_t l, const uint64_t *moduli, const uint64_t *const *irp,
                    double scale, size_t sparse_slots, double *out_re, double *out_im)
{
    or_encoder *e = (or_encoder *)h;
    size_t n = e->n;
    uint64_t *cp = (uint64_t *)malloc(l * n * sizeof(uint64_t));
    memcpy(cp, plain, l * n * sizeof(uint64_t));
    for (size_t j = 0; j < l; j++) or_ntt_inverse(cp + j * n, e->logn, moduli[j], irp[j]);
    uint64_t *mp = (uint64_t *)malloc(l * n * sizeof(uint64_t));
    or_crt_compose(cp, l, n, moduli, mp);
    if (sparse_slots && sparse_slots != e->slots) {
        size_t sparsity = e->slots / sparse_slots;
        for (size_t i = 0; i < n; i++)
            if (((i - 1) & (sparsity - 1)) != sparsity - 1)
                for (size_t j = 0; j < l; j++) mp[i * l + j] = 0;
    } else {
        sparse_slots = e->slots;
    }
    /* Q and upper_half_threshold = ceil(Q/2) */
    uint64_t *Q = (uint64_t *)calloc(l + 1, sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)calloc(l + 1, sizeof(uint64_t));
    Q[0] = 1;
    for (size_t i = 0; i < l; i++) {
        mp_mul_word(Q, l, moduli[i], tmp);
        memcpy(Q, tmp, l * sizeof(uint64_t));
    }
    uint64_t *thr = (uint64_t *)calloc(l, sizeof(uint64_t));
    /* (Q + 1) / 2 */
    {
        uint64_t one[64] = {0};
        one[0] = 1;
        memcpy(thr, Q, l * sizeof(uint64_t));
        mp_add(thr, one, l);
        for (size_t i = 0; i < l; i++) thr[i] = (thr[i] >> 1) | (i + 1 < l ? (thr[i + 1] << 63) : 0);
    }
    double inv_scale = 1.0 / scale, two64 = pow(2.0, 64);
    double complex *res = (double complex *)malloc(n * sizeof(double complex));
    for (size_t i = 0; i < n; i++) {
        double r = 0.0;
        const uint64_t *x = mp + i * l;
        double s64 = inv_scale;
        if (mp_geq(x, thr, l)) {
            for (size_t j = 0; j < l; j++, s64 *= two64) {
                if (x[j] > Q[j]) {
                    uint64_t d = x[j] - Q[j];
                    r += d ? (double)d * s64 : 0.0;
                } else {
                    uint64_t d = Q[j] - x[j];
                    r -= d ? (double)d * s64 : 0.0;
                }
            }
        } else {
            for (size_t j = 0; j < l; j++, s64 *= two64) r += x[j] ? (double)x[j] * s64 : 0.0;
        }
        res[i] = r;
    }
    fft_to_rev(res, e->logn, e->roots);
    for (size_t i = 0; i < sparse_slots; i++) {
        double complex v = res[e->index_map[i]];
        out_re[i] = creal(v);
        if (out_im) out_im[i] = cimag(v);
    }
    free(res);
    free(thr);
    free(tmp);
    free(Q);
    free(mp);
    free(cp);
}
