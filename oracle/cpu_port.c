/*
 * cpu_port.c -- performance-faithful CPU port of the SEAL-4.1-bs arithmetic
 * that the reference's CPU path (src/ over thirdparty/SEAL-4.1-bs) runs, used
 * ONLY as the timed CPU baseline of bench.py (and checked word for word
 * against ckks_oracle.c by tests/test_oracle_port.py).  Never linked into the
 * product.
 *
 * Where ckks_oracle.c is written for clarity (every product reduced with a
 * 128-bit '%'), this file uses the methods SEAL itself uses, so that its
 * timings stand for SEAL's CPU cost:
 *   - Harvey's lazy NTT butterflies with Shoup-precomputed twiddles (values
 *     kept in [0, 4q), reduced once at the end; util/dwthandler.h:94-360,
 *     util/ntt.cpp:394-477);
 *   - Shoup multiplication by precomputed operands (util/uintarithsmallmod.h
 *     MultiplyUIntModOperand) and Barrett reduction of 128-bit sums with
 *     floor(2^128 / q) (Barrett 1986; util/uintarithsmallmod.h:167-330);
 *   - key switching in SEAL's form: one digit per data prime, one special
 *     prime, lazy 128-bit inner products (evaluator.cpp:2724-3021);
 *   - rescale with the q_last / 2 rounding (util/rns.cpp:830-905).
 * Single-threaded like SEAL (which has no internal threading); concurrency
 * comes from independent calls on separate threads.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

typedef struct {
    uint64_t q, two_q, rh, rl; /* floor(2^128 / q) = rh 2^64 + rl */
    uint64_t *rp, *rps;        /* forward twiddles (bit-reversed) and their Shoup quotients */
    uint64_t *irp, *irps;      /* inverse twiddles ("scrambled" order) and quotients */
    uint64_t ninv, ninvs;      /* n^-1 and its quotient */
} port_mod;

typedef struct {
    int logn;
    size_t n, count; /* count moduli: L data primes, then the special prime */
    port_mod *m;
} port_ctx;

/* from ckks_oracle.c: the SEAL root tables */
int or_ntt_tables(int logn, uint64_t q, uint64_t *root_powers, uint64_t *inv_root_powers);
uint64_t or_invmod(uint64_t a, uint64_t q);

static inline uint64_t shoup_of(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }
static inline uint64_t hi64(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) >> 64); }

/* y * w mod q in [0, 2q) for y < 2^64 (Shoup) */
static inline uint64_t mul_shoup_lazy(uint64_t y, uint64_t w, uint64_t ws, uint64_t q)
{
    return y * w - hi64(y, ws) * q;
}

/* x mod q for a 128-bit x, Barrett with floor(2^128 / q): the quotient
 * estimate floor(x r / 2^128) (taken mod 2^64, which is all the remainder
 * needs) is at most 2 short, so two conditional subtractions finish it. */
static inline uint64_t barrett128(u128 x, const port_mod *m)
{
    const uint64_t xl = (uint64_t)x, xh = (uint64_t)(x >> 64);
    u128 s = ((u128)xl * m->rl) >> 64;
    s += (u128)xh * m->rl;
    const u128 c = (u128)xl * m->rh;
    s += c; /* any carry out of 2^128 only changes the quotient by a multiple of 2^64 */
    const uint64_t qe = xh * m->rh + (uint64_t)(s >> 64);
    uint64_t r = xl - qe * m->q;
    if (r >= m->q) r -= m->q;
    if (r >= m->q) r -= m->q;
    return r;
}

static inline uint64_t barrett64(uint64_t x, const port_mod *m)
{
    /* floor(x rh / 2^64) with rh = floor(2^128/q) >> 64 = floor(2^64 / q) */
    uint64_t r = x - hi64(x, m->rh) * m->q; /* the estimate is at most 2 short */
    if (r >= m->q) r -= m->q;
    return r >= m->q ? r - m->q : r;
}

void *or_port_create(int logn, const uint64_t *moduli, size_t count)
{
    port_ctx *c = (port_ctx *)calloc(1, sizeof(port_ctx));
    c->logn = logn;
    c->n = (size_t)1 << logn;
    c->count = count;
    c->m = (port_mod *)calloc(count, sizeof(port_mod));
    for (size_t i = 0; i < count; i++) {
        port_mod *m = &c->m[i];
        const uint64_t q = moduli[i];
        m->q = q;
        m->two_q = 2 * q;
        const u128 full = ~(u128)0; /* 2^128 - 1: floor((2^128 - 1) / q) = floor(2^128 / q) for odd q */
        const u128 ratio = full / q;
        m->rh = (uint64_t)(ratio >> 64);
        m->rl = (uint64_t)ratio;
        m->rp = (uint64_t *)malloc(c->n * 8);
        m->rps = (uint64_t *)malloc(c->n * 8);
        m->irp = (uint64_t *)malloc(c->n * 8);
        m->irps = (uint64_t *)malloc(c->n * 8);
        or_ntt_tables(logn, q, m->rp, m->irp);
        for (size_t k = 0; k < c->n; k++) {
            m->rps[k] = shoup_of(m->rp[k], q);
            m->irps[k] = shoup_of(m->irp[k], q);
        }
        m->ninv = or_invmod((uint64_t)(c->n % q), q);
        m->ninvs = shoup_of(m->ninv, q);
    }
    return c;
}

void or_port_destroy(void *h)
{
    port_ctx *c = (port_ctx *)h;
    if (!c) return;
    for (size_t i = 0; i < c->count; i++) {
        free(c->m[i].rp);
        free(c->m[i].rps);
        free(c->m[i].irp);
        free(c->m[i].irps);
    }
    free(c->m);
    free(c);
}

/* forward negacyclic NTT, Cooley-Tukey with Harvey's lazy butterflies: input
 * in [0, 4q) (any residue below 4q), output in [0, 4q) when lazy, else [0, q) */
static void ntt_fwd(const port_ctx *c, const port_mod *m, uint64_t *a, int lazy)
{
    const size_t n = c->n;
    const uint64_t q = m->q, tq = m->two_q;
    size_t gap = n >> 1, idx = 0;
    for (size_t mm = 1; mm < n; mm <<= 1) {
        for (size_t i = 0; i < mm; i++) {
            ++idx;
            const uint64_t w = m->rp[idx], ws = m->rps[idx];
            uint64_t *x = a + 2 * i * gap, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                uint64_t u = x[j];
                u = u >= tq ? u - tq : u;
                const uint64_t v = mul_shoup_lazy(y[j], w, ws, q);
                x[j] = u + v;
                y[j] = u + tq - v;
            }
        }
        gap >>= 1;
    }
    if (!lazy)
        for (size_t j = 0; j < n; j++) {
            uint64_t v = a[j];
            v = v >= tq ? v - tq : v;
            a[j] = v >= q ? v - q : v;
        }
}

/* inverse negacyclic NTT, Gentleman-Sande with lazy butterflies, then the
 * scaling by n^-1; input in [0, 2q), output in [0, q) */
static void ntt_inv(const port_ctx *c, const port_mod *m, uint64_t *a)
{
    const size_t n = c->n;
    const uint64_t q = m->q, tq = m->two_q;
    size_t gap = 1, idx = 0;
    for (size_t mm = n >> 1; mm >= 1; mm >>= 1) {
        for (size_t i = 0; i < mm; i++) {
            ++idx;
            const uint64_t w = m->irp[idx], ws = m->irps[idx];
            uint64_t *x = a + 2 * i * gap, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                const uint64_t u = x[j], v = y[j];
                uint64_t s = u + v;
                x[j] = s >= tq ? s - tq : s;
                y[j] = mul_shoup_lazy(u + tq - v, w, ws, q);
            }
        }
        gap <<= 1;
        if (mm == 1) break;
    }
    for (size_t j = 0; j < n; j++) {
        uint64_t v = mul_shoup_lazy(a[j], m->ninv, m->ninvs, q);
        a[j] = v >= q ? v - q : v;
    }
}

void or_port_ntt_forward(void *h, size_t prime, uint64_t *a)
{
    port_ctx *c = (port_ctx *)h;
    ntt_fwd(c, &c->m[prime], a, 0);
}

void or_port_ntt_inverse(void *h, size_t prime, uint64_t *a)
{
    port_ctx *c = (port_ctx *)h;
    ntt_inv(c, &c->m[prime], a);
}

/* ct (2 x l x n, NTT) += switch(target (l x n, NTT)) with key
 * [L][2][L+1][n] over the count = L + 1 moduli (one digit per data prime) */
void or_port_switch_key(void *h, uint64_t *ct, const uint64_t *target, size_t l, const uint64_t *key)
{
    port_ctx *c = (port_ctx *)h;
    const size_t n = c->n, L1 = c->count, L = L1 - 1;
    uint64_t *tt = (uint64_t *)malloc(l * n * 8);
    memcpy(tt, target, l * n * 8);
    for (size_t j = 0; j < l; j++) ntt_inv(c, &c->m[j], tt + j * n);
    uint64_t *tn = (uint64_t *)malloc(l * n * 8);
    u128 *acc0 = (u128 *)malloc(n * sizeof(u128)), *acc1 = (u128 *)malloc(n * sizeof(u128));
    uint64_t *res = (uint64_t *)malloc(2 * (l + 1) * n * 8);
    for (size_t I = 0; I <= l; I++) {
        const size_t ki = I == l ? L : I;
        const port_mod *mI = &c->m[ki];
        /* the digits lifted to prime I, in NTT form (lazy: products of values
         * below 4q and key words below q stay below 2^124) */
        for (size_t j = 0; j < l; j++) {
            uint64_t *d = tn + j * n;
            if (j == I) {
                memcpy(d, target + j * n, n * 8);
                continue;
            }
            const uint64_t *s = tt + j * n;
            if (c->m[j].q <= mI->q)
                memcpy(d, s, n * 8);
            else
                for (size_t k = 0; k < n; k++) d[k] = barrett64(s[k], mI);
            ntt_fwd(c, mI, d, 1);
        }
        memset(acc0, 0, n * sizeof(u128));
        memset(acc1, 0, n * sizeof(u128));
        for (size_t j = 0; j < l; j++) {
            const uint64_t *d = tn + j * n;
            const uint64_t *k0 = key + ((j * 2 + 0) * L1 + ki) * n, *k1 = key + ((j * 2 + 1) * L1 + ki) * n;
            for (size_t k = 0; k < n; k++) {
                acc0[k] += (u128)d[k] * k0[k];
                acc1[k] += (u128)d[k] * k1[k];
            }
        }
        for (size_t k = 0; k < n; k++) {
            res[(0 * (l + 1) + I) * n + k] = barrett128(acc0[k], mI);
            res[(1 * (l + 1) + I) * n + k] = barrett128(acc1[k], mI);
        }
    }
    /* ModDown with rounding: (acc_i - [acc_p + (p-1)/2]_qi + [(p-1)/2]_qi) p^-1 */
    const port_mod *mp = &c->m[L];
    const uint64_t half = mp->q >> 1;
    uint64_t *t = (uint64_t *)malloc(n * 8);
    for (size_t poly = 0; poly < 2; poly++) {
        uint64_t *last = res + (poly * (l + 1) + l) * n;
        ntt_inv(c, mp, last);
        for (size_t k = 0; k < n; k++) {
            uint64_t v = last[k] + half;
            last[k] = v >= mp->q ? v - mp->q : v;
        }
        for (size_t i = 0; i < l; i++) {
            const port_mod *mi = &c->m[i];
            const uint64_t hq = barrett64(half, mi);
            for (size_t k = 0; k < n; k++) {
                uint64_t v = barrett64(last[k], mi) + mi->q - hq;
                t[k] = v >= mi->q ? v - mi->q : v;
            }
            ntt_fwd(c, mi, t, 0);
            const uint64_t pinv = or_invmod(barrett64(mp->q, mi), mi->q), pinvs = shoup_of(pinv, mi->q);
            const uint64_t *ai = res + (poly * (l + 1) + i) * n;
            uint64_t *ci = ct + (poly * l + i) * n;
            for (size_t k = 0; k < n; k++) {
                uint64_t d = ai[k] + mi->q - t[k];
                uint64_t v = mul_shoup_lazy(d, pinv, pinvs, mi->q);
                v = v >= mi->q ? v - mi->q : v;
                v += ci[k];
                ci[k] = v >= mi->q ? v - mi->q : v;
            }
        }
    }
    free(t);
    free(res);
    free(acc0);
    free(acc1);
    free(tn);
    free(tt);
}

/* one polynomial of l NTT-form limbs -> round(x / q_(l-1)) on l - 1 limbs, in place */
void or_port_rescale(void *h, uint64_t *poly, size_t l)
{
    port_ctx *c = (port_ctx *)h;
    const size_t n = c->n;
    const port_mod *ml = &c->m[l - 1];
    const uint64_t half = ml->q >> 1;
    uint64_t *last = poly + (l - 1) * n;
    ntt_inv(c, ml, last);
    for (size_t k = 0; k < n; k++) {
        uint64_t v = last[k] + half;
        last[k] = v >= ml->q ? v - ml->q : v;
    }
    uint64_t *t = (uint64_t *)malloc(n * 8);
    for (size_t i = 0; i + 1 < l; i++) {
        const port_mod *mi = &c->m[i];
        const uint64_t hq = barrett64(half, mi);
        for (size_t k = 0; k < n; k++) {
            uint64_t v = barrett64(last[k], mi) + mi->q - hq;
            t[k] = v >= mi->q ? v - mi->q : v;
        }
        ntt_fwd(c, mi, t, 0);
        const uint64_t inv = or_invmod(barrett64(ml->q, mi), mi->q), invs = shoup_of(inv, mi->q);
        uint64_t *ci = poly + i * n;
        for (size_t k = 0; k < n; k++) {
            uint64_t v = mul_shoup_lazy(ci[k] + mi->q - t[k], inv, invs, mi->q);
            ci[k] = v >= mi->q ? v - mi->q : v;
        }
    }
    free(t);
}

/* out[p][i] = a[p][i] * b[i] (polys x l limbs) */
void or_port_multiply_plain(void *h, const uint64_t *a, const uint64_t *b, size_t polys, size_t l, uint64_t *out)
{
    port_ctx *c = (port_ctx *)h;
    const size_t n = c->n;
    for (size_t p = 0; p < polys; p++)
        for (size_t i = 0; i < l; i++) {
            const port_mod *m = &c->m[i];
            const uint64_t *x = a + (p * l + i) * n, *y = b + i * n;
            uint64_t *o = out + (p * l + i) * n;
            for (size_t k = 0; k < n; k++) o[k] = barrett128((u128)x[k] * y[k], m);
        }
}

/* (a0 b0, a0 b1 + a1 b0, a1 b1) */
void or_port_tensor(void *h, const uint64_t *a, const uint64_t *b, size_t l, uint64_t *out)
{
    port_ctx *c = (port_ctx *)h;
    const size_t n = c->n;
    for (size_t i = 0; i < l; i++) {
        const port_mod *m = &c->m[i];
        for (size_t k = 0; k < n; k++) {
            const uint64_t a0 = a[i * n + k], a1 = a[(l + i) * n + k], b0 = b[i * n + k], b1 = b[(l + i) * n + k];
            out[i * n + k] = barrett128((u128)a0 * b0, m);
            out[(l + i) * n + k] = barrett128((u128)a0 * b1 + (u128)a1 * b0, m);
            out[(2 * l + i) * n + k] = barrett128((u128)a1 * b1, m);
        }
    }
}

void or_port_add(void *h, const uint64_t *a, const uint64_t *b, size_t polys, size_t l, uint64_t *out)
{
    port_ctx *c = (port_ctx *)h;
    const size_t n = c->n;
    for (size_t p = 0; p < polys; p++)
        for (size_t i = 0; i < l; i++) {
            const uint64_t q = c->m[i].q;
            const uint64_t *x = a + (p * l + i) * n, *y = b + (p * l + i) * n;
            uint64_t *o = out + (p * l + i) * n;
            for (size_t k = 0; k < n; k++) {
                const uint64_t s = x[k] + y[k];
                o[k] = s >= q ? s - q : s;
            }
        }
}

/* Residues and NTT of rounded encoder coefficients below 2^63 in magnitude
 * (the FFT and the rounding are the encoder's, ckks.h:457-620): out l x n */
void or_port_encode_residues(void *h, const double *coeffs, size_t l, uint64_t *out)
{
    port_ctx *c = (port_ctx *)h;
    const size_t n = c->n;
    for (size_t i = 0; i < l; i++) {
        const port_mod *m = &c->m[i];
        uint64_t *o = out + i * n;
        for (size_t k = 0; k < n; k++) {
            const double v = coeffs[k];
            const double r = v < 0 ? -floor(-v + 0.5) : floor(v + 0.5); /* std::round */
            const uint64_t a = (uint64_t)(r < 0 ? -r : r);
            const uint64_t x = barrett64(a, m);
            o[k] = (r < 0 && x) ? m->q - x : x;
        }
        ntt_fwd(c, m, o, 0);
    }
}
