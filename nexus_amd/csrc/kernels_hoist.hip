// Hoisted key switching: one ModUp of c1 shared by several Galois
// automorphisms (reference Phantom hoisting_inplace,
// cuda/thirdparty/phantom-fhe/src/evaluate.cu:1727-1920).  The automorphism is
// applied to the ModUp digits inside the inner product: digit values are
// gathered through the NTT-domain permutation of each element, so the
// permuted digits are never written out.
#include "fpmod.h"
#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

constexpr int kBlock = 256;

struct OutPrimes {
    uint8_t p[kMaxPrimes];
};

struct RotArgs {
    const uint64_t *key[kMaxHoist];
    const uint32_t *perm[kMaxHoist];
    const uint64_t *t[kMaxHoist];      // k_ks_inner_rot: the ModUp digits of rotation r's input
    const uint64_t *aseed[kMaxHoist];  // non-null: the key's uniform half regenerated (ctr_uniform)
};

// acc[r][k][I][c] = sum_j T[I][j][perm_r[c]] * key_r[j][k][kp(I)][c]  (one rotation r per grid row group).
// qd non-null (every prime below 2^51): exact double-precision products
// (fp_mulmod), full-rate FMAs instead of 64-bit integer products; same words.
__global__ void k_ks_inner_rot(uint64_t *acc, RotArgs args, const ModConst *mod, const double2 *qd, OutPrimes outp,
                               uint32_t l, uint32_t key_limbs, uint32_t nout, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t ri = blockIdx.y;  // r * nout + I
    const uint32_t r = ri / nout, I = ri - r * nout;
    const uint32_t kp = outp.p[I];
    const ModConst m = mod[kp];
    const uint32_t src = args.perm[r][c];
    const uint64_t *t = args.t[r] + (uint64_t)I * l * n + src;
    const uint64_t *k0 = args.key[r] + (uint64_t)kp * n + c;
    const uint64_t kstride = (uint64_t)key_limbs * n;
    const uint64_t *aseed = args.aseed[r];
    const uint64_t ctr = (uint64_t)kp * n + c;
    uint64_t *out = acc + (uint64_t)r * 2 * nout * n;
    if (qd) {
        const double q = qd[kp].x, qinv = qd[kp].y;
        double s0 = 0.0, s1 = 0.0;
        for (uint32_t j = 0; j < l; j++) {
            const double tv = (double)t[(uint64_t)j * n];
            const uint64_t *kj = k0 + (uint64_t)j * 2 * kstride;
            const double a = (double)kj[0];
            const double b = (double)(aseed ? ctr_uniform(aseed[j], ctr, m.q) : kj[kstride]);
            s0 += fp_mulmod(tv, a, q, qinv);
            s1 += fp_mulmod(tv, b, q, qinv);
            if (j & 1) {  // |s| stays below 4 q < 2^53
                s0 = fp_reduce(s0, q, qinv);
                s1 = fp_reduce(s1, q, qinv);
            }
        }
        out[(uint64_t)I * n + c] = fp_final(s0, q, qinv);
        out[((uint64_t)nout + I) * n + c] = fp_final(s1, q, qinv);
        return;
    }
    uint64_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
    for (uint32_t j = 0; j < l; j++) {
        const uint64_t tv = t[(uint64_t)j * n];
        const uint64_t *kj = k0 + (uint64_t)j * 2 * kstride;
        const uint64_t a = kj[0], b = aseed ? ctr_uniform(aseed[j], ctr, m.q) : kj[kstride];
        uint64_t plo = tv * a, phi = mulhi64(tv, a);
        lo0 += plo;
        hi0 += phi + (lo0 < plo);
        plo = tv * b;
        phi = mulhi64(tv, b);
        lo1 += plo;
        hi1 += phi + (lo1 < plo);
    }
    out[(uint64_t)I * n + c] = barrett128(hi0, lo0, m);
    out[((uint64_t)nout + I) * n + c] = barrett128(hi1, lo1, m);
}

// acc[k][I][c] (+)= sum_r sum_j T[I][j][perm_r[c]] * key_r[j][k][kp(I)][c]:
// every product is below 2^120, so the 128-bit sums are folded every 128 terms.
__global__ void k_ks_inner_rot_sum(uint64_t *acc, const uint64_t *T, RotArgs args, uint32_t count,
                                   const ModConst *mod, const double2 *qd, OutPrimes outp, uint32_t l,
                                   uint32_t key_limbs, uint32_t nout, uint32_t n, int accumulate, uint64_t t_stride)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t I = blockIdx.y;
    const uint32_t kp = outp.p[I];
    const ModConst m = mod[kp];
    const uint64_t kstride = (uint64_t)key_limbs * n;
    if (qd) {  // every prime below 2^51: exact double-precision products
        const double q = qd[kp].x, qinv = qd[kp].y;
        double s0 = 0.0, s1 = 0.0;
        if (accumulate) {
            s0 = (double)acc[(uint64_t)I * n + c];
            s1 = (double)acc[((uint64_t)nout + I) * n + c];
        }
        const uint64_t ctr = (uint64_t)kp * n + c;
        for (uint32_t r = 0; r < count; r++) {
            const uint32_t src = args.perm[r][c];
            const uint64_t *t = T + r * t_stride + (uint64_t)I * l * n + src;
            const uint64_t *k0 = args.key[r] + (uint64_t)kp * n + c;
            const uint64_t *aseed = args.aseed[r];
            for (uint32_t j = 0; j < l; j++) {
                const double tv = (double)t[(uint64_t)j * n];
                const uint64_t *kj = k0 + (uint64_t)j * 2 * kstride;
                const double a = (double)kj[0];
                const double b = (double)(aseed ? ctr_uniform(aseed[j], ctr, m.q) : kj[kstride]);
                // two products a term: |s| <= q/2 (1 + eps) + 2.5 q after each pair, below 2^53
                s0 = fp_reduce(s0 + fp_mulmod(tv, a, q, qinv), q, qinv);
                s1 = fp_reduce(s1 + fp_mulmod(tv, b, q, qinv), q, qinv);
            }
        }
        acc[(uint64_t)I * n + c] = fp_final(s0, q, qinv);
        acc[((uint64_t)nout + I) * n + c] = fp_final(s1, q, qinv);
        return;
    }
    uint64_t r0 = 0, r1 = 0;  // folded partial sums (reduced)
    if (accumulate) {
        r0 = acc[(uint64_t)I * n + c];
        r1 = acc[((uint64_t)nout + I) * n + c];
    }
    uint64_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
    uint32_t terms = 0;
    for (uint32_t r = 0; r < count; r++) {
        const uint32_t src = args.perm[r][c];
        const uint64_t *t = T + r * t_stride + (uint64_t)I * l * n + src;
        const uint64_t *k0 = args.key[r] + (uint64_t)kp * n + c;
        const uint64_t *aseed = args.aseed[r];
        const uint64_t ctr = (uint64_t)kp * n + c;
        for (uint32_t j = 0; j < l; j++) {
            const uint64_t tv = t[(uint64_t)j * n];
            const uint64_t *kj = k0 + (uint64_t)j * 2 * kstride;
            const uint64_t a = kj[0], b = aseed ? ctr_uniform(aseed[j], ctr, m.q) : kj[kstride];
            uint64_t plo = tv * a, phi = mulhi64(tv, a);
            lo0 += plo;
            hi0 += phi + (lo0 < plo);
            plo = tv * b;
            phi = mulhi64(tv, b);
            lo1 += plo;
            hi1 += phi + (lo1 < plo);
            if (++terms == 128) {
                r0 = add_mod(r0, barrett128(hi0, lo0, m), m.q);
                r1 = add_mod(r1, barrett128(hi1, lo1, m), m.q);
                lo0 = hi0 = lo1 = hi1 = 0;
                terms = 0;
            }
        }
    }
    acc[(uint64_t)I * n + c] = add_mod(r0, barrett128(hi0, lo0, m), m.q);
    acc[((uint64_t)nout + I) * n + c] = add_mod(r1, barrett128(hi1, lo1, m), m.q);
}

// out[0] = sum_i sigma_i(a_i[0]) (a_i[0] where perm_i is null), out[1] = sum of
// a_i[1] over the inputs without a permutation: the data-basis part of a sum
// of rotations before the switched c1 terms are added.
__global__ void k_rot_base(uint64_t *out, RotBaseArgs args, uint32_t count, const ModConst *mod, uint32_t l,
                           uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // poly * l + limb
    const uint32_t poly = row / l, limb = row - poly * l;
    const uint64_t q = mod[limb].q;
    uint64_t sum = 0;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t *a = args.in[i] + (uint64_t)row * n;
        const uint32_t *perm = args.perm[i];
        if (perm) {
            if (poly == 0) sum = add_mod(sum, a[perm[c]], q);
        } else {
            sum = add_mod(sum, a[c], q);
        }
    }
    out[(uint64_t)row * n + c] = sum;
}

OutPrimes make_out(const uint8_t *p, uint32_t nout)
{
    OutPrimes o;
    for (uint32_t i = 0; i < kMaxPrimes; i++) o.p[i] = i < nout ? p[i] : 0;
    return o;
}

RotArgs make_args(const uint64_t *const *keys, const uint32_t *const *perms, uint32_t count,
                  const uint64_t *const *aseeds, const uint64_t *T = nullptr, const uint64_t *const *Ts = nullptr)
{
    if (count > (uint32_t)kMaxHoist) throw std::invalid_argument("hoisted inner product: too many rotations");
    RotArgs a;
    for (uint32_t i = 0; i < kMaxHoist; i++) {
        a.key[i] = i < count ? keys[i] : nullptr;
        a.perm[i] = i < count ? perms[i] : nullptr;
        a.t[i] = i < count ? (Ts ? Ts[i] : T) : nullptr;
        a.aseed[i] = (i < count && aseeds) ? aseeds[i] : nullptr;
    }
    return a;
}

// {q, 1/q} when every prime is below 2^51 (the double-precision products), else null
const double2 *fp_primes(const DeviceTables &t) { return t.max_bits <= 51 ? t.qd : nullptr; }

// key words read per (output prime, digit) of one rotation: 2, or 1 with the
// uniform half regenerated
double key_words_read(const uint64_t *const *aseeds, uint32_t count)
{
    if (!aseeds || !count) return 2.0;
    uint32_t seeded = 0;
    for (uint32_t i = 0; i < count; i++) seeded += aseeds[i] ? 1u : 0u;
    return 2.0 - (double)seeded / count;
}

}  // namespace

void ks_inner_product_rot(const DeviceTables &t, uint64_t *acc, const uint64_t *T, const uint64_t *const *keys,
                          const uint32_t *const *perms, uint32_t count, uint32_t l, uint32_t key_limbs,
                          const uint8_t *out_primes_host, uint32_t nout, hipStream_t s, const uint64_t *const *Ts,
                          const uint64_t *const *aseeds)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("ks_inner_rot", s,
                   (double)n * count * nout * ((8.0 + 8.0 * key_words_read(aseeds, count)) * l + 16.0 + 4.0));
    hipLaunchKernelGGL(k_ks_inner_rot, dim3((n + kBlock - 1) / kBlock, count * nout), dim3(kBlock), 0, s, acc,
                       make_args(keys, perms, count, aseeds, T, Ts), t.mod, fp_primes(t),
                       make_out(out_primes_host, nout), l, key_limbs, nout, n);
}

void ks_inner_product_rot_sum(const DeviceTables &t, uint64_t *acc, const uint64_t *T, const uint64_t *const *keys,
                              const uint32_t *const *perms, uint32_t count, uint32_t l, uint32_t key_limbs,
                              const uint8_t *out_primes_host, uint32_t nout, bool accumulate, hipStream_t s,
                              uint64_t t_stride, const uint64_t *const *aseeds)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("ks_inner_rot_sum", s,
                   (double)n * nout * (count * ((8.0 + 8.0 * key_words_read(aseeds, count)) * l + 4.0) + 16.0));
    hipLaunchKernelGGL(k_ks_inner_rot_sum, dim3((n + kBlock - 1) / kBlock, nout), dim3(kBlock), 0, s, acc, T,
                       make_args(keys, perms, count, aseeds), count, t.mod, fp_primes(t),
                       make_out(out_primes_host, nout), l, key_limbs, nout, n, accumulate ? 1 : 0, t_stride);
}

void rot_base(const DeviceTables &t, uint64_t *out, const uint64_t *const *in, const uint32_t *const *perms,
              uint32_t count, uint32_t l, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    RotBaseArgs a;
    for (uint32_t i = 0; i < kMaxHoist; i++) {
        a.in[i] = i < count ? in[i] : nullptr;
        a.perm[i] = i < count ? perms[i] : nullptr;
    }
    ProfScope prof("add", s, 8.0 * n * l * (2.0 + 2.0 * count));
    hipLaunchKernelGGL(k_rot_base, dim3((n + kBlock - 1) / kBlock, 2 * l), dim3(kBlock), 0, s, out, a, count, t.mod,
                       l, n);
}

}  // namespace nx
