// Dyadic (coefficient-wise) RNS kernels: HBM-bound integer work.  Grid y walks
// (poly, limb) pairs so the modulus is uniform per workgroup; grid x covers the
// n coefficients of a limb, one coefficient per lane.
#include "fpmod.h"
#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

constexpr int kBlock = 256;

struct PrimeMap {
    uint8_t p[kMaxPrimes];
};

struct LimbWords {
    uint64_t v[kMaxPrimes];
};

struct LimbShoup {
    ulonglong2 w[kMaxPrimes];
};

inline dim3 grid_for(int logn, uint32_t rows)
{
    uint32_t n = 1u << logn;
    return dim3((n + kBlock - 1) / kBlock, rows);
}

__global__ void k_add(uint64_t *out, const uint64_t *a, const uint64_t *b, const ModConst *mod, uint32_t limbs,
                      uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t q = mod[limb].q;
    const uint64_t idx = (uint64_t)row * n + c;
    out[idx] = add_mod(a[idx], b[idx], q);
}

// out = a with pt added to (or subtracted from) its first polynomial; the
// other polynomials are copied in the same launch
__global__ void k_add_plain_first(uint64_t *out, const uint64_t *a, const uint64_t *pt, const ModConst *mod,
                                  uint32_t limbs, uint32_t n, int sub)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t idx = (uint64_t)row * n + c;
    uint64_t v = a[idx];
    if (row < limbs) {
        const uint64_t q = mod[limb].q, p = pt[idx];
        v = sub ? sub_mod(v, p, q) : add_mod(v, p, q);
    }
    out[idx] = v;
}

__global__ void k_sub(uint64_t *out, const uint64_t *a, const uint64_t *b, const ModConst *mod, uint32_t limbs,
                      uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t q = mod[limb].q;
    const uint64_t idx = (uint64_t)row * n + c;
    out[idx] = sub_mod(a[idx], b[idx], q);
}

__global__ void k_negate(uint64_t *out, const uint64_t *a, const ModConst *mod, uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t q = mod[limb].q;
    const uint64_t idx = (uint64_t)row * n + c;
    out[idx] = neg_mod(a[idx], q);
}

__global__ void k_mul_plain(uint64_t *out, const uint64_t *a, const uint64_t *pt, const ModConst *mod,
                            uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const ModConst m = mod[limb];
    const uint64_t idx = (uint64_t)row * n + c;
    out[idx] = mul_mod(a[idx], pt[(uint64_t)limb * n + c], m);
}

__global__ void k_tensor(uint64_t *out, const uint64_t *a, const uint64_t *b, const ModConst *mod, uint32_t limbs,
                         uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t limb = blockIdx.y;
    const ModConst m = mod[limb];
    const uint64_t ps = (uint64_t)limbs * n;
    const uint64_t idx = (uint64_t)limb * n + c;
    const uint64_t a0 = a[idx], a1 = a[ps + idx], b0 = b[idx], b1 = b[ps + idx];
    out[idx] = mul_mod(a0, b0, m);
    // a0*b1 + a1*b0 as one 129-bit sum reduced once
    uint64_t lo1 = a0 * b1, hi1 = mulhi64(a0, b1);
    uint64_t lo2 = a1 * b0, hi2 = mulhi64(a1, b0);
    uint64_t lo = lo1 + lo2;
    uint64_t hi = hi1 + hi2 + (lo < lo1);
    out[ps + idx] = barrett128(hi, lo, m);
    out[2 * ps + idx] = mul_mod(a1, b1, m);
}

// Tensor products of `count` pairs: c01[r][2][l][n] = (a0 b0, a0 b1 + a1 b0),
// c2[r][l][n] = a1 b1, from ciphertexts given by pointer.
__global__ void k_tensor_many(uint64_t *c01, uint64_t *c2, PairPtrs p, const ModConst *mod, uint32_t limbs,
                              uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // r * limbs + limb
    const uint32_t r = row / limbs, limb = row - r * limbs;
    const ModConst m = mod[limb];
    const uint64_t ps = (uint64_t)limbs * n;
    const uint64_t idx = (uint64_t)limb * n + c;
    const uint64_t *a = p.x[r], *b = p.y[r];
    const uint64_t a0 = a[idx], a1 = a[ps + idx], b0 = b[idx], b1 = b[ps + idx];
    uint64_t *o = c01 + (uint64_t)r * 2 * ps;
    o[idx] = mul_mod(a0, b0, m);
    uint64_t lo1 = a0 * b1, hi1 = mulhi64(a0, b1);
    uint64_t lo2 = a1 * b0, hi2 = mulhi64(a1, b0);
    uint64_t lo = lo1 + lo2;
    uint64_t hi = hi1 + hi2 + (lo < lo1);
    o[ps + idx] = barrett128(hi, lo, m);
    c2[(uint64_t)r * ps + idx] = mul_mod(a1, b1, m);
}

// out[0..2][limb] (+)= sum_r a_r (x) b_r (the size-3 tensor of each pair),
// count <= kPtrMax pairs of [2][limbs][n] ciphertexts: 128-bit sums of the
// products (each below 2^120; at most 2 count <= 64 of them per output word)
// and one Barrett reduction
__global__ void k_tensor_sum(uint64_t *out, PairPtrs p, uint32_t count, const ModConst *mod, uint32_t limbs,
                             uint32_t n, int accumulate)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t limb = blockIdx.y;
    const ModConst m = mod[limb];
    const uint64_t ps = (uint64_t)limbs * n;
    const uint64_t idx = (uint64_t)limb * n + c;
    uint64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    auto mac = [&](int k, uint64_t x, uint64_t y) {
        const uint64_t pl = x * y, ph = mulhi64(x, y);
        lo[k] += pl;
        hi[k] += ph + (lo[k] < pl);
    };
    for (uint32_t r = 0; r < count; r++) {
        const uint64_t *a = p.x[r], *b = p.y[r];
        const uint64_t a0 = a[idx], a1 = a[ps + idx], b0 = b[idx], b1 = b[ps + idx];
        mac(0, a0, b0);
        mac(1, a0, b1);
        mac(1, a1, b0);
        mac(2, a1, b1);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        uint64_t v = barrett128(hi[k], lo[k], m);
        if (accumulate) v = add_mod(v, out[k * ps + idx], m.q);
        out[k * ps + idx] = v;
    }
}

// out_r[p][limb] = (x_r[p][limb] - tmp[r * polys + p][limb]) inv[limb] for the
// first l_out limbs; x_r has x_limbs limbs per poly (pointer lists p.y = x, p.x
// unused, outs = out pointers)
__global__ void k_add_limb_scaled(uint64_t *acc, uint64_t acc_stride, const uint64_t *d, uint64_t d_stride,
                                  uint32_t limb, ulonglong2 w, const ModConst *mod, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint64_t p = blockIdx.y;
    const uint64_t q = mod[limb].q;
    uint64_t *a = acc + p * acc_stride + (uint64_t)limb * n + c;
    *a = add_mod(*a, shoup(d[p * d_stride + (uint64_t)limb * n + c], w.x, w.y, q), q);
}

struct LinArgs {
    const uint64_t *x[kLinMax];
    uint64_t x_stride[kLinMax];
    uint64_t w[kLinMax][kMaxPrimes];
};

// out[p][i][c] (+)= sum_k x_k[p][i][c] w_k[i] mod q_i over the first `limbs`
// limbs of each x_k (poly stride x_stride[k]): 128-bit sums, one reduction
__global__ void k_lincomb(uint64_t *out, LinArgs a, uint32_t terms, const ModConst *mod, uint32_t limbs, uint32_t n,
                          int accumulate)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // p * limbs + i
    const uint32_t p = row / limbs, i = row - p * limbs;
    uint64_t lo = 0, hi = 0;
    for (uint32_t k = 0; k < terms; k++) {
        const uint64_t x = a.x[k][p * a.x_stride[k] + (uint64_t)i * n + c];
        const uint64_t w = a.w[k][i];
        const uint64_t plo = x * w, phi = mulhi64(x, w);
        lo += plo;
        hi += phi + (lo < plo);
    }
    const ModConst m = mod[i];
    uint64_t r = barrett128(hi, lo, m);
    uint64_t *o = out + (uint64_t)row * n + c;
    if (accumulate) r = add_mod(r, *o, m.q);
    *o = r;
}

// 16 bytes per thread: words 2i, 2i+1 of poly blockIdx.y (words = limbs n, even)
__global__ void k_copy_polys(uint64_t *dst, uint64_t dst_stride, const uint64_t *src, uint64_t src_stride,
                             uint64_t pairs)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= pairs) return;
    const uint64_t p = blockIdx.y;
    const ulonglong2 v = reinterpret_cast<const ulonglong2 *>(src + p * src_stride)[i];
    reinterpret_cast<ulonglong2 *>(dst + p * dst_stride)[i] = v;
}

__global__ void k_mul_scalar(uint64_t *out, const uint64_t *a, LimbShoup scal, const ModConst *mod, uint32_t limbs,
                             uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t q = mod[limb].q;
    const ulonglong2 w = scal.w[limb];
    const uint64_t idx = (uint64_t)row * n + c;
    out[idx] = shoup(a[idx], w.x, w.y, q);
}

__global__ void k_permute(uint64_t *out, const uint64_t *a, const uint32_t *perm, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint64_t row = blockIdx.y;
    out[row * n + c] = a[row * n + perm[c]];
}

__global__ void k_fill_limbs(uint64_t *out, LimbWords value, uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    out[(uint64_t)row * n + c] = value.v[limb];
}

__global__ void k_add_scalar(uint64_t *out, const uint64_t *a, LimbWords v, const ModConst *mod, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t limb = blockIdx.y;
    const uint64_t idx = (uint64_t)limb * n + c;
    out[idx] = add_mod(a[idx], v.v[limb], mod[limb].q);
}

// ---- rescale / ModDown
__global__ void k_center_last(uint64_t *tmp, const uint64_t *last, uint64_t last_stride, const ModConst *mod,
                              uint32_t last_prime, uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;
    const uint32_t poly = row / limbs, limb = row - poly * limbs;
    const uint64_t p = mod[last_prime].q;
    const uint64_t half = p >> 1;
    uint64_t v = last[poly * last_stride + c] + half;
    v = v >= p ? v - p : v;  // [0, p): (x + (p-1)/2) mod p
    const ModConst mi = mod[limb];
    const uint64_t r = sub_mod(barrett64(v, mi), barrett64(half, mi), mi.q);
    tmp[(uint64_t)row * n + c] = r;
}

// ---- key switching
__global__ void k_modup_expand(uint64_t *T, const uint64_t *coef, const ModConst *mod, PrimeMap outp, uint32_t l,
                               uint32_t nout, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // (b * nout + I) * l + j
    const uint32_t bi = row / l, j = row - bi * l;
    const uint32_t b = bi / nout, I = bi - b * nout;
    const ModConst m = mod[outp.p[I]];
    const uint64_t v = coef[((uint64_t)b * l + j) * n + c];
    T[(uint64_t)row * n + c] = v >= m.q ? barrett64(v, m) : v;
}

// aseed non-null: poly 1 of key digit j (uniform) is ctr_uniform(aseed[j], kp n + c, q_kp), regenerated
// instead of read
// qd non-null (every prime below 2^51): exact double-precision products (fp_mulmod); same words
__global__ void k_ks_inner(uint64_t *acc, const uint64_t *T, const uint64_t *key, const uint64_t *aseed,
                           const ModConst *mod, const double2 *qd, PrimeMap outp, uint32_t l, uint32_t key_limbs,
                           uint32_t nout, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t bi = blockIdx.y;  // b * nout + I
    const uint32_t b = bi / nout, I = bi - b * nout;
    const uint32_t kp = outp.p[I];
    const ModConst m = mod[kp];
    uint64_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
    const uint64_t *t = T + (uint64_t)bi * l * n + c;
    acc += (uint64_t)b * 2 * nout * n;
    const uint64_t *k0 = key + (uint64_t)kp * n + c;
    const uint64_t kstride = (uint64_t)key_limbs * n;  // between poly 0 and 1 of a digit
    const uint64_t ctr = (uint64_t)kp * n + c;
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
        acc[(uint64_t)I * n + c] = fp_final(s0, q, qinv);
        acc[((uint64_t)nout + I) * n + c] = fp_final(s1, q, qinv);
        return;
    }
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
    acc[(uint64_t)I * n + c] = barrett128(hi0, lo0, m);
    acc[((uint64_t)nout + I) * n + c] = barrett128(hi1, lo1, m);
}

// ---- sampling (counter-based; deterministic given the seed)
__device__ __forceinline__ uint64_t rand64(uint64_t seed, uint64_t ctr) { return ctr_rand64(seed, ctr); }

__global__ void k_sample_uniform(uint64_t *out, const ModConst *mod, PrimeMap pm, uint32_t limbs, uint32_t n,
                                 uint64_t seed)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t q = mod[pm.p[limb]].q;
    const uint64_t idx = (uint64_t)row * n + c;
    out[idx] = ctr_uniform(seed, idx, q);
}

__global__ void k_sample_cbd(uint64_t *out, const ModConst *mod, PrimeMap pm, uint32_t limbs, uint32_t n,
                             uint64_t seed)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint64_t r = rand64(seed, c);
    const int e = __popcll(r & 0x1fffffull) - __popcll((r >> 21) & 0x1fffffull);
    for (uint32_t i = 0; i < limbs; i++) {
        const uint64_t q = mod[pm.p[i]].q;
        out[(uint64_t)i * n + c] = e >= 0 ? (uint64_t)e : q - (uint64_t)(-e);
    }
}

__global__ void k_small_to_rns(uint64_t *out, const int8_t *coef, const ModConst *mod, PrimeMap pm, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t limb = blockIdx.y;
    const uint64_t q = mod[pm.p[limb]].q;
    const int v = coef[c];
    out[(uint64_t)limb * n + c] = v >= 0 ? (uint64_t)v : q - (uint64_t)(-v);
}

__global__ void k_i64_to_rns(uint64_t *out, const int64_t *coef, const ModConst *mod, PrimeMap pm, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t limb = blockIdx.y;
    const ModConst m = mod[pm.p[limb]];
    const int64_t v = coef[c];
    const uint64_t a = barrett64(v >= 0 ? (uint64_t)v : (uint64_t)(-v), m);
    out[(uint64_t)limb * n + c] = v >= 0 ? a : neg_mod(a, m.q);
}

__global__ void k_mul_add(uint64_t *out, const uint64_t *a, const uint64_t *b, const uint64_t *cc,
                          const ModConst *mod, PrimeMap pm, uint32_t n, int negate_product)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t limb = blockIdx.y;
    const ModConst m = mod[pm.p[limb]];
    const uint64_t idx = (uint64_t)limb * n + c;
    uint64_t r = mul_mod(a[idx], b[idx], m);
    if (negate_product) r = neg_mod(r, m.q);
    if (cc) r = add_mod(r, cc[idx], m.q);
    out[idx] = r;
}

PrimeMap make_map(const uint8_t *p, uint32_t count)
{
    PrimeMap m;
    for (uint32_t i = 0; i < kMaxPrimes; i++) m.p[i] = i < count ? p[i] : 0;
    return m;
}

}  // namespace

void poly_add(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, uint32_t polys,
              uint32_t limbs, hipStream_t s)
{
    hipLaunchKernelGGL(k_add, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, a, b, t.mod, limbs,
                       1u << t.logn);
}

void poly_sub(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, uint32_t polys,
              uint32_t limbs, hipStream_t s)
{
    hipLaunchKernelGGL(k_sub, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, a, b, t.mod, limbs,
                       1u << t.logn);
}

void poly_add_plain_first(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *pt, uint32_t polys,
                          uint32_t limbs, bool sub, hipStream_t s)
{
    ProfScope prof("add", s, 8.0 * (1u << t.logn) * limbs * (2.0 * polys + 1.0));
    hipLaunchKernelGGL(k_add_plain_first, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, a, pt, t.mod,
                       limbs, 1u << t.logn, sub ? 1 : 0);
}

void poly_negate(const DeviceTables &t, uint64_t *out, const uint64_t *a, uint32_t polys, uint32_t limbs,
                 hipStream_t s)
{
    hipLaunchKernelGGL(k_negate, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, a, t.mod, limbs,
                       1u << t.logn);
}

void poly_mul_plain(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *pt, uint32_t polys,
                    uint32_t limbs, hipStream_t s)
{
    ProfScope prof("mul_plain", s, 8.0 * (1u << t.logn) * ((double)polys * limbs * 2 + limbs));
    hipLaunchKernelGGL(k_mul_plain, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, a, pt, t.mod, limbs,
                       1u << t.logn);
}

void poly_tensor(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, uint32_t limbs,
                 hipStream_t s)
{
    ProfScope prof("tensor", s, 56.0 * (1u << t.logn) * limbs);
    hipLaunchKernelGGL(k_tensor, grid_for(t.logn, limbs), dim3(kBlock), 0, s, out, a, b, t.mod, limbs,
                       1u << t.logn);
}

void poly_tensor_many(const DeviceTables &t, uint64_t *c01, uint64_t *c2, const uint64_t *const *a,
                      const uint64_t *const *b, uint32_t count, uint32_t limbs, hipStream_t s)
{
    PairPtrs p;
    for (uint32_t i = 0; i < kPtrMax; i++) {
        p.x[i] = i < count ? a[i] : nullptr;
        p.y[i] = i < count ? b[i] : nullptr;
    }
    ProfScope prof("tensor", s, 56.0 * (1u << t.logn) * limbs * count);
    hipLaunchKernelGGL(k_tensor_many, grid_for(t.logn, count * limbs), dim3(kBlock), 0, s, c01, c2, p, t.mod, limbs,
                       1u << t.logn);
}

void poly_tensor_sum(const DeviceTables &t, uint64_t *out, const uint64_t *const *a, const uint64_t *const *b,
                     uint32_t count, uint32_t limbs, bool accumulate, hipStream_t s)
{
    if (count == 0 || count > (uint32_t)kPtrMax) throw std::invalid_argument("poly_tensor_sum: 1..32 pairs");
    if (t.max_bits > 60) throw std::invalid_argument("poly_tensor_sum: primes of at most 60 bits");
    PairPtrs p;
    for (uint32_t i = 0; i < kPtrMax; i++) {
        p.x[i] = i < count ? a[i] : nullptr;
        p.y[i] = i < count ? b[i] : nullptr;
    }
    ProfScope prof("tensor", s, 8.0 * (1u << t.logn) * limbs * (4.0 * count + (accumulate ? 6.0 : 3.0)));
    hipLaunchKernelGGL(k_tensor_sum, dim3(((1u << t.logn) + kBlock - 1) / kBlock, limbs), dim3(kBlock), 0, s, out, p,
                       count, t.mod, limbs, 1u << t.logn, accumulate ? 1 : 0);
}

void rns_add_limb_scaled(const DeviceTables &t, uint64_t *acc, uint64_t acc_stride, const uint64_t *d,
                         uint64_t d_stride, uint32_t limb, ulonglong2 w, uint32_t polys, hipStream_t s)
{
    ProfScope prof("add_limb_scaled", s, 24.0 * (1u << t.logn) * polys);
    hipLaunchKernelGGL(k_add_limb_scaled, dim3(((1u << t.logn) + kBlock - 1) / kBlock, polys), dim3(kBlock), 0, s,
                       acc, acc_stride, d, d_stride, limb, w, t.mod, 1u << t.logn);
}

void poly_lincomb(const DeviceTables &t, uint64_t *out, const uint64_t *const *x, const uint64_t *x_stride,
                  const uint64_t *const *w, uint32_t terms, uint32_t limbs, bool accumulate, hipStream_t s)
{
    if (terms == 0 || terms > (uint32_t)kLinMax || limbs > (uint32_t)kMaxPrimes)
        throw std::invalid_argument("poly_lincomb: terms or limbs out of range");
    LinArgs a;
    for (uint32_t k = 0; k < (uint32_t)kLinMax; k++) {
        a.x[k] = k < terms ? x[k] : nullptr;
        a.x_stride[k] = k < terms ? x_stride[k] : 0;
        for (uint32_t i = 0; i < (uint32_t)kMaxPrimes; i++) a.w[k][i] = (k < terms && i < limbs) ? w[k][i] : 0;
    }
    ProfScope prof("lincomb", s, 8.0 * (1u << t.logn) * 2 * limbs * (terms + (accumulate ? 2.0 : 1.0)));
    hipLaunchKernelGGL(k_lincomb, grid_for(t.logn, 2 * limbs), dim3(kBlock), 0, s, out, a, terms, t.mod, limbs,
                       1u << t.logn, accumulate ? 1 : 0);
}

void poly_copy_limbs(const DeviceTables &t, uint64_t *dst, uint64_t dst_stride, const uint64_t *src,
                     uint64_t src_stride, uint32_t polys, uint32_t limbs, hipStream_t s)
{
    const uint64_t pairs = ((uint64_t)limbs << t.logn) / 2;
    if (!polys || !pairs) return;
    ProfScope prof("copy", s, 16.0 * (double)pairs * 2.0 * polys);
    hipLaunchKernelGGL(k_copy_polys, dim3((uint32_t)((pairs + kBlock - 1) / kBlock), polys), dim3(kBlock), 0, s, dst,
                       dst_stride, src, src_stride, pairs);
}

void poly_mul_scalar(const DeviceTables &t, uint64_t *out, const uint64_t *a, const ulonglong2 *host_w,
                     uint32_t polys, uint32_t limbs, hipStream_t s)
{
    LimbShoup w;
    for (uint32_t i = 0; i < kMaxPrimes; i++) w.w[i] = i < limbs ? host_w[i] : make_ulonglong2(0, 0);
    hipLaunchKernelGGL(k_mul_scalar, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, a, w, t.mod, limbs,
                       1u << t.logn);
}

void poly_permute(uint64_t *out, const uint64_t *a, const uint32_t *perm, uint32_t polys, uint32_t limbs, int logn,
                  hipStream_t s)
{
    hipLaunchKernelGGL(k_permute, grid_for(logn, polys * limbs), dim3(kBlock), 0, s, out, a, perm, 1u << logn);
}

void poly_fill_limbs(uint64_t *out, const uint64_t *host_value, uint32_t polys, uint32_t limbs, int logn,
                     hipStream_t s)
{
    LimbWords v;
    for (uint32_t i = 0; i < kMaxPrimes; i++) v.v[i] = i < limbs ? host_value[i] : 0;
    hipLaunchKernelGGL(k_fill_limbs, grid_for(logn, polys * limbs), dim3(kBlock), 0, s, out, v, limbs, 1u << logn);
}

void poly_add_scalar(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *host_v,
                     uint32_t limbs, hipStream_t s)
{
    LimbWords v;
    for (uint32_t i = 0; i < kMaxPrimes; i++) v.v[i] = i < limbs ? host_v[i] : 0;
    hipLaunchKernelGGL(k_add_scalar, grid_for(t.logn, limbs), dim3(kBlock), 0, s, out, a, v, t.mod, 1u << t.logn);
}

void rns_center_last(const DeviceTables &t, uint64_t *tmp, const uint64_t *last, uint64_t last_stride,
                     uint32_t last_prime, uint32_t polys, uint32_t limbs, hipStream_t s)
{
    hipLaunchKernelGGL(k_center_last, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, tmp, last, last_stride,
                       t.mod, last_prime, limbs, 1u << t.logn);
}

void ks_modup_expand(const DeviceTables &t, uint64_t *T, const uint64_t *coef, uint32_t B, uint32_t l,
                     const uint8_t *out_primes_host, uint32_t nout, hipStream_t s)
{
    ProfScope prof("ks_modup", s, 8.0 * (1u << t.logn) * B * ((double)l + (double)nout * l));
    hipLaunchKernelGGL(k_modup_expand, grid_for(t.logn, B * nout * l), dim3(kBlock), 0, s, T, coef, t.mod,
                       make_map(out_primes_host, nout), l, nout, 1u << t.logn);
}

void ks_inner_product(const DeviceTables &t, uint64_t *acc, const uint64_t *T, const uint64_t *key, uint32_t B,
                      uint32_t l, uint32_t key_limbs, const uint8_t *out_primes_host, uint32_t nout, hipStream_t s,
                      const uint64_t *aseed)
{
    // T limbs (8 n) and the key limbs read (16 n, or 8 n with a regenerated
    // uniform half) per (output prime, digit), two outputs per prime
    ProfScope prof("ks_inner", s, (double)(1u << t.logn) * B * nout * ((aseed ? 16.0 : 24.0) * l + 16.0));
    hipLaunchKernelGGL(k_ks_inner, grid_for(t.logn, B * nout), dim3(kBlock), 0, s, acc, T, key, aseed, t.mod,
                       t.max_bits <= 51 ? t.qd : nullptr, make_map(out_primes_host, nout), l, key_limbs, nout,
                       1u << t.logn);
}

void sample_uniform(const DeviceTables &t, uint64_t *out, uint32_t polys, uint32_t limbs, const uint8_t *primes_host,
                    uint64_t seed, hipStream_t s)
{
    hipLaunchKernelGGL(k_sample_uniform, grid_for(t.logn, polys * limbs), dim3(kBlock), 0, s, out, t.mod,
                       make_map(primes_host, limbs), limbs, 1u << t.logn, seed);
}

void sample_cbd(const DeviceTables &t, uint64_t *out, uint32_t limbs, const uint8_t *primes_host, uint64_t seed,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_sample_cbd, grid_for(t.logn, 1), dim3(kBlock), 0, s, out, t.mod,
                       make_map(primes_host, limbs), limbs, 1u << t.logn, seed);
}

void small_to_rns(const DeviceTables &t, uint64_t *out, const int8_t *coef, uint32_t limbs,
                  const uint8_t *primes_host, hipStream_t s)
{
    hipLaunchKernelGGL(k_small_to_rns, grid_for(t.logn, limbs), dim3(kBlock), 0, s, out, coef, t.mod,
                       make_map(primes_host, limbs), 1u << t.logn);
}

void i64_to_rns(const DeviceTables &t, uint64_t *out, const int64_t *coef, uint32_t limbs,
                const uint8_t *primes_host, hipStream_t s)
{
    hipLaunchKernelGGL(k_i64_to_rns, grid_for(t.logn, limbs), dim3(kBlock), 0, s, out, coef, t.mod,
                       make_map(primes_host, limbs), 1u << t.logn);
}

void poly_mul_add(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, const uint64_t *c,
                  uint32_t limbs, const uint8_t *primes_host, int negate_product, hipStream_t s)
{
    hipLaunchKernelGGL(k_mul_add, grid_for(t.logn, limbs), dim3(kBlock), 0, s, out, a, b, c, t.mod,
                       make_map(primes_host, limbs), 1u << t.logn, negate_product);
}

}  // namespace nx

namespace nx {

namespace {

__global__ void k_limb_add_scaled(uint64_t *out, const uint64_t *a, ulonglong2 w, const ModConst *mod, uint32_t prime,
                                  uint32_t n)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const uint64_t q = mod[prime].q;
    out[c] = add_mod(out[c], shoup(a[c], w.x, w.y, q), q);
}

// out[(i + k) mod n] = +-a[i]: coefficient i moves to i + k, wrapping with a
// sign flip each time it passes x^n (util/polyarithsmallmod.h negacyclic_shift_poly_coeffmod).
__global__ void k_negacyclic_shift(uint64_t *out, const uint64_t *a, const ModConst *mod, uint32_t k, uint32_t limbs,
                                   uint32_t n)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const uint64_t q = mod[limb].q;
    const uint64_t v = a[(uint64_t)row * n + i];
    const uint32_t pos = (i + k) & (2 * n - 1);
    const bool flip = pos >= n;
    out[(uint64_t)row * n + (pos & (n - 1))] = flip ? neg_mod(v, q) : v;
}

}  // namespace

void limb_add_scaled(const DeviceTables &t, uint64_t *out, const uint64_t *a, ulonglong2 w, uint32_t prime,
                     hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    hipLaunchKernelGGL(k_limb_add_scaled, dim3((n + 255) / 256), dim3(256), 0, s, out, a, w, t.mod, prime, n);
}

void poly_negacyclic_shift(const DeviceTables &t, uint64_t *out, const uint64_t *a, uint32_t k, uint32_t polys,
                           uint32_t limbs, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    hipLaunchKernelGGL(k_negacyclic_shift, dim3((n + 255) / 256, polys * limbs), dim3(256), 0, s, out, a, t.mod, k,
                       limbs, n);
}

}  // namespace nx

namespace nx {

namespace {

__global__ void k_mul2_add(uint64_t *out, const uint64_t *a, const uint64_t *P, const uint64_t *b, const uint64_t *Q,
                           const ModConst *mod, uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y, limb = row % limbs;
    const ModConst m = mod[limb];
    const uint64_t idx = (uint64_t)row * n + c, pidx = (uint64_t)limb * n + c;
    const uint64_t x = a[idx], y = P[pidx], u = b[idx], v = Q[pidx];
    uint64_t lo1 = x * y, hi1 = mulhi64(x, y);
    uint64_t lo2 = u * v, hi2 = mulhi64(u, v);
    uint64_t lo = lo1 + lo2;
    uint64_t hi = hi1 + hi2 + (lo < lo1);
    out[idx] = barrett128(hi, lo, m);
}

// One lane per (output column, poly, limb, coefficient); 128-bit accumulation,
// folded back below q every 32 products (each product < 2^122, so 33 terms
// stay below 2^128).
__global__ void k_mm_acc(uint64_t *out, const uint64_t *E, const uint64_t *A, const ModConst *mod, uint32_t inner,
                         uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // (i * 2 + p) * limbs + l
    const uint32_t limb = row % limbs;
    const uint32_t ip = row / limbs, i = ip >> 1, p = ip & 1;
    const ModConst m = mod[limb];
    const uint64_t ct_words = 2ull * limbs * n;
    const uint64_t *e = E + (uint64_t)i * inner * ct_words + ((uint64_t)p * limbs + limb) * n + c;
    const uint64_t *aa = A + (uint64_t)limb * n + c;
    const uint64_t a_stride = (uint64_t)limbs * n;
    uint64_t lo = 0, hi = 0;
    for (uint32_t j = 0; j < inner; j++) {
        const uint64_t x = e[(uint64_t)j * ct_words], y = aa[(uint64_t)j * a_stride];
        const uint64_t plo = x * y, phi = mulhi64(x, y);
        lo += plo;
        hi += phi + (lo < plo);
        if ((j & 31) == 31) {
            lo = barrett128(hi, lo, m);
            hi = 0;
        }
    }
    out[(uint64_t)row * n + c] = barrett128(hi, lo, m);
}

}  // namespace

void poly_mul2_add(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *P, const uint64_t *b,
                   const uint64_t *Q, uint32_t polys, uint32_t limbs, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("mul2_add", s, 8.0 * n * ((double)polys * limbs * 3 + 2.0 * limbs));
    hipLaunchKernelGGL(k_mul2_add, dim3((n + 255) / 256, polys * limbs), dim3(256), 0, s, out, a, P, b, Q, t.mod,
                       limbs, n);
}

void mm_accumulate(const DeviceTables &t, uint64_t *out, const uint64_t *E, const uint64_t *A, uint32_t cols,
                   uint32_t inner, uint32_t limbs, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("mm_acc", s, 8.0 * n * limbs * ((double)cols * inner * 2 + inner + cols * 2.0));
    hipLaunchKernelGGL(k_mm_acc, dim3((n + 255) / 256, cols * 2 * limbs), dim3(256), 0, s, out, E, A, t.mod, inner,
                       limbs, n);
}

}  // namespace nx
