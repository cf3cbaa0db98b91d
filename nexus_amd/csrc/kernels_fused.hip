// Fused forms of the hybrid base conversions and of the plaintext
// multiply-accumulate of BSGS linear transforms.
//
// The conversions keep each input word in registers and loop over a chunk of
// output primes: the per-input Shoup product [x Dhat^-1]_q is computed once
// per chunk instead of once per output prime.  Conversion constants are
// uniform across a wavefront (scalar loads).
#include <atomic>

#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kRegAlpha = 16;  // digits / special primes held in registers
constexpr uint32_t kOutChunk = 4;   // output primes per thread (more waves in flight)

// prime index of each input limb of a ModDown conversion
struct InPrimes {
    uint8_t p[kRegAlpha + 1];
};

// Words, not bytes: a uniform index into the kernel argument then takes a
// scalar load, which does not wait on the thread's outstanding stores.
struct OutMap {
    uint32_t p[kMaxPrimes];
};

// thread (b, j, chunk, c): T[b][I][j][c] for the output primes I of its chunk
__global__ void k_modup_conv_fused(uint64_t *T, const uint64_t *coef, const ulonglong2 *w, const uint64_t *conv,
                                   const ModConst *mod, OutMap outp, uint32_t l, uint32_t alpha, uint32_t digits,
                                   uint32_t nout, uint32_t key_limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t chunks = (nout + kOutChunk - 1) / kOutChunk;
    const uint32_t row = blockIdx.y;  // (b * digits + j) * chunks + chunk
    const uint32_t bj = row / chunks, chunk = row - bj * chunks;
    const uint32_t b = bj / digits, j = bj - b * digits;
    const uint32_t first = j * alpha;
    const uint32_t s = min(alpha, l - first);
    const uint32_t base = (j * alpha + s - 1) * alpha;
    const uint64_t *x = coef + (uint64_t)b * l * n + (uint64_t)first * n + c;
    uint64_t y[kRegAlpha];
#pragma unroll
    for (uint32_t i = 0; i < kRegAlpha; i++) {
        if (i < s) {
            const ulonglong2 wi = w[base + i];
            y[i] = shoup(x[(uint64_t)i * n], wi.x, wi.y, mod[first + i].q);
        }
    }
    uint64_t *out = T + ((uint64_t)b * nout * digits + j) * n + c;
    const uint32_t I_end = min(nout, (chunk + 1) * kOutChunk);
    for (uint32_t I = chunk * kOutChunk; I < I_end; I++) {
        const uint32_t kp = outp.p[I];
        const uint64_t *f = conv + (uint64_t)base * key_limbs + kp;
        uint64_t lo = 0, hi = 0;
#pragma unroll
        for (uint32_t i = 0; i < kRegAlpha; i++) {
            if (i < s) {
                const uint64_t fi = f[(uint64_t)i * key_limbs];
                const uint64_t plo = y[i] * fi, phi = mulhi64(y[i], fi);
                lo += plo;
                hi += phi + (lo < plo);
            }
        }
        out[(uint64_t)I * digits * n] = barrett128(hi, lo, mod[kp]);
    }
}

// Sum of up to 16 products of factors below 2^58 kept as three 64-bit partial
// sums of 32x32-bit products (value = lo32 + mid 2^32 + hi 2^64): four
// multiplies per product and no carry chain -- the generic form needs seven.
// Bounds: lo32 < 16 2^32; mid < 16 (2^32 + 2^59) < 2^64; hi < 16 2^52.
struct Sum3 {
    uint64_t lo32 = 0, mid = 0, hi = 0;
    __device__ __forceinline__ void mac(uint64_t a, uint64_t b)
    {
        const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
        const uint64_t p00 = a0 * b0;
        lo32 += (uint32_t)p00;
        mid += (p00 >> 32) + a0 * b1 + a1 * b0;
        hi += a1 * b1;
    }
    __device__ __forceinline__ void fold(uint64_t &h, uint64_t &l) const
    {
        const uint64_t ml = mid << 32;
        l = lo32 + ml;
        h = hi + (mid >> 32) + (l < ml);
    }
};

// k_modup_conv_fused for alpha = A (compile time: no per-term branches) with
// every prime below 2^58 (Sum3 products); a partial last digit zero-fills.
// Exact products in double precision for a < 2^51 and w < q < 2^51:
// h = a w, e = fma(a, w, -h) = a w - h, Q = rint(h / q) within 1.25 of a w / q,
// r = fma(-Q, q, h) + e = a w - Q q with |r| <= 1.25 q (every intermediate an
// integer below 2^53); reduce_f maps |z| < 2^52 to |z| <= q/2 (1 + eps).
__device__ __forceinline__ double mulmod_f(double a, double w, double q, double qinv)
{
    const double h = a * w;
    const double e = __builtin_fma(a, w, -h);
    const double Q = __builtin_rint(h * qinv);
    return __builtin_fma(-Q, q, h) + e;
}

__device__ __forceinline__ double reduce_f(double z, double q, double qinv)
{
    return __builtin_fma(-__builtin_rint(z * qinv), q, z);
}

constexpr uint64_t kFpPrimeBound = 1ull << 51;
constexpr uint32_t kConvChunk = 12;  // output primes per thread of k_modup_conv_a

template <int A>
__global__ __launch_bounds__(kBlock) void k_modup_conv_a(uint64_t *__restrict__ T, const uint64_t *__restrict__ coef,
                                                        const ulonglong2 *__restrict__ w,
                                                        const uint64_t *__restrict__ conv,
                                                        const double *__restrict__ convd,
                                                        const ModConst *__restrict__ mod,
                                                        const double2 *__restrict__ qd, OutMap outp,
                                                        uint32_t l, uint32_t digits, uint32_t nout,
                                                        uint32_t key_limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t chunks = (nout + kConvChunk - 1) / kConvChunk;
    const uint32_t row = blockIdx.y;  // (b * digits + j) * chunks + chunk
    const uint32_t bj = row / chunks, chunk = row - bj * chunks;
    const uint32_t b = bj / digits, j = bj - b * digits;
    const uint32_t first = j * A;
    const uint32_t s = min((uint32_t)A, l - first);
    const uint32_t base = (j * A + s - 1) * A;
    const uint64_t *x = coef + (uint64_t)b * l * n + (uint64_t)first * n + c;
    // the digit's residues are below 2^51 whenever its primes are
    bool digit_fp = convd != nullptr;
#pragma unroll
    for (int i = 0; i < A; i++)
        if ((uint32_t)i < s && mod[first + i].q >= kFpPrimeBound) digit_fp = false;
    uint64_t y[A];
    double yd[A];
#pragma unroll
    for (int i = 0; i < A; i++) {
        y[i] = 0;
        yd[i] = 0.0;
        if ((uint32_t)i < s) {
            const ulonglong2 wi = w[base + i];
            const uint64_t qi = mod[first + i].q;
            if (digit_fp) {
                // [x Dhat^-1]_q in double precision, mapped to [0, q)
                const double q = qd[first + i].x, qinv = qd[first + i].y;
                double v = reduce_f(mulmod_f((double)x[(uint64_t)i * n], (double)wi.x, q, qinv), q, qinv);
                v = v < 0 ? v + q : v;
                yd[i] = v;
                y[i] = (uint64_t)v;
            } else {
                y[i] = shoup(x[(uint64_t)i * n], wi.x, wi.y, qi);
                yd[i] = (double)y[i];
            }
        }
    }
    uint64_t *out = T + ((uint64_t)b * nout * digits + j) * n + c;
    const uint32_t I_end = min(nout, (chunk + 1) * kConvChunk);
    for (uint32_t I = chunk * kConvChunk; I < I_end; I++) {
        if (I - first < s) continue;  // the digit's own primes: copied in NTT form
        const uint32_t kp = outp.p[I];
        const uint64_t qk = mod[kp].q;
        // Rows i >= s of a digit's factor block are zero in the tables (and y, yd
        // are zero there), so every term is formed without a branch and the
        // factor loads issue together.
        if (digit_fp && qk < kFpPrimeBound) {
            // sum of products in double precision, reduced every second term
            const double *fd = convd + (uint64_t)base * key_limbs + kp;
            const double q = qd[kp].x, qinv = qd[kp].y;
            double f[A];
#pragma unroll
            for (int i = 0; i < A; i++) f[i] = fd[(uint64_t)i * key_limbs];
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < A; i++) {
                acc += mulmod_f(yd[i], f[i], q, qinv);
                if (i & 1) acc = reduce_f(acc, q, qinv);
            }
            acc = reduce_f(acc, q, qinv);
            acc = acc < 0 ? acc + q : acc;
            out[(uint64_t)I * digits * n] = (uint64_t)acc;
            continue;
        }
        const uint64_t *f = conv + (uint64_t)base * key_limbs + kp;
        uint64_t fv[A];
#pragma unroll
        for (int i = 0; i < A; i++) fv[i] = f[(uint64_t)i * key_limbs];
        Sum3 acc;
#pragma unroll
        for (int i = 0; i < A; i++) acc.mac(y[i], fv[i]);
        uint64_t hi, lo;
        acc.fold(hi, lo);
        out[(uint64_t)I * digits * n] = barrett128(hi, lo, mod[kp]);
    }
}

constexpr uint32_t kMdChunk = 8;  // output primes per thread of k_moddown_conv_a

// k_moddown_conv_fused for A input primes, every prime below 2^58 (Sum3
// products).  When convd is given (every input prime below 2^51), output
// primes below 2^51 take exact double-precision products (mulmod_f), the sum
// reduced every second term.
template <int A>
__global__ __launch_bounds__(kBlock) void k_moddown_conv_a(uint64_t *__restrict__ z,
                                                          const uint64_t *__restrict__ acc, uint64_t acc_stride,
                                                          const ulonglong2 *__restrict__ w,
                                                          const uint64_t *__restrict__ conv,
                                                          const double *__restrict__ convd,
                                                          const ModConst *__restrict__ mod,
                                                          const double2 *__restrict__ qd, InPrimes inp, uint32_t l,
                                                          uint32_t L, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t chunks = (l + kMdChunk - 1) / kMdChunk;
    const uint32_t p = blockIdx.y / chunks, chunk = blockIdx.y - p * chunks;
    const uint64_t *y = acc + (uint64_t)p * acc_stride + (uint64_t)l * n + c;
    uint64_t v[A];
    double frac = 0.0;
#pragma unroll
    for (int k = 0; k < A; k++) {
        const ulonglong2 wk = w[k];
        const uint64_t pk = mod[inp.p[k]].q;
        v[k] = shoup(y[(uint64_t)k * n], wk.x, wk.y, pk);
        frac += (double)v[k] / (double)pk;
    }
    const uint64_t r = (uint64_t)floor(frac + 0.5);  // <= A
    uint64_t *out = z + (uint64_t)p * l * n + c;
    const uint32_t i_end = min(l, (chunk + 1) * kMdChunk);
    for (uint32_t i = chunk * kMdChunk; i < i_end; i++) {
        const uint64_t qi = mod[i].q;
        if (convd && qi < kFpPrimeBound) {
            const double q = qd[i].x, qinv = qd[i].y;
            double f[A + 1];  // loaded together, ahead of the products
#pragma unroll
            for (int k = 0; k <= A; k++) f[k] = convd[(uint64_t)k * L + i];
            // - r (P mod q_i) as r (q_i - P mod q_i)
            double s = mulmod_f((double)(uint32_t)r, f[A], q, qinv);
#pragma unroll
            for (int k = 0; k < A; k++) {
                s += mulmod_f((double)v[k], f[k], q, qinv);
                if (k & 1) s = reduce_f(s, q, qinv);  // |s| stays below 4q < 2^53
            }
            s = reduce_f(s, q, qinv);
            s = s < 0 ? s + q : s;
            out[(uint64_t)i * n] = (uint64_t)s;
            continue;
        }
        uint64_t fv[A];
#pragma unroll
        for (int k = 0; k < A; k++) fv[k] = conv[(uint64_t)k * L + i];
        Sum3 s3;
#pragma unroll
        for (int k = 0; k < A; k++) s3.mac(v[k], fv[k]);
        uint64_t hi, lo;
        s3.fold(hi, lo);
        // + r (P mod q_i)'s negative: r < 2^5, so the product fits one word
        const uint64_t add = r * (qi - conv[(uint64_t)A * L + i]);
        lo += add;
        hi += lo < add;
        out[(uint64_t)i * n] = barrett128(hi, lo, mod[i]);
    }
}

// thread (p, chunk, c): z[p][i][c] for the data primes i of its chunk
__global__ void k_moddown_conv_fused(uint64_t *z, const uint64_t *acc, uint64_t acc_stride, const ulonglong2 *w,
                                     const uint64_t *conv, const ModConst *mod, uint32_t l, uint32_t alpha,
                                     uint32_t L, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t chunks = (l + kOutChunk - 1) / kOutChunk;
    const uint32_t p = blockIdx.y / chunks, chunk = blockIdx.y - p * chunks;
    const uint64_t *y = acc + (uint64_t)p * acc_stride + (uint64_t)l * n + c;
    uint64_t v[kRegAlpha];
    double frac = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < kRegAlpha; k++) {
        if (k < alpha) {
            const ulonglong2 wk = w[k];
            const uint64_t pk = mod[L + k].q;
            v[k] = shoup(y[(uint64_t)k * n], wk.x, wk.y, pk);
            frac += (double)v[k] / (double)pk;
        }
    }
    const uint64_t r = (uint64_t)floor(frac + 0.5);
    uint64_t *out = z + (uint64_t)p * l * n + c;
    const uint32_t i_end = min(l, (chunk + 1) * kOutChunk);
    for (uint32_t i = chunk * kOutChunk; i < i_end; i++) {
        const uint64_t qi = mod[i].q;
        uint64_t lo = 0, hi = 0;
#pragma unroll
        for (uint32_t k = 0; k < kRegAlpha; k++) {
            if (k < alpha) {
                const uint64_t f = conv[(uint64_t)k * L + i];
                const uint64_t plo = v[k] * f, phi = mulhi64(v[k], f);
                lo += plo;
                hi += phi + (lo < plo);
            }
        }
        const uint64_t negP = qi - conv[(uint64_t)alpha * L + i];
        const uint64_t plo = r * negP, phi = mulhi64(r, negP);
        lo += plo;
        hi += phi + (lo < plo);
        out[(uint64_t)i * n] = barrett128(hi, lo, mod[i]);
    }
}

struct DotArgs {
    const uint64_t *ct[kMaxDot];
    const uint64_t *pt[kMaxDot];
};

// out[k][limb][c] = sum_r ct_r[k][limb][c] * pt_r[limb][c]  (k < 2), folded
// every 32 products.  A thread owns two adjacent coefficients of one limb for
// both polynomials: every plaintext word is read once, all loads are 16 bytes.
__global__ __launch_bounds__(kBlock) void k_dot_plain(uint64_t *out, DotArgs args, uint32_t count,
                                                      const ModConst *mod, uint32_t limbs, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;  // coefficient pair
    if (2 * i >= n) return;
    const uint32_t limb = blockIdx.y;
    const ModConst m = mod[limb];
    const uint64_t poff = (uint64_t)limb * n + 2 * i, off1 = poff + (uint64_t)limbs * n;
    uint64_t lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
    for (uint32_t r = 0; r < count; r++) {
        const ulonglong2 y = *reinterpret_cast<const ulonglong2 *>(args.pt[r] + poff);
        const ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(args.ct[r] + poff);
        const ulonglong2 b = *reinterpret_cast<const ulonglong2 *>(args.ct[r] + off1);
        const uint64_t xs[4] = {a.x, a.y, b.x, b.y}, ys[4] = {y.x, y.y, y.x, y.y};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t plo = xs[k] * ys[k], phi = mulhi64(xs[k], ys[k]);
            lo[k] += plo;
            hi[k] += phi + (lo[k] < plo);
        }
        if ((r & 31) == 31) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                lo[k] = barrett128(hi[k], lo[k], m);
                hi[k] = 0;
            }
        }
    }
    ulonglong2 o0, o1;
    o0.x = barrett128(hi[0], lo[0], m);
    o0.y = barrett128(hi[1], lo[1], m);
    o1.x = barrett128(hi[2], lo[2], m);
    o1.y = barrett128(hi[3], lo[3], m);
    *reinterpret_cast<ulonglong2 *>(out + poff) = o0;
    *reinterpret_cast<ulonglong2 *>(out + off1) = o1;
}

struct MultiDotArgs {
    const uint64_t *ct[kMultiDotK];
    const uint64_t *pt[kMultiDotG][kMultiDotK];  // null: no term
    uint64_t *out[kMultiDotG];                    // null: unused row
};

// out_g[k][limb][c] (+)= sum_r ct_r[k][limb][c] * pt_g,r[limb][c] for the rows g
// of one launch (k < 2, count <= kMultiDotK, so no intermediate fold): every
// ciphertext word is read once for all rows.  Two adjacent coefficients of one
// limb per thread, 16-byte loads, as k_dot_plain.
__global__ __launch_bounds__(kBlock) void k_dot_plain_multi(MultiDotArgs args, uint32_t count, const ModConst *mod,
                                                            uint32_t limbs, uint32_t n, int accumulate)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;  // coefficient pair
    if (2 * i >= n) return;
    const uint32_t limb = blockIdx.y;
    const ModConst m = mod[limb];
    const uint64_t poff = (uint64_t)limb * n + 2 * i, off1 = poff + (uint64_t)limbs * n;
    uint64_t lo[kMultiDotG][4] = {}, hi[kMultiDotG][4] = {};
    for (uint32_t r = 0; r < count; r++) {
        const ulonglong2 a = *reinterpret_cast<const ulonglong2 *>(args.ct[r] + poff);
        const ulonglong2 b = *reinterpret_cast<const ulonglong2 *>(args.ct[r] + off1);
        const uint64_t xs[4] = {a.x, a.y, b.x, b.y};
#pragma unroll
        for (int g = 0; g < kMultiDotG; g++) {
            const uint64_t *p = args.pt[g][r];
            if (!p) continue;  // uniform across the launch
            const ulonglong2 y = *reinterpret_cast<const ulonglong2 *>(p + poff);
            const uint64_t ys[4] = {y.x, y.y, y.x, y.y};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t plo = xs[k] * ys[k], phi = mulhi64(xs[k], ys[k]);
                lo[g][k] += plo;
                hi[g][k] += phi + (lo[g][k] < plo);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < kMultiDotG; g++) {
        uint64_t *o = args.out[g];
        if (!o) continue;
        ulonglong2 o0, o1;
        o0.x = barrett128(hi[g][0], lo[g][0], m);
        o0.y = barrett128(hi[g][1], lo[g][1], m);
        o1.x = barrett128(hi[g][2], lo[g][2], m);
        o1.y = barrett128(hi[g][3], lo[g][3], m);
        if (accumulate) {
            const ulonglong2 p0 = *reinterpret_cast<const ulonglong2 *>(o + poff);
            const ulonglong2 p1 = *reinterpret_cast<const ulonglong2 *>(o + off1);
            o0.x = add_mod(o0.x, p0.x, m.q);
            o0.y = add_mod(o0.y, p0.y, m.q);
            o1.x = add_mod(o1.x, p1.x, m.q);
            o1.y = add_mod(o1.y, p1.y, m.q);
        }
        *reinterpret_cast<ulonglong2 *>(o + poff) = o0;
        *reinterpret_cast<ulonglong2 *>(o + off1) = o1;
    }
}

// One coefficient of one polynomial per thread (grid y over poly x limb): four
// times the threads of k_dot_plain, for launches with few limbs.
__global__ __launch_bounds__(kBlock) void k_dot_plain1(uint64_t *out, DotArgs args, uint32_t count,
                                                       const ModConst *mod, uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // k * limbs + limb
    const uint32_t limb = row % limbs;
    const ModConst m = mod[limb];
    const uint64_t off = (uint64_t)row * n + c, poff = (uint64_t)limb * n + c;
    uint64_t lo = 0, hi = 0;
    for (uint32_t r = 0; r < count; r++) {
        const uint64_t x = args.ct[r][off], y = args.pt[r][poff];
        const uint64_t plo = x * y, phi = mulhi64(x, y);
        lo += plo;
        hi += phi + (lo < plo);
        if ((r & 31) == 31) {
            lo = barrett128(hi, lo, m);
            hi = 0;
        }
    }
    out[off] = barrett128(hi, lo, m);
}

// layout per limb-count bucket (<= 2, <= 4, <= 8, <= 16, more), 0 = default
std::atomic<int> g_dot_layout[5];

int dot_bucket(uint32_t limbs)
{
    return limbs <= 2 ? 0 : limbs <= 4 ? 1 : limbs <= 8 ? 2 : limbs <= 16 ? 3 : 4;
}

}  // namespace

void dot_plain_set_layout(uint32_t limbs, int layout)
{
    if (layout == 0 || layout == 1) g_dot_layout[dot_bucket(limbs)].store(layout);
}

int dot_plain_get_layout(uint32_t limbs) { return g_dot_layout[dot_bucket(limbs)].load(); }

bool ks_conv_fused_ok(uint32_t alpha) { return alpha <= kRegAlpha; }

void ks_modup_convert_fused(const DeviceTables &t, uint64_t *T, const uint64_t *coef, const ulonglong2 *w,
                            const uint64_t *conv, uint32_t B, uint32_t l, uint32_t alpha, uint32_t key_limbs,
                            const uint8_t *out_primes_host, uint32_t nout, hipStream_t s)
{
    const uint32_t n = 1u << t.logn, digits = (l + alpha - 1) / alpha;
    OutMap o;
    for (uint32_t i = 0; i < kMaxPrimes; i++) o.p[i] = i < nout ? out_primes_host[i] : 0;
    ProfScope prof("ks_modup_conv", s, 8.0 * n * B * ((double)l + (double)nout * digits));
    const uint32_t chunks = (nout + kOutChunk - 1) / kOutChunk;
    const dim3 grid((n + kBlock - 1) / kBlock, B * digits * chunks);
    const dim3 grid_a((n + kBlock - 1) / kBlock, B * digits * ((nout + kConvChunk - 1) / kConvChunk));
    if (t.max_bits <= 58) {
        switch (alpha) {
#define NX_A(A)                                                                                                     \
    case A:                                                                                                         \
        hipLaunchKernelGGL(k_modup_conv_a<A>, grid_a, dim3(kBlock), 0, s, T, coef, w, conv, t.modup_conv_d, t.mod, t.qd, o, l, \
                           digits, nout, key_limbs, n);                                                             \
        return;
            NX_A(2) NX_A(3) NX_A(4) NX_A(5) NX_A(6) NX_A(7) NX_A(8) NX_A(9) NX_A(10) NX_A(11) NX_A(12) NX_A(13)
            NX_A(14) NX_A(15) NX_A(16)
#undef NX_A
        default: break;
        }
    }
    hipLaunchKernelGGL(k_modup_conv_fused, grid, dim3(kBlock), 0, s, T, coef, w, conv, t.mod, o, l, alpha, digits,
                       nout, key_limbs, n);
}

// convd only when every input prime is below 2^51
static const double *inputs_fp(const DeviceTables &t, const double *convd, const InPrimes &inp, uint32_t A)
{
    for (uint32_t k = 0; k < A; k++)
        if (t.mod_host[inp.p[k]] >= kFpPrimeBound) return nullptr;
    return convd;
}

void ks_moddown_convert_fused(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                              const ulonglong2 *w, const uint64_t *conv, const double *convd, uint32_t polys,
                              uint32_t l, uint32_t alpha, uint32_t L, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("ks_moddown_conv", s, 8.0 * n * polys * ((double)alpha + l));
    InPrimes inp;
    for (uint32_t k = 0; k <= kRegAlpha; k++) inp.p[k] = (uint8_t)(L + (k < alpha ? k : 0));
    const uint32_t chunks = (l + kOutChunk - 1) / kOutChunk;
    const dim3 grid((n + kBlock - 1) / kBlock, polys * chunks);
    const dim3 grid_a((n + kBlock - 1) / kBlock, polys * ((l + kMdChunk - 1) / kMdChunk));
    const double *cd = inputs_fp(t, convd, inp, alpha);
    if (t.max_bits <= 58) {
        switch (alpha) {
#define NX_A(A)                                                                                                  \
    case A:                                                                                                      \
        hipLaunchKernelGGL(k_moddown_conv_a<A>, grid_a, dim3(kBlock), 0, s, z, acc, acc_stride, w, conv, cd,     \
                           t.mod, t.qd, inp, l, L, n);                                                           \
        return;
            NX_A(2) NX_A(3) NX_A(4) NX_A(5) NX_A(6) NX_A(7) NX_A(8) NX_A(9) NX_A(10) NX_A(11) NX_A(12) NX_A(13)
            NX_A(14) NX_A(15) NX_A(16)
#undef NX_A
        default: break;
        }
    }
    hipLaunchKernelGGL(k_moddown_conv_fused, grid, dim3(kBlock), 0, s, z, acc, acc_stride, w, conv, t.mod, l, alpha,
                       L, n);
}

void poly_dot_plain_multi(const DeviceTables &t, uint64_t *const *outs, const uint64_t *const *cts,
                          const uint64_t *const *pts, uint32_t rows, uint32_t count, uint32_t limbs, bool accumulate,
                          hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    if (rows == 0 || rows > (uint32_t)kMultiDotG || count > (uint32_t)kMultiDotK)
        throw std::invalid_argument("poly_dot_plain_multi: rows or terms out of range");
    MultiDotArgs a;
    uint32_t terms = 0;
    for (uint32_t r = 0; r < (uint32_t)kMultiDotK; r++) a.ct[r] = r < count ? cts[r] : nullptr;
    for (uint32_t g = 0; g < (uint32_t)kMultiDotG; g++) {
        a.out[g] = g < rows ? outs[g] : nullptr;
        for (uint32_t r = 0; r < (uint32_t)kMultiDotK; r++) {
            a.pt[g][r] = (g < rows && r < count) ? pts[(size_t)g * count + r] : nullptr;
            terms += a.pt[g][r] ? 1u : 0u;
        }
    }
    ProfScope prof("dot_plain", s,
                   8.0 * n * limbs * (2.0 * count + terms + 2.0 * rows * (accumulate ? 2.0 : 1.0)));
    hipLaunchKernelGGL(k_dot_plain_multi, dim3((n / 2 + kBlock - 1) / kBlock, limbs), dim3(kBlock), 0, s, a, count,
                       t.mod, limbs, n, accumulate ? 1 : 0);
}

void poly_dot_plain(const DeviceTables &t, uint64_t *out, const uint64_t *const *cts, const uint64_t *const *pts,
                    uint32_t count, uint32_t limbs, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    DotArgs a;
    for (uint32_t i = 0; i < kMaxDot; i++) {
        a.ct[i] = i < count ? cts[i] : nullptr;
        a.pt[i] = i < count ? pts[i] : nullptr;
    }
    ProfScope prof("dot_plain", s, 8.0 * n * limbs * (3.0 * count + 2.0));
    if (g_dot_layout[dot_bucket(limbs)].load() == 1)
        hipLaunchKernelGGL(k_dot_plain1, dim3((n + kBlock - 1) / kBlock, 2 * limbs), dim3(kBlock), 0, s, out, a,
                           count, t.mod, limbs, n);
    else  // two coefficients per thread (n is a power of two >= 2)
        hipLaunchKernelGGL(k_dot_plain, dim3((n / 2 + kBlock - 1) / kBlock, limbs), dim3(kBlock), 0, s, out, a,
                           count, t.mod, limbs, n);
}


struct TargetPtrs {
    const uint64_t *p[kNttSrcMax];
};

__global__ void k_modup_copy_own(uint64_t *T, const uint64_t *targets, TargetPtrs list, uint32_t l, uint32_t alpha,
                                 uint32_t digits, uint32_t nout, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t b = blockIdx.y / l, I = blockIdx.y - b * l;
    const uint64_t *src = (targets ? targets + (uint64_t)b * l * n : list.p[b]) + (uint64_t)I * n;
    T[(((uint64_t)b * nout + I) * digits + I / alpha) * n + c] = src[c];
}

void ks_modup_copy_own(const DeviceTables &t, uint64_t *T, const uint64_t *targets,
                       const uint64_t *const *target_list, uint32_t B, uint32_t l, uint32_t alpha, uint32_t digits,
                       uint32_t nout, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    TargetPtrs list;
    for (uint32_t i = 0; i < (uint32_t)kNttSrcMax; i++)
        list.p[i] = (!targets && i < B) ? target_list[i] : nullptr;
    if (!targets && B > (uint32_t)kNttSrcMax) throw std::invalid_argument("modup_copy_own: too many targets");
    ProfScope prof("modup_copy", s, 16.0 * n * B * l);
    hipLaunchKernelGGL(k_modup_copy_own, dim3((n + kBlock - 1) / kBlock, B * l), dim3(kBlock), 0, s, T, targets, list,
                       l, alpha, digits, nout, n);
}

bool ks_moddown_rescale_ok(const DeviceTables &t, uint32_t alpha) { return t.max_bits <= 58 && alpha + 1 <= 16; }

void ks_moddown_rescale_convert(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                                const ulonglong2 *w, const uint64_t *conv, const double *convd, uint32_t polys,
                                uint32_t l_out, uint32_t alpha, uint32_t L, hipStream_t s)
{
    const uint32_t n = 1u << t.logn, A = alpha + 1;
    ProfScope prof("ks_moddown_conv", s, 8.0 * n * polys * ((double)A + l_out));
    InPrimes inp;
    for (uint32_t k = 0; k <= kRegAlpha; k++) inp.p[k] = (uint8_t)(k == 0 ? l_out : L + (k - 1 < alpha ? k - 1 : 0));
    const uint32_t chunks = (l_out + kMdChunk - 1) / kMdChunk;
    const dim3 grid((n + kBlock - 1) / kBlock, polys * chunks);
    const double *cd = inputs_fp(t, convd, inp, A);
    switch (A) {
#define NX_A(AA)                                                                                                  \
    case AA:                                                                                                      \
        hipLaunchKernelGGL(k_moddown_conv_a<AA>, grid, dim3(kBlock), 0, s, z, acc, acc_stride, w, conv, cd,       \
                           t.mod, t.qd, inp, l_out, L, n);                                                        \
        return;
        NX_A(3) NX_A(4) NX_A(5) NX_A(6) NX_A(7) NX_A(8) NX_A(9) NX_A(10) NX_A(11) NX_A(12) NX_A(13) NX_A(14)
        NX_A(15) NX_A(16)
#undef NX_A
    default: break;
    }
}
}  // namespace nx
