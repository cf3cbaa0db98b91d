// Hybrid key-switching base conversions (reference Phantom modup /
// moddown_from_NTT, cuda/thirdparty/phantom-fhe/src/rns_bconv.cu:455-634,
// 776-828): fast basis extension between a digit's primes (or the special
// primes) and every other prime of the key-switching basis.
//
//   modup:   T[I][j] = sum_{i < s_j} [x_{j alpha + i} Dhat_i^-1]_{q_{j alpha + i}} * (Dhat_i mod t_I)  mod t_I
//   moddown: z[p][i] = sum_{k < alpha} v_k * (Phat_k mod q_i) - r * (P mod q_i)  mod q_i,
//            v_k = [y_k Phat_k^-1]_{p_k},  r = round(sum_k v_k / p_k)
// The ModDown conversion is centred: sum_k v_k Phat_k - r P is the
// representative of y in (-P/2, P/2] whenever the float64 rounding r is right,
// i.e. unless sum_k v_k / p_k lies within about 2^-40 of a half-integer (then
// the representative is off by P and the rounding of x / P by one), so
// (x - z) / P = round(x / P) as in the single-special-prime path.  Each term is below 2^122 and at most 33
// terms are summed, so one 128-bit accumulator and one Barrett reduction per
// output word suffice.
#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

constexpr int kBlock = 256;

struct OutMap {
    uint8_t p[kMaxPrimes];
};

__global__ void k_modup_conv(uint64_t *T, const uint64_t *coef, const ulonglong2 *w, const uint64_t *conv,
                             const ModConst *mod, OutMap outp, uint32_t l, uint32_t alpha, uint32_t digits,
                             uint32_t nout, uint32_t key_limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // (b * nout + I) * digits + j
    const uint32_t bi = row / digits, j = row - bi * digits;
    const uint32_t b = bi / nout, I = bi - b * nout;
    coef += (uint64_t)b * l * n;
    const uint32_t kp = outp.p[I];
    const uint32_t first = j * alpha;
    const uint32_t s = min(alpha, l - first);
    const uint32_t base = (j * alpha + s - 1) * alpha;
    uint64_t lo = 0, hi = 0;
    for (uint32_t i = 0; i < s; i++) {
        const uint32_t g = first + i;
        const ulonglong2 wi = w[base + i];
        const uint64_t y = shoup(coef[(uint64_t)g * n + c], wi.x, wi.y, mod[g].q);
        const uint64_t f = conv[(uint64_t)(base + i) * key_limbs + kp];
        const uint64_t plo = y * f, phi = mulhi64(y, f);
        lo += plo;
        hi += phi + (lo < plo);
    }
    T[(uint64_t)row * n + c] = barrett128(hi, lo, mod[kp]);
}

__global__ void k_moddown_conv(uint64_t *z, const uint64_t *acc, uint64_t acc_stride, const ulonglong2 *w,
                               const uint64_t *conv, const ModConst *mod, uint32_t l, uint32_t alpha, uint32_t L,
                               uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // poly * l + i
    const uint32_t p = row / l, i = row - p * l;
    const uint64_t *y = acc + (uint64_t)p * acc_stride + (uint64_t)l * n + c;  // special limbs follow the l data limbs
    const uint64_t qi = mod[i].q;
    uint64_t lo = 0, hi = 0;
    double frac = 0.0;
    for (uint32_t k = 0; k < alpha; k++) {
        const ulonglong2 wk = w[k];
        const uint64_t pk = mod[L + k].q;
        const uint64_t v = shoup(y[(uint64_t)k * n], wk.x, wk.y, pk);
        frac += (double)v / (double)pk;
        const uint64_t f = conv[(uint64_t)k * L + i];
        const uint64_t plo = v * f, phi = mulhi64(v, f);
        lo += plo;
        hi += phi + (lo < plo);
    }
    // subtract r * (P mod q_i): add r * (q_i - P mod q_i)
    const uint64_t r = (uint64_t)floor(frac + 0.5);
    const uint64_t negP = qi - conv[(uint64_t)alpha * L + i];
    const uint64_t plo = r * negP, phi = mulhi64(r, negP);
    lo += plo;
    hi += phi + (lo < plo);
    z[(uint64_t)row * n + c] = barrett128(hi, lo, mod[i]);
}

}  // namespace

void ks_modup_convert(const DeviceTables &t, uint64_t *T, const uint64_t *coef, const ulonglong2 *w,
                      const uint64_t *conv, uint32_t B, uint32_t l, uint32_t alpha, uint32_t key_limbs,
                      const uint8_t *out_primes_host, uint32_t nout, hipStream_t s)
{
    const uint32_t n = 1u << t.logn, digits = (l + alpha - 1) / alpha;
    OutMap o;
    for (uint32_t i = 0; i < kMaxPrimes; i++) o.p[i] = i < nout ? out_primes_host[i] : 0;
    ProfScope prof("ks_modup_conv", s, 8.0 * n * B * ((double)l + (double)nout * digits));
    hipLaunchKernelGGL(k_modup_conv, dim3((n + kBlock - 1) / kBlock, B * nout * digits), dim3(kBlock), 0, s, T, coef,
                       w, conv, t.mod, o, l, alpha, digits, nout, key_limbs, n);
}

void ks_moddown_convert(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                        const ulonglong2 *w, const uint64_t *conv, uint32_t polys, uint32_t l, uint32_t alpha,
                        uint32_t L, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("ks_moddown_conv", s, 8.0 * n * polys * ((double)alpha + l));
    hipLaunchKernelGGL(k_moddown_conv, dim3((n + kBlock - 1) / kBlock, polys * l), dim3(kBlock), 0, s, z, acc,
                       acc_stride, w, conv, t.mod, l, alpha, L, n);
}

}  // namespace nx
