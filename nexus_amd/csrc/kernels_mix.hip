// Periodic plaintexts and block mixing.
//
// A slot vector that repeats every `period` slots is a polynomial in
// X^(n / B), B = 2 period (rotation by `period` is the automorphism
// X -> X^(5^period), whose fixed ring is spanned by the powers of X^(n / B)).
// In the bit-reversed NTT order the value at position c is the evaluation at
// psi^(2 br(c) + 1), and a polynomial in X^(n / B) only sees that exponent mod
// 2B, i.e. the top log2(B) bits of c: the NTT form is constant over B blocks
// of n / B positions.  A set of such plaintexts is therefore held as
// [entry][limb][block] words, and a product W (.) x costs one word of W per
// block instead of one per position.
//
// block_mix computes out[m] = sum_k W[p(m, k)] (.) in[k] for many outputs at
// once: a workgroup stages a 64-position tile of every input polynomial
// through LDS and each wave accumulates whole outputs from it, so every input
// word is read from HBM once however many outputs use it.
#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 64;  // positions per workgroup (one wave's lanes)

// work[e][i][c] = coef[e][i][c / stride] where c is a multiple of stride, else 0;
// flags cleared
__global__ void k_periodic_scatter(uint64_t *work, uint8_t *flags, const uint64_t *coef, uint32_t blocks,
                                   uint32_t log_stride, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint64_t row = blockIdx.y;  // e * limbs + i
    const uint32_t stride = 1u << log_stride;
    work[row * n + c] = (c & (stride - 1)) ? 0 : coef[row * blocks + (c >> log_stride)];
    if (c < blocks) flags[row * blocks + c] = 0;
}

// table[e][i][b] = work[e][i][b * stride]; flags[e][i][b] = 1 where a position
// of block b differs from its first
__global__ void k_periodic_gather(uint64_t *table, uint8_t *flags, const uint64_t *work, uint32_t blocks,
                                  uint32_t log_stride, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint64_t row = blockIdx.y;
    const uint32_t b = c >> log_stride;
    const uint64_t v = work[row * n + c], first = work[row * n + ((uint64_t)b << log_stride)];
    if (v != first) flags[row * blocks + b] = 1;
    if ((c & ((1u << log_stride) - 1)) == 0) table[row * blocks + b] = v;
}

// out[i][c] = table[index][i][c / stride]
__global__ void k_periodic_expand(uint64_t *out, const uint64_t *table, uint32_t limbs, uint32_t blocks,
                                  uint32_t log_stride, uint32_t n)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    const uint32_t i = blockIdx.y;
    out[(uint64_t)i * n + c] = table[(uint64_t)i * blocks + (c >> log_stride)];
}

// Three-word accumulator of 64x64-bit products from four 32x32-bit products:
// lo32 + 2^32 mid + 2^64 hi.  With both factors below 2^58, 16 products keep
// every word below 2^64.
struct Acc3 {
    uint64_t lo32 = 0, mid = 0, hi = 0;
    __device__ __forceinline__ void mac(uint64_t a, uint64_t b)
    {
        const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
        const uint64_t p00 = a0 * b0;
        lo32 += (uint32_t)p00;
        mid += (p00 >> 32) + a0 * b1 + a1 * b0;
        hi += a1 * b1;
    }
    __device__ __forceinline__ uint64_t reduce(const ModConst &m) const
    {
        const uint64_t ml = mid << 32;
        const uint64_t l = lo32 + ml;
        const uint64_t h = hi + (mid >> 32) + (l < ml);
        return barrett128(h, l, m);
    }
};

// Exact products in double precision for 0 <= a, w < q < 2^51 (see the NTT):
// r = a w - rint(a w / q) q with |r| <= 1.25 q; reduce_f maps |z| < 2^52 to
// |z| <= q/2 (1 + eps).
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

// One (poly, limb) row and a tile of 64 positions per workgroup; LDS holds the
// tile of all K inputs and the weights of the tile's blocks.  runs[first[o] ..
// first[o + 1]) list the terms of output o as runs (k0, entry0, len, dentry):
// terms (k0 + j, entry0 + j dentry) for j < len, so the inner loop computes its
// LDS addresses instead of loading them.  FAST: every prime below 2^58
// (products folded every 16 terms through Acc3); otherwise 128-bit sums folded
// every 32 terms.
template <bool FAST>
__global__ __launch_bounds__(kBlock) void k_block_mix(MixPtrs ptrs, const uint64_t *__restrict__ table,
                                                      const int4 *__restrict__ runs,
                                                      const uint32_t *__restrict__ first, const ModConst *mod,
                                                      uint32_t K, uint32_t M, uint32_t limbs, uint32_t count,
                                                      uint32_t blocks, uint32_t log_stride, uint32_t n)
{
    extern __shared__ uint64_t lds[];
    const uint32_t row = blockIdx.y;  // poly * limbs + limb
    const uint32_t limb = row % limbs;
    const uint32_t pos0 = blockIdx.x * kTile;
    const uint32_t bpt = log_stride >= 6 ? 1u : (uint32_t)kTile >> log_stride;  // blocks per tile
    const uint32_t b0 = pos0 >> log_stride;
    uint64_t *xs = lds;               // [K][64]
    uint64_t *ws = lds + K * kTile;   // [count][bpt]
    const uint64_t roff = (uint64_t)row * n + pos0;
    const ModConst m = mod[limb];
    const uint32_t lane = threadIdx.x % kTile;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);
    const uint32_t lb = log_stride >= 6 ? 0u : lane >> log_stride;
    if (m.q < kFpPrimeBound) {
        // double precision: the tile and the weights as doubles, one exact
        // product per term, the sum reduced every second term (|sum| < 3 q)
        double *xd = reinterpret_cast<double *>(xs), *wd = reinterpret_cast<double *>(ws);
        // wave w stages input rows w, w + 4, ...: the row (and its pointer) is
        // uniform across the wave, eight loads in flight
        {
            const uint32_t w0 = wave;
            uint32_t k = w0;
            for (; k + 28 < K; k += 32) {
                uint64_t v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = ptrs.in[k + 4 * u][roff + lane];
#pragma unroll
                for (int u = 0; u < 8; u++) xd[(k + 4 * u) * kTile + lane] = (double)v[u];
            }
            for (; k < K; k += 4) xd[k * kTile + lane] = (double)ptrs.in[k][roff + lane];
        }
        for (uint32_t e = threadIdx.x; e < count * bpt; e += kBlock) {
            const uint32_t p = e / bpt, b = e - p * bpt;
            wd[e] = (double)table[((uint64_t)p * limbs + limb) * blocks + b0 + b];
        }
        __syncthreads();
        const double q = (double)m.q, qinv = 1.0 / q;
        for (uint32_t o = wave; o < M; o += kBlock / kTile) {
            // two independent sums (each reduced after every second term of its
            // own) so consecutive products do not wait on one another
            double a0 = 0.0, a1 = 0.0;
            for (uint32_t ri = first[o]; ri < first[o + 1]; ri++) {
                // a run: k = k0 + j, entry = p0 + j dp for j < len (indices computed, not loaded)
                const int4 rn = runs[ri];
                const double *xp = xd + rn.x * kTile + lane;
                const int xstep = kTile, wstep = rn.w * (int)bpt;
                const double *wp = wd + rn.y * (int)bpt + lb;
                int j = 0;
                for (; j + 4 <= rn.z; j += 4) {
                    a0 += mulmod_f(xp[0], wp[0], q, qinv);
                    a1 += mulmod_f(xp[xstep], wp[wstep], q, qinv);
                    a0 += mulmod_f(xp[2 * xstep], wp[2 * wstep], q, qinv);
                    a1 += mulmod_f(xp[3 * xstep], wp[3 * wstep], q, qinv);
                    a0 = reduce_f(a0, q, qinv);
                    a1 = reduce_f(a1, q, qinv);
                    xp += 4 * xstep;
                    wp += 4 * wstep;
                }
                for (; j < rn.z; j++) {
                    a0 = reduce_f(a0 + mulmod_f(xp[0], wp[0], q, qinv), q, qinv);
                    xp += xstep;
                    wp += wstep;
                }
            }
            double acc = reduce_f(a0 + a1, q, qinv);
            acc = acc < 0 ? acc + q : acc;
            ptrs.out[o][roff + lane] = (uint64_t)acc;
        }
        return;
    }
    {
        uint32_t k = wave;
        for (; k + 28 < K; k += 32) {
            uint64_t v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = ptrs.in[k + 4 * u][roff + lane];
#pragma unroll
            for (int u = 0; u < 8; u++) xs[(k + 4 * u) * kTile + lane] = v[u];
        }
        for (; k < K; k += 4) xs[k * kTile + lane] = ptrs.in[k][roff + lane];
    }
    for (uint32_t e = threadIdx.x; e < count * bpt; e += kBlock) {
        const uint32_t p = e / bpt, b = e - p * bpt;
        ws[e] = table[((uint64_t)p * limbs + limb) * blocks + b0 + b];
    }
    __syncthreads();
    for (uint32_t o = wave; o < M; o += kBlock / kTile) {
        uint64_t r = 0;
        Acc3 acc;
        uint64_t lo = 0, hi = 0;
        uint32_t cnt = 0;
        for (uint32_t ri = first[o]; ri < first[o + 1]; ri++) {
            const int4 rn = runs[ri];
            for (int j = 0; j < rn.z; j++) {
                const uint64_t x = xs[(rn.x + j) * kTile + lane], w = ws[(rn.y + j * rn.w) * (int)bpt + lb];
                if (FAST) {
                    acc.mac(x, w);
                    if (++cnt == 16) {
                        r = add_mod(r, acc.reduce(m), m.q);
                        acc = Acc3();
                        cnt = 0;
                    }
                } else {
                    const uint64_t plo = x * w, phi = mulhi64(x, w);
                    lo += plo;
                    hi += phi + (lo < plo);
                    if (++cnt == 32) {
                        r = add_mod(r, barrett128(hi, lo, m), m.q);
                        lo = hi = 0;
                        cnt = 0;
                    }
                }
            }
        }
        r = FAST ? add_mod(r, acc.reduce(m), m.q) : add_mod(r, barrett128(hi, lo, m), m.q);
        ptrs.out[o][roff + lane] = r;
    }
}

}  // namespace

void periodic_scatter(const DeviceTables &t, uint64_t *work, uint8_t *flags, const uint64_t *coef, uint32_t rows,
                      uint32_t blocks, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    const uint32_t log_stride = (uint32_t)t.logn - (uint32_t)__builtin_ctz(blocks);
    hipLaunchKernelGGL(k_periodic_scatter, dim3((n + kBlock - 1) / kBlock, rows), dim3(kBlock), 0, s, work, flags,
                       coef, blocks, log_stride, n);
}

void periodic_gather(const DeviceTables &t, uint64_t *table, uint8_t *flags, const uint64_t *work, uint32_t rows,
                     uint32_t blocks, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    const uint32_t log_stride = (uint32_t)t.logn - (uint32_t)__builtin_ctz(blocks);
    hipLaunchKernelGGL(k_periodic_gather, dim3((n + kBlock - 1) / kBlock, rows), dim3(kBlock), 0, s, table, flags,
                       work, blocks, log_stride, n);
}

void periodic_expand(const DeviceTables &t, uint64_t *out, const uint64_t *table, uint32_t limbs, uint32_t blocks,
                     hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    const uint32_t log_stride = (uint32_t)t.logn - (uint32_t)__builtin_ctz(blocks);
    hipLaunchKernelGGL(k_periodic_expand, dim3((n + kBlock - 1) / kBlock, limbs), dim3(kBlock), 0, s, out, table,
                       limbs, blocks, log_stride, n);
}

size_t block_mix_lds_bytes(uint32_t K, uint32_t count, uint32_t blocks, int logn)
{
    const uint32_t log_stride = (uint32_t)logn - (uint32_t)__builtin_ctz(blocks);
    const uint32_t bpt = log_stride >= 6 ? 1u : (uint32_t)kTile >> log_stride;
    return 8 * ((size_t)K * kTile + (size_t)count * bpt);
}

void block_mix(const DeviceTables &t, const MixPtrs &p, const uint64_t *table, const int32_t *terms,
               const uint32_t *first, uint32_t K, uint32_t M, uint32_t nterms, uint32_t limbs, uint32_t count,
               uint32_t blocks, hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    const uint32_t log_stride = (uint32_t)t.logn - (uint32_t)__builtin_ctz(blocks);
    const size_t lds = block_mix_lds_bytes(K, count, blocks, t.logn);
    // inputs read once, outputs written once (both polynomials of every limb)
    ProfScope prof("block_mix", s, 16.0 * n * limbs * ((double)K + (double)M));
    (void)nterms;
    const dim3 grid(n / kTile, 2 * limbs);
    if (t.max_bits <= 58)
        hipLaunchKernelGGL(k_block_mix<true>, grid, dim3(kBlock), lds, s, p, table,
                           reinterpret_cast<const int4 *>(terms), first, t.mod, K, M, limbs, count, blocks,
                           log_stride, n);
    else
        hipLaunchKernelGGL(k_block_mix<false>, grid, dim3(kBlock), lds, s, p, table,
                           reinterpret_cast<const int4 *>(terms), first, t.mod, K, M, limbs, count, blocks,
                           log_stride, n);
}

}  // namespace nx
