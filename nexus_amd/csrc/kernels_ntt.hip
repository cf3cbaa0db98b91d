// Negacyclic NTT / INTT over RNS limbs for gfx950.
//
// Transform definition (SEAL, util/ntt.cpp:241-477 and util/dwthandler.h):
//  forward: Cooley-Tukey stages m = 1, 2, ..., n/2 with gap n/(2m); butterfly
//           block i of stage m uses fwd[m + i]; natural order in, bit-reversed out.
//  inverse: Gentleman-Sande stages m = n/2, ..., 1 with gap n/(2m); block i of
//           stage m uses inv[n - 2m + 1 + i]; followed by multiplication by n^-1.
// Outputs are fully reduced, so they are bit-identical to SEAL's however the
// lazy reductions inside are scheduled.
//
// Mapping.  n <= 2^13: one workgroup holds a whole limb in LDS.  n >= 2^14: the
// limb is viewed as an N1 x N2 matrix (index = t * N2 + c).  The first log2(N1)
// forward stages only pair elements of the same column c (gap >= N2), the last
// log2(N2) stages only pair elements of the same row t; the inverse runs rows
// first, then columns.  Each pass stages a tile of 16 columns (or 8 contiguous
// rows) through LDS and runs radix-8/4/2 register butterflies between
// barriers, 256 threads (four wave64) per workgroup.  Row twiddles for row r at
// local stage s are the contiguous block fwd[2^s (N1 + r) ...].
//
// Arithmetic.  Limbs of primes q < 2^51 run in double precision, where the
// FMA is full rate: with |a| <= q (1 + eps) and 0 <= w < q,
//     h = a w (rounded), e = fma(a, w, -h) = a w - h exactly,
//     Q = rint(h / q) (|Q - a w / q| <= 1.25), r = fma(-Q, q, h) + e = a w - Q q
// is exact (every intermediate is an integer below 2^53) with |r| <= 1.25 q;
// reduce(z) = fma(-rint(z / q), q, z) brings |z| < 2^52 to |z| <= q/2 (1 + eps).
// Butterfly outputs are reduced, loads are reduced, and stores are mapped to
// [0, q), so the output words equal the integer path's.  Larger primes use
// 64-bit Shoup butterflies (ten 32-bit multiplies each) on [0, 4q).  This
// file is compiled without floating-point contraction.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <stdexcept>

#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

enum Role { kWhole = 0, kCol = 1, kRow = 2 };

struct NttKArgs {
    uint64_t *data;
    uint64_t poly_stride;
    const uint64_t *src;  // input of this pass (null: in place)
    uint64_t src_poly_stride;
    uint32_t nsrcs;       // > 0: poly p of this pass reads srcs[p]
    const uint64_t *srcs[kNttSrcMax];
    double tw_bytes;      // mean twiddle bytes per entry over the batch (profiler accounting)
    uint32_t limbs;
    uint32_t logn;
    uint32_t prime_by_poly;
    uint32_t poly_period;
    const ulonglong2 *tw;
    const ModConst *mod;
    const ulonglong2 *ninv;
    const double *twd;    // the same twiddles as doubles
    const double *ninvd;  // n^-1 mod q as doubles
    uint32_t skip_l, skip_alpha, skipped;  // NttBatch::skip_*; skipped: limbs left out
    NttEpilogue epi;                        // applied by the forward transform's last pass
    uint32_t epi_int;                       // double-precision path: the epilogue in 64-bit integers
    uint8_t prime_of[kMaxPrimes];
};

__host__ __device__ inline bool epi_on(const NttEpilogue &e) { return e.out != nullptr || e.outs[0] != nullptr; }

// NttEpilogue's store of the transform value t (in [0, q)) of poly p, limb i
// of a batch of l limbs
__device__ __forceinline__ void epi_store(const NttEpilogue &e, uint32_t p, uint32_t i, uint64_t gidx, uint64_t t,
                                          uint64_t q, uint32_t l, uint32_t n)
{
    const uint64_t off = (uint64_t)i * n + gidx;
    const ulonglong2 w = e.inv[i];
    const uint32_t ct = p / e.polys_per, k = p - ct * e.polys_per;
    const uint64_t xv = e.xs[0] ? e.xs[ct][((uint64_t)k * e.x_limbs + i) * n + gidx] : e.x[p * e.x_stride + off];
    uint64_t r = shoup(sub_mod(xv, t, q), w.x, w.y, q);
    if (e.out) {
        uint64_t *o = e.out + p * e.out_stride + off;
        if (e.accumulate == 1 || (e.accumulate == 2 && (p & 1) == 0))
            r = add_mod(e.base ? e.base[p * e.base_stride + off] : *o, r, q);
        *o = r;
        return;
    }
    const uint64_t *a0 = e.a0s[0] ? e.a0s[ct] : e.a0;
    if (a0 && k == 0) r = add_mod(r, a0[(uint64_t)i * n + e.perm[ct][gidx]], q);
    if (e.d) {
        const ulonglong2 w2 = e.inv2[i];
        r = add_mod(r, shoup(e.d[p * e.d_stride + off], w2.x, w2.y, q), q);
    }
    e.outs[ct][((uint64_t)k * l + i) * n + gidx] = r;
}

constexpr int kThreads = 256;
constexpr int kTuneLogNMin = 14, kTuneLogN = 4, kTuneBuckets = 9;

// CW: sub-transforms per workgroup -- columns of the column pass (8 by
// default: a 16 KB tile keeps twice the workgroups of a 16-wide one resident,
// and once a batch outgrows the last-level cache the strided column reads of
// 8-wide tiles stream faster), and an upper bound on the rows of the row pass.
template <int LOGS, int ROLE, int CW>
struct Geometry {
    static constexpr int S = 1 << LOGS;
    static constexpr int ROWS = S >= 2048 ? 1 : 2048 / S;
    static constexpr int C = ROLE == kCol ? CW : (ROLE == kRow ? (ROWS < CW ? ROWS : CW) : 1);
    static constexpr int TILE = S * C;
};

// LDS address of local position `pos` of sub-transform `cc`.
template <int ROLE, int S, int C>
__device__ __forceinline__ int lds_addr(int cc, int pos)
{
    return ROLE == kCol ? pos * C + cc : cc * S + pos;
}

// Forward radix-2^K step over local stages s .. s+K-1.
template <int LOGS, int ROLE, int K, int CW>
__device__ __forceinline__ void fwd_step(uint64_t *lds, int s, const ulonglong2 *tw, uint64_t q, uint32_t m0,
                                         uint32_t row0)
{
    using G = Geometry<LOGS, ROLE, CW>;
    constexpr int R = 1 << K;
    constexpr int GROUPS_PER = G::S / R;
    constexpr int GROUPS = G::TILE / R;
    const uint64_t two_q = 2 * q;
    const int B = G::S >> s;
    const int D = B >> K;
    for (int g = threadIdx.x; g < GROUPS; g += kThreads) {
        int cc, gg;
        if (ROLE == kCol) {
            cc = g % G::C;
            gg = g / G::C;
        } else {
            cc = g / GROUPS_PER;
            gg = g % GROUPS_PER;
        }
        const int blockid = gg / D;
        const int o = gg - blockid * D;
        const uint32_t mrow = (ROLE == kRow) ? (m0 + row0 + cc) : m0;
        uint64_t v[R];
        int addr[R];
#pragma unroll
        for (int t = 0; t < R; t++) {
            addr[t] = lds_addr<ROLE, G::S, G::C>(cc, blockid * B + o + t * D);
            v[t] = lds[addr[t]];
        }
#pragma unroll
        for (int d = 0; d < K; d++) {
            const int h = 1 << (K - 1 - d);
#pragma unroll
            for (int j = 0; j < (1 << d); j++) {
                const uint32_t b = (uint32_t)blockid * (1u << d) + j;
                const ulonglong2 w = tw[((uint32_t)1 << (s + d)) * mrow + b];
#pragma unroll
                for (int u = 0; u < h; u++) {
                    const int t = j * 2 * h + u;
                    uint64_t x = v[t];
                    x = x >= two_q ? x - two_q : x;
                    const uint64_t y = shoup_lazy(v[t + h], w.x, w.y, q);
                    v[t] = x + y;
                    v[t + h] = x - y + two_q;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < R; t++) lds[addr[t]] = v[t];
    }
}

// Inverse radix-2^K step over local stages s .. s+K-1 (gaps 2^s .. 2^(s+K-1)).
template <int LOGS, int ROLE, int K, int CW>
__device__ __forceinline__ void inv_step(uint64_t *lds, int s, const ulonglong2 *tw, uint64_t q, uint32_t n,
                                         uint32_t mx, uint32_t row0)
{
    using G = Geometry<LOGS, ROLE, CW>;
    constexpr int R = 1 << K;
    constexpr int GROUPS_PER = G::S / R;
    constexpr int GROUPS = G::TILE / R;
    const uint64_t two_q = 2 * q;
    const int so = 1 << s;
    for (int g = threadIdx.x; g < GROUPS; g += kThreads) {
        int cc, gg;
        if (ROLE == kCol) {
            cc = g % G::C;
            gg = g / G::C;
        } else {
            cc = g / GROUPS_PER;
            gg = g % GROUPS_PER;
        }
        const int o = gg & (so - 1);
        const int blockid = gg >> s;
        const int base = blockid * (so << K) + o;
        const uint32_t r = row0 + cc;
        uint64_t v[R];
        int addr[R];
#pragma unroll
        for (int t = 0; t < R; t++) {
            addr[t] = lds_addr<ROLE, G::S, G::C>(cc, base + t * so);
            v[t] = lds[addr[t]];
        }
#pragma unroll
        for (int d = 0; d < K; d++) {
            const int sd = s + d;
            const uint32_t first = n - (mx >> sd) + 1 + (ROLE == kRow ? r * ((uint32_t)G::S >> (sd + 1)) : 0u);
#pragma unroll
            for (int j = 0; j < (R >> (d + 1)); j++) {
                const uint32_t b = (uint32_t)blockid * (uint32_t)(R >> (d + 1)) + j;
                const ulonglong2 w = tw[first + b];
#pragma unroll
                for (int u = 0; u < (1 << d); u++) {
                    const int t = j * (2 << d) + u;
                    const uint64_t x = v[t], y = v[t + (1 << d)];
                    uint64_t sum = x + y;
                    v[t] = sum >= two_q ? sum - two_q : sum;
                    v[t + (1 << d)] = shoup_lazy(x - y + two_q, w.x, w.y, q);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < R; t++) lds[addr[t]] = v[t];
    }
}

constexpr uint64_t kFpPrimeBound = 1ull << 51;

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

// epi_store in double precision (every prime of the batch below 2^51): the
// transform value t in [0, q) as a double, the Shoup constants' values (below
// q) as factors of exact products.  |x - t| < q, each product is within 1.25 q
// and a term in [0, q) is added at most twice, so the sum stays below 2^53 and
// the stored residue equals epi_store's.
__device__ __forceinline__ void epi_store_f(const NttEpilogue &e, uint32_t p, uint32_t i, uint64_t gidx, double t,
                                            double q, double qinv, uint32_t l, uint32_t n)
{
    const uint64_t off = (uint64_t)i * n + gidx;
    const uint32_t ct = p / e.polys_per, k = p - ct * e.polys_per;
    const uint64_t xv = e.xs[0] ? e.xs[ct][((uint64_t)k * e.x_limbs + i) * n + gidx] : e.x[p * e.x_stride + off];
    double r = mulmod_f((double)xv - t, (double)e.inv[i].x, q, qinv);
    uint64_t *o;
    if (e.out) {
        o = e.out + p * e.out_stride + off;
        if (e.accumulate == 1 || (e.accumulate == 2 && (p & 1) == 0))
            r += (double)(e.base ? e.base[p * e.base_stride + off] : *o);
    } else {
        const uint64_t *a0 = e.a0s[0] ? e.a0s[ct] : e.a0;
        if (a0 && k == 0) r += (double)a0[(uint64_t)i * n + e.perm[ct][gidx]];
        if (e.d) r += mulmod_f((double)e.d[p * e.d_stride + off], (double)e.inv2[i].x, q, qinv);
        o = e.outs[ct] + ((uint64_t)k * l + i) * n + gidx;
    }
    r = reduce_f(r, q, qinv);
    *o = (uint64_t)(r < 0 ? r + q : r);
}

// The butterflies of one forward radix-2^K group in registers: local stages
// s .. s+K-1 of block `blockid`, twiddle row mrow (1 for columns).
template <int K, bool LAZY>
__device__ __forceinline__ void fwd_group_f(double *v, int s, uint32_t blockid, uint32_t mrow, const double *tw,
                                            double q, double qinv)
{
#pragma unroll
    for (int d = 0; d < K; d++) {
        const int h = 1 << (K - 1 - d);
#pragma unroll
        for (int j = 0; j < (1 << d); j++) {
            const uint32_t b = blockid * (1u << d) + j;
            const double w = tw[((uint32_t)1 << (s + d)) * mrow + b];
#pragma unroll
            for (int u = 0; u < h; u++) {
                const int t = j * 2 * h + u;
                const double x = v[t];
                const double y = mulmod_f(v[t + h], w, q, qinv);
                if (LAZY) {
                    v[t] = x + y;
                    v[t + h] = x - y;
                } else {
                    v[t] = reduce_f(x + y, q, qinv);
                    v[t + h] = reduce_f(x - y, q, qinv);
                }
            }
        }
    }
}

// Forward radix-2^K step in double precision (values |v| <= q/2 (1 + eps)).
// LAZY: the sums and differences are left unreduced (see ntt_tile_f).
template <int LOGS, int ROLE, int K, int CW, bool LAZY, int NT = kThreads>
__device__ __forceinline__ void fwd_step_f(double *lds, int s, const double *tw, double q, double qinv, uint32_t m0,
                                           uint32_t row0)
{
    using G = Geometry<LOGS, ROLE, CW>;
    constexpr int R = 1 << K;
    constexpr int GROUPS_PER = G::S / R;
    constexpr int GROUPS = G::TILE / R;
    const int B = G::S >> s;
    const int D = B >> K;
    for (int g = threadIdx.x; g < GROUPS; g += NT) {
        int cc, gg;
        if (ROLE == kCol) {
            cc = g % G::C;
            gg = g / G::C;
        } else {
            cc = g / GROUPS_PER;
            gg = g % GROUPS_PER;
        }
        const int blockid = gg / D;
        const int o = gg - blockid * D;
        const uint32_t mrow = (ROLE == kRow) ? (m0 + row0 + cc) : m0;
        double v[R];
        int addr[R];
#pragma unroll
        for (int t = 0; t < R; t++) {
            addr[t] = lds_addr<ROLE, G::S, G::C>(cc, blockid * B + o + t * D);
            v[t] = lds[addr[t]];
        }
        fwd_group_f<K, LAZY>(v, s, (uint32_t)blockid, mrow, tw, q, qinv);
#pragma unroll
        for (int t = 0; t < R; t++) lds[addr[t]] = v[t];
    }
}

// Inverse radix-2^K step in double precision.  LAZY: the products are left
// unreduced (|v| <= 1.07 q; see ntt_tile_f).
template <int LOGS, int ROLE, int K, int CW, bool LAZY, int NT = kThreads>
__device__ __forceinline__ void inv_step_f(double *lds, int s, const double *tw, double q, double qinv, uint32_t n,
                                           uint32_t mx, uint32_t row0)
{
    using G = Geometry<LOGS, ROLE, CW>;
    constexpr int R = 1 << K;
    constexpr int GROUPS_PER = G::S / R;
    constexpr int GROUPS = G::TILE / R;
    const int so = 1 << s;
    for (int g = threadIdx.x; g < GROUPS; g += NT) {
        int cc, gg;
        if (ROLE == kCol) {
            cc = g % G::C;
            gg = g / G::C;
        } else {
            cc = g / GROUPS_PER;
            gg = g % GROUPS_PER;
        }
        const int o = gg & (so - 1);
        const int blockid = gg >> s;
        const int base = blockid * (so << K) + o;
        const uint32_t r = row0 + cc;
        double v[R];
        int addr[R];
#pragma unroll
        for (int t = 0; t < R; t++) {
            addr[t] = lds_addr<ROLE, G::S, G::C>(cc, base + t * so);
            v[t] = lds[addr[t]];
        }
#pragma unroll
        for (int d = 0; d < K; d++) {
            const int sd = s + d;
            const uint32_t first = n - (mx >> sd) + 1 + (ROLE == kRow ? r * ((uint32_t)G::S >> (sd + 1)) : 0u);
#pragma unroll
            for (int j = 0; j < (R >> (d + 1)); j++) {
                const uint32_t b = (uint32_t)blockid * (uint32_t)(R >> (d + 1)) + j;
                const double w = tw[first + b];
#pragma unroll
                for (int u = 0; u < (1 << d); u++) {
                    const int t = j * (2 << d) + u;
                    const double x = v[t], y = v[t + (1 << d)];
                    v[t] = reduce_f(x + y, q, qinv);
                    const double p = mulmod_f(x - y, w, q, qinv);
                    v[t + (1 << d)] = LAZY ? p : reduce_f(p, q, qinv);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < R; t++) lds[addr[t]] = v[t];
    }
}

// The radix steps of one pass in double precision, barrier after each.
template <int LOGS, int ROLE, bool FWD, int CW, bool LAZY, int NT = kThreads, bool R16 = false>
__device__ __forceinline__ void stages_f(double *lds, const double *tw, double q, double qinv, uint32_t n,
                                         uint32_t N1, uint32_t first)
{
    if constexpr (R16) {
        // radix-16 steps (half the LDS passes and barriers of radix-8), then
        // the remaining 1..3 stages in one step
        int s = 0;
        if (FWD) {
            const uint32_t m0 = ROLE == kRow ? N1 : 1u;
#pragma unroll
            for (int st = 0; st < LOGS / 4; st++) {
                fwd_step_f<LOGS, ROLE, 4, CW, LAZY, NT>(lds, s, tw, q, qinv, m0, first);
                s += 4;
                __syncthreads();
            }
            if constexpr (LOGS % 4 != 0) {
                fwd_step_f<LOGS, ROLE, LOGS % 4, CW, LAZY, NT>(lds, s, tw, q, qinv, m0, first);
                __syncthreads();
            }
        } else {
            const uint32_t mx = ROLE == kCol ? N1 : n;
#pragma unroll
            for (int st = 0; st < LOGS / 4; st++) {
                inv_step_f<LOGS, ROLE, 4, CW, LAZY, NT>(lds, s, tw, q, qinv, n, mx, first);
                s += 4;
                __syncthreads();
            }
            if constexpr (LOGS % 4 != 0) {
                inv_step_f<LOGS, ROLE, LOGS % 4, CW, LAZY, NT>(lds, s, tw, q, qinv, n, mx, first);
                __syncthreads();
            }
        }
        return;
    }
    int s = 0;
    if (FWD) {
        const uint32_t m0 = ROLE == kRow ? N1 : 1u;
        if constexpr (LOGS >= 3) {
#pragma unroll
            for (int st = 0; st < LOGS / 3; st++) {
                fwd_step_f<LOGS, ROLE, 3, CW, LAZY, NT>(lds, s, tw, q, qinv, m0, first);
                s += 3;
                __syncthreads();
            }
        }
        if constexpr (LOGS % 3 == 2) {
            fwd_step_f<LOGS, ROLE, 2, CW, LAZY, NT>(lds, s, tw, q, qinv, m0, first);
            __syncthreads();
        } else if constexpr (LOGS % 3 == 1) {
            fwd_step_f<LOGS, ROLE, 1, CW, LAZY, NT>(lds, s, tw, q, qinv, m0, first);
            __syncthreads();
        }
    } else {
        const uint32_t mx = ROLE == kCol ? N1 : n;
        if constexpr (LOGS >= 3) {
#pragma unroll
            for (int st = 0; st < LOGS / 3; st++) {
                inv_step_f<LOGS, ROLE, 3, CW, LAZY, NT>(lds, s, tw, q, qinv, n, mx, first);
                s += 3;
                __syncthreads();
            }
        }
        if constexpr (LOGS % 3 == 2) {
            inv_step_f<LOGS, ROLE, 2, CW, LAZY, NT>(lds, s, tw, q, qinv, n, mx, first);
            __syncthreads();
        } else if constexpr (LOGS % 3 == 1) {
            inv_step_f<LOGS, ROLE, 1, CW, LAZY, NT>(lds, s, tw, q, qinv, n, mx, first);
            __syncthreads();
        }
    }
}

// The whole transform body of one tile in double precision: load (reduced),
// the radix steps, store in [0, q) (times n^-1 on the inverse's last pass).
//
// Lazy reduction, for primes with q (1/2 + 5/4 LOGS) <= 2^50.  Let B bound |v|.
// mulmod_f(a, w) for |a| <= 2^50, 0 <= w < q: |h / q - a w / q| <= 3.3 2^-53
// |a w / q| <= 0.41, so |Q - a w / q| <= 0.91 and the result is below 0.91 q
// (every intermediate stays an integer below 2^53, as for reduced inputs).
// Forward (Cooley-Tukey): x +- mulmod_f(y, w) grows B by at most 0.91 q a
// stage from q/2 after the load, so B <= q (1/2 + 0.91 LOGS) <= 2^50 throughout
// and no sum needs reducing; the store reduces once.  Inverse
// (Gentleman-Sande): the sums are reduced (|x + y| <= 1.82 q), the products
// are not (<= 0.91 q); the store (or the n^-1 product) reduces once.  The
// stored words are the fully reduced residues either way.
template <int LOGS, int ROLE, bool FWD, int CW, bool R16>
__device__ __forceinline__ void ntt_tile_f(const NttKArgs &a, uint64_t *lds_words, uint64_t *data,
                                           const uint64_t *src, uint32_t prime, uint32_t n, uint32_t N2, uint32_t N1,
                                           uint32_t first, uint32_t poly, uint32_t limb)
{
    using G = Geometry<LOGS, ROLE, CW>;
    const bool epi = FWD && ROLE != kCol && epi_on(a.epi);
    double *lds = reinterpret_cast<double *>(lds_words);
    const double q = (double)a.mod[prime].q;
    const double qinv = 1.0 / q;
    const double *tw = a.twd + (uint64_t)prime * n;
    const bool lazy = q * (0.5 + 1.25 * LOGS) <= 0x1p50;  // uniform: one prime per workgroup
    if constexpr (FWD && ROLE == kRow && LOGS == 8 && G::C == 8 && !R16) {
        // rows of 256, 8 rows per tile, 256 threads: each thread loads one
        // radix-8 group of the first three stages (row g / 32, positions
        // g % 32 + 32 t) and runs it in registers; one LDS pass and one barrier
        // fewer, the rest as below
        {
            const int g = threadIdx.x, cc = g / 32, o = g % 32;
            double v[8];
#pragma unroll
            for (int t = 0; t < 8; t++)
                v[t] = reduce_f((double)src[(uint64_t)(first + cc) * G::S + o + 32 * t], q, qinv);
            if (lazy)
                fwd_group_f<3, true>(v, 0, 0, N1 + first + cc, tw, q, qinv);
            else
                fwd_group_f<3, false>(v, 0, 0, N1 + first + cc, tw, q, qinv);
#pragma unroll
            for (int t = 0; t < 8; t++) lds[cc * G::S + o + 32 * t] = v[t];
        }
        __syncthreads();
        if (lazy) {
            fwd_step_f<LOGS, ROLE, 3, CW, true>(lds, 3, tw, q, qinv, N1, first);
            __syncthreads();
            fwd_step_f<LOGS, ROLE, 2, CW, true>(lds, 6, tw, q, qinv, N1, first);
        } else {
            fwd_step_f<LOGS, ROLE, 3, CW, false>(lds, 3, tw, q, qinv, N1, first);
            __syncthreads();
            fwd_step_f<LOGS, ROLE, 2, CW, false>(lds, 6, tw, q, qinv, N1, first);
        }
        __syncthreads();
    } else {
        for (int e = threadIdx.x; e < G::TILE; e += kThreads) {
            uint64_t gidx;
            if (ROLE == kCol) {
                const int k = e / G::C, cc = e % G::C;
                gidx = (uint64_t)k * N2 + first + cc;
            } else if (ROLE == kRow) {
                gidx = (uint64_t)first * G::S + e;
            } else {
                gidx = e;
            }
            lds[e] = reduce_f((double)src[gidx], q, qinv);
        }
        __syncthreads();
        if (lazy)
            stages_f<LOGS, ROLE, FWD, CW, true, kThreads, R16>(lds, tw, q, qinv, n, N1, first);
        else
            stages_f<LOGS, ROLE, FWD, CW, false, kThreads, R16>(lds, tw, q, qinv, n, N1, first);
    }
    const bool scale_out = !FWD && ROLE != kRow;  // the inverse's last pass
    const double ni = a.ninvd[prime];
    for (int e = threadIdx.x; e < G::TILE; e += kThreads) {
        uint64_t gidx;
        if (ROLE == kCol) {
            const int k = e / G::C, cc = e % G::C;
            gidx = (uint64_t)k * N2 + first + cc;
        } else if (ROLE == kRow) {
            gidx = (uint64_t)first * G::S + e;
        } else {
            gidx = e;
        }
        double v = lds[e];
        if (scale_out)
            v = reduce_f(mulmod_f(v, ni, q, qinv), q, qinv);
        else if (lazy)
            v = reduce_f(v, q, qinv);
        v = v < 0 ? v + q : v;
        if (epi) {
            if (a.epi_int)
                epi_store(a.epi, poly, limb, gidx, (uint64_t)v, (uint64_t)q, a.limbs, n);
            else
                epi_store_f(a.epi, poly, limb, gidx, v, q, qinv, a.limbs, n);
        } else {
            data[gidx] = (uint64_t)v;
        }
    }
}

template <int LOGS, int ROLE, bool FWD, int CW, bool R16>
__device__ __forceinline__ void ntt_pass(const NttKArgs &a)
{
    using G = Geometry<LOGS, ROLE, CW>;
    __shared__ uint64_t lds[G::TILE];
    const uint32_t bidx = blockIdx.y;
    const uint32_t poly = bidx / a.limbs, limb = bidx - poly * a.limbs;
    if (a.skip_alpha) {  // uniform per workgroup, before any barrier
        const uint32_t I = poly % a.poly_period;
        if (I < a.skip_l && I / a.skip_alpha == limb) return;
    }
    const uint32_t prime =
        a.prime_by_poly ? a.prime_of[a.poly_period ? poly % a.poly_period : poly] : a.prime_of[limb];
    const uint32_t n = 1u << a.logn;
    uint64_t *data = a.data + poly * a.poly_stride + (uint64_t)limb * n;
    const uint64_t *src = a.nsrcs ? a.srcs[poly] + (uint64_t)limb * n
                                  : (a.src ? a.src + poly * a.src_poly_stride + (uint64_t)limb * n : data);
    const ulonglong2 *tw = a.tw + (uint64_t)prime * n;
    const uint64_t q = a.mod[prime].q;
    const uint32_t N2 = ROLE == kCol ? (n >> LOGS) : (ROLE == kRow ? (uint32_t)G::S : n);
    const uint32_t N1 = n / N2;
    // Column tiles share 128-byte lines with their neighbours (CW x 8 bytes per
    // row).  Workgroups go to the 8 XCDs round-robin, so consecutive tiles are
    // given to workgroups on the same XCD (one L2) when the grid allows.
    uint32_t tile = blockIdx.x;
    if (ROLE == kCol && (gridDim.x & 7) == 0) tile = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t first = tile * G::C;  // first column / row of the tile
    if (q < kFpPrimeBound) {
        ntt_tile_f<LOGS, ROLE, FWD, CW, R16>(a, lds, data, src, prime, n, N2, N1, first, poly, limb);
        return;
    }

    // ---- load
    for (int e = threadIdx.x; e < G::TILE; e += kThreads) {
        uint64_t gidx;
        if (ROLE == kCol) {
            const int k = e / G::C, cc = e % G::C;
            gidx = (uint64_t)k * N2 + first + cc;
        } else if (ROLE == kRow) {
            gidx = (uint64_t)first * G::S + e;
        } else {
            gidx = e;
        }
        lds[e] = src[gidx];
    }
    __syncthreads();

    if (FWD) {
        const uint32_t m0 = ROLE == kRow ? N1 : 1u;
        int s = 0;
        if constexpr (LOGS >= 3) {
#pragma unroll
            for (int st = 0; st < LOGS / 3; st++) {
                fwd_step<LOGS, ROLE, 3, CW>(lds, s, tw, q, m0, first);
                s += 3;
                __syncthreads();
            }
        }
        if constexpr (LOGS % 3 == 2) {
            fwd_step<LOGS, ROLE, 2, CW>(lds, s, tw, q, m0, first);
            __syncthreads();
        } else if constexpr (LOGS % 3 == 1) {
            fwd_step<LOGS, ROLE, 1, CW>(lds, s, tw, q, m0, first);
            __syncthreads();
        }
    } else {
        const uint32_t mx = ROLE == kCol ? N1 : n;
        int s = 0;
        if constexpr (LOGS >= 3) {
#pragma unroll
            for (int st = 0; st < LOGS / 3; st++) {
                inv_step<LOGS, ROLE, 3, CW>(lds, s, tw, q, n, mx, first);
                s += 3;
                __syncthreads();
            }
        }
        if constexpr (LOGS % 3 == 2) {
            inv_step<LOGS, ROLE, 2, CW>(lds, s, tw, q, n, mx, first);
            __syncthreads();
        } else if constexpr (LOGS % 3 == 1) {
            inv_step<LOGS, ROLE, 1, CW>(lds, s, tw, q, n, mx, first);
            __syncthreads();
        }
    }

    // ---- store (final reduction on the last pass)
    const bool last_pass = FWD ? (ROLE != kCol) : (ROLE != kRow);
    const bool epi = FWD && ROLE != kCol && epi_on(a.epi);
    const ulonglong2 ni = a.ninv[prime];
    for (int e = threadIdx.x; e < G::TILE; e += kThreads) {
        uint64_t gidx;
        if (ROLE == kCol) {
            const int k = e / G::C, cc = e % G::C;
            gidx = (uint64_t)k * N2 + first + cc;
        } else if (ROLE == kRow) {
            gidx = (uint64_t)first * G::S + e;
        } else {
            gidx = e;
        }
        uint64_t v = lds[e];
        if (last_pass) {
            if (FWD) {
                v = v >= 2 * q ? v - 2 * q : v;
                v = v >= q ? v - q : v;
            } else {
                v = shoup(v, ni.x, ni.y, q);
            }
        }
        if (epi)
            epi_store(a.epi, poly, limb, gidx, v, q, a.limbs, n);
        else
            data[gidx] = v;
    }
}

template <int LOGS, int ROLE, bool FWD, int CW, bool R16>
__global__ __launch_bounds__(kThreads) void k_ntt(NttKArgs a)
{
    ntt_pass<LOGS, ROLE, FWD, CW, R16>(a);
}

// ---- ModUp base conversion fused into the forward column pass
//
// Workgroup (tile, b, digit j, chunk): the digit's A input limbs are read once
// over the tile (8 columns x 2^LOGS rows) and their products y_k = [x_k
// Dhat_k^-1]_(q_k) kept in registers; then for each output prime I of the
// chunk (the primes outside the digit) the converted values sum_k y_k f_(k,I)
// are written to LDS and transformed by the column stages of prime I, and the
// column-transformed limb is stored into T[b][I][j].  The conversion of
// T's limb is never written and read back.  Every prime below 2^51: the
// products are the exact double-precision ones of the unfused conversion
// (k_modup_conv_a), and the column stages those of the double-precision pass,
// so the stored words equal the unfused conversion followed by the column pass.
// The column stages of one tile whose values v (this thread's positions
// e = threadIdx.x + NT t, reduced) are in registers, stored column-transformed
// into out (positions gidx).  For 2^8-row columns, 8-column tiles and 256
// threads each thread's 8 positions are exactly one radix-8 group of the first
// three stages, so those run in registers before the values go to LDS, and
// the last two stages run in registers on the way out: three LDS passes and
// three barriers instead of eight passes and four.  The words are the same.
template <int LOGS, int CW, int NT, int PER, bool LAZY>
__device__ __forceinline__ void col_stages_store(double *v, double *lds, const double *tw, double q, double qinv,
                                                 uint32_t n, uint32_t N1, uint32_t N2, uint32_t col0, uint64_t *out,
                                                 const uint32_t *gidx)
{
    if constexpr (LOGS == 8 && CW == 8 && NT == 256 && PER == 8) {
        fwd_group_f<3, LAZY>(v, 0, 0, 1, tw, q, qinv);  // rows gg + 32 t of column threadIdx.x % 8
#pragma unroll
        for (int t = 0; t < PER; t++) lds[threadIdx.x + t * NT] = v[t];
        __syncthreads();
        fwd_step_f<LOGS, kCol, 3, CW, LAZY, NT>(lds, 3, tw, q, qinv, 1, col0);
        __syncthreads();
        double w[2][4];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int g = threadIdx.x + k * NT, cc = g % CW, gg = g / CW;
#pragma unroll
            for (int t = 0; t < 4; t++) w[k][t] = lds[(4 * gg + t) * CW + cc];
        }
        __syncthreads();  // every read of this tile done before the next one is written
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int g = threadIdx.x + k * NT, cc = g % CW, gg = g / CW;
            fwd_group_f<2, LAZY>(w[k], 6, (uint32_t)gg, 1, tw, q, qinv);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                double x = w[k][t];
                if (LAZY) x = reduce_f(x, q, qinv);
                x = x < 0 ? x + q : x;
                out[(uint64_t)(4 * gg + t) * N2 + col0 + cc] = (uint64_t)x;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < PER; t++) lds[threadIdx.x + t * NT] = v[t];
        __syncthreads();
        stages_f<LOGS, kCol, true, CW, LAZY, NT>(lds, tw, q, qinv, n, N1, col0);
#pragma unroll
        for (int t = 0; t < PER; t++) {
            double x = lds[threadIdx.x + t * NT];
            if (LAZY) x = reduce_f(x, q, qinv);
            x = x < 0 ? x + q : x;
            out[gidx[t]] = (uint64_t)x;
        }
        // each thread rewrites only the LDS words it has just read: the next
        // tile's stages start after the barrier that follows those writes
    }
}

struct ModUpColArgs {
    uint64_t *T;
    const uint64_t *coef;
    const ulonglong2 *w;
    const double *convd;
    const double2 *qd;
    const double *twd;
    uint32_t l, digits, nout, key_limbs, logn, chunk, chunks;
    uint32_t outp[kMaxPrimes];  // words: uniform indices take scalar loads
};

template <int LOGS, int A, int CW, int NT>
__global__ __launch_bounds__(NT) void k_modup_col(ModUpColArgs a)
{
    using G = Geometry<LOGS, kCol, CW>;
    constexpr int PER = G::TILE / NT;
    __shared__ double lds[G::TILE];
    const uint32_t n = 1u << a.logn;
    const uint32_t N2 = n >> LOGS, N1 = 1u << LOGS;
    const uint32_t row = blockIdx.y;  // (b * digits + j) * chunks + chunk
    const uint32_t bj = row / a.chunks, chunk = row - bj * a.chunks;
    const uint32_t b = bj / a.digits, j = bj - b * a.digits;
    const uint32_t first = j * A;
    const uint32_t s = min((uint32_t)A, a.l - first);
    const uint32_t count = a.nout - s;  // output primes outside the digit
    const uint32_t r0 = chunk * a.chunk;
    if (r0 >= count) return;  // uniform, before any barrier
    const uint32_t r1 = min(count, r0 + a.chunk);
    uint32_t tile = blockIdx.x;
    if ((gridDim.x & 7) == 0) tile = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t col0 = tile * G::C;
    uint32_t gidx[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = threadIdx.x + t * NT;
        gidx[t] = (uint32_t)(e / G::C) * N2 + col0 + (uint32_t)(e % G::C);
    }
    // the digit's products [x Dhat^-1]_q in [0, q) (rows i >= s stay zero)
    const uint32_t base = (j * A + s - 1) * A;
    double y[A][PER];
#pragma unroll
    for (int i = 0; i < A; i++) {
        const bool on = (uint32_t)i < s;
        const uint32_t p = on ? first + i : first;
        const double q = a.qd[p].x, qinv = a.qd[p].y;
        const double wi = (double)a.w[base + (on ? i : 0)].x;
        const uint64_t *x = a.coef + ((uint64_t)b * a.l + p) * n;
#pragma unroll
        for (int t = 0; t < PER; t++) {
            double v = reduce_f(mulmod_f((double)x[gidx[t]], wi, q, qinv), q, qinv);
            v = v < 0 ? v + q : v;
            y[i][t] = on ? v : 0.0;
        }
    }
    for (uint32_t r = r0; r < r1; r++) {
        const uint32_t I = r < first ? r : r + s;
        const uint32_t kp = a.outp[I];
        const double q = a.qd[kp].x, qinv = a.qd[kp].y;
        const double *fd = a.convd + (uint64_t)base * a.key_limbs + kp;
        double f[A];
#pragma unroll
        for (int i = 0; i < A; i++) f[i] = fd[(uint64_t)i * a.key_limbs];
        double cv[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < A; i++) {
                acc += mulmod_f(y[i][t], f[i], q, qinv);
                if (i & 1) acc = reduce_f(acc, q, qinv);
            }
            cv[t] = reduce_f(acc, q, qinv);  // |v| <= q/2 (1 + eps)
        }
        const double *tw = a.twd + (uint64_t)kp * n;
        uint64_t *out = a.T + (((uint64_t)b * a.nout + I) * a.digits + j) * n;
        if (q * (0.5 + 1.25 * LOGS) <= 0x1p50)  // uniform
            col_stages_store<LOGS, CW, NT, PER, true>(cv, lds, tw, q, qinv, n, N1, N2, col0, out, gidx);
        else
            col_stages_store<LOGS, CW, NT, PER, false>(cv, lds, tw, q, qinv, n, N1, N2, col0, out, gidx);
    }
}

// ---- ModDown base conversion fused into the forward column pass
//
// Workgroup (tile, poly p, chunk): the A input limbs at limbs lo .. lo + A - 1
// of poly p of acc (coefficient form) are read once over the tile; per
// position v_k = [y_k w_k]_(p_k) (Shoup products, as k_moddown_conv_a) and the
// centring quotient r = floor(sum_k v_k / p_k + 1/2) stay in registers; each
// output prime q_i (i < lo) of the chunk gets sum_k v_k f_(k,i) - r (P mod q_i)
// in exact double-precision products, transformed by the column stages of q_i
// and stored into z[p][i].  The words equal the unfused conversion followed
// by the column pass.  Every prime below 2^51.
// ---- Rescale centring fused into the forward column pass
//
// Workgroup (tile, poly p, chunk): the INTT'd last limb of poly p (prime p_l,
// [0, p_l)) is read once over the tile; each output prime q_i of the chunk
// gets (x + (p_l - 1)/2) mod p_l - (p_l - 1)/2 (SEAL's centred remainder,
// an integer of magnitude below 2^51, exact in double precision) reduced mod
// q_i, transformed by the column stages of q_i, stored into z[p][i].  The
// words equal k_center_last followed by the column pass.  Every prime below 2^51.
struct CenterColArgs {
    uint64_t *z;
    const uint64_t *last;
    uint64_t last_stride;
    const double2 *qd;
    const double *twd;
    uint64_t p, half;
    uint32_t limbs, logn, chunk, chunks;
};

template <int LOGS>
__global__ __launch_bounds__(kThreads) void k_center_col(CenterColArgs a)
{
    using G = Geometry<LOGS, kCol, 8>;
    constexpr int PER = G::TILE / kThreads;
    __shared__ double lds[G::TILE];
    const uint32_t n = 1u << a.logn;
    const uint32_t N2 = n >> LOGS, N1 = 1u << LOGS;
    const uint32_t p = blockIdx.y / a.chunks, chunk = blockIdx.y - p * a.chunks;
    const uint32_t i0 = chunk * a.chunk;
    if (i0 >= a.limbs) return;  // uniform, before any barrier
    const uint32_t i1 = min(a.limbs, i0 + a.chunk);
    uint32_t tile = blockIdx.x;
    if ((gridDim.x & 7) == 0) tile = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t col0 = tile * G::C;
    uint32_t gidx[PER];
    double v[PER];
    const uint64_t *x = a.last + (uint64_t)p * a.last_stride;
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = threadIdx.x + t * kThreads;
        gidx[t] = (uint32_t)(e / G::C) * N2 + col0 + (uint32_t)(e % G::C);
        uint64_t u = x[gidx[t]] + a.half;
        u = u >= a.p ? u - a.p : u;
        v[t] = (double)u - (double)a.half;
    }
    for (uint32_t i = i0; i < i1; i++) {
        const double q = a.qd[i].x, qinv = a.qd[i].y;
        double cv[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) cv[t] = reduce_f(v[t], q, qinv);
        const double *tw = a.twd + (uint64_t)i * n;
        uint64_t *out = a.z + ((uint64_t)p * a.limbs + i) * n;
        if (q * (0.5 + 1.25 * LOGS) <= 0x1p50)  // uniform
            col_stages_store<LOGS, 8, kThreads, PER, true>(cv, lds, tw, q, qinv, n, N1, N2, col0, out, gidx);
        else
            col_stages_store<LOGS, 8, kThreads, PER, false>(cv, lds, tw, q, qinv, n, N1, N2, col0, out, gidx);
    }
}

struct ModDownColArgs {
    uint64_t *z;
    const uint64_t *acc;
    uint64_t acc_stride;
    const ulonglong2 *w;
    const double *convd;
    const double2 *qd;
    const ModConst *mod;
    const double *twd;
    uint32_t lo, L, logn, chunk, chunks;
    uint32_t inp[kMaxPrimes];
};

template <int LOGS, int A, int CW, int NT>
__global__ __launch_bounds__(NT) void k_moddown_col(ModDownColArgs a)
{
    using G = Geometry<LOGS, kCol, CW>;
    constexpr int PER = G::TILE / NT;
    __shared__ double lds[G::TILE];
    const uint32_t n = 1u << a.logn;
    const uint32_t N2 = n >> LOGS, N1 = 1u << LOGS;
    const uint32_t p = blockIdx.y / a.chunks, chunk = blockIdx.y - p * a.chunks;
    const uint32_t i0 = chunk * a.chunk;
    if (i0 >= a.lo) return;  // uniform, before any barrier
    const uint32_t i1 = min(a.lo, i0 + a.chunk);
    uint32_t tile = blockIdx.x;
    if ((gridDim.x & 7) == 0) tile = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t col0 = tile * G::C;
    uint32_t gidx[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = threadIdx.x + t * NT;
        gidx[t] = (uint32_t)(e / G::C) * N2 + col0 + (uint32_t)(e % G::C);
    }
    const uint64_t *y = a.acc + (uint64_t)p * a.acc_stride + (uint64_t)a.lo * n;
    double v[A][PER], rr[PER];
    double frac[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) frac[t] = 0.0;
#pragma unroll
    for (int k = 0; k < A; k++) {
        const ulonglong2 wk = a.w[k];
        const uint64_t pk = a.mod[a.inp[k]].q;
#pragma unroll
        for (int t = 0; t < PER; t++) {
            const uint64_t x = shoup(y[(uint64_t)k * n + gidx[t]], wk.x, wk.y, pk);
            frac[t] += (double)x / (double)pk;
            v[k][t] = (double)x;
        }
    }
#pragma unroll
    for (int t = 0; t < PER; t++) rr[t] = (double)(uint32_t)(uint64_t)floor(frac[t] + 0.5);  // <= A
    for (uint32_t i = i0; i < i1; i++) {
        const double q = a.qd[i].x, qinv = a.qd[i].y;
        double f[A + 1];
#pragma unroll
        for (int k = 0; k <= A; k++) f[k] = a.convd[(uint64_t)k * a.L + i];
        double cv[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) {
            // - r (P mod q_i) as r (q_i - P mod q_i)
            double s = mulmod_f(rr[t], f[A], q, qinv);
#pragma unroll
            for (int k = 0; k < A; k++) {
                s += mulmod_f(v[k][t], f[k], q, qinv);
                if (k & 1) s = reduce_f(s, q, qinv);
            }
            cv[t] = reduce_f(s, q, qinv);  // |v| <= q/2 (1 + eps)
        }
        const double *tw = a.twd + (uint64_t)i * n;
        uint64_t *out = a.z + ((uint64_t)p * a.lo + i) * n;
        if (q * (0.5 + 1.25 * LOGS) <= 0x1p50)  // uniform
            col_stages_store<LOGS, CW, NT, PER, true>(cv, lds, tw, q, qinv, n, N1, N2, col0, out, gidx);
        else
            col_stages_store<LOGS, CW, NT, PER, false>(cv, lds, tw, q, qinv, n, N1, N2, col0, out, gidx);
    }
}

// The forward column pass with 8-wide tiles held to 64 VGPRs: 8 waves per SIMD
// instead of 7 (the other passes reach 8 as they are, or would spill).
template <int LOGS, int ROLE, bool FWD, int CW, bool R16>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_ntt_w8(NttKArgs a)
{
    ntt_pass<LOGS, ROLE, FWD, CW, R16>(a);
}

template <int LOGS, int ROLE, bool FWD, int CW = 16, bool R16 = false>
void launch_one(const NttKArgs &a, uint32_t batch, hipStream_t s)
{
    using G = Geometry<LOGS, ROLE, CW>;
    const uint32_t n = 1u << a.logn;
    const uint32_t tiles = ROLE == kWhole ? 1u : (n / G::TILE);
    dim3 grid(tiles, batch);
    static const char *names[2][3] = {{"ntt_inv_whole", "ntt_inv_col", "ntt_inv_row"},
                                      {"ntt_fwd_whole", "ntt_fwd_col", "ntt_fwd_row"}};
    // algorithmic bytes: each limb read + written once (16 n) plus the pass's
    // twiddle entries: {w, w'} (16 B) on the 64-bit path, w (8 B) in double
    // precision
    const double tw = ROLE == kWhole ? (double)n : (ROLE == kCol ? (double)G::S : (double)(n - n / G::S));
    // with an epilogue the last pass also reads x (and base or out) per word
    const bool epi = FWD && ROLE != kCol && epi_on(a.epi);
    const double epi_reads =
        a.epi.out ? (a.epi.accumulate ? 1.0 : 0.0) : ((a.epi.a0 || a.epi.a0s[0]) ? 0.5 : 0.0) + (a.epi.d ? 1.0 : 0.0);
    const double epi_bytes = epi ? 8.0 * n * (1.0 + epi_reads) : 0.0;
    ProfScope prof(epi ? "ntt_fwd_sub_mul" : names[FWD ? 1 : 0][ROLE], s,
                   (double)(batch - a.skipped) * (16.0 * n + a.tw_bytes * tw + epi_bytes));
    if constexpr (ROLE == kCol && FWD && CW == 8 && !R16)
        hipLaunchKernelGGL((k_ntt_w8<LOGS, ROLE, FWD, CW, R16>), grid, dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL((k_ntt<LOGS, ROLE, FWD, CW, R16>), grid, dim3(kThreads), 0, s, a);
}

template <bool FWD>
void launch_whole(const NttKArgs &a, uint32_t batch, hipStream_t s)
{
    switch (a.logn) {
#define NX_CASE(L) \
    case L: launch_one<L, kWhole, FWD>(a, batch, s); break;
        NX_CASE(1) NX_CASE(2) NX_CASE(3) NX_CASE(4) NX_CASE(5) NX_CASE(6) NX_CASE(7)
        NX_CASE(8) NX_CASE(9) NX_CASE(10) NX_CASE(11) NX_CASE(12) NX_CASE(13)
#undef NX_CASE
    default: break;
    }
}

// Epilogue arithmetic of the double-precision path (0: exact double-precision
// products, 1: 64-bit integer Shoup products; the same words).  Set by the autotuner.
std::atomic<int> g_epi_int{0};

// Tile widths chosen by the autotuner: [log n - 14][pass][direction][batch bucket],
// 0 = the default (tiles of 8).  Written by ntt_set_tile, read at every launch.
std::atomic<int> g_tile[kTuneLogN][2][2][kTuneBuckets];
// Radix of the double-precision passes, same indexing: 0 = radix-8 steps, 1 =
// radix-16 steps (the integer path always takes radix 8).  Set by the autotuner.
std::atomic<int> g_radix[kTuneLogN][2][2][kTuneBuckets];

int tune_bucket(uint32_t batch)
{
    int b = 0;
    while (b + 1 < kTuneBuckets && (2u << b) <= batch) b++;
    return b;  // floor(log2(batch)), capped
}

template <int LOGS, int ROLE, bool FWD, bool R16>
void launch_width(int cw, const NttKArgs &a, uint32_t batch, hipStream_t s)
{
    if (cw == 4)
        launch_one<LOGS, ROLE, FWD, 4, R16>(a, batch, s);
    else if (cw == 8)
        launch_one<LOGS, ROLE, FWD, 8, R16>(a, batch, s);
    else
        launch_one<LOGS, ROLE, FWD, 16, R16>(a, batch, s);
}

template <int LOGS, int ROLE, bool FWD>
void launch_sized(const NttKArgs &a, uint32_t batch, hipStream_t s)
{
    int cw = 0, r16 = 0;
    const int li = (int)a.logn - kTuneLogNMin;
    if (li >= 0 && li < kTuneLogN) {
        cw = g_tile[li][ROLE == kCol ? 0 : 1][FWD ? 1 : 0][tune_bucket(batch)].load();
        r16 = g_radix[li][ROLE == kCol ? 0 : 1][FWD ? 1 : 0][tune_bucket(batch)].load();
    }
    if (!cw) cw = 8;
    if (r16)
        launch_width<LOGS, ROLE, FWD, true>(cw, a, batch, s);
    else
        launch_width<LOGS, ROLE, FWD, false>(cw, a, batch, s);
}

template <int ROLE, bool FWD>
void launch_pass(int logs, const NttKArgs &a, uint32_t batch, hipStream_t s)
{
    switch (logs) {
    case 7: launch_sized<7, ROLE, FWD>(a, batch, s); break;
    case 8: launch_sized<8, ROLE, FWD>(a, batch, s); break;
    case 9: launch_sized<9, ROLE, FWD>(a, batch, s); break;
    default: break;
    }
}

NttKArgs make_args(const DeviceTables &t, const NttBatch &b, bool fwd)
{
    NttKArgs a;
    a.data = b.data;
    a.poly_stride = b.poly_stride;
    a.src = b.src;
    a.src_poly_stride = b.src_poly_stride;
    a.nsrcs = 0;
    for (int i = 0; i < kNttSrcMax; i++) a.srcs[i] = nullptr;
    {
        // limbs whose prime takes the double-precision path read 8-byte twiddles
        const uint32_t m = b.prime_by_poly ? (b.poly_period ? b.poly_period : b.polys) : b.limbs;
        const uint32_t cnt = m < (uint32_t)kMaxPrimes ? m : (uint32_t)kMaxPrimes;
        uint32_t fp = 0;
        for (uint32_t i = 0; i < cnt; i++) fp += t.mod_host[b.prime_of[i]] < kFpPrimeBound ? 1u : 0u;
        a.tw_bytes = cnt ? (8.0 * fp + 16.0 * (cnt - fp)) / cnt : 16.0;
    }
    if (b.src_list) {
        // a source per poly: the batch must fit the argument array
        if (b.polys > (uint32_t)kNttSrcMax) throw std::invalid_argument("ntt: too many sources in one batch");
        a.nsrcs = b.polys;
        for (uint32_t i = 0; i < b.polys; i++) a.srcs[i] = b.src_list[i];
    }
    a.epi = b.epi;
    a.epi_int = (uint32_t)g_epi_int.load();
    if (!fwd) {
        a.epi.out = nullptr;
        a.epi.outs[0] = nullptr;
    }
    a.skip_l = a.skip_alpha = a.skipped = 0;
    if (b.skip_alpha && b.prime_by_poly && b.poly_period) {
        a.skip_l = b.skip_l;
        a.skip_alpha = b.skip_alpha;
        for (uint32_t I = 0; I < b.skip_l && I < b.poly_period; I++)
            if (I / b.skip_alpha < b.limbs) a.skipped += b.polys / b.poly_period;
    }
    a.limbs = b.limbs;
    a.logn = (uint32_t)t.logn;
    a.prime_by_poly = b.prime_by_poly;
    a.poly_period = b.poly_period;
    a.tw = fwd ? t.fwd : t.inv;
    a.mod = t.mod;
    a.ninv = t.ninv;
    a.twd = fwd ? t.fwd_d : t.inv_d;
    a.ninvd = t.ninv_d;
    for (int i = 0; i < kMaxPrimes; i++) a.prime_of[i] = b.prime_of[i];
    return a;
}

}  // namespace

// Split of log2(n) into column (N1) and row (N2) passes for n >= 2^14.
static inline void split_logn(int logn, int &l1, int &l2)
{
    l1 = logn / 2;
    l2 = logn - l1;
}

void ntt_set_epilogue_int(bool on) { g_epi_int.store(on ? 1 : 0); }
bool ntt_get_epilogue_int() { return g_epi_int.load() != 0; }

void ntt_set_tile(int logn, int pass, bool fwd, uint32_t batch, int width)
{
    const int li = logn - kTuneLogNMin;
    if (li < 0 || li >= kTuneLogN || pass < 0 || pass > 1) return;
    if (width != 0 && width != 4 && width != 8 && width != 16) return;
    g_tile[li][pass][fwd ? 1 : 0][tune_bucket(batch)].store(width);
}

void ntt_set_radix16(int logn, int pass, bool fwd, uint32_t batch, bool on)
{
    const int li = logn - kTuneLogNMin;
    if (li < 0 || li >= kTuneLogN || pass < 0 || pass > 1) return;
    g_radix[li][pass][fwd ? 1 : 0][tune_bucket(batch)].store(on ? 1 : 0);
}

bool ntt_get_radix16(int logn, int pass, bool fwd, uint32_t batch)
{
    const int li = logn - kTuneLogNMin;
    if (li < 0 || li >= kTuneLogN || pass < 0 || pass > 1) return false;
    return g_radix[li][pass][fwd ? 1 : 0][tune_bucket(batch)].load() != 0;
}

int ntt_get_tile(int logn, int pass, bool fwd, uint32_t batch)
{
    const int li = logn - kTuneLogNMin;
    if (li < 0 || li >= kTuneLogN || pass < 0 || pass > 1) return 0;
    return g_tile[li][pass][fwd ? 1 : 0][tune_bucket(batch)].load();
}

void ntt_forward(const DeviceTables &t, const NttBatch &b, hipStream_t s)
{
    const uint32_t batch = b.polys * b.limbs;
    if (!batch) return;
    NttKArgs a = make_args(t, b, true);
    if (t.logn <= 13) {
        launch_whole<true>(a, batch, s);
    } else {
        int l1, l2;
        split_logn(t.logn, l1, l2);
        if (!b.rows_only) launch_pass<kCol, true>(l1, a, batch, s);
        a.src = nullptr;  // the second pass runs in place on the output
        a.nsrcs = 0;
        launch_pass<kRow, true>(l2, a, batch, s);
    }
}

// Algorithmic double-precision operations (an FMA counts two, rint one) of the
// fused column kernels: an exact product (mul, fma, mul, rint, fma, add) is 8,
// a reduction (mul, rint, fma) 4; a conversion to one output word over A
// inputs is A products, A additions, a reduction every second term and a
// final one; a column pass is half a butterfly (a product, an addition and a
// subtraction) per word and stage plus the store's reduction.
constexpr double kFlopsProduct = 8.0, kFlopsReduce = 4.0;
inline double conv_flops(uint32_t A) { return A * (kFlopsProduct + 1.0) + (A / 2) * kFlopsReduce + kFlopsReduce; }
inline double col_flops(int logs) { return logs * (kFlopsProduct + 2.0) / 2.0 + kFlopsReduce; }

// Fused ModUp column pass per power-of-two bucket of the batch size B:
// 0 = on (outputs per workgroup by the default rule), -1 = off (conversion and
// column pass as separate launches), k > 0 = on with k outputs per workgroup.
// Process-wide; set by the autotuner.
static std::atomic<int> g_modup_col[kTuneBuckets];
constexpr int kColNarrow = 256;  // mode bit: 4-column tiles
constexpr int kColMaxA = 12;   // largest digit / special-prime count of the fused conversions
constexpr int kColWide = 512;    // mode bit: 512 threads per workgroup (8-column tiles, half the registers each)

void ks_modup_col_set(uint32_t B, int mode) { g_modup_col[tune_bucket(B)].store(mode); }

int ks_modup_col_get(uint32_t B) { return g_modup_col[tune_bucket(B)].load(); }

bool ks_modup_col_ok(const DeviceTables &t, uint32_t alpha)
{
    return t.max_bits <= 51 && alpha >= 2 && alpha <= kColMaxA && t.logn >= 14 && t.logn <= 17 &&
           t.modup_conv_d != nullptr;
}

void ks_modup_convert_col(const DeviceTables &t, uint64_t *T, const uint64_t *coef, const ulonglong2 *w,
                          uint32_t B, uint32_t l, uint32_t alpha, uint32_t key_limbs, const uint8_t *out_primes_host,
                          uint32_t nout, hipStream_t s)
{
    if (!ks_modup_col_ok(t, alpha)) throw std::invalid_argument("ks_modup_convert_col: unsupported parameters");
    if (nout > (uint32_t)kMaxPrimes || l < 1 || nout <= alpha) throw std::invalid_argument("ks_modup_convert_col: bad sizes");
    ModUpColArgs a;
    a.T = T;
    a.coef = coef;
    a.w = w;
    a.convd = t.modup_conv_d;
    a.qd = t.qd;
    a.twd = t.fwd_d;
    a.l = l;
    a.digits = (l + alpha - 1) / alpha;
    a.nout = nout;
    a.key_limbs = key_limbs;
    a.logn = (uint32_t)t.logn;
    for (int i = 0; i < kMaxPrimes; i++) a.outp[i] = i < (int)nout ? out_primes_host[i] : 0;
    int l1, l2;
    split_logn(t.logn, l1, l2);
    const uint32_t n = 1u << t.logn;
    // mode: outputs per workgroup in its low byte (0: default rule), 4-column
    // tiles (half the registers per thread) when bit 8 is set
    const int mode = g_modup_col[tune_bucket(B)].load();
    const uint32_t cw = (mode > 0 && (mode & kColNarrow)) ? 4u : 8u;
    const uint32_t nt = (mode > 0 && (mode & kColWide)) ? 512u : 256u;
    const uint32_t tiles = n / ((1u << l1) * cw);
    // outputs per workgroup: enough workgroups to fill the device, at least 4
    // outputs per read of the digit's inputs
    const uint32_t count = nout - std::min(alpha, l);  // most outputs of a digit (a partial digit has more)
    const uint32_t want = (2048u + tiles * B * a.digits - 1) / (tiles * B * a.digits);
    uint32_t chunk = (count + want - 1) / want;
    chunk = std::max(chunk, std::min(4u, count));
    if (mode > 0 && (mode & 255)) chunk = std::min((uint32_t)(mode & 255), count);
    a.chunk = chunk;
    a.chunks = (nout - 1 + chunk - 1) / chunk;  // covers a partial last digit (nout - s <= nout - 1 outputs)
    uint32_t outs = 0;
    for (uint32_t j = 0; j < a.digits; j++) outs += nout - std::min(alpha, l - j * alpha);
    ProfScope prof("ks_modup_col", s,
                   8.0 * n * ((double)B * l + (double)B * outs) + 8.0 * (double)(1u << l1) * B * outs,
                   (double)n * B * ((double)l * kFlopsProduct + outs * (conv_flops(alpha) + col_flops(l1))));
    const dim3 grid(tiles, B * a.digits * a.chunks);
#define NX_MC(LOGS, A)                                                                          \
    if (l1 == LOGS && alpha == A) {                                                           \
        if (cw == 4)                                                                          \
            hipLaunchKernelGGL((k_modup_col<LOGS, A, 4, 256>), grid, dim3(256), 0, s, a);      \
        else if (nt == 512 || A > 8)                                                          \
            hipLaunchKernelGGL((k_modup_col<LOGS, A, 8, 512>), grid, dim3(512), 0, s, a);      \
        else                                                                                  \
            hipLaunchKernelGGL((k_modup_col<LOGS, A, 8, (A > 8 ? 512 : 256)>), grid, dim3(A > 8 ? 512 : 256), 0, s, a); \
        return;                                                                               \
    }
    NX_MC(7, 2) NX_MC(7, 3) NX_MC(7, 4) NX_MC(7, 5) NX_MC(7, 6) NX_MC(7, 7) NX_MC(7, 8) NX_MC(7, 9) NX_MC(7, 10)
    NX_MC(7, 11) NX_MC(7, 12)
    NX_MC(8, 2) NX_MC(8, 3) NX_MC(8, 4) NX_MC(8, 5) NX_MC(8, 6) NX_MC(8, 7) NX_MC(8, 8) NX_MC(8, 9) NX_MC(8, 10)
    NX_MC(8, 11) NX_MC(8, 12)
#undef NX_MC
    throw std::invalid_argument("ks_modup_convert_col: unsupported split");
}

bool center_col_ok(const DeviceTables &t) { return t.max_bits <= 51 && t.logn >= 14 && t.logn <= 17; }

void rns_center_last_col(const DeviceTables &t, uint64_t *z, const uint64_t *last, uint64_t last_stride,
                         uint32_t last_prime, uint32_t polys, uint32_t limbs, hipStream_t s)
{
    if (!center_col_ok(t) || limbs < 1) throw std::invalid_argument("rns_center_last_col: unsupported parameters");
    CenterColArgs a;
    a.z = z;
    a.last = last;
    a.last_stride = last_stride;
    a.qd = t.qd;
    a.twd = t.fwd_d;
    a.p = t.mod_host[last_prime];
    a.half = a.p >> 1;
    a.limbs = limbs;
    a.logn = (uint32_t)t.logn;
    int l1, l2;
    split_logn(t.logn, l1, l2);
    const uint32_t n = 1u << t.logn;
    const uint32_t tiles = n / ((1u << l1) * 8u);
    const uint32_t want = (2048u + tiles * polys - 1) / (tiles * polys);
    uint32_t chunk = (limbs + want - 1) / want;
    chunk = std::max(chunk, std::min(4u, limbs));
    a.chunk = chunk;
    a.chunks = (limbs + chunk - 1) / chunk;
    ProfScope prof("rescale_col", s, 8.0 * n * polys * (1.0 + limbs) + 8.0 * (double)(1u << l1) * polys * limbs,
                   (double)n * polys * limbs * (kFlopsReduce + col_flops(l1)));
    const dim3 grid(tiles, polys * a.chunks);
    if (l1 == 7)
        hipLaunchKernelGGL((k_center_col<7>), grid, dim3(kThreads), 0, s, a);
    else if (l1 == 8)
        hipLaunchKernelGGL((k_center_col<8>), grid, dim3(kThreads), 0, s, a);
    else
        throw std::invalid_argument("rns_center_last_col: unsupported split");
}

static std::atomic<int> g_moddown_col[kTuneBuckets];

void ks_moddown_col_set(uint32_t polys, int mode) { g_moddown_col[tune_bucket(polys)].store(mode); }

int ks_moddown_col_get(uint32_t polys) { return g_moddown_col[tune_bucket(polys)].load(); }

bool ks_moddown_col_ok(const DeviceTables &t, uint32_t A)
{
    return t.max_bits <= 51 && A >= 2 && A <= kColMaxA + 1 && t.logn >= 14 && t.logn <= 17;
}

void ks_moddown_convert_col(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                            const ulonglong2 *w, const double *convd, const uint8_t *in_primes_host, uint32_t A,
                            uint32_t polys, uint32_t lo, uint32_t L, hipStream_t s)
{
    if (!ks_moddown_col_ok(t, A) || !convd) throw std::invalid_argument("ks_moddown_convert_col: unsupported parameters");
    if (lo < 1 || lo + A > (uint32_t)kMaxPrimes) throw std::invalid_argument("ks_moddown_convert_col: bad sizes");
    ModDownColArgs a;
    a.z = z;
    a.acc = acc;
    a.acc_stride = acc_stride;
    a.w = w;
    a.convd = convd;
    a.qd = t.qd;
    a.mod = t.mod;
    a.twd = t.fwd_d;
    a.lo = lo;
    a.L = L;
    a.logn = (uint32_t)t.logn;
    for (int i = 0; i < kMaxPrimes; i++) a.inp[i] = i < (int)A ? in_primes_host[i] : 0;
    int l1, l2;
    split_logn(t.logn, l1, l2);
    const uint32_t n = 1u << t.logn;
    const int mode = g_moddown_col[tune_bucket(polys)].load();
    const uint32_t cw = (mode > 0 && (mode & kColNarrow)) ? 4u : 8u;
    const uint32_t nt = (mode > 0 && (mode & kColWide)) ? 512u : 256u;
    const uint32_t tiles = n / ((1u << l1) * cw);
    const uint32_t want = (2048u + tiles * polys - 1) / (tiles * polys);
    uint32_t chunk = (lo + want - 1) / want;
    chunk = std::max(chunk, std::min(4u, lo));
    if (mode > 0 && (mode & 255)) chunk = std::min((uint32_t)(mode & 255), lo);
    a.chunk = chunk;
    a.chunks = (lo + chunk - 1) / chunk;
    ProfScope prof("ks_moddown_col", s,
                   8.0 * n * polys * ((double)A + lo) + 8.0 * (double)(1u << l1) * polys * lo,
                   (double)n * polys * ((double)A * 2.0 + lo * (conv_flops(A + 1) + col_flops(l1))));
    const dim3 grid(tiles, polys * a.chunks);
#define NX_MD(LOGS, AA)                                                                         \
    if (l1 == LOGS && A == AA) {                                                              \
        if (cw == 4)                                                                          \
            hipLaunchKernelGGL((k_moddown_col<LOGS, AA, 4, 256>), grid, dim3(256), 0, s, a);   \
        else if (nt == 512 || AA > 9)                                                         \
            hipLaunchKernelGGL((k_moddown_col<LOGS, AA, 8, 512>), grid, dim3(512), 0, s, a);   \
        else                                                                                  \
            hipLaunchKernelGGL((k_moddown_col<LOGS, AA, 8, (AA > 9 ? 512 : 256)>), grid, dim3(AA > 9 ? 512 : 256), 0, s, a); \
        return;                                                                               \
    }
    NX_MD(7, 2) NX_MD(7, 3) NX_MD(7, 4) NX_MD(7, 5) NX_MD(7, 6) NX_MD(7, 7) NX_MD(7, 8) NX_MD(7, 9) NX_MD(7, 10)
    NX_MD(7, 11) NX_MD(7, 12) NX_MD(7, 13)
    NX_MD(8, 2) NX_MD(8, 3) NX_MD(8, 4) NX_MD(8, 5) NX_MD(8, 6) NX_MD(8, 7) NX_MD(8, 8) NX_MD(8, 9) NX_MD(8, 10)
    NX_MD(8, 11) NX_MD(8, 12) NX_MD(8, 13)
#undef NX_MD
    throw std::invalid_argument("ks_moddown_convert_col: unsupported split");
}

void ntt_inverse(const DeviceTables &t, const NttBatch &b, hipStream_t s)
{
    const uint32_t batch = b.polys * b.limbs;
    if (!batch) return;
    NttKArgs a = make_args(t, b, false);
    if (t.logn <= 13) {
        launch_whole<false>(a, batch, s);
    } else {
        int l1, l2;
        split_logn(t.logn, l1, l2);
        launch_pass<kRow, false>(l2, a, batch, s);
        a.src = nullptr;
        a.nsrcs = 0;
        launch_pass<kCol, false>(l1, a, batch, s);
    }
}

}  // namespace nx
