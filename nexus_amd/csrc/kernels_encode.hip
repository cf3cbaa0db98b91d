// CKKS encoding on the GPU with SEAL's rounding (ckks.h:457-620 encode_internal,
// util/dwthandler.h transform_from_rev).  Every butterfly performs the same
// IEEE double operations in the same order as SEAL's host transform -- u + t,
// and (u - t) * w as (dr wr - di wi, dr wi + di wr); in the last stage
// (u + t) * s and (u - t) * (w s) -- and this file is compiled with
// -ffp-contract=off, so the rounded coefficients are bit-identical to SEAL's.
// The residues are exact (the integer value of each rounded coefficient mod q),
// so they equal SEAL's 64-bit, 128-bit and multi-word decompositions.
#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsLog = 11;  // stages 0..10 run on 2048-point chunks held in LDS (32 KiB)

__device__ __forceinline__ void gs_butterfly(double2 &x, double2 &y, double2 w, bool last, double scalar)
{
    const double2 u = x, t = y;
    const double dr = u.x - t.x, di = u.y - t.y;
    if (!last) {
        x = make_double2(u.x + t.x, u.y + t.y);
        const double ac = dr * w.x, bd = di * w.y, ad = dr * w.y, bc = di * w.x;
        y = make_double2(ac - bd, ad + bc);
    } else {
        const double sr = w.x * scalar, si = w.y * scalar;
        const double ar = u.x + t.x, ai = u.y + t.y;
        x = make_double2(ar * scalar, ai * scalar);
        const double ac = dr * sr, bd = di * si, ad = dr * si, bc = di * sr;
        y = make_double2(ac - bd, ad + bc);
    }
}

__global__ void k_enc_zero(double2 *v, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) v[i] = make_double2(0.0, 0.0);
}

// v[index_map[i]] = z_i, v[index_map[i + n/2]] = conj(z_i)
__global__ void k_enc_scatter(double2 *v, const double *re, const double *im, const uint32_t *index_map,
                              uint32_t count, uint32_t slots)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const double x = re[i], y = im ? im[i] : 0.0;
    v[index_map[i]] = make_double2(x, y);
    v[index_map[i + slots]] = make_double2(x, -y);
}

// Stages 0..stages-1 (gap 1 .. 2^(stages-1)) on one chunk of 2^stages points:
// every butterfly group of these stages lies inside one chunk.  Group i of
// stage s (gap 2^s) uses roots[n - n/2^s + i + 1].
__global__ __launch_bounds__(kBlock) void k_fft_low(double2 *v, const double2 *roots, uint32_t n, uint32_t logn,
                                                     uint32_t stages, double scalar)
{
    __shared__ double2 sh[1 << kLdsLog];
    const uint32_t C = 1u << stages;
    const uint32_t base = blockIdx.x * C;
    for (uint32_t k = threadIdx.x; k < C; k += kBlock) sh[k] = v[base + k];
    __syncthreads();
    for (uint32_t s = 0; s < stages; s++) {
        const bool last = s + 1 == logn;
        const uint32_t gap = 1u << s;
        for (uint32_t b = threadIdx.x; b < C / 2; b += kBlock) {
            const uint32_t il = b >> s, j = b & (gap - 1);
            const uint32_t off = il * 2 * gap;
            const uint32_t ig = (base >> (s + 1)) + il;
            gs_butterfly(sh[off + j], sh[off + gap + j], roots[n - (n >> s) + ig + 1], last, scalar);
        }
        __syncthreads();
    }
    for (uint32_t k = threadIdx.x; k < C; k += kBlock) v[base + k] = sh[k];
}

// one stage with gap 2^s over the whole vector
__global__ void k_fft_stage(double2 *v, const double2 *roots, uint32_t n, uint32_t s, int last, double scalar)
{
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= n / 2) return;
    const uint32_t gap = 1u << s;
    const uint32_t i = b >> s, j = b & (gap - 1);
    const uint32_t off = i * 2 * gap;
    double2 x = v[off + j], y = v[off + gap + j];
    gs_butterfly(x, y, roots[n - (n >> s) + i + 1], last != 0, scalar);
    v[off + j] = x;
    v[off + gap + j] = y;
}

// round(real part) mod every prime, written as [limb][n]; block b writes the
// largest |real part| of its coefficients to blockmax[b] (as the IEEE bit
// pattern, which orders non-negative doubles).
__global__ __launch_bounds__(kBlock) void k_enc_residues(uint64_t *out, const double2 *v, const ModConst *mod,
                                                         uint32_t limbs, uint32_t n, uint64_t *blockmax)
{
    __shared__ uint64_t red[kBlock];
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    uint64_t mine = 0;
    if (c < n) {
        const double x = v[c].x;
        mine = (uint64_t)__double_as_longlong(fabs(x));
        const double r = round(x);
        const bool neg = signbit(r);
        const double a = fabs(r);
        // |r| = mant * 2^sh with mant < 2^64 exactly
        uint64_t mant;
        int sh = 0;
        if (a < 18446744073709551616.0) {
            mant = (uint64_t)a;
        } else {
            int e;
            const double f = frexp(a, &e);  // a = f 2^e, f in [0.5, 1)
            mant = (uint64_t)ldexp(f, 53);
            sh = e - 53;
        }
        for (uint32_t j = 0; j < limbs; j++) {
            const ModConst m = mod[j];
            uint64_t res = barrett64(mant, m);
            for (int left = sh; left > 0;) {
                const int k = left < 63 ? left : 63;
                // res < q, so res * 2^k < q 2^64
                res = barrett128(res >> (64 - k), res << k, m);
                left -= k;
            }
            out[(uint64_t)j * n + c] = neg ? neg_mod(res, m.q) : res;
        }
    }
    red[threadIdx.x] = mine;
    __syncthreads();
    for (uint32_t w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) blockmax[blockIdx.x] = red[0];
}

}  // namespace

uint32_t encode_blocks(int logn) { return ((1u << logn) + kBlock - 1) / kBlock; }

void encode_fft(const DeviceTables &t, uint64_t *out, double2 *work, const double *re, const double *im,
                uint32_t count, const uint32_t *index_map, const double2 *roots, double scalar, uint32_t limbs,
                uint64_t *blockmax, hipStream_t s)
{
    const uint32_t logn = (uint32_t)t.logn, n = 1u << logn, slots = n / 2;
    ProfScope prof("encode", s, 16.0 * n * 2 + 8.0 * (count * 2.0) + 8.0 * n * limbs);
    if (count < slots) hipLaunchKernelGGL(k_enc_zero, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, work, n);
    if (count)
        hipLaunchKernelGGL(k_enc_scatter, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, work, re, im,
                           index_map, count, slots);
    const uint32_t low = logn < (uint32_t)kLdsLog ? logn : (uint32_t)kLdsLog;
    hipLaunchKernelGGL(k_fft_low, dim3(n >> low), dim3(kBlock), 0, s, work, roots, n, logn, low, scalar);
    for (uint32_t st = low; st < logn; st++)
        hipLaunchKernelGGL(k_fft_stage, dim3((n / 2 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, work, roots, n, st,
                           st + 1 == logn ? 1 : 0, scalar);
    hipLaunchKernelGGL(k_enc_residues, dim3(encode_blocks(t.logn)), dim3(kBlock), 0, s, out, work, t.mod, limbs, n,
                       blockmax);
}

}  // namespace nx
