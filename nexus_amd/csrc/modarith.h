// Exact 64-bit modular arithmetic for RNS limbs, shared by host set-up code and
// gfx950 kernels.  Moduli are SEAL-style NTT primes below 2^61 (q = 1 mod 2N).
//
//  * Shoup multiplication by a precomputed constant w (w' = floor(w * 2^64 / q)):
//    one 64x64 high product + two low products; the lazy form returns [0, 2q).
//  * Barrett reduction of a 128-bit value with the 128-bit ratio floor(2^128/q)
//    (the same constant SEAL keeps in Modulus::const_ratio, modulus.h).
//  * Product reduction a*b mod q for a, b < 4q uses the 128-bit Barrett path.
//
// On the device the 64x64->128 products lower to v_mad_u64_u32 chains on the
// integer VALU; nothing here touches MFMA.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define NX_HD __host__ __device__ __forceinline__
#else
#define NX_HD inline
#endif

namespace nx {

typedef unsigned __int128 u128;

// Per-modulus constants, identical layout on host and device.
struct ModConst {
    uint64_t q;       // the prime
    uint64_t ratio0;  // floor(2^128 / q), low word
    uint64_t ratio1;  // floor(2^128 / q), high word
    uint64_t two_q;
};

NX_HD uint64_t mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((u128)a * b) >> 64);
#endif
}

// Counter-based generator of the key sampler: the residue of sample_uniform's
// word `ctr` under `seed` is ctr_uniform(seed, ctr, q).  Key-switching keys
// generated here regenerate their uniform half from it instead of reading it.
NX_HD uint64_t ctr_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

NX_HD uint64_t ctr_rand64(uint64_t seed, uint64_t ctr) { return ctr_mix64(seed ^ ctr_mix64(ctr + 0x9e3779b97f4a7c15ull)); }

NX_HD uint64_t ctr_uniform(uint64_t seed, uint64_t ctr, uint64_t q) { return mulhi64(ctr_rand64(seed, ctr), q); }

// Shoup lazy: a * w mod q in [0, 2q) for any a < 2^64.
NX_HD uint64_t shoup_lazy(uint64_t a, uint64_t w, uint64_t wp, uint64_t q)
{
    uint64_t t = mulhi64(a, wp);
    return a * w - t * q;
}

NX_HD uint64_t shoup(uint64_t a, uint64_t w, uint64_t wp, uint64_t q)
{
    uint64_t r = shoup_lazy(a, w, wp, q);
    return r >= q ? r - q : r;
}

NX_HD uint64_t csub(uint64_t a, uint64_t q) { return a >= q ? a - q : a; }

NX_HD uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q)
{
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}

NX_HD uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q)
{
    return a >= b ? a - b : a + q - b;
}

NX_HD uint64_t neg_mod(uint64_t a, uint64_t q) { return a ? q - a : 0; }

// 128-bit value (hi, lo) -> [0, q).  The quotient estimate is
// floor(x * ratio / 2^128) computed exactly (all cross-term carries kept), so it
// is at most one below floor(x / q) and one conditional subtraction finishes.
NX_HD uint64_t barrett128(uint64_t hi, uint64_t lo, const ModConst &m)
{
    uint64_t c0 = mulhi64(lo, m.ratio0);
    // lo * ratio1
    uint64_t a_lo = lo * m.ratio1, a_hi = mulhi64(lo, m.ratio1);
    // hi * ratio0
    uint64_t b_lo = hi * m.ratio0, b_hi = mulhi64(hi, m.ratio0);
    uint64_t s1 = a_lo + c0;
    uint64_t carry1 = s1 < c0;
    uint64_t s2 = s1 + b_lo;
    uint64_t carry2 = s2 < s1;
    uint64_t qhat = hi * m.ratio1 + a_hi + b_hi + carry1 + carry2;
    uint64_t r = lo - qhat * m.q;
    return r >= m.q ? r - m.q : r;
}

NX_HD uint64_t barrett64(uint64_t a, const ModConst &m) { return barrett128(0, a, m); }

NX_HD uint64_t mul_mod(uint64_t a, uint64_t b, const ModConst &m)
{
    return barrett128(mulhi64(a, b), a * b, m);
}

// ---------------------------------------------------------------- host helpers
inline uint64_t host_mul_mod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

inline uint64_t host_pow_mod(uint64_t a, uint64_t e, uint64_t q)
{
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = host_mul_mod(r, a, q);
        a = host_mul_mod(a, a, q);
        e >>= 1;
    }
    return r;
}

inline uint64_t host_inv_mod(uint64_t a, uint64_t q)
{
    // q prime: Fermat
    return host_pow_mod(a % q, q - 2, q);
}

inline uint64_t shoup_quotient(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

inline ModConst make_modconst(uint64_t q)
{
    ModConst m;
    m.q = q;
    // floor(2^128 / q) = floor((2^128 - 1) / q) for q not a power of two
    u128 all = ~(u128)0;
    u128 r = all / q;
    m.ratio0 = (uint64_t)r;
    m.ratio1 = (uint64_t)(r >> 64);
    m.two_q = 2 * q;
    return m;
}

}  // namespace nx
