// Launchers for the gfx950 kernels.  Every launcher is stream-ordered on the
// stream it is given and performs no host synchronisation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>

#include "modarith.h"

namespace nx {

constexpr int kMaxPrimes = 64;

// Device-resident per-prime tables of one context.
struct DeviceTables {
    int logn;
    int num_primes;
    int max_bits = 64;        // bit size of the largest prime of the context
    const ulonglong2 *fwd;   // [prime][n] {psi power, Shoup quotient}, SEAL bit-reversed order
    const ulonglong2 *inv;   // [prime][n] {inverse psi power, Shoup quotient}, SEAL scrambled order
    const ulonglong2 *ninv;  // [prime] {n^-1 mod q, Shoup quotient}
    const double *fwd_d;     // [prime][n] the fwd powers as doubles (exact below 2^53)
    const double *inv_d;     // [prime][n] the inv powers as doubles
    const double *ninv_d;    // [prime] n^-1 mod q as doubles
    const double *modup_conv_d = nullptr;  // hybrid ModUp factors as doubles (Context::d_modup_conv)
    const ModConst *mod;     // [prime]
    const double2 *qd = nullptr;  // [prime] {q, 1 / q} in double precision (1 / q correctly rounded)
    const uint64_t *mod_host = nullptr;  // [prime] the primes, host copy (launch-time choices)
};

// A batch of in-place NTTs over polys x limbs RNS limbs of n words each.
// Limb `i` of poly `p` lives at data + p * poly_stride + i * n.  Its prime is
// prime_of[i] (prime_by_poly = 0) or prime_of[p'] (prime_by_poly = 1) with
// p' = p, or p mod poly_period when poly_period > 0 (batches of ciphertexts
// that each carry the same per-poly prime pattern).
// With src set, the transform reads its input from src (limb i of poly p at
// src + p * src_poly_stride + i * n) and writes data: out of place, src unchanged.
// Optional last step of a forward NTT (NttBatch::epi): instead of the
// transform t of poly p, limb i (prime i), the last pass stores
//   v = (x[p][i] - t) inv[i]
// (strides per poly, limb i at i n) in one of two forms:
// * out non-null: out[p][i] = v (+ base[p][i], or + out[p][i] without base,
//   where accumulate asks: 0 = never, 1 = every poly, 2 = even polys);
// * outs[0] non-null: poly p = per r + k (per = polys_per, 2 by default) of a
//   batch of l limbs goes to outs[r] + (k l + i) n, plus a0[i n + perm[r][j]]
//   at index j when k = 0 and a0 is set, plus d[p][i] inv2[i] when d is set;
//   with xs[0] set, x[p][i] is read at xs[r] + (k x_limbs + i) n instead.
// The batch's data buffer keeps only the first pass's intermediate.  These are
// the rescale and ModDown tails folded into the transform.
constexpr int kEpiMax = 32;
struct NttEpilogue {
    uint64_t *out = nullptr;
    uint64_t out_stride = 0;
    const uint64_t *x = nullptr;
    uint64_t x_stride = 0;
    const ulonglong2 *inv = nullptr;
    int accumulate = 0;
    const uint64_t *base = nullptr;
    uint64_t base_stride = 0;
    uint64_t *outs[kEpiMax] = {};
    const uint32_t *perm[kEpiMax] = {};
    const uint64_t *a0 = nullptr;
    const uint64_t *a0s[kEpiMax] = {};  // per output ciphertext: its input's c0 (else a0)
    const uint64_t *d = nullptr;
    uint64_t d_stride = 0;
    const ulonglong2 *inv2 = nullptr;
    const uint64_t *xs[kEpiMax] = {};
    uint32_t x_limbs = 0;
    uint32_t polys_per = 2;
};

struct NttBatch {
    uint64_t *data;
    uint64_t poly_stride;
    uint32_t polys;
    uint32_t limbs;
    uint32_t prime_by_poly;
    uint32_t poly_period = 0;
    uint8_t prime_of[kMaxPrimes];
    const uint64_t *src = nullptr;
    uint64_t src_poly_stride = 0;
    // or one source per poly (host array of polys <= kNttSrcMax device pointers;
    // limb i of poly p at src_list[p] + i * n)
    const uint64_t *const *src_list = nullptr;
    // with prime_by_poly and poly_period: limb j of poly p is left as it is
    // when I = p % poly_period < skip_l and I / skip_alpha == j (0: none)
    uint32_t skip_l = 0, skip_alpha = 0;
    NttEpilogue epi;  // forward transforms only
    // forward, n >= 2^14: the column pass has already run (ks_modup_convert_col);
    // only the row pass is launched
    uint32_t rows_only = 0;
};
constexpr int kNttSrcMax = 32;

// Forward negacyclic NTT: natural order in, SEAL bit-reversed order out, [0, q).
void ntt_forward(const DeviceTables &t, const NttBatch &b, hipStream_t s);
// Inverse: bit-reversed in, natural order out, multiplied by n^-1, [0, q).
void ntt_inverse(const DeviceTables &t, const NttBatch &b, hipStream_t s);
// Tile width of the two-pass NTT (n >= 2^14) for a pass (0 = column, 1 = row),
// a direction and the power-of-two bucket of a batch size: 4, 8 or 16 (rows:
// 4 or 8), 0 = the default rule.  Process-wide; set by the autotuner.
void ntt_set_tile(int logn, int pass, bool fwd, uint32_t batch, int width);
// The epilogue of the double-precision path in 64-bit integer products (on) or
// in exact double-precision products (off, the default); the same words.
// Process-wide; set by the autotuner.
void ntt_set_epilogue_int(bool on);
bool ntt_get_epilogue_int();
int ntt_get_tile(int logn, int pass, bool fwd, uint32_t batch);
// Radix of the double-precision passes (n >= 2^14) for a pass, direction and
// batch bucket: radix-16 steps (on) or radix-8 (off, the default).  Process-wide;
// set by the autotuner.
void ntt_set_radix16(int logn, int pass, bool fwd, uint32_t batch, bool on);
bool ntt_get_radix16(int logn, int pass, bool fwd, uint32_t batch);

// ---------------------------------------------------------------- dyadic ops
// All pointers address [polys][limbs][n] arrays; limb i uses prime i.
void poly_add(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, uint32_t polys,
              uint32_t limbs, hipStream_t s);
void poly_sub(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, uint32_t polys,
              uint32_t limbs, hipStream_t s);
void poly_negate(const DeviceTables &t, uint64_t *out, const uint64_t *a, uint32_t polys, uint32_t limbs,
                 hipStream_t s);
// out = a (polys polynomials) with pt[limbs][n] added to / subtracted from the first
void poly_add_plain_first(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *pt, uint32_t polys,
                          uint32_t limbs, bool sub, hipStream_t s);
// out[p][i] = a[p][i] * pt[i]
void poly_mul_plain(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *pt, uint32_t polys,
                    uint32_t limbs, hipStream_t s);
// (a0,a1) x (b0,b1) -> (a0b0, a0b1+a1b0, a1b1); out may alias neither input.
void poly_tensor(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, uint32_t limbs,
                 hipStream_t s);
// Pointer lists of batched element-wise kernels (count <= kPtrMax).
constexpr int kPtrMax = 32;
struct PairPtrs {
    const uint64_t *x[kPtrMax];
    const uint64_t *y[kPtrMax];
};
// c01[r][2][l][n] = (a_r0 b_r0, a_r0 b_r1 + a_r1 b_r0), c2[r][l][n] = a_r1 b_r1
void poly_tensor_many(const DeviceTables &t, uint64_t *c01, uint64_t *c2, const uint64_t *const *a,
                      const uint64_t *const *b, uint32_t count, uint32_t limbs, hipStream_t s);
// out[3][limbs][n] (+)= sum_r a_r (x) b_r, the size-3 tensors of count <= kPtrMax
// pairs of [2][limbs][n] ciphertexts (primes of at most 60 bits)
void poly_tensor_sum(const DeviceTables &t, uint64_t *out, const uint64_t *const *a, const uint64_t *const *b,
                     uint32_t count, uint32_t limbs, bool accumulate, hipStream_t s);
// acc[p][limb][c] += d[p][limb][c] * w (Shoup constant of prime `limb`) for polys p
void rns_add_limb_scaled(const DeviceTables &t, uint64_t *acc, uint64_t acc_stride, const uint64_t *d,
                         uint64_t d_stride, uint32_t limb, ulonglong2 w, uint32_t polys, hipStream_t s);
// out[p][i] = a[p][i] * w[i] with per-limb Shoup constants {w, w'} (host array, passed by value)
void poly_mul_scalar(const DeviceTables &t, uint64_t *out, const uint64_t *a, const ulonglong2 *host_w,
                     uint32_t polys, uint32_t limbs, hipStream_t s);
// out[p][i][c] = a[p][i][perm[c]]
void poly_permute(uint64_t *out, const uint64_t *a, const uint32_t *perm, uint32_t polys, uint32_t limbs, int logn,
                  hipStream_t s);
// Fill limb i of each poly with host_value[i] (host array, passed by value).
void poly_fill_limbs(uint64_t *out, const uint64_t *host_value, uint32_t polys, uint32_t limbs, int logn,
                     hipStream_t s);
// out[p][i] (+)= sum_k x[k][p][i] w[k][i] for polys p < 2, limbs i < limbs,
// terms <= kLinMax (x[k] poly stride x_stride[k] words, at least `limbs` limbs;
// w[k] host arrays of residues): constant multiples summed in one pass
constexpr int kLinMax = 4;
void poly_lincomb(const DeviceTables &t, uint64_t *out, const uint64_t *const *x, const uint64_t *x_stride,
                  const uint64_t *const *w, uint32_t terms, uint32_t limbs, bool accumulate, hipStream_t s);
// dst[p][0 .. limbs n) = src[p][0 .. limbs n) for p < polys (poly strides in
// words, multiples of 2; n >= 2): the first limbs of each poly, e.g. a mod switch
void poly_copy_limbs(const DeviceTables &t, uint64_t *dst, uint64_t dst_stride, const uint64_t *src,
                     uint64_t src_stride, uint32_t polys, uint32_t limbs, hipStream_t s);
// out[i][c] = a[i][c] + v[i] on one poly (host array of residues, passed by value)
void poly_add_scalar(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *host_v,
                     uint32_t limbs, hipStream_t s);

// ---------------------------------------------------------------- RNS ops
// Rescale / ModDown tail, SEAL-exact rounding.
//  last: coefficient-domain limb (already INTT'd) of prime `last_prime`, one per poly
//  out[p][i] = (x[p][i] - NTT_i(centered(last[p]) mod q_i)) * inv[i]   for i < limbs
// This writes the centred residues mod q_i into tmp[p][i]; the caller's forward
// NTT of tmp combines them with x through an NttEpilogue.
void rns_center_last(const DeviceTables &t, uint64_t *tmp, const uint64_t *last, uint64_t last_stride,
                     uint32_t last_prime, uint32_t polys, uint32_t limbs, hipStream_t s);
// rns_center_last fused into the forward NTT's column pass: tmp[p][i] holds the
// column-transformed centred residues; the caller's forward NTT runs its row
// pass only (NttBatch::rows_only).  Needs center_col_ok: every prime below
// 2^51, 2^14 <= n <= 2^17.
bool center_col_ok(const DeviceTables &t);
void rns_center_last_col(const DeviceTables &t, uint64_t *tmp, const uint64_t *last, uint64_t last_stride,
                         uint32_t last_prime, uint32_t polys, uint32_t limbs, hipStream_t s);
// Key-switch ModUp over a batch of B targets: for ciphertext b, every output
// prime I in out_primes (count = nout) and digit j < l,
// T[b][I][j] = coef[b][j] mod q_I  (coef = INTT of the target, [B][l][n]).
void ks_modup_expand(const DeviceTables &t, uint64_t *T, const uint64_t *coef, uint32_t B, uint32_t l,
                     const uint8_t *out_primes_host, uint32_t nout, hipStream_t s);
// Inner product: acc[b][k][I][c] = sum_j T[b][I][j][c] * key[j][k][kp(I)][c] mod q_kp(I)
// key layout [digits][2][key_limbs][n], shared by the batch.
// aseed (device array, one seed per digit) non-null: poly 1 of every key digit
// is the sampler's uniform residues, ctr_uniform(aseed[j], kp n + c, q_kp),
// regenerated in the kernel instead of read (keys generated by this engine).
void ks_inner_product(const DeviceTables &t, uint64_t *acc, const uint64_t *T, const uint64_t *key, uint32_t B,
                      uint32_t l, uint32_t key_limbs, const uint8_t *out_primes_host, uint32_t nout, hipStream_t s,
                      const uint64_t *aseed = nullptr);

// Hybrid key switching (alpha special primes): fast basis extension of the
// INTT'd limbs coef[B][l][n] digit by digit into T[B][I][j][n] for every output
// prime I of (q_0..q_{l-1}, p_0..p_{alpha-1}); and of the special limbs of acc
// (limbs l..l+alpha-1 of each poly, poly stride acc_stride) into z[polys][l][n].
// Tables as Context::d_modup_* / d_moddown_*.
void ks_modup_convert(const DeviceTables &t, uint64_t *T, const uint64_t *coef, const ulonglong2 *w,
                      const uint64_t *conv, uint32_t B, uint32_t l, uint32_t alpha, uint32_t key_limbs,
                      const uint8_t *out_primes_host, uint32_t nout, hipStream_t s);
void ks_moddown_convert(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                        const ulonglong2 *w, const uint64_t *conv, uint32_t polys, uint32_t l, uint32_t alpha,
                        uint32_t L, hipStream_t s);

// Fused forms (one thread per input word, looping the outputs); usable while
// alpha <= 16.
bool ks_conv_fused_ok(uint32_t alpha);
void ks_modup_convert_fused(const DeviceTables &t, uint64_t *T, const uint64_t *coef, const ulonglong2 *w,
                            const uint64_t *conv, uint32_t B, uint32_t l, uint32_t alpha, uint32_t key_limbs,
                            const uint8_t *out_primes_host, uint32_t nout, hipStream_t s);
// convd: the factors as doubles (Context::d_moddown_conv_d); used when every
// special prime is below 2^51
void ks_moddown_convert_fused(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                              const ulonglong2 *w, const uint64_t *conv, const double *convd, uint32_t polys,
                              uint32_t l, uint32_t alpha, uint32_t L, hipStream_t s);
// ModDown by P' = P q_(l-1) in one rounding: the limbs (q_(l-1), p_0..p_(alpha-1))
// of each poly of acc (starting at limb l_out = l - 1, poly stride acc_stride,
// coefficient form) converted exactly (centred) to q_0..q_(l_out - 1) into
// z[polys][l_out][n], with the per-level tables Context::d_mdr_*.  Needs
// ks_moddown_rescale_ok (every prime below 2^58, alpha + 1 <= 16).
bool ks_moddown_rescale_ok(const DeviceTables &t, uint32_t alpha);
// T[b][I][I / alpha] = target_b[I] (NTT form) for I < l: the limbs of ModUp
// digit j on its own primes, which the conversion reproduces exactly.
// targets [B][l][n], or target_list (B <= kNttSrcMax pointers) when non-null.
void ks_modup_copy_own(const DeviceTables &t, uint64_t *T, const uint64_t *targets,
                       const uint64_t *const *target_list, uint32_t B, uint32_t l, uint32_t alpha, uint32_t digits,
                       uint32_t nout, hipStream_t s);
// ModUp conversion fused into the forward NTT's column pass: T[b][I][j] for
// the output primes I outside digit j, column-transformed (the caller runs the
// row pass with NttBatch::rows_only); the digit's own limbs are not written
// (ks_modup_copy_own).  Needs ks_modup_col_ok: every prime below 2^51,
// 2 <= alpha <= 8, 2^14 <= n <= 2^17, the double-precision factor table.
bool ks_modup_col_ok(const DeviceTables &t, uint32_t alpha);
// per power-of-two bucket of B: -1 = unfused; otherwise fused, with k =
// mode & 255 outputs per workgroup (0: the default rule), 4-column tiles
// instead of 8 when mode & 256, 512 threads per workgroup (8-column tiles,
// half the registers per thread) when mode & 512
void ks_modup_col_set(uint32_t B, int mode);
int ks_modup_col_get(uint32_t B);
void ks_modup_convert_col(const DeviceTables &t, uint64_t *T, const uint64_t *coef, const ulonglong2 *w,
                          uint32_t B, uint32_t l, uint32_t alpha, uint32_t key_limbs, const uint8_t *out_primes_host,
                          uint32_t nout, hipStream_t s);
// ModDown conversion fused into the forward column pass: the A coefficient-form
// limbs lo .. lo + A - 1 of each poly of acc (primes in_primes_host, Shoup
// constants w, double factor table convd with rows k <= A of stride L as in
// Context::d_moddown_conv_d / d_mdr_conv_d) converted, centred, to the primes
// q_0 .. q_(lo-1) and column-transformed into z[polys][lo][n]; the caller runs
// the row pass (NttBatch::rows_only), with its epilogue.  Needs
// ks_moddown_col_ok: every prime below 2^51, 2 <= A <= 9, 2^14 <= n <= 2^17.
bool ks_moddown_col_ok(const DeviceTables &t, uint32_t A);
void ks_moddown_convert_col(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                            const ulonglong2 *w, const double *convd, const uint8_t *in_primes_host, uint32_t A,
                            uint32_t polys, uint32_t lo, uint32_t L, hipStream_t s);
// per power-of-two bucket of the poly count: as ks_modup_col_set
void ks_moddown_col_set(uint32_t polys, int mode);
int ks_moddown_col_get(uint32_t polys);
void ks_moddown_rescale_convert(const DeviceTables &t, uint64_t *z, const uint64_t *acc, uint64_t acc_stride,
                                const ulonglong2 *w, const uint64_t *conv, const double *convd, uint32_t polys,
                                uint32_t l_out, uint32_t alpha, uint32_t L, hipStream_t s);
// out[k][limb] = sum_r cts[r][k][limb] * pts[r][limb] for k < 2 (count <= kMaxDot):
// the plaintext multiply-accumulate of a BSGS giant step in one pass.
constexpr int kMaxDot = 64;
// out_g[k][limb] (+)= sum_r cts[r][k][limb] * pts[g count + r][limb] for rows
// g < rows <= kMultiDotG, k < 2, count <= kMultiDotK (pts entries may be null:
// no term); each ciphertext is read once for all rows.  accumulate adds into
// the outputs instead of writing them.
constexpr int kMultiDotG = 4, kMultiDotK = 32;
void poly_dot_plain_multi(const DeviceTables &t, uint64_t *const *outs, const uint64_t *const *cts,
                          const uint64_t *const *pts, uint32_t rows, uint32_t count, uint32_t limbs, bool accumulate,
                          hipStream_t s);
void poly_dot_plain(const DeviceTables &t, uint64_t *out, const uint64_t *const *cts, const uint64_t *const *pts,
                    uint32_t count, uint32_t limbs, hipStream_t s);
// Layout of k_dot_plain for launches of at most `limbs` limbs: 0 = default
// (two coefficients of both polynomials per thread), 1 = one coefficient of
// one polynomial per thread (four times the threads).  Set by the autotuner.
void dot_plain_set_layout(uint32_t limbs, int layout);
int dot_plain_get_layout(uint32_t limbs);

// Periodic plaintexts (kernels_mix.hip): `rows` = entries x limbs polynomials
// whose NTT values are constant over `blocks` blocks of n / blocks positions.
// scatter: work[row][c] = coef[row][c / stride] at multiples of stride, else 0
// (and flags[row][*] = 0); gather: table[row][b] = work[row][b stride], flags
// set where a block is not constant; expand: one entry's full NTT form.
void periodic_scatter(const DeviceTables &t, uint64_t *work, uint8_t *flags, const uint64_t *coef, uint32_t rows,
                      uint32_t blocks, hipStream_t s);
void periodic_gather(const DeviceTables &t, uint64_t *table, uint8_t *flags, const uint64_t *work, uint32_t rows,
                     uint32_t blocks, hipStream_t s);
void periodic_expand(const DeviceTables &t, uint64_t *out, const uint64_t *table, uint32_t limbs, uint32_t blocks,
                     hipStream_t s);
// out[m] = sum over terms (k, e) of output m of table[e] (.) in[k], both
// polynomials, for K inputs and M outputs of `limbs` limbs (K, M <= kMixMax):
// terms as runs[first[m] .. first[m + 1]) of int32 quads (k0, e0, len, de),
// meaning (k0 + j, e0 + j de) for j < len (16-byte aligned).  Needs n >= 64
// and block_mix_lds_bytes(...) <= 64 KiB.
constexpr int kMixMax = 128;
struct MixPtrs {
    const uint64_t *in[kMixMax];
    uint64_t *out[kMixMax];
};
size_t block_mix_lds_bytes(uint32_t K, uint32_t count, uint32_t blocks, int logn);
void block_mix(const DeviceTables &t, const MixPtrs &p, const uint64_t *table, const int32_t *terms,
               const uint32_t *first, uint32_t K, uint32_t M, uint32_t nterms, uint32_t limbs, uint32_t count,
               uint32_t blocks, hipStream_t s);

// Hoisted variants over one shared ModUp T[nout][l][n] (B = 1): rotation r
// reads digit values through the NTT-domain Galois permutation perms[r]
// (T[I][j][perms[r][c]]) and uses key keys[r].  _rot writes one result per
// rotation, acc[r][2][nout][n]; _rot_sum writes (or adds to acc) their sum,
// acc[2][nout][n].
constexpr int kMaxHoist = 32;
void ks_inner_product_rot(const DeviceTables &t, uint64_t *acc, const uint64_t *T, const uint64_t *const *keys,
                          const uint32_t *const *perms, uint32_t count, uint32_t l, uint32_t key_limbs,
                          const uint8_t *out_primes_host, uint32_t nout, hipStream_t s,
                          const uint64_t *const *Ts = nullptr, const uint64_t *const *aseeds = nullptr);
// t_stride > 0: rotation r reads its own ModUp at T + r t_stride (a sum of
// rotations of different inputs) instead of the shared T.
void ks_inner_product_rot_sum(const DeviceTables &t, uint64_t *acc, const uint64_t *T, const uint64_t *const *keys,
                              const uint32_t *const *perms, uint32_t count, uint32_t l, uint32_t key_limbs,
                              const uint8_t *out_primes_host, uint32_t nout, bool accumulate, hipStream_t s,
                              uint64_t t_stride = 0, const uint64_t *const *aseeds = nullptr);
// out[2][l][n] = sum_i (sigma_i(in_i[0]), 0), or (in_i[0], in_i[1]) where
// perms[i] is null (count <= kMaxHoist; in_i are [2][l][n] ciphertexts).
struct RotBaseArgs {
    const uint64_t *in[kMaxHoist];
    const uint32_t *perm[kMaxHoist];
};
void rot_base(const DeviceTables &t, uint64_t *out, const uint64_t *const *in, const uint32_t *const *perms,
              uint32_t count, uint32_t l, hipStream_t s);

// ---------------------------------------------------------------- keys
// Deterministic counter-based sampling (not SEAL's PRNG).
void sample_uniform(const DeviceTables &t, uint64_t *out, uint32_t polys, uint32_t limbs, const uint8_t *primes_host,
                    uint64_t seed, hipStream_t s);
// Centered binomial (eta = 21) small error, written as residues in all limbs.
void sample_cbd(const DeviceTables &t, uint64_t *out, uint32_t limbs, const uint8_t *primes_host, uint64_t seed,
                hipStream_t s);
// Signed small coefficients (int8) -> residues for each limb.
void small_to_rns(const DeviceTables &t, uint64_t *out, const int8_t *coef, uint32_t limbs,
                  const uint8_t *primes_host, hipStream_t s);
// Signed 64-bit coefficients -> residues.
void i64_to_rns(const DeviceTables &t, uint64_t *out, const int64_t *coef, uint32_t limbs,
                const uint8_t *primes_host, hipStream_t s);
// out = a * b + c  per limb (c may be null), all NTT domain, limb i uses primes[i]
void poly_mul_add(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *b, const uint64_t *c,
                  uint32_t limbs, const uint8_t *primes_host, int negate_product, hipStream_t s);
// out[c] += a[c] * w (mod q_prime) for one limb; w = {value, Shoup quotient}
void limb_add_scaled(const DeviceTables &t, uint64_t *out, const uint64_t *a, ulonglong2 w, uint32_t prime,
                     hipStream_t s);
// out[p][i] = a[p][i] * P[i] + b[p][i] * Q[i]  (NTT-domain, limb i uses prime i)
void poly_mul2_add(const DeviceTables &t, uint64_t *out, const uint64_t *a, const uint64_t *P, const uint64_t *b,
                   const uint64_t *Q, uint32_t polys, uint32_t limbs, hipStream_t s);
// out[i][p][l] = sum_{j < inner} E[i * inner + j][p][l] * A[j][l]  for i < cols, p < 2
// (E: [cols*inner][2][limbs][n] expanded ciphertexts, A: [inner][limbs][n] plaintexts)
void mm_accumulate(const DeviceTables &t, uint64_t *out, const uint64_t *E, const uint64_t *A, uint32_t cols,
                   uint32_t inner, uint32_t limbs, hipStream_t s);
// ModRaise: in0 = coefficient-domain q0 limb of each poly ([polys][n]) -> out [polys][limbs][n]
void mod_raise(const DeviceTables &t, uint64_t *out, const uint64_t *in0, uint32_t polys, uint32_t limbs,
               hipStream_t s);
// Negacyclic shift of coefficient-domain limbs: out = a * x^k mod (x^n + 1), 0 <= k < 2n.
void poly_negacyclic_shift(const DeviceTables &t, uint64_t *out, const uint64_t *a, uint32_t k, uint32_t polys,
                           uint32_t limbs, hipStream_t s);

// ---------------------------------------------------------------- encode
// CKKS encode of count <= n/2 complex values (re, im device arrays; im may be
// null) to the rounded coefficients' residues out[limbs][n] (coefficient
// domain), bit-identical to SEAL's host encoder.  work holds n double2;
// index_map / roots are SEAL's matrix_reps_index_map and inverse root powers;
// scalar = scale / n.  blockmax[encode_blocks(logn)] receives per-block maxima
// of |coefficient| as IEEE bit patterns.
uint32_t encode_blocks(int logn);
void encode_fft(const DeviceTables &t, uint64_t *out, double2 *work, const double *re, const double *im,
                uint32_t count, const uint32_t *index_map, const double2 *roots, double scalar, uint32_t limbs,
                uint64_t *blockmax, hipStream_t s);

// VALU issue-rate probe on the current device: kind 0 = 64 x 64 -> 128-bit
// products per second, kind 1 = double-precision FMAs per second
double probe_valu_rate(int kind, hipStream_t s);

}  // namespace nx
