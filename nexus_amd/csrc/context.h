// Host-side engine state: CKKS parameter chain, device tables, device memory,
// ciphertext / plaintext / key objects.  Internal to libnexus_ckks.so; the
// public surface is the C ABI in include/nexus_ckks.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "modarith.h"

namespace nx {

// ---------------------------------------------------------------- errors
enum ErrCode { kOk = 0, kInvalidArgument = -1, kLogicError = -2, kOutOfRange = -3, kDeviceError = -4,
               kNoMemory = -5 };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }
inline void require(bool ok, const char *msg, int code = kInvalidArgument)
{
    if (!ok) fail(code, msg);
}

void hip_check(hipError_t e, const char *what);
void set_last_error(const char *msg);  // thread-local message behind nx_last_error()
#define NX_HIP(x) ::nx::hip_check((x), #x)

// ---------------------------------------------------------------- memory
// Size-bucketed cache over hipMalloc.  All work of one context is ordered on
// one stream, so a block released after its last enqueued use can be handed to
// the next allocation without waiting: the reuse is ordered after it.
class DeviceCache {
public:
    ~DeviceCache();
    void *get(size_t bytes);      // throws when the device is out of memory
    void *try_get(size_t bytes);  // null when the device is out of memory
    void put(void *p, size_t bytes);
    void trim();
    static size_t size_class(size_t bytes);
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_cached() const { return cached_; }

private:
    std::unordered_map<size_t, std::vector<void *>> free_;
    size_t in_use_ = 0, cached_ = 0;
    std::mutex mu_;
};

class Context;

// A lane: one HIP stream (default priority) and the memory cache of the work
// ordered on it.  A thread binds a lane of a context to run that context's
// operations on the lane's stream; threads on different lanes overlap on the
// device while sharing the context's keys and tables.  A buffer returns to the
// cache of the lane that allocated it, so a ciphertext used on another lane
// must be released only after that lane has synchronised.
struct Lane {
    hipStream_t stream = nullptr;
    std::shared_ptr<DeviceCache> cache;
};

// Stream of the calling thread for one context: its bound lane's stream, else
// the context's default stream.  Converts to hipStream_t wherever a stream is
// passed.
class StreamRef {
public:
    explicit StreamRef(const Context *c) : c_(c) {}
    operator hipStream_t() const;
    hipStream_t default_stream() const { return def_; }

private:
    friend class Context;
    const Context *c_;
    hipStream_t def_ = nullptr;
};

// Device buffer from a context's cache.  The buffer shares ownership of the
// cache, so releasing it stays valid even after its context is gone.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(Context *ctx, size_t words);
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept { *this = std::move(o); }
    DevBuf &operator=(DevBuf &&o) noexcept;
    void reset(Context *ctx, size_t words);  // reallocate when the size changes
    void release();
    uint64_t *get() const { return p_; }
    size_t words() const { return words_; }

private:
    std::shared_ptr<DeviceCache> cache_;
    uint64_t *p_ = nullptr;
    size_t words_ = 0;
};

// ---------------------------------------------------------------- objects
struct Ciphertext {
    Context *ctx = nullptr;
    size_t size = 0;         // number of polynomials (2 or 3)
    size_t limbs = 0;        // RNS limbs (data primes q_0..q_{limbs-1})
    double scale = 1.0;
    bool ntt = true;
    DevBuf data;             // [size][limbs][n]
};

struct Plaintext {
    Context *ctx = nullptr;
    size_t limbs = 0;
    double scale = 1.0;
    bool ntt = true;
    DevBuf data;             // [limbs][n]
};

struct SecretKey {
    Context *ctx = nullptr;
    std::vector<int8_t> coeff;  // ternary coefficients (host copy, for export)
    DevBuf ntt;                 // [L+1][n], NTT form at the key level
};

// Key-switching keys: one entry per Galois element (0 = relinearisation).
// Layout per entry: [L digits][2 polys][L+1 limbs][n], NTT form.
struct KSwitchKeys {
    Context *ctx = nullptr;
    std::map<uint32_t, DevBuf> keys;
};

// ---------------------------------------------------------------- context
class Context {
public:
    // The last `special` primes of bit_sizes are the key-switching special
    // primes p_0..p_{alpha-1}; with one, key switching is SEAL's (one digit
    // per data prime); with alpha > 1 it is hybrid: digit j is the alpha
    // consecutive data primes q_{j alpha}.. (Phantom's special_modulus_size).
    Context(int logn, const std::vector<int> &bit_sizes, uint32_t hamming_weight, size_t sparse_slots,
            size_t special = 1);
    ~Context();

    int logn;
    size_t n;
    size_t L;                          // data primes at the top level
    size_t alpha = 1;                  // special primes
    size_t dnum = 0;                   // key-switching digits at the top level, ceil(L / alpha)
    std::vector<uint64_t> moduli;      // q_0..q_{L-1}, p_0..p_{alpha-1}
    std::vector<ModConst> mods;
    uint32_t hamming_weight;
    size_t sparse_slots;
    StreamRef stream{this};                                              // the calling thread's stream
    std::shared_ptr<DeviceCache> cache = std::make_shared<DeviceCache>();  // default lane's cache

    // lanes (created on demand, index >= 1; 0 is the default stream) and the
    // calling thread's binding
    Lane *lane(size_t index);
    void bind_lane(size_t index);   // 0 unbinds
    Lane *bound_lane() const;
    void trim_caches();  // free the cached blocks of the default lane and every lane
    const std::shared_ptr<DeviceCache> &current_cache() const;

    // device tables
    DeviceTables tables;
    ulonglong2 *d_fwd = nullptr, *d_inv = nullptr, *d_ninv = nullptr;
    double *d_fwd_d = nullptr, *d_inv_d = nullptr, *d_ninv_d = nullptr;
    double2 *d_qd = nullptr;  // [prime] {q, 1 / q} in double precision
    ModConst *d_mod = nullptr;
    // rescale: d_rescale_inv[l] = {q_{l-1}^-1 mod q_i, Shoup} for i < l-1 (row l of a (L+1)x(L+1) table)
    ulonglong2 *d_rescale_inv = nullptr;
    // ModDown: {P^-1 mod q_i, Shoup}, P = p_0 ... p_{alpha-1}
    ulonglong2 *d_p_inv = nullptr;

    // chain helpers
    size_t chain_index_of(size_t limbs) const { return L + 1 - limbs; }
    size_t limbs_of(size_t chain_index) const
    {
        require(chain_index >= 1 && chain_index <= L, "chain_index out of range", kOutOfRange);
        return L + 1 - chain_index;
    }
    int total_bits(size_t limbs) const;  // bit count of q_0 ... q_{limbs-1}
    size_t key_limbs() const { return L + alpha; }
    size_t digits(size_t limbs) const { return (limbs + alpha - 1) / alpha; }
    size_t key_words() const { return dnum * 2 * key_limbs() * n; }
    // hybrid ModUp / ModDown base-conversion tables (alpha > 1)
    //   modup_w[(j*alpha + s-1)*alpha + i]: [Dhat_i^-1]_{q_{j alpha + i}} (Shoup), digit j cut to s primes
    //   modup_conv[((j*alpha + s-1)*alpha + i) * (L+alpha) + o]: Dhat_i mod (prime o)
    //   moddown_w[k]: [Phat_k^-1]_{p_k} (Shoup); moddown_conv[k*L + i]: Phat_k mod q_i,
    //   moddown_conv[alpha*L + i]: P mod q_i
    ulonglong2 *d_modup_w = nullptr, *d_moddown_w = nullptr;
    uint64_t *d_modup_conv = nullptr, *d_moddown_conv = nullptr;
    double *d_modup_conv_d = nullptr;  // d_modup_conv as doubles
    // ModDown factors as doubles (row alpha: q_i - P mod q_i), and the same per
    // level for the ModDown by P q_(l-1) ((alpha + 2) L doubles a level)
    double *d_moddown_conv_d = nullptr, *d_mdr_conv_d = nullptr;
    // ModDown by P' = P q_(l-1) at level l (hybrid relinearisation + rescale in one
    // rounding), base (q_(l-1), p_0..p_(alpha-1)):
    //   mdr_w[l (alpha+1) + k] = [(P'/b_k)^-1]_(b_k) (Shoup)
    //   mdr_conv[(l (alpha+2) + k) L + i] = P'/b_k mod q_i, row alpha+1: P' mod q_i
    //   mdr_inv[l L + i] = {P'^-1 mod q_i, Shoup};  mdr_pmod[l] = P mod q_(l-1) (host, Shoup)
    ulonglong2 *d_mdr_w = nullptr, *d_mdr_inv = nullptr;
    uint64_t *d_mdr_conv = nullptr;
    std::vector<ulonglong2> mdr_pmod;
    const ulonglong2 *rescale_inv(size_t limbs) const { return d_rescale_inv + limbs * moduli.size(); }

    // Galois permutation tables (NTT domain), built on first use
    const uint32_t *galois_table(uint32_t elt);
    uint32_t elt_from_step(int step) const;

    // encoder tables on the device: matrix_reps_index_map (n) and the inverse
    // complex root powers (n) of the encoding FFT
    uint32_t *d_enc_index = nullptr;
    double2 *d_enc_roots = nullptr;

    // host copies of SEAL-ordered root tables (encoder uses complex roots, not these)
    std::vector<std::vector<uint64_t>> fwd_host, inv_host;
    // NTT-domain images of monomials x^k, keyed by (k, limbs)
    std::map<std::pair<uint32_t, size_t>, DevBuf> xk_cache;
    std::mutex xk_mu;

private:
    void build_hybrid_tables();
    std::vector<std::unique_ptr<Lane>> lanes_;
    std::mutex lane_mu_;
    std::map<uint32_t, uint32_t *> galois_;
    std::mutex mu_;
};

// ---------------------------------------------------------------- operations
// NTT of a [polys][limbs][n] buffer whose limb i uses prime i.
void ntt_fwd(Context &c, uint64_t *data, size_t polys, size_t limbs, size_t poly_stride);
// forward NTT whose last pass stores through epi instead of into data
// rows_only: the column pass has already run on data (a fused conversion)
void ntt_fwd_epi(Context &c, uint64_t *data, size_t polys, size_t limbs, size_t poly_stride, const NttEpilogue &epi,
                 bool rows_only = false);
void ntt_inv(Context &c, uint64_t *data, size_t polys, size_t limbs, size_t poly_stride);
// out of place: src (poly stride src_stride) -> dst (poly stride dst_stride)
void ntt_fwd_from(Context &c, uint64_t *dst, const uint64_t *src, size_t polys, size_t limbs, size_t dst_stride,
                  size_t src_stride);
void ntt_inv_from(Context &c, uint64_t *dst, const uint64_t *src, size_t polys, size_t limbs, size_t dst_stride,
                  size_t src_stride);

void ct_resize(Ciphertext &ct, Context *c, size_t size, size_t limbs);
void ct_copy(Ciphertext &dst, const Ciphertext &src);
// copy `limbs` limbs of each of `polys` polys between strided buffers (stream-ordered)
void copy_limbs(Context &c, uint64_t *dst, size_t dst_stride, const uint64_t *src, size_t src_stride, size_t polys,
                size_t limbs);
void pt_copy(Plaintext &dst, const Plaintext &src);

void op_add(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst);
void op_sub(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst);
void op_negate(const Ciphertext &a, Ciphertext &dst);
void op_add_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst);
void op_sub_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst);
void op_multiply_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst);
void op_multiply(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst);
void op_square(const Ciphertext &a, Ciphertext &dst);
void op_relinearize(const Ciphertext &a, const KSwitchKeys &rk, Ciphertext &dst);
void op_rescale_to_next(const Ciphertext &a, Ciphertext &dst);
// batched forms, word for word the per-ciphertext results
void op_rescale_many(const std::vector<const Ciphertext *> &in, const std::vector<Ciphertext *> &out);
void op_multiply_relin_rescale_many(const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                                    const KSwitchKeys &rk, const std::vector<Ciphertext *> &out);
// the same products with relinearisation and rescale in one rounding (hybrid
// key switching): each coefficient within 1 of the form above
void op_multiply_relin_rescale_fused_many(const std::vector<const Ciphertext *> &a,
                                          const std::vector<const Ciphertext *> &b, const KSwitchKeys &rk,
                                          const std::vector<Ciphertext *> &out);
// SEAL-4.1-bs add/sub/multiply_inplace_reduced_error (evaluator.cpp:420-594):
// dst = a (op) b after aligning the levels and scales; Mul relinearises
enum class ReducedOp { Add, Sub, Mul };
void op_reduced_error(const Ciphertext &a, const Ciphertext &b, ReducedOp op, const KSwitchKeys *rk, Ciphertext &dst);
// allocation-free rescale (noalloc.hip): [polys][limbs][n] NTT-form residues at
// `in` -> [polys][limbs-1][n] at `out`, both caller device memory, scratch in a
// static device array whose address is taken per device at context creation
void noalloc_init_current_device();
size_t noalloc_scratch_words();
void rescale_noalloc(Context &c, const uint64_t *in, size_t polys, size_t limbs, uint64_t *out);
void op_mod_switch_to_next(const Ciphertext &a, Ciphertext &dst);
void op_mod_switch_to(const Ciphertext &a, size_t chain_index, Ciphertext &dst);
void op_plain_mod_switch_to(const Plaintext &a, size_t chain_index, Plaintext &dst);
void op_apply_galois(const Ciphertext &a, uint32_t elt, const KSwitchKeys &gk, Ciphertext &dst);
void op_rotate(const Ciphertext &a, int steps, const KSwitchKeys &gk, Ciphertext &dst);
void op_conjugate(const Ciphertext &a, const KSwitchKeys &gk, Ciphertext &dst);
void op_to_ntt(const Ciphertext &a, Ciphertext &dst);
void op_from_ntt(const Ciphertext &a, Ciphertext &dst);
void op_add_many(const std::vector<const Ciphertext *> &in, Ciphertext &dst);
// sum_r cts[r] * pts[r] in one pass (size-2 ciphertexts of one level)
void op_dot_plain(const std::vector<const Ciphertext *> &cts, const std::vector<const Plaintext *> &pts,
                  Ciphertext &dst);
// sum_k xs[k] * encode_constant(values[k], scales[k]) at `limbs` limbs (each x_k
// read at its first limbs), scale xs[0].scale * scales[0]
void op_linear_combination(const std::vector<const Ciphertext *> &xs, const std::vector<double> &values,
                           const std::vector<double> &scales, size_t limbs, Ciphertext &dst);
// outs[g] = sum_r pts[g K + r] * cts[r] (K = cts.size(), null: no term), each
// ciphertext read once for every few rows
void op_dot_plain_multi(const std::vector<const Ciphertext *> &cts, const std::vector<const Plaintext *> &pts,
                        const std::vector<Ciphertext *> &outs);
void op_multiply_power_of_x(const Ciphertext &a, int64_t exponent, Ciphertext &dst);
void op_mod_raise(const Ciphertext &a, Ciphertext &dst);
// hoisted rotations (one ModUp of c1): one output per step / their sum
void op_rotate_hoisted(const Ciphertext &a, const std::vector<int> &steps, const KSwitchKeys &gk,
                       std::vector<Ciphertext> &dst);
void op_hoisting(const Ciphertext &a, const std::vector<int> &steps, const KSwitchKeys &gk, Ciphertext &dst);
// every input rotated by every step, hoisted per input, batched across inputs:
// dst[b][k] = op_rotate_hoisted(in[b], steps)[k] word for word
void op_rotate_hoisted_many(const std::vector<const Ciphertext *> &in, const std::vector<int> &steps,
                            const KSwitchKeys &gk, std::vector<std::vector<Ciphertext>> &dst);
// *dst[g] = op_rotate_sum(groups[g], steps[g]) word for word, the ModDowns batched
void op_rotate_sum_many(const std::vector<std::vector<const Ciphertext *>> &groups,
                        const std::vector<std::vector<int>> &steps, const KSwitchKeys &gk,
                        const std::vector<Ciphertext *> &dst);
void op_rotate_sum(const std::vector<const Ciphertext *> &in, const std::vector<int> &steps, const KSwitchKeys &gk,
                   Ciphertext &dst);

// key switch of `target` (limbs x n, NTT) under key `key` added into ct (2 polys)
void key_switch_into(Context &c, Ciphertext &ct, const uint64_t *target, const DevBuf &key);
// Key-switch pieces.  ModUp of B targets [B][l][n] -> T [B][l+1][l][n] (NTT,
// output primes q_0..q_{l-1}, p); ModDown of acc [B][2][l+1][n] into cts
// [B][2][l][n]: accumulated (zero_c1 false) or c0 accumulated and c1 written.
void ks_out_primes(const Context &c, size_t l, uint8_t *outp);
DevBuf ks_modup(Context &c, const uint64_t *targets, size_t B, size_t l);
// ModUp of B <= kNttSrcMax targets given by pointer ([l][n] each)
DevBuf ks_modup_list(Context &c, const uint64_t *const *targets, size_t B, size_t l);
// ModDown up to the final combination: INTT of the special limbs of acc
// [B][2][l+alpha][n], conversion, NTT; returns tmp [B][2][l][n] for
// out = (acc - tmp) P^-1.
// col_done non-null: the conversion may run fused into the forward column
// pass, and *col_done says whether it did (the caller's forward NTT of the
// result then runs its row pass only).  transform: the whole forward NTT runs
// here.
DevBuf ks_moddown_prepare(Context &c, uint64_t *acc, size_t B, size_t l, bool transform = true,
                          bool *col_done = nullptr);
// base (optional, [B][2][l][n]): the accumulated terms are read from base instead of cts.
void ks_moddown(Context &c, uint64_t *cts, uint64_t *acc, size_t B, size_t l, bool zero_c1,
                const uint64_t *base = nullptr);
// B key switches in one pass: targets [B][l][n] into cts [B][2][l][n]; with
// zero_c1 the c1 of each result is written (not accumulated); with base, cts =
// base + key switch.
void key_switch_batch(Context &c, uint64_t *cts, const uint64_t *targets, size_t B, size_t l, const DevBuf &key,
                      bool zero_c1, const uint64_t *base = nullptr);

// A batch of B ciphertexts of one shape (size 2) stored back to back: [B][2][limbs][n].
struct CtBatch {
    Context *ctx = nullptr;
    size_t B = 0, limbs = 0;
    double scale = 1.0;
    DevBuf data;
    uint64_t *ct(size_t i) const { return data.get() + i * 2 * limbs * ctx->n; }
};

// SEAL 4.1 serialization (compression "none") and parms_id -- serialize.cpp
void blake2b(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void parms_id(const Context &c, size_t limbs, uint64_t out[4]);  // limbs 0 = key level
std::vector<uint8_t> save_ciphertext(const Ciphertext &ct);
void load_ciphertext(Context &c, const uint8_t *data, size_t n, Ciphertext &out);
std::vector<uint8_t> save_plaintext(const Plaintext &pt);
void load_plaintext(Context &c, const uint8_t *data, size_t n, Plaintext &out);
std::vector<uint8_t> save_kswitch_keys(const KSwitchKeys &keys, bool galois);
void load_kswitch_keys(Context &c, const uint8_t *data, size_t n, bool galois, KSwitchKeys &out);

// Periodic plaintexts -- periodic.cpp.  `count` real slot vectors repeating
// every `period` slots, encoded exactly as polynomials in X^(n / (2 period))
// (coefficients round(scale c_k) of the real inverse embedding) and held in
// NTT form as `blocks` = 2 period words per limb: table [count][limbs][blocks].
struct PeriodicPlain {
    Context *ctx = nullptr;
    size_t count = 0, period = 0, blocks = 0, limbs = 0;
    double scale = 1.0;
    DevBuf table;
    // device term lists of block_mix, per index pattern (M, K, index)
    struct Terms {
        DevBuf words;  // first[M + 1] (u32) then (k, entry) int32 pairs, 8-byte aligned
        uint32_t M = 0, nterms = 0;
    };
    std::map<std::vector<int32_t>, Terms> terms;
    std::mutex mu;
};
void periodic_create(Context &c, const double *values, size_t period, size_t count, size_t limbs, double scale,
                     PeriodicPlain &out);
void periodic_expand_entry(const PeriodicPlain &p, size_t index, Plaintext &out);
// out[m] = sum_k W[index[m K + k]] (.) in[k] (index < 0: no term)
void op_periodic_mix(const std::vector<const Ciphertext *> &in, PeriodicPlain &w, const int32_t *index, size_t M,
                     const std::vector<Ciphertext *> &out);

// launch-shape autotuner -- autotune.cpp; returns a JSON report
std::string autotune(Context &c, int reps);

// encoder (device FFT and NTT) -- encoder.cpp
void upload_encoder_tables(Context &c);
void encode_vector(Context &c, const double *re, const double *im, size_t count, size_t limbs, double scale,
                   Plaintext &out);
void encode_constant(Context &c, double value, size_t limbs, double scale, Plaintext &out);
// residues of round(value * scale) mod q_0 .. q_(limbs-1) (the constant encode_constant fills in)
void constant_residues(Context &c, double value, size_t limbs, double scale, uint64_t *res);
void decode(Context &c, const Plaintext &pt, double *re, double *im, size_t count);

// keys -- keys.cpp
void secret_key_generate(Context &c, uint64_t seed, SecretKey &sk);
void secret_key_from_coeffs(Context &c, const int8_t *coeff, SecretKey &sk);
void relin_key_generate(Context &c, const SecretKey &sk, uint64_t seed, KSwitchKeys &out);
void galois_keys_generate(Context &c, const SecretKey &sk, const uint32_t *elts, size_t count, uint64_t seed,
                          KSwitchKeys &out);
void encrypt_symmetric(Context &c, const SecretKey &sk, const Plaintext &pt, uint64_t seed, Ciphertext &out);
void decrypt(Context &c, const SecretKey &sk, const Ciphertext &ct, Plaintext &out);

inline bool are_close(double a, double b)
{
    double f = std::max(std::max(std::fabs(a), std::fabs(b)), 1.0);
    return std::fabs(a - b) < 2.220446049250313e-16 * f;
}

}  // namespace nx
