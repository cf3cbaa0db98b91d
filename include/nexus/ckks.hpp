// C++ drop-in for NEXUS's CKKSEvaluator facade (reference
// cuda/src/ckks_evaluator.cuh:11-499 and src/ckks_evaluator.h:12-103), written
// over the C ABI in nexus_ckks.h.  Ciphertexts and plaintexts are value types
// with deep copy, as PhantomCiphertext / seal::Ciphertext are; every failure of
// the C layer is rethrown as the reference's exception type
// (std::invalid_argument / std::logic_error / std::out_of_range).
//
// Semantics follow the SEAL-4.1-bs path of the reference (src/): encode(double)
// is SEAL's constant encoder, multiply_const / add_const encode at the
// ciphertext's scale (SEAL evaluator.cpp:395-418).
#pragma once

#include <complex>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../nexus_ckks.h"

namespace nexus {

void throw_status(int rc);
inline void check(int rc)
{
    if (rc != NX_OK) throw_status(rc);
}

class CKKSContext {
public:
    CKKSContext(int log_n, const std::vector<int> &bit_sizes, uint32_t hamming_weight = 0, size_t sparse_slots = 0,
                size_t special_primes = 1);
    explicit CKKSContext(nx_context *borrowed);  // non-owning view of a C handle
    ~CKKSContext();
    CKKSContext(const CKKSContext &) = delete;
    CKKSContext &operator=(const CKKSContext &) = delete;
    nx_context *get() const { return h_; }
    int log_n() const { return log_n_; }
    size_t poly_degree() const { return size_t(1) << log_n_; }
    size_t data_primes() const { return L_; }
    size_t special_primes() const { return alpha_; }
    const std::vector<uint64_t> &moduli() const { return moduli_; }
    // modulus q_{k-1} counted from the top of a ciphertext at chain_index (k = 1: last limb)
    uint64_t modulus_from_top(size_t chain_index, int k) const { return moduli_[L_ + 1 - chain_index - k]; }
    uint32_t galois_elt(int step) const;
    void synchronize() const { check(nx_context_synchronize(h_)); }

private:
    void load_info();
    nx_context *h_ = nullptr;
    bool owns_ = true;
    int log_n_;
    size_t L_, alpha_ = 1;
    std::vector<uint64_t> moduli_;
};

class Ciphertext {
public:
    Ciphertext() = default;
    explicit Ciphertext(const CKKSContext &ctx);
    Ciphertext(const Ciphertext &o);
    Ciphertext(Ciphertext &&o) noexcept;
    Ciphertext &operator=(const Ciphertext &o);
    Ciphertext &operator=(Ciphertext &&o) noexcept;
    ~Ciphertext();
    nx_ciphertext *get() const { return h_; }
    const CKKSContext *context() const { return ctx_; }
    void bind(const CKKSContext &ctx);  // allocate the handle when first written
    static Ciphertext copy_of(const CKKSContext &ctx, const nx_ciphertext *h);
    size_t size() const;
    size_t params_id() const;  // chain index (Phantom naming)
    size_t chain_index() const { return params_id(); }
    size_t coeff_modulus_size() const;
    double scale() const;
    void set_scale(double s);

private:
    const CKKSContext *ctx_ = nullptr;
    nx_ciphertext *h_ = nullptr;
};

class Plaintext {
public:
    Plaintext() = default;
    explicit Plaintext(const CKKSContext &ctx);
    Plaintext(const Plaintext &o);
    Plaintext(Plaintext &&o) noexcept;
    Plaintext &operator=(const Plaintext &o);
    Plaintext &operator=(Plaintext &&o) noexcept;
    ~Plaintext();
    nx_plaintext *get() const { return h_; }
    void bind(const CKKSContext &ctx);
    size_t params_id() const;
    double scale() const;
    void set_scale(double s);

private:
    const CKKSContext *ctx_ = nullptr;
    nx_plaintext *h_ = nullptr;
};

class SecretKey {
public:
    SecretKey(const CKKSContext &ctx, uint64_t seed);
    explicit SecretKey(nx_secret_key *borrowed) : h_(borrowed), owns_(false) {}
    ~SecretKey();
    SecretKey(const SecretKey &) = delete;
    SecretKey &operator=(const SecretKey &) = delete;
    nx_secret_key *get() const { return h_; }

private:
    nx_secret_key *h_ = nullptr;
    bool owns_ = true;
};

// Device-resident key-switching keys (relinearisation key = element 0).
class KSwitchKeys {
public:
    explicit KSwitchKeys(const CKKSContext &ctx);
    explicit KSwitchKeys(nx_kswitch_keys *borrowed) : h_(borrowed), owns_(false) {}
    ~KSwitchKeys();
    KSwitchKeys(const KSwitchKeys &) = delete;
    KSwitchKeys &operator=(const KSwitchKeys &) = delete;
    nx_kswitch_keys *get() const { return h_; }
    bool has(uint32_t elt) const;

private:
    nx_kswitch_keys *h_ = nullptr;
    bool owns_ = true;
};
typedef KSwitchKeys RelinKeys;
typedef KSwitchKeys GaloisKeys;

// `count` real slot vectors repeating every `period` slots (values
// [count][period]), encoded exactly and held as 2 period NTT words per limb
// (nx_periodic_plain_create).
class PeriodicPlaintexts {
public:
    PeriodicPlaintexts(const CKKSContext &ctx, const std::vector<double> &values, size_t period, size_t chain_index,
                       double scale);
    ~PeriodicPlaintexts();
    PeriodicPlaintexts(const PeriodicPlaintexts &) = delete;
    PeriodicPlaintexts &operator=(const PeriodicPlaintexts &) = delete;
    nx_periodic_plain *get() const { return h_; }
    size_t count() const { return count_; }
    void expand(size_t index, Plaintext &out) const;
    // the full plaintext of entry `index`, expanded once and kept with this set
    const Plaintext &expanded(size_t index) const;

private:
    const CKKSContext *ctx_;
    nx_periodic_plain *h_ = nullptr;
    size_t count_ = 0;
    mutable std::map<size_t, Plaintext> expanded_;
    mutable std::mutex mu_;  // guards expanded_ (sets are shared between lanes)
};

class Encoder {
public:
    Encoder() = default;
    explicit Encoder(const CKKSContext *ctx) : ctx_(ctx) {}
    size_t slot_count() const { return ctx_->poly_degree() / 2; }
    // vector inputs (FFT encoder, SEAL ckks.h:457-620)
    void encode(const std::vector<double> &values, size_t chain_index, double scale, Plaintext &plain) const;
    void encode(const std::vector<double> &values, double scale, Plaintext &plain) const;
    void encode(const std::vector<std::complex<double>> &values, size_t chain_index, double scale,
                Plaintext &plain) const;
    void encode(const std::vector<std::complex<double>> &values, double scale, Plaintext &plain) const;
    // constant inputs (SEAL ckks.cpp:73-200: the constant in every slot)
    void encode(double value, size_t chain_index, double scale, Plaintext &plain) const;
    void encode(double value, double scale, Plaintext &plain) const;
    void decode(const Plaintext &plain, std::vector<double> &values) const;
    void decode(const Plaintext &plain, std::vector<std::complex<double>> &values) const;

private:
    const CKKSContext *ctx_ = nullptr;
};

class Encryptor {
public:
    Encryptor() = default;
    Encryptor(const CKKSContext *ctx, const SecretKey *sk) : ctx_(ctx), sk_(sk) {}
    void encrypt(const Plaintext &plain, Ciphertext &ct);

private:
    const CKKSContext *ctx_ = nullptr;
    const SecretKey *sk_ = nullptr;
    uint64_t counter_ = 0x6e657875;
};

class Decryptor {
public:
    Decryptor() = default;
    Decryptor(const CKKSContext *ctx, const SecretKey *sk) : ctx_(ctx), sk_(sk) {}
    void decrypt(const Ciphertext &ct, Plaintext &plain) const;
    void create_galois_keys_from_steps(const std::vector<int> &steps, GaloisKeys &keys) const;
    void create_galois_keys_from_elts(const std::vector<uint32_t> &elts, GaloisKeys &keys) const;
    void create_relin_keys(RelinKeys &keys) const;

private:
    const CKKSContext *ctx_ = nullptr;
    const SecretKey *sk_ = nullptr;
};

class Evaluator {
public:
    Evaluator() = default;
    explicit Evaluator(const CKKSContext *ctx) : ctx_(ctx) {}
    const CKKSContext *context() const { return ctx_; }

    // Per-op mode: every multi-ciphertext entry point below (hoisted
    // rotations, sums of rotations, multiply-accumulates, linear combinations,
    // block mixes, batched products and rescales) runs as the sequence of
    // single-ciphertext calls the reference's evaluators issue to their
    // library (rotate_vector, multiply_plain, add, multiply, relinearize,
    // rescale_to_next per term), and the lockstep programs built on them
    // (gelu_many, bootstrap_many) run input by input.  Results decrypt alike;
    // this is the plain per-call drop-in the batched engine is measured
    // against.
    void set_per_op(bool on) { per_op_ = on; }
    bool per_op() const { return per_op_; }

    void mod_switch_to_next_inplace(Ciphertext &ct) const;
    void mod_switch_to_inplace(Ciphertext &ct, size_t chain_index) const;
    void mod_switch_to(const Ciphertext &ct, size_t chain_index, Ciphertext &dest) const;
    void mod_switch_to_inplace(Plaintext &pt, size_t chain_index) const;
    void rescale_to_next_inplace(Ciphertext &ct) const;
    void relinearize_inplace(Ciphertext &ct, const RelinKeys &relin_keys) const;

    void square(const Ciphertext &ct, Ciphertext &dest) const;
    void square_inplace(Ciphertext &ct) const;
    void multiply(const Ciphertext &ct1, const Ciphertext &ct2, Ciphertext &dest) const;
    void multiply_inplace(Ciphertext &ct1, const Ciphertext &ct2) const;
    void multiply_plain(const Ciphertext &ct, const Plaintext &plain, Ciphertext &dest) const;
    void multiply_plain_inplace(Ciphertext &ct, const Plaintext &plain) const;

    // sum_r cts[r] * pts[r] in one pass
    void multiply_plain_accumulate(const std::vector<const Ciphertext *> &cts, const std::vector<const Plaintext *> &pts,
                                   Ciphertext &dest) const;
    // dest = sum_k cts[k] * encode(values[k], chain_index, scales[k]) at chain_index,
    // each input read at its first limbs: word for word mod_switch + multiply_plain
    // + add per term, at scale cts[0].scale * scales[0]
    void linear_combination(const std::vector<const Ciphertext *> &cts, const std::vector<double> &values,
                            const std::vector<double> &scales, size_t chain_index, Ciphertext &dest) const;
    // dest[g] = sum_r pts[g K + r] * cts[r] (K = cts.size(), null: no term): the
    // giant steps of one BSGS output at once, each ciphertext read once per 4 rows
    void multiply_plain_accumulate_multi(const std::vector<const Ciphertext *> &cts,
                                         const std::vector<const Plaintext *> &pts,
                                         std::vector<Ciphertext> &dest) const;
    // batched: dest[i] = rescale(cts[i]) / rescale(relinearize(a[i] * b[i])),
    // word for word the per-ciphertext results (dest may alias the inputs);
    // fused: relinearisation and rescale in one rounding (hybrid key
    // switching), each coefficient within 1 of the unfused result
    void rescale_to_next_many(const std::vector<const Ciphertext *> &cts, std::vector<Ciphertext> &dest) const;
    void multiply_relin_rescale_many(const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                                     const RelinKeys &rk, std::vector<Ciphertext> &dest, bool fused = false) const;
    // dest = sum_i a[i] * b[i] (size 3): word for word multiply and add per pair
    void multiply_sum(const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                      Ciphertext &dest) const;
    // out[m] = sum_k w[index[m K + k]] (.) cts[k] (index -1: no term), M = index.size() / K
    void multiply_periodic_mix(const std::vector<const Ciphertext *> &cts, const PeriodicPlaintexts &w,
                               const std::vector<int32_t> &index, std::vector<Ciphertext> &out) const;
    void add_plain(const Ciphertext &ct, const Plaintext &plain, Ciphertext &dest) const;
    void add_plain_inplace(Ciphertext &ct, const Plaintext &plain) const;
    void add(const Ciphertext &ct1, const Ciphertext &ct2, Ciphertext &dest) const;
    void add_inplace(Ciphertext &ct1, const Ciphertext &ct2) const;
    void add_many(const std::vector<Ciphertext> &cts, Ciphertext &dest) const;
    void sub_plain(const Ciphertext &ct, const Plaintext &plain, Ciphertext &dest) const;
    void sub_plain_inplace(Ciphertext &ct, const Plaintext &plain) const;
    void sub(const Ciphertext &ct1, const Ciphertext &ct2, Ciphertext &dest) const;
    void sub_inplace(Ciphertext &ct1, const Ciphertext &ct2) const;
    void negate(const Ciphertext &ct, Ciphertext &dest) const;
    void negate_inplace(Ciphertext &ct) const;

    void rotate_vector(const Ciphertext &ct, int steps, const GaloisKeys &keys, Ciphertext &dest) const;
    void rotate_vector_inplace(Ciphertext &ct, int steps, const GaloisKeys &keys) const;
    void apply_galois(const Ciphertext &ct, uint32_t elt, const GaloisKeys &keys, Ciphertext &dest) const;
    void apply_galois_inplace(Ciphertext &ct, uint32_t elt, const GaloisKeys &keys) const;
    // hoisted rotations (one ModUp shared by all steps; Phantom hoisting_inplace)
    void rotate_vector_hoisted(const Ciphertext &ct, const std::vector<int> &steps, const GaloisKeys &keys,
                               std::vector<Ciphertext> &dest) const;
    void hoisting_inplace(Ciphertext &ct, const GaloisKeys &keys, const std::vector<int> &steps) const;
    // every input rotated by every step, batched across inputs: dest[b][k] = rot(cts[b], steps[k])
    void rotate_vector_hoisted_many(const std::vector<const Ciphertext *> &cts, const std::vector<int> &steps,
                                    const GaloisKeys &keys, std::vector<std::vector<Ciphertext>> &dest) const;
    // dest[g] = sum_i rot(groups[g][i], steps[g][i]), the ModDowns of all groups batched
    void rotate_sum_many(const std::vector<std::vector<const Ciphertext *>> &groups,
                         const std::vector<std::vector<int>> &steps, const GaloisKeys &keys,
                         std::vector<Ciphertext> &dest) const;
    // sum_i rot(cts[i], steps[i]) with a single ModDown (step 0 = the input)
    void rotate_sum(const std::vector<const Ciphertext *> &cts, const std::vector<int> &steps, const GaloisKeys &keys,
                    Ciphertext &dest) const;
    void complex_conjugate(const Ciphertext &ct, const GaloisKeys &keys, Ciphertext &dest) const;
    void complex_conjugate_inplace(Ciphertext &ct, const GaloisKeys &keys) const;

    void transform_to_ntt(const Ciphertext &ct, Ciphertext &dest) const;
    void transform_to_ntt_inplace(Ciphertext &ct) const;
    void transform_from_ntt(const Ciphertext &ct, Ciphertext &dest) const;
    void transform_from_ntt_inplace(Ciphertext &ct) const;
    void multiply_power_of_x(const Ciphertext &ct, int64_t exponent, Ciphertext &dest) const;

    // SEAL-4.1-bs extensions (evaluator.cpp:395-594)
    void multiply_const(const Ciphertext &ct, double value, Ciphertext &dest) const;
    void multiply_const_inplace(Ciphertext &ct, double value) const;
    void add_const(const Ciphertext &ct, double value, Ciphertext &dest) const;
    void add_const_inplace(Ciphertext &ct, double value) const;
    void add_reduced_error(const Ciphertext &ct1, const Ciphertext &ct2, Ciphertext &dest) const;
    void add_inplace_reduced_error(Ciphertext &ct1, const Ciphertext &ct2) const;
    void sub_reduced_error(const Ciphertext &ct1, const Ciphertext &ct2, Ciphertext &dest) const;
    void sub_inplace_reduced_error(Ciphertext &ct1, const Ciphertext &ct2) const;
    void multiply_reduced_error(const Ciphertext &ct1, const Ciphertext &ct2, const RelinKeys &relin_keys,
                                Ciphertext &dest) const;
    void multiply_inplace_reduced_error(Ciphertext &ct1, const Ciphertext &ct2, const RelinKeys &relin_keys) const;
    void double_inplace(Ciphertext &ct) const;
    void multiply_vector_reduced_error(const Ciphertext &ct, const std::vector<double> &values,
                                       Ciphertext &dest) const;
    void multiply_vector_reduced_error(const Ciphertext &ct, const std::vector<std::complex<double>> &values,
                                       Ciphertext &dest) const;

private:
    const CKKSContext *ctx_ = nullptr;
    bool per_op_ = false;
};

// NEXUS's helper layer (src/ckks_evaluator.{h,cpp}): polynomial programs over
// the evaluator.
class CKKSEvaluator {
public:
    CKKSEvaluator(const CKKSContext *context, const SecretKey *sk, RelinKeys *relin_keys, GaloisKeys *galois_keys,
                  double scale, std::vector<uint32_t> galois_elts = {});

    const CKKSContext *context;
    RelinKeys *relin_keys;
    GaloisKeys *galois_keys;
    std::vector<uint32_t> galois_elts;

    Encoder encoder;
    Encryptor encryptor;
    Evaluator evaluator;
    Decryptor decryptor;

    size_t degree;
    double scale;
    size_t slot_count;

    std::vector<double> init_vec_with_value(double value) const { return std::vector<double>(slot_count, value); }
    // vector-encoded constant (FFT path), cached per (value, level, scale)
    const Plaintext &vector_constant(double value, size_t chain_index, double scale);

    Ciphertext sgn_eval(const Ciphertext &x, int d_g, int d_f, double sgn_factor = 0.5);
    Ciphertext invert_sqrt(const Ciphertext &x, int d_newt = 5, int d_gold = 2);
    Ciphertext exp(const Ciphertext &x);
    Ciphertext inverse(const Ciphertext &x, int iter = 4);
    // exp / inverse over inputs of one level and scale in lockstep: the
    // single-input operation sequence, every multiplication one batched call
    std::vector<Ciphertext> exp_many(const std::vector<const Ciphertext *> &x);
    std::vector<Ciphertext> inverse_many(const std::vector<const Ciphertext *> &x, int iter = 4);
    void eval_odd_deg9_poly(const std::vector<double> &a, const Ciphertext &x, Ciphertext &dest);
    // the same steps over a batch of inputs of one level and scale in lockstep
    // (every multiplication with its relinearisation and rescale one batched
    // call); word for word the single-input results
    std::vector<Ciphertext> sgn_eval_many(const std::vector<const Ciphertext *> &x, int d_g, int d_f,
                                          double sgn_factor = 0.5);
    void eval_odd_deg9_poly_many(const std::vector<double> &a, std::vector<Ciphertext> &x);
    double calculate_MAE(const std::vector<double> &y_true, const Ciphertext &ct, size_t n);

private:
    Ciphertext init_guess(const Ciphertext &x);
    Ciphertext eval_line(Ciphertext x, const Plaintext &m, const Plaintext &c);
    Ciphertext newton_iter(Ciphertext x, Ciphertext res, int iter);
    std::pair<Ciphertext, Ciphertext> goldschmidt_iter(Ciphertext v, Ciphertext y, int d);
    uint64_t get_modulus(const Ciphertext &x, int k) const;

    std::map<std::tuple<double, size_t, double>, Plaintext> vconst_;
};

// Evaluators (reference src/{gelu,layer_norm,softmax,argmax}.cpp)
class GELUEvaluator {
public:
    explicit GELUEvaluator(CKKSEvaluator &ckks) : ckks_(&ckks) {}
    void gelu(const Ciphertext &x, Ciphertext &res);
    // gelu of every input (one level and scale) in lockstep; word for word gelu()
    void gelu_many(const std::vector<const Ciphertext *> &x, std::vector<Ciphertext> &res);

private:
    CKKSEvaluator *ckks_;
};

class LNEvaluator {
public:
    explicit LNEvaluator(CKKSEvaluator &ckks) : ckks_(&ckks) {}
    void layer_norm(const Ciphertext &x, Ciphertext &res, int len);

private:
    CKKSEvaluator *ckks_;
};

class SoftmaxEvaluator {
public:
    explicit SoftmaxEvaluator(CKKSEvaluator &ckks) : ckks_(&ckks) {}
    void softmax(const Ciphertext &x, Ciphertext &res, int len);

private:
    CKKSEvaluator *ckks_;
};

}  // namespace nexus
