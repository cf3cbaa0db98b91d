// Encrypted BERT-base encoder layer over the evaluator programs (GELU,
// softmax, LayerNorm helpers of src/ckks_evaluator.cpp, the Bootstrapper) and
// plaintext-weight linear maps.  The reference ships these pieces but no
// composed layer; the layer is:
//   q, k, v = x Wq / 8, x Wk, x Wv        (12 heads of 64 features)
//   o       = softmax(q_h k_h^T) v_h       per head
//   z       = LN(x + o Wo)
//   out     = LN(z + GELU(z W1) W2)
// with LN(r) = r * invert_sqrt(mean(r^2)) on r * sqrt(30) (the range the
// reference's inverse-sqrt initial guess is tuned for), softmax as
// exp(x) ~ (1 + x/128)^128 and 1/sum from the reference's Newton product on
// sum / s, GELU as src/gelu.cpp.  No biases; LayerNorm gamma = 1, beta = 0.
//
// Packing ("token groups"): a token matrix (tokens x features) is held in
// features / 256 ciphertexts; token i owns slots [256 i, 256 i + 256) of every
// ciphertext and ciphertext c holds features [256 c, 256 c + 256) there, so a
// ciphertext carries 4 heads and slots / 256 tokens (128 at N = 2^16).
// Rotating by 256 t moves every token by t (cyclically over all tokens).
//
// Levels: linear maps run at two limbs and end at one; their outputs are
// bootstrapped (the range factor of each bootstrap is folded into the weights
// and into the last SlotToCoeff level), so attention, LayerNorm and GELU each
// start at the top compute level.  The chain needs 19 compute levels above q0
// besides the bootstrapping levels, e.g. {51, 46 x 19, 51 x 14, 58 x 3}.
#pragma once
#include <chrono>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "bootstrapper.hpp"
#include "ckks.hpp"

namespace nexus {

struct BertShape {
    int hidden = 768;  // multiple of 256
    int heads = 12;    // hidden / heads = 64
    int ffn = 3072;    // multiple of 256
};

class BertLayer {
public:
    static constexpr int kGroup = 256;   // slots per token
    static constexpr int kHeadDim = 64;
    enum Matrix { Wq = 0, Wk, Wv, Wo, W1, W2, kMatrices };

    struct Options {
        int baby = 32;                       // baby steps of the linear maps (power of two)
        // token rotations rot(x, 256 t) and within-head offset rotations rot(x, d)
        // hoisted directly from one ModUp of x (a key per step: about 250 more
        // Galois keys, 50 GB at N = 2^16) instead of baby and giant steps; the
        // layer call measured the same either way (the key switches, not the
        // ModUps, dominate these stages), so the baby-giant form is the default
        bool direct_rotations = false;
        double ln_prescale = 5.477225575051661;  // sqrt(30)
        bool stage_timing = false;           // synchronise and time each stage of forward()
    };

    // Weights are row-major [in][out] (x @ W), copied.  The context's chain
    // must hold 19 compute levels below the bootstrapping levels.
    BertLayer(CKKSEvaluator &ck, Bootstrapper &bt, const BertShape &shape, const std::vector<const double *> &w,
              const Options &opt);
    // A view for another lane (thread): shares the weights and their encoded
    // diagonals (encoded once, under a lock, then read-only) and runs on its
    // own evaluator and bootstrapper.
    BertLayer(const BertLayer &shared, CKKSEvaluator &ck, Bootstrapper &bt);

    // Rotation steps the layer (without the bootstrapper) needs.
    static std::vector<int> rotation_steps(size_t slots, const BertShape &shape, const Options &opt);
    size_t tokens_per_ct() const { return T_; }
    // stage name -> milliseconds of the last forward() (stage_timing only)
    const std::vector<std::pair<std::string, double>> &stage_ms() const { return stage_ms_; }
    void set_stage_timing(bool on) { opt_.stage_timing = on; }

    // A run of tokens forming one sequence: rows [start, start + len).
    struct Segment {
        int start, len;
    };

    // x: hidden / 256 ciphertexts at >= 2 limbs and one scale near the
    // evaluator's; rows >= seq_len must be zero.  Returns the layer output in
    // the same packing at 2 limbs (its scale is tracked exactly, near the
    // evaluator's).
    std::vector<Ciphertext> forward(const std::vector<Ciphertext> &x, int seq_len);
    // Several sequences packed in one set of ciphertexts: attention stays
    // within each segment, so every sequence's rows are what it would get
    // alone.  Rows outside the segments must be zero.
    std::vector<Ciphertext> forward(const std::vector<Ciphertext> &x, const std::vector<Segment> &segments);
    // Several sets in lockstep: every stage runs over all sets before the next
    // one, and each bootstrapping step refreshes the ciphertexts of every set
    // together, so real messages pair up across sets (two per bootstrap) where
    // one set alone leaves an odd one.  Each set's output equals forward() of
    // that set up to the bootstrap pairing (the pairs differ, so the words do).
    std::vector<std::vector<Ciphertext>> forward_sets(const std::vector<std::vector<Ciphertext>> &xs,
                                                      const std::vector<std::vector<Segment>> &segments);

    // Continuous batching: a sequence encrypted at rows [0, len) moved to rows
    // [offset, offset + len) (token rotation, no level), and rows [offset,
    // offset + len) of a packed result moved back to [0, len) with every other
    // row cleared (one level).
    std::vector<Ciphertext> shift_rows(const std::vector<Ciphertext> &x, int offset);
    std::vector<Ciphertext> extract_rows(const std::vector<Ciphertext> &x, int offset, int len);
    const BertShape &shape() const { return shape_; }
    const CKKSContext *context() const { return ck_.context; }

    // stages (exposed for tests)
    // y = factor * x W at one limb and exactly the evaluator's scale (x taken
    // to two limbs first)
    std::vector<Ciphertext> linear(Matrix m, const std::vector<Ciphertext> &x, double factor);
    // per-head softmax(q k^T) v of bootstrapped q (already / 8), k, v
    std::vector<Ciphertext> attention(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                                      const std::vector<Ciphertext> &v, int seq_len);
    std::vector<Ciphertext> attention(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                                      const std::vector<Ciphertext> &v, const std::vector<Segment> &segments);
    // r * invert_sqrt(mean(r^2)) per token (r already pre-scaled)
    std::vector<Ciphertext> layer_norm(const std::vector<Ciphertext> &r);

private:
    // attention up to the probabilities (before their bootstrap), then from
    // the refreshed probabilities to the output
    struct AttnState {
        std::vector<int> seg, tlo, thi, active;
        int halves = 0;
        std::vector<std::vector<Ciphertext>> Pe;  // [ciphertext][half]
    };
    AttnState attention_probs(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                              const std::vector<Segment> &segments);
    std::vector<Ciphertext> attention_values(AttnState &st, const std::vector<Ciphertext> &v);
    // plaintext diagonals of one matrix at one (factor, level, scale): every
    // diagonal is 256-periodic, so the diagonals of output ciphertext o over
    // the inputs of group gi ([gi group, (gi + 1) group)) are one periodic set,
    // entry ((c - gi group) n_a + a - a_min) baby + b, and index[o groups + gi]
    // is its block-mix index (giant step a by (input, baby step); -1: no term)
    struct Encoded {
        int group = 0, groups = 0;
        std::vector<std::unique_ptr<PeriodicPlaintexts>> sets;  // [o * groups + gi]
        std::vector<std::vector<int32_t>> index;
        std::vector<char> giant_any;  // [o * n_a + a - a_min]: the giant step has a term
    };
    struct Weights {
        int n_in = 0, n_out = 0, a_min = 0, a_max = 0;
        std::vector<double> w;  // [in][out], as given
        // at most two encodings are kept; callers hold theirs while in use
        std::vector<std::pair<std::tuple<double, size_t, double>, std::shared_ptr<const Encoded>>> enc;
    };
    struct Shared {
        Weights m[kMatrices];
        std::mutex mu;
    };
    std::shared_ptr<const Encoded> diagonals(Matrix m, double factor, size_t chain_index, double scale);
    Ciphertext rotate_tokens(const Ciphertext &x, int t);
    std::vector<Ciphertext> token_rotations(const Ciphertext &x, int t0, int count);
    // rot(x, 256 t) for every x of xs and t of ts (distinct offsets in [0, T)), batched
    std::vector<std::vector<Ciphertext>> token_rotations_many(const std::vector<const Ciphertext *> &xs,
                                                              const std::vector<int> &ts);
    std::vector<Ciphertext> offset_rotations(const Ciphertext &x, int dmin, int dmax);
    std::vector<std::vector<Ciphertext>> offset_rotations_many(const std::vector<const Ciphertext *> &xs, int dmin,
                                                               int dmax);
    // the 64 one-hot masks A_p (1 at slots = p mod 64), period 64, as periodic plaintexts
    const PeriodicPlaintexts &offset_masks(size_t limbs, double scale);
    // per-token segment index (-1 outside every segment) and its 1 / length
    std::vector<int> segment_of(const std::vector<Segment> &segments) const;
    const Plaintext &key_mask(int eta, int tc, const std::vector<int> &seg, size_t limbs);
    const Plaintext &row_mask(int tag, const std::vector<int> &seg, const std::vector<Segment> &segments, size_t limbs,
                              double scale, int period);
    // bootstrap real messages in place, two per bootstrap, times factor
    void bootstrap_real(const std::vector<Ciphertext *> &cts, double factor);
    Ciphertext scaled(const Ciphertext &x, double value);  // x * value at the same scale, one level
    size_t chain_of(size_t limbs) const { return ck_.context->data_primes() + 1 - limbs; }
    uint64_t prime(size_t limbs) const { return ck_.context->moduli()[limbs - 1]; }
    // stage_timing: synchronise and add the time since the last mark to `name`
    void mark(const char *name);

    CKKSEvaluator &ck_;
    Bootstrapper &bt_;
    BertShape shape_;
    Options opt_;
    size_t slots_, T_;
    std::shared_ptr<Shared> shared_;
    std::map<std::tuple<int, size_t, double, std::vector<int>>, Plaintext> masks_;
    std::map<std::pair<size_t, double>, std::unique_ptr<PeriodicPlaintexts>> offset_masks_;
    std::vector<std::pair<std::string, double>> stage_ms_;
    std::chrono::steady_clock::time_point t_mark_;
};

}  // namespace nexus
