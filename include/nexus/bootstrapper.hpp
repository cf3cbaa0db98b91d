// CKKS bootstrapping over the C++ drop-in (reference Bootstrapper,
// src/bootstrapping/Bootstrapper.{h,cpp}; cuda/src/bootstrapping/Bootstrapper.cu).
//
// Pipeline (as the reference's bootstrap_sparse_3 / bootstrap_full_3,
// cuda/src/bootstrapping/Bootstrapper.cu:2987-3113):
//   ModRaise -> SubSum (sparse packing) -> CoeffToSlot (3 BSGS levels) ->
//   EvalMod (Chebyshev-basis fit of cos(2 pi (K u - 1/4) / 2^r) on the
//   windows |K u - i| <= 2^-log_width, then r double-angle steps) ->
//   SlotToCoeff (3 BSGS levels).
// The linear transforms are the special-FFT factorisation of the slot map
// (butterfly stages with twiddles exp(2 pi i 5^j / 4 len)) grouped into three
// levels in diagonal form; the bit-reversal permutations of CoeffToSlot and
// SlotToCoeff cancel because EvalMod acts slot-wise.  Every level lands on an
// exactly known scale: plaintext diagonals and constants are encoded at the
// scale that compensates the prime dropped by the following rescale.
#pragma once

#include <complex>
#include <map>
#include <tuple>
#include <vector>

#include "ckks.hpp"

namespace nexus {

// One level of a BSGS linear transform: y = sum_t D_t (.) rot(x, t * stride),
// with t = g * j + i (baby i in [0, g), giant j).  Diagonals are length-N/2
// slot vectors, already pre-rotated by -g*j*stride.
struct LinearLevel {
    int stride = 1;
    int g = 1;
    std::map<std::pair<int, int>, std::vector<std::complex<double>>> diag;  // key (j, i)
};

// Host-side plan of the three CoeffToSlot and three SlotToCoeff levels
// (reference genfftcoeff_3 / geninvfftcoeff_3, Bootstrapper.cu:1137,1545).
struct BootPlan {
    int logn = 0, logNh = 0;
    double K = 25;
    std::vector<LinearLevel> cts, stc;  // in application order
    std::vector<int> rotation_steps;    // sorted, unique; 0 = conjugation
};
BootPlan make_boot_plan(int logNh, int logn, double K);

// Chebyshev coefficients of the EvalMod polynomial and the levels its
// evaluation consumes (host only).
std::vector<double> eval_mod_coeffs(int K, int degree, int double_angles, int log_width);
int chebyshev_depth(const std::vector<double> &c);

class Bootstrapper {
public:
    struct Params {
        int logn = 0;            // log2 of the slot count (logn < logNh: sparse packing)
        int K = 25;              // bound of the ModRaise overflow |I| (reference boundary_K)
        int cheb_degree = 59;    // EvalMod polynomial degree (reference sin_cos_deg; 59 fits depth 6)
        int double_angles = 2;   // r (reference scale_factor = 2)
        int log_width = 10;      // fit windows |K u - i| <= 2^-log_width (reference loge)
        double final_scale = 0;  // output scale (0: the input's scale)
    };

    Bootstrapper(CKKSEvaluator &ckks, const Params &p);

    // Rotation steps whose Galois keys bootstrapping needs (0 = conjugation),
    // as the reference's addLeftRotKeys_Linear_to_vector_3 (Bootstrapper.cu:82-178).
    const std::vector<int> &rotation_steps() const { return plan_.rotation_steps; }
    // Levels bootstrapping consumes from the top data level.
    int depth() const;
    const BootPlan &plan() const { return plan_; }
    const Params &params() const { return p_; }
    const std::vector<double> &cheb_coeffs() const { return cheb_; }

    // The output holds factor x the input message (factor folded into the
    // last SlotToCoeff level: no extra level, error scaled by factor).
    void bootstrap(const Ciphertext &in, Ciphertext &out, double factor = 1.0);
    // bootstrap of several ciphertexts (out[i] = factors[i] x in[i]'s message):
    // bootstrap_mixed with this bootstrapper for every input
    void bootstrap_many(const std::vector<const Ciphertext *> &in, std::vector<Ciphertext> &out,
                        const std::vector<double> &factors);
    // bootstrap of inputs with their own bootstrappers (slot counts) on one
    // context, e.g. dense and sparse packings: out[i] = factors[i] x in[i]'s
    // message through bts[i]'s transforms; inputs at any level (each is
    // brought to q_0 by ModRaise's mod switch); every CoeffToSlot and
    // SlotToCoeff level batched over the inputs of one bootstrapper, one
    // EvalMod over every input's halves (the bootstrappers need the same
    // EvalMod parameters).  Word for word each input's own bootstrap with one
    // special prime.
    static void bootstrap_mixed(const std::vector<Bootstrapper *> &bts, const std::vector<const Ciphertext *> &in,
                                std::vector<Ciphertext> &out, const std::vector<double> &factors);
    // Bootstrapper::set_final_scale (reference Bootstrapper.h)
    void set_final_scale(double s) { p_.final_scale = s; }

    // stages (exposed for tests)
    void mod_raise(const Ciphertext &in, Ciphertext &out) const;
    void sub_sum(Ciphertext &ct) const;
    // sparse: one output u (real, 2n values per period); dense: re and im parts
    void coeff_to_slot(const Ciphertext &in, Ciphertext &u_re, Ciphertext &u_im);
    void eval_mod(const Ciphertext &u, Ciphertext &out);
    // EvalMod of inputs at one level and scale, in lockstep
    void eval_mod_many(const std::vector<const Ciphertext *> &u, std::vector<Ciphertext> &out);
    void slot_to_coeff(const Ciphertext &f_re, const Ciphertext *f_im, double input_scale, Ciphertext &out,
                       double factor = 1.0);

private:
    Ciphertext apply_level(const LinearLevel &lv, int tag, const Ciphertext &x, double out_scale, double value_factor);
    std::vector<Ciphertext> apply_level_many(const LinearLevel &lv, int tag, const std::vector<const Ciphertext *> &xs,
                                             const std::vector<double> &out_scale,
                                             const std::vector<double> &value_factor);
    Ciphertext cheb_eval(const std::vector<double> &c, size_t level, double scale);
    const Ciphertext &basis(int k);
    Ciphertext match(const Ciphertext &ct, size_t level, double scale);
    Ciphertext mul_exact(const Ciphertext &a, const Ciphertext &b);
    // the same steps over a batch in lockstep (one scale and level per batch)
    using Batch = std::vector<Ciphertext>;
    Batch cheb_eval_many(const std::vector<double> &c, size_t level, double scale);
    const Batch &basis_many(int k);
    Batch match_many(const Batch &ct, size_t level, double scale);
    Batch mul_exact_many(const Batch &a, const Batch &b);
    size_t chain_of(size_t level) const { return ck_.context->data_primes() + 1 - level; }
    double prime(size_t level) const { return (double)ck_.context->moduli()[level - 1]; }

    CKKSEvaluator &ck_;
    Params p_;
    BootPlan plan_;
    std::vector<double> cheb_;
    std::map<int, Ciphertext> basis_;  // Chebyshev T_k of the current EvalMod input
    std::map<int, Batch> basis_many_;  // the same for a batch of inputs
    std::map<std::tuple<int, int, int, size_t, double>, Plaintext> ptcache_;
};

// ArgmaxEvaluator (reference src/argmax.{h,cpp}, cuda/src/argmax.cu): log2(len)
// rounds of max(a, b) = (a + b)/2 + (a - b) sgn(a - b)/2 with a bootstrap per
// round, then the one-hot indicator sgn(x - max) + 1.
class ArgmaxEvaluator {
public:
    ArgmaxEvaluator(CKKSEvaluator &ckks, Bootstrapper &bootstrapper) : ckks_(&ckks), bt_(&bootstrapper) {}
    void argmax(Ciphertext &x, Ciphertext &res, int len);
    void bootstrap(Ciphertext &x);

private:
    CKKSEvaluator *ckks_;
    Bootstrapper *bt_;
};

}  // namespace nexus
