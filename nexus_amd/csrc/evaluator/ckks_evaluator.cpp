// NEXUS's polynomial helper programs (reference src/ckks_evaluator.cpp) and the
// GELU / LayerNorm / Softmax evaluators (src/gelu.cpp, src/layer_norm.cpp,
// src/softmax.cpp), restated over the C++ drop-in.  The sequence of
// operations, the constants and the scale bookkeeping follow the SEAL path of
// the reference line by line, so that the ciphertext residues are the ones
// SEAL's evaluator would produce from the same inputs and keys.
#include <cmath>
#include <tuple>

#include "../../../include/nexus/ckks.hpp"

namespace nexus {

namespace {
// sign-function composite coefficients (src/ckks_evaluator.h:22-27)
const double F4[10] = {0, 315, 0, -420, 0, 378, 0, -180, 0, 35};
const double G4[10] = {0, 5850, 0, -34974, 0, 97015, 0, -113492, 0, 46623};
const int F4_SCALE = 1 << 7, G4_SCALE = 1 << 10;
}  // namespace

CKKSEvaluator::CKKSEvaluator(const CKKSContext *ctx, const SecretKey *sk, RelinKeys *rk, GaloisKeys *gk,
                             double scale_, std::vector<uint32_t> elts)
    : context(ctx), relin_keys(rk), galois_keys(gk), galois_elts(std::move(elts)), encoder(ctx), encryptor(ctx, sk),
      evaluator(ctx), decryptor(ctx, sk), scale(scale_)
{
    slot_count = ctx->poly_degree() / 2;
    degree = ctx->poly_degree();
}

const Plaintext &CKKSEvaluator::vector_constant(double value, size_t chain_index, double sc)
{
    auto key = std::make_tuple(value, chain_index, sc);
    auto it = vconst_.find(key);
    if (it != vconst_.end()) return it->second;
    Plaintext p;
    encoder.encode(init_vec_with_value(value), chain_index, sc, p);
    return vconst_.emplace(key, std::move(p)).first->second;
}

uint64_t CKKSEvaluator::get_modulus(const Ciphertext &x, int k) const
{
    // src/ckks_evaluator.cpp:511-515: modulus[size - k] of x's level
    const size_t sz = x.coeff_modulus_size();
    return context->moduli()[sz - k];
}

// src/ckks_evaluator.cpp:181-212
Ciphertext CKKSEvaluator::sgn_eval(const Ciphertext &x, int d_g, int d_f, double sgn_factor)
{
    std::vector<double> f4(10), g4(10), f4_last(10), g4_last(10);
    for (int i = 0; i <= 9; i++) {
        f4[i] = F4[i] / F4_SCALE;
        f4_last[i] = f4[i] * sgn_factor;
        g4[i] = G4[i] / G4_SCALE;
        g4_last[i] = g4[i] * sgn_factor;
    }
    Ciphertext dest = x;
    for (int i = 0; i < d_g; i++) eval_odd_deg9_poly(i == d_g - 1 ? g4_last : g4, dest, dest);
    for (int i = 0; i < d_f; i++) eval_odd_deg9_poly(i == d_f - 1 ? f4_last : f4, dest, dest);
    return dest;
}

// src/ckks_evaluator.cpp:214-364: errorless odd degree-9 evaluation
// P(x) = (a3 + a5 x^2) x^3 + (a7 x + a9 x^3) x^6 + a1 x, depth 4.
void CKKSEvaluator::eval_odd_deg9_poly(const std::vector<double> &a, const Ciphertext &x_in, Ciphertext &dest)
{
    const Evaluator &ev = evaluator;
    Ciphertext x = x_in;
    const double D = x.scale();
    uint64_t p = get_modulus(x, 2), q = get_modulus(x, 3), r = get_modulus(x, 4);

    Ciphertext x2, x3, x6;
    ev.square(x, x2);
    ev.relinearize_inplace(x2, *relin_keys);
    ev.rescale_to_next_inplace(x2);
    ev.mod_switch_to_next_inplace(x);
    ev.multiply(x2, x, x3);
    ev.relinearize_inplace(x3, *relin_keys);
    ev.rescale_to_next_inplace(x3);
    ev.square(x3, x6);
    ev.relinearize_inplace(x6, *relin_keys);
    ev.rescale_to_next_inplace(x6);

    Plaintext a1, a3, a5, a7, a9;
    // T1
    Ciphertext T1;
    const double a5_scale = D / x2.scale() * p / x3.scale() * q;
    encoder.encode(a[5], x2.params_id(), a5_scale, a5);
    ev.multiply_plain(x2, a5, T1);
    ev.rescale_to_next_inplace(T1);
    encoder.encode(a[3], T1.params_id(), T1.scale(), a3);
    ev.add_plain_inplace(T1, a3);
    ev.multiply_inplace(T1, x3);
    ev.relinearize_inplace(T1, *relin_keys);
    ev.rescale_to_next_inplace(T1);
    // T2
    Ciphertext T2;
    const double a9_scale = D / x3.scale() * r / x6.scale() * q;
    encoder.encode(a[9], x3.params_id(), a9_scale, a9);
    ev.multiply_plain(x3, a9, T2);
    ev.rescale_to_next_inplace(T2);
    Ciphertext a7x;
    const double a7_scale = T2.scale() / x.scale() * p;
    encoder.encode(a[7], x.params_id(), a7_scale, a7);
    ev.multiply_plain(x, a7, a7x);
    ev.rescale_to_next_inplace(a7x);
    const double mid_scale = (T2.scale() + a7x.scale()) / 2;
    T2.set_scale(mid_scale);
    a7x.set_scale(mid_scale);
    {
        // T2 + a7x at T2's level: a7x read in place (mod switch + add in one pass)
        Ciphertext sum;
        ev.linear_combination({&T2, &a7x}, {1.0, 1.0}, {1.0, 1.0}, T2.params_id(), sum);
        T2 = std::move(sum);
    }
    ev.multiply_inplace(T2, x6);
    ev.relinearize_inplace(T2, *relin_keys);
    ev.rescale_to_next_inplace(T2);
    // T3
    Ciphertext T3;
    encoder.encode(a[1], x.params_id(), (double)p, a1);
    ev.multiply_plain(x, a1, T3);
    ev.rescale_to_next_inplace(T3);

    const double mid3 = (T1.scale() + T2.scale() + T3.scale()) / 3;
    T1.set_scale(mid3);
    T2.set_scale(mid3);
    T3.set_scale(mid3);
    // T2 + T1 + T3 at T2's level: T1 and T3 read in place
    Ciphertext out;
    ev.linear_combination({&T2, &T1, &T3}, {1.0, 1.0, 1.0}, {1.0, 1.0, 1.0}, T2.params_id(), out);
    dest = std::move(out);
}

// ---- lockstep batches of the evaluators above
namespace {

std::vector<const Ciphertext *> ptrs_of(const std::vector<Ciphertext> &v)
{
    std::vector<const Ciphertext *> p;
    for (const auto &x : v) p.push_back(&x);
    return p;
}

void rescale_all(const Evaluator &ev, std::vector<Ciphertext> &v) { ev.rescale_to_next_many(ptrs_of(v), v); }

}  // namespace

void CKKSEvaluator::eval_odd_deg9_poly_many(const std::vector<double> &a, std::vector<Ciphertext> &X)
{
    const Evaluator &ev = evaluator;
    const size_t B = X.size();
    const double D = X[0].scale();
    const uint64_t p = get_modulus(X[0], 2), q = get_modulus(X[0], 3), r = get_modulus(X[0], 4);
    std::vector<Ciphertext> x2, x3, x6, xs(B);
    ev.multiply_relin_rescale_many(ptrs_of(X), ptrs_of(X), *relin_keys, x2, true);
    for (size_t i = 0; i < B; i++) ev.mod_switch_to(X[i], X[i].params_id() + 1, xs[i]);
    ev.multiply_relin_rescale_many(ptrs_of(x2), ptrs_of(xs), *relin_keys, x3, true);
    ev.multiply_relin_rescale_many(ptrs_of(x3), ptrs_of(x3), *relin_keys, x6, true);
    Plaintext a1, a3, a5, a7, a9;
    // T1
    std::vector<Ciphertext> T1(B);
    encoder.encode(a[5], x2[0].params_id(), D / x2[0].scale() * p / x3[0].scale() * q, a5);
    for (size_t i = 0; i < B; i++) ev.multiply_plain(x2[i], a5, T1[i]);
    rescale_all(ev, T1);
    encoder.encode(a[3], T1[0].params_id(), T1[0].scale(), a3);
    for (auto &t : T1) ev.add_plain_inplace(t, a3);
    ev.multiply_relin_rescale_many(ptrs_of(T1), ptrs_of(x3), *relin_keys, T1, true);
    // T2
    std::vector<Ciphertext> T2(B), a7x(B);
    encoder.encode(a[9], x3[0].params_id(), D / x3[0].scale() * r / x6[0].scale() * q, a9);
    for (size_t i = 0; i < B; i++) ev.multiply_plain(x3[i], a9, T2[i]);
    rescale_all(ev, T2);
    encoder.encode(a[7], xs[0].params_id(), T2[0].scale() / xs[0].scale() * p, a7);
    for (size_t i = 0; i < B; i++) ev.multiply_plain(xs[i], a7, a7x[i]);
    rescale_all(ev, a7x);
    for (size_t i = 0; i < B; i++) {
        // T2 + a7x at T2's level and the mean scale: a7x read in place
        const double mid = (T2[i].scale() + a7x[i].scale()) / 2;
        T2[i].set_scale(mid);
        a7x[i].set_scale(mid);
        Ciphertext sum;
        ev.linear_combination({&T2[i], &a7x[i]}, {1.0, 1.0}, {1.0, 1.0}, T2[i].params_id(), sum);
        T2[i] = std::move(sum);
    }
    ev.multiply_relin_rescale_many(ptrs_of(T2), ptrs_of(x6), *relin_keys, T2, true);
    // T3
    std::vector<Ciphertext> T3(B);
    encoder.encode(a[1], xs[0].params_id(), (double)p, a1);
    for (size_t i = 0; i < B; i++) ev.multiply_plain(xs[i], a1, T3[i]);
    rescale_all(ev, T3);
    for (size_t i = 0; i < B; i++) {
        // T2 + T1 + T3 at T2's level and the mean scale: T1, T3 read in place
        const double mid3 = (T1[i].scale() + T2[i].scale() + T3[i].scale()) / 3;
        T1[i].set_scale(mid3);
        T2[i].set_scale(mid3);
        T3[i].set_scale(mid3);
        Ciphertext sum;
        ev.linear_combination({&T2[i], &T1[i], &T3[i]}, {1.0, 1.0, 1.0}, {1.0, 1.0, 1.0}, T2[i].params_id(), sum);
        T2[i] = std::move(sum);
    }
    X = std::move(T2);
}

std::vector<Ciphertext> CKKSEvaluator::sgn_eval_many(const std::vector<const Ciphertext *> &x, int d_g, int d_f,
                                                     double sgn_factor)
{
    std::vector<double> f4(10), g4(10), f4_last(10), g4_last(10);
    for (int i = 0; i <= 9; i++) {
        f4[i] = F4[i] / F4_SCALE;
        f4_last[i] = f4[i] * sgn_factor;
        g4[i] = G4[i] / G4_SCALE;
        g4_last[i] = g4[i] * sgn_factor;
    }
    std::vector<Ciphertext> dest;
    for (const Ciphertext *c : x) dest.push_back(*c);
    for (int i = 0; i < d_g; i++) eval_odd_deg9_poly_many(i == d_g - 1 ? g4_last : g4, dest);
    for (int i = 0; i < d_f; i++) eval_odd_deg9_poly_many(i == d_f - 1 ? f4_last : f4, dest);
    return dest;
}

void GELUEvaluator::gelu_many(const std::vector<const Ciphertext *> &xin, std::vector<Ciphertext> &res)
{
    CKKSEvaluator &ck = *ckks_;
    const Evaluator &ev = ck.evaluator;
    const size_t B = xin.size();
    if (!B) throw std::invalid_argument("gelu_many: no inputs");
    if (ev.per_op()) {
        // per-op mode: the reference's GELU input by input
        std::vector<Ciphertext> out(B);
        for (size_t i = 0; i < B; i++) gelu(*xin[i], out[i]);
        res = std::move(out);
        return;
    }
    for (const Ciphertext *c : xin)
        if (c->params_id() != xin[0]->params_id() || c->scale() != xin[0]->scale())
            throw std::invalid_argument("gelu_many: inputs must share the level and scale");
    const Ciphertext &x0 = *xin[0];
    const Plaintext &p0 = ck.vector_constant(-3.5, x0.params_id(), x0.scale());
    const Plaintext &p1 = ck.vector_constant(3.5, x0.params_id(), x0.scale());
    const Plaintext &delta = ck.vector_constant(1.0 / 8.5, x0.params_id(), x0.scale());
    // both sign arguments of every input in one batch: b0 = [0, B), b1 = [B, 2B)
    std::vector<Ciphertext> b(2 * B);
    for (size_t i = 0; i < B; i++) {
        ev.sub_plain(*xin[i], p0, b[i]);
        ev.multiply_plain_inplace(b[i], delta);
        ev.sub_plain(*xin[i], p1, b[B + i]);
        ev.multiply_plain_inplace(b[B + i], delta);
    }
    rescale_all(ev, b);
    b = ck.sgn_eval_many(ptrs_of(b), 2, 2);
    const Plaintext &half = ck.vector_constant(0.5, b[B].params_id(), b[B].scale());
    std::vector<Ciphertext> a1(B), a2(B);
    for (size_t i = 0; i < B; i++) {
        ev.sub(b[i], b[B + i], a1[i]);
        ev.add_plain(b[B + i], half, a2[i]);
    }
    // powers: x2, x4, then (x6 = x2 x4, x8 = x4^2), then (x10 = x2 x8, x12 = x6^2),
    // x2 switched down to each product's level as gelu() does
    std::vector<Ciphertext> x2, x4, p68, p1012;
    ev.multiply_relin_rescale_many(xin, xin, *ck.relin_keys, x2, true);
    ev.multiply_relin_rescale_many(ptrs_of(x2), ptrs_of(x2), *ck.relin_keys, x4, true);
    for (auto &c : x2) ev.mod_switch_to_inplace(c, x4[0].params_id());
    {
        std::vector<const Ciphertext *> l, r;
        for (size_t i = 0; i < B; i++) l.push_back(&x2[i]), r.push_back(&x4[i]);
        for (size_t i = 0; i < B; i++) l.push_back(&x4[i]), r.push_back(&x4[i]);
        ev.multiply_relin_rescale_many(l, r, *ck.relin_keys, p68, true);  // x6 = [0, B), x8 = [B, 2B)
    }
    for (auto &c : x2) ev.mod_switch_to_inplace(c, p68[B].params_id());
    {
        std::vector<const Ciphertext *> l, r;
        for (size_t i = 0; i < B; i++) l.push_back(&x2[i]), r.push_back(&p68[B + i]);
        for (size_t i = 0; i < B; i++) l.push_back(&p68[i]), r.push_back(&p68[i]);
        ev.multiply_relin_rescale_many(l, r, *ck.relin_keys, p1012, true);  // x10, x12
    }
    const double A[] = {2.25775755e-04, 0.5, 3.96880960e-01, -6.37042698e-02, 8.38841647e-03, -7.17830961e-04,
                        3.49617829e-05, -7.26059653e-07};
    std::vector<Plaintext> coeff_A(1);
    ck.encoder.encode(A[0], ck.scale, coeff_A[0]);
    // term j of input i: A_j x^(2j - 2) for j >= 2, A_1 x for j = 1; each the
    // power times the constant A_j encoded at the evaluator's scale, read in place
    std::vector<std::vector<Ciphertext>> terms(8, std::vector<Ciphertext>(B));
    for (size_t j = 1; j < 8; j++) {
        for (size_t i = 0; i < B; i++) {
            const Ciphertext *pw[8] = {nullptr, xin[i], &x2[i], &x4[i], &p68[i], &p68[B + i], &p1012[i],
                                       &p1012[B + i]};
            ev.linear_combination({pw[j]}, {A[j]}, {ck.scale}, pw[j]->params_id(), terms[j][i]);
        }
        rescale_all(ev, terms[j]);
        for (auto &t : terms[j]) t.set_scale(ck.scale);
    }
    // Ax = sum of the seven terms at the level of the last, each read in place
    std::vector<Ciphertext> Ax(B);
    for (size_t i = 0; i < B; i++) {
        std::vector<const Ciphertext *> xs{&terms[7][i]};
        for (size_t j = 1; j < 7; j++) xs.push_back(&terms[j][i]);
        ev.linear_combination(xs, std::vector<double>(xs.size(), 1.0), std::vector<double>(xs.size(), 1.0),
                              terms[7][i].params_id(), Ax[i]);
    }
    ev.mod_switch_to_inplace(coeff_A[0], Ax[0].params_id());
    for (auto &c : Ax) ev.add_plain_inplace(c, coeff_A[0]);
    std::vector<Ciphertext> xs(B), s;
    std::vector<const Ciphertext *> l, r;
    for (size_t i = 0; i < B; i++) {
        ev.mod_switch_to_inplace(Ax[i], a1[i].params_id());
        ev.mod_switch_to(*xin[i], a2[i].params_id(), xs[i]);
        l.push_back(&Ax[i]);
        r.push_back(&a1[i]);
    }
    for (size_t i = 0; i < B; i++) l.push_back(&xs[i]), r.push_back(&a2[i]);
    ev.multiply_relin_rescale_many(l, r, *ck.relin_keys, s, true);  // s1 = [0, B), s2 = [B, 2B)
    std::vector<Ciphertext> out(B);
    for (size_t i = 0; i < B; i++) {
        s[i].set_scale(ck.scale);
        s[B + i].set_scale(ck.scale);
        ev.mod_switch_to_inplace(s[B + i], s[i].params_id());
        ev.add(s[i], s[B + i], out[i]);
    }
    res = std::move(out);
}

// src/ckks_evaluator.cpp:366-422
Ciphertext CKKSEvaluator::newton_iter(Ciphertext x, Ciphertext res, int iter)
{
    const Evaluator &ev = evaluator;
    for (int i = 0; i < iter; i++) {
        Plaintext three_half, neg_half;
        encoder.encode(1.5, scale, three_half);
        encoder.encode(-0.5, scale, neg_half);
        Ciphertext res_sq;
        ev.square(res, res_sq);
        ev.relinearize_inplace(res_sq, *relin_keys);
        ev.rescale_to_next_inplace(res_sq);
        Ciphertext res_x;
        ev.mod_switch_to_inplace(neg_half, x.params_id());
        ev.multiply_plain(x, neg_half, res_x);
        ev.rescale_to_next_inplace(res_x);
        // bring both to the lower of the two levels
        if (res.coeff_modulus_size() < res_x.coeff_modulus_size())
            ev.mod_switch_to_inplace(res_x, res.params_id());
        else
            ev.mod_switch_to_inplace(res, res_x.params_id());
        ev.multiply_inplace(res_x, res);
        ev.relinearize_inplace(res_x, *relin_keys);
        ev.rescale_to_next_inplace(res_x);
        ev.mod_switch_to_inplace(res_sq, res_x.params_id());
        ev.multiply_inplace(res_x, res_sq);
        ev.relinearize_inplace(res_x, *relin_keys);
        ev.rescale_to_next_inplace(res_x);
        ev.mod_switch_to_inplace(three_half, res.params_id());
        ev.multiply_plain_inplace(res, three_half);
        ev.rescale_to_next_inplace(res);
        ev.mod_switch_to_inplace(res, res_x.params_id());
        res_x.set_scale(scale);
        res.set_scale(scale);
        ev.add_inplace(res, res_x);
    }
    return res;
}

// src/ckks_evaluator.cpp:424-478
std::pair<Ciphertext, Ciphertext> CKKSEvaluator::goldschmidt_iter(Ciphertext v, Ciphertext y, int d)
{
    const Evaluator &ev = evaluator;
    Ciphertext x, h, r, temp;
    Plaintext constant;
    encoder.encode(0.5, scale, constant);
    ev.mod_switch_to_inplace(v, y.params_id());
    ev.multiply(v, y, x);
    ev.relinearize_inplace(x, *relin_keys);
    ev.rescale_to_next_inplace(x);
    ev.mod_switch_to_inplace(constant, y.params_id());
    ev.multiply_plain(y, constant, h);
    ev.rescale_to_next_inplace(h);
    for (int i = 0; i < d; i++) {
        encoder.encode(0.5, scale, constant);
        ev.multiply(x, h, r);
        ev.relinearize_inplace(r, *relin_keys);
        ev.rescale_to_next_inplace(r);
        r.set_scale(scale);
        ev.negate(r, temp);
        ev.mod_switch_to_inplace(constant, temp.params_id());
        ev.add_plain(temp, constant, r);
        ev.mod_switch_to_inplace(x, r.params_id());
        ev.multiply(x, r, temp);
        ev.relinearize_inplace(temp, *relin_keys);
        ev.rescale_to_next_inplace(temp);
        x.set_scale(scale);
        temp.set_scale(scale);
        ev.mod_switch_to_inplace(x, temp.params_id());
        ev.add_inplace(x, temp);
        ev.mod_switch_to_inplace(h, r.params_id());
        ev.multiply(h, r, temp);
        ev.relinearize_inplace(temp, *relin_keys);
        ev.rescale_to_next_inplace(temp);
        h.set_scale(scale);
        temp.set_scale(scale);
        ev.mod_switch_to_inplace(h, temp.params_id());
        ev.add_inplace(h, temp);
    }
    encoder.encode(2.0, scale, constant);
    ev.mod_switch_to_inplace(constant, h.params_id());
    ev.multiply_plain_inplace(h, constant);
    ev.rescale_to_next_inplace(h);
    return {x, h};
}

// src/ckks_evaluator.cpp:517-532
Ciphertext CKKSEvaluator::eval_line(Ciphertext x, const Plaintext &m_in, const Plaintext &c_in)
{
    const Evaluator &ev = evaluator;
    Plaintext m = m_in, c = c_in;
    ev.mod_switch_to_inplace(m, x.params_id());
    ev.multiply_plain_inplace(x, m);
    ev.rescale_to_next_inplace(x);
    ev.mod_switch_to_inplace(c, x.params_id());
    x.set_scale(scale);
    ev.add_plain_inplace(x, c);
    return x;
}

Ciphertext CKKSEvaluator::init_guess(const Ciphertext &x)
{
    Plaintext A, B;
    encoder.encode(-1.29054537e-04, scale, A);
    encoder.encode(1.29054537e-01, scale, B);
    return eval_line(x, A, B);
}

// src/ckks_evaluator.cpp:480-485
Ciphertext CKKSEvaluator::invert_sqrt(const Ciphertext &x, int d_newt, int d_gold)
{
    Ciphertext res = init_guess(x);
    Ciphertext y = newton_iter(x, res, d_newt);
    return goldschmidt_iter(x, y, d_gold).second;
}

// src/ckks_evaluator.cpp:487-509
Ciphertext CKKSEvaluator::inverse(const Ciphertext &x, int iter)
{
    const Evaluator &ev = evaluator;
    Ciphertext y, tmp, res;
    Plaintext one;
    encoder.encode(1.0, x.params_id(), x.scale(), one);
    ev.sub_plain(x, one, y);
    ev.negate_inplace(y);
    ev.add_plain(y, one, tmp);
    res = tmp;
    for (int i = 0; i < iter; i++) {
        ev.square_inplace(y);
        ev.relinearize_inplace(y, *relin_keys);
        ev.rescale_to_next_inplace(y);
        encoder.encode(1.0, y.params_id(), y.scale(), one);
        ev.add_plain(y, one, tmp);
        ev.mod_switch_to_inplace(res, tmp.params_id());
        ev.multiply_inplace(res, tmp);
        ev.relinearize_inplace(res, *relin_keys);
        ev.rescale_to_next_inplace(res);
    }
    return res;
}

// src/ckks_evaluator.cpp:534-548: exp(x) ~ (1 + x/128)^128
Ciphertext CKKSEvaluator::exp(const Ciphertext &x_in)
{
    const Evaluator &ev = evaluator;
    Ciphertext x = x_in;
    Plaintext one, inv128;
    encoder.encode(0.0078125, x.params_id(), x.scale(), inv128);
    ev.multiply_plain_inplace(x, inv128);
    ev.rescale_to_next_inplace(x);
    encoder.encode(1.0, x.params_id(), x.scale(), one);
    ev.add_plain_inplace(x, one);
    for (int i = 0; i < 7; i++) {
        ev.square_inplace(x);
        ev.relinearize_inplace(x, *relin_keys);
        ev.rescale_to_next_inplace(x);
    }
    return x;
}

// exp() over inputs of one level and scale in lockstep: exp()'s operation
// sequence, every squaring of all inputs one batched multiply-relinearise-rescale
std::vector<Ciphertext> CKKSEvaluator::exp_many(const std::vector<const Ciphertext *> &xs)
{
    const Evaluator &ev = evaluator;
    std::vector<Ciphertext> x(xs.size());
    if (xs.empty()) return x;
    for (size_t k = 0; k < xs.size(); k++) {
        if (xs[k]->params_id() != xs[0]->params_id() || xs[k]->scale() != xs[0]->scale())
            throw std::invalid_argument("exp_many: inputs must share the level and scale");
        x[k] = *xs[k];
    }
    Plaintext one, inv128;
    encoder.encode(0.0078125, x[0].params_id(), x[0].scale(), inv128);
    for (auto &c : x) ev.multiply_plain_inplace(c, inv128);
    std::vector<const Ciphertext *> p;
    for (auto &c : x) p.push_back(&c);
    std::vector<Ciphertext> y;
    ev.rescale_to_next_many(p, y);
    encoder.encode(1.0, y[0].params_id(), y[0].scale(), one);
    for (auto &c : y) ev.add_plain_inplace(c, one);
    for (int i = 0; i < 7; i++) {
        std::vector<const Ciphertext *> q;
        for (auto &c : y) q.push_back(&c);
        std::vector<Ciphertext> z;
        ev.multiply_relin_rescale_many(q, q, *relin_keys, z);
        y = std::move(z);
    }
    return y;
}

// inverse() over inputs of one level and scale in lockstep: inverse()'s
// operation sequence, every multiplication one batched call
std::vector<Ciphertext> CKKSEvaluator::inverse_many(const std::vector<const Ciphertext *> &xs, int iter)
{
    const Evaluator &ev = evaluator;
    const size_t K = xs.size();
    std::vector<Ciphertext> y(K), res(K);
    if (!K) return res;
    for (size_t k = 0; k < K; k++)
        if (xs[k]->params_id() != xs[0]->params_id() || xs[k]->scale() != xs[0]->scale())
            throw std::invalid_argument("inverse_many: inputs must share the level and scale");
    Plaintext one;
    encoder.encode(1.0, xs[0]->params_id(), xs[0]->scale(), one);
    for (size_t k = 0; k < K; k++) {
        ev.sub_plain(*xs[k], one, y[k]);
        ev.negate_inplace(y[k]);
        ev.add_plain(y[k], one, res[k]);
    }
    for (int i = 0; i < iter; i++) {
        std::vector<const Ciphertext *> yp;
        for (auto &c : y) yp.push_back(&c);
        std::vector<Ciphertext> y2;
        ev.multiply_relin_rescale_many(yp, yp, *relin_keys, y2);
        y = std::move(y2);
        encoder.encode(1.0, y[0].params_id(), y[0].scale(), one);
        std::vector<Ciphertext> tmp(K);
        for (size_t k = 0; k < K; k++) {
            ev.add_plain(y[k], one, tmp[k]);
            ev.mod_switch_to_inplace(res[k], tmp[k].params_id());
        }
        std::vector<const Ciphertext *> rp, tp;
        for (size_t k = 0; k < K; k++) {
            rp.push_back(&res[k]);
            tp.push_back(&tmp[k]);
        }
        std::vector<Ciphertext> r2;
        ev.multiply_relin_rescale_many(rp, tp, *relin_keys, r2);
        res = std::move(r2);
    }
    return res;
}

double CKKSEvaluator::calculate_MAE(const std::vector<double> &y_true, const Ciphertext &ct, size_t n)
{
    Plaintext p;
    decryptor.decrypt(ct, p);
    std::vector<double> y;
    encoder.decode(p, y);
    double s = 0;
    for (size_t i = 0; i < n; i++) s += std::fabs(y_true[i] - y[i]);
    return s / n;
}

// ---------------------------------------------------------------- GELU
// src/gelu.cpp:5-111:  GELU(x) = A(x) (b0 - b1) + x (b1 + 1/2),
// b0 = sgn((x + 3.5) / 8.5) / 2, b1 = sgn((x - 3.5) / 8.5) / 2, A a degree-12
// even polynomial plus x/2.
void GELUEvaluator::gelu(const Ciphertext &x_in, Ciphertext &res)
{
    CKKSEvaluator &ck = *ckks_;
    const Evaluator &ev = ck.evaluator;
    Ciphertext x = x_in;
    Ciphertext b0, b1;
    const Plaintext &p0 = ck.vector_constant(-3.5, x.params_id(), x.scale());
    const Plaintext &p1 = ck.vector_constant(3.5, x.params_id(), x.scale());
    const Plaintext &delta = ck.vector_constant(1.0 / 8.5, x.params_id(), x.scale());
    ev.sub_plain(x, p0, b0);
    ev.multiply_plain_inplace(b0, delta);
    ev.rescale_to_next_inplace(b0);
    ev.sub_plain(x, p1, b1);
    ev.multiply_plain_inplace(b1, delta);
    ev.rescale_to_next_inplace(b1);
    b0 = ck.sgn_eval(b0, 2, 2);
    b1 = ck.sgn_eval(b1, 2, 2);
    const Plaintext &half = ck.vector_constant(0.5, b1.params_id(), b1.scale());
    Ciphertext a1, a2;
    ev.sub(b0, b1, a1);
    ev.add_plain(b1, half, a2);

    Ciphertext x_2, x_4, x_6, x_8, x_10, x_12;
    ev.square(x, x_2);
    ev.relinearize_inplace(x_2, *ck.relin_keys);
    ev.rescale_to_next_inplace(x_2);
    ev.square(x_2, x_4);
    ev.relinearize_inplace(x_4, *ck.relin_keys);
    ev.rescale_to_next_inplace(x_4);
    ev.mod_switch_to_inplace(x_2, x_4.params_id());
    ev.multiply(x_2, x_4, x_6);
    ev.relinearize_inplace(x_6, *ck.relin_keys);
    ev.rescale_to_next_inplace(x_6);
    ev.square(x_4, x_8);
    ev.relinearize_inplace(x_8, *ck.relin_keys);
    ev.rescale_to_next_inplace(x_8);
    ev.mod_switch_to_inplace(x_2, x_8.params_id());
    ev.multiply(x_2, x_8, x_10);
    ev.relinearize_inplace(x_10, *ck.relin_keys);
    ev.rescale_to_next_inplace(x_10);
    ev.square(x_6, x_12);
    ev.relinearize_inplace(x_12, *ck.relin_keys);
    ev.rescale_to_next_inplace(x_12);

    const double A[] = {2.25775755e-04, 0.5, 3.96880960e-01, -6.37042698e-02, 8.38841647e-03, -7.17830961e-04,
                        3.49617829e-05, -7.26059653e-07};
    std::vector<Plaintext> coeff_A(8);
    for (size_t i = 0; i < 8; i++) ck.encoder.encode(A[i], ck.scale, coeff_A[i]);
    std::vector<Ciphertext> cts(8);
    cts[1] = x;
    cts[2] = x_2;
    cts[3] = x_4;
    cts[4] = x_6;
    cts[5] = x_8;
    cts[6] = x_10;
    cts[7] = x_12;
    for (size_t i = 1; i < 8; i++) {
        ev.mod_switch_to_inplace(coeff_A[i], cts[i].params_id());
        ev.multiply_plain_inplace(cts[i], coeff_A[i]);
        ev.rescale_to_next_inplace(cts[i]);
        cts[i].set_scale(ck.scale);
    }
    Ciphertext Ax = cts[7];
    for (size_t i = 1; i < 7; i++) {
        ev.mod_switch_to_inplace(cts[i], Ax.params_id());
        ev.add_inplace(Ax, cts[i]);
    }
    ev.mod_switch_to_inplace(coeff_A[0], Ax.params_id());
    ev.add_plain_inplace(Ax, coeff_A[0]);

    Ciphertext s1, s2;
    ev.mod_switch_to_inplace(Ax, a1.params_id());
    ev.multiply(Ax, a1, s1);
    ev.relinearize_inplace(s1, *ck.relin_keys);
    ev.rescale_to_next_inplace(s1);
    ev.mod_switch_to_inplace(x, a2.params_id());
    ev.multiply(x, a2, s2);
    ev.relinearize_inplace(s2, *ck.relin_keys);
    ev.rescale_to_next_inplace(s2);
    s1.set_scale(ck.scale);
    s2.set_scale(ck.scale);
    ev.mod_switch_to_inplace(s2, s1.params_id());
    ev.add(s1, s2, res);
}

// ---------------------------------------------------------------- LayerNorm
// src/layer_norm.cpp:6-33: x * (mean(x^2))^(-1/2) over rows of `len` slots
// (gamma = 1, beta = 0, no mean subtraction), rows packed with a rotated copy.
void LNEvaluator::layer_norm(const Ciphertext &x_in, Ciphertext &res, int len)
{
    CKKSEvaluator &ck = *ckks_;
    const Evaluator &ev = ck.evaluator;
    Ciphertext x = x_in, tmp, x2;
    const int log_step = (int)std::log2((double)len);
    ev.rotate_vector(x, -len, *ck.galois_keys, tmp);
    ev.add_inplace(x, tmp);
    ev.square(x, x2);
    ev.relinearize_inplace(x2, *ck.relin_keys);
    ev.rescale_to_next_inplace(x2);
    tmp = x2;
    for (int i = 0; i < log_step; i++) {
        ev.rotate_vector(tmp, 1 << i, *ck.galois_keys, res);
        ev.add_inplace(res, tmp);
        tmp = res;
    }
    Plaintext delta;
    ck.encoder.encode(1.0 / 768, res.params_id(), res.scale(), delta);
    ev.multiply_plain_inplace(res, delta);
    ev.rescale_to_next_inplace(res);
    res = ck.invert_sqrt(res, 4, 2);
    ev.mod_switch_to_inplace(x, res.params_id());
    ev.multiply(res, x, res);
    ev.relinearize_inplace(res, *ck.relin_keys);
    ev.rescale_to_next_inplace(res);
}

// ---------------------------------------------------------------- Softmax
// src/softmax.cpp:6-42: exp(x) / sum(exp(x)) over rows of `len` slots, with
// exp(x) ~ (1 + x/128)^128 and 1/s from Newton's inverse on s/100.
void SoftmaxEvaluator::softmax(const Ciphertext &x_in, Ciphertext &res, int len)
{
    CKKSEvaluator &ck = *ckks_;
    const Evaluator &ev = ck.evaluator;
    Ciphertext x = x_in, tmp, exp_x;
    const int log_step = (int)std::log2((double)len);
    ev.rotate_vector(x, -len, *ck.galois_keys, tmp);
    ev.add_inplace(x, tmp);
    exp_x = ck.exp(x);
    tmp = exp_x;
    for (int i = 0; i < log_step; i++) {
        ev.rotate_vector(tmp, 1 << i, *ck.galois_keys, res);
        ev.add_inplace(res, tmp);
        tmp = res;
    }
    Plaintext delta;
    ck.encoder.encode(0.01, res.params_id(), res.scale(), delta);
    ev.multiply_plain_inplace(res, delta);
    ev.rescale_to_next_inplace(res);
    res = ck.inverse(res);
    ck.encoder.encode(0.01, res.params_id(), res.scale(), delta);
    ev.multiply_plain_inplace(res, delta);
    ev.rescale_to_next_inplace(res);
    ev.mod_switch_to_inplace(exp_x, res.params_id());
    ev.multiply(res, exp_x, res);
    ev.relinearize_inplace(res, *ck.relin_keys);
    ev.rescale_to_next_inplace(res);
}

}  // namespace nexus
