// RAII value types and the Encoder / Encryptor / Decryptor / Evaluator classes
// of the C++ drop-in (include/nexus/ckks.hpp), all over the C ABI.
#include <cmath>
#include <cstring>

#include "../../../include/nexus/ckks.hpp"

namespace nexus {

void throw_status(int rc)
{
    const std::string msg = nx_last_error();
    switch (rc) {
    case NX_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case NX_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    case NX_ERR_LOGIC: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
    }
}

// ---------------------------------------------------------------- context
CKKSContext::CKKSContext(int log_n, const std::vector<int> &bits, uint32_t hw, size_t sparse, size_t special)
    : log_n_(log_n)
{
    check(nx_context_create_hybrid(log_n, bits.data(), bits.size(), special, hw, sparse, &h_));
    load_info();
}

CKKSContext::CKKSContext(nx_context *borrowed) : h_(borrowed), owns_(false), log_n_(0) { load_info(); }

void CKKSContext::load_info()
{
    size_t slots = 0;
    check(nx_context_info(h_, &log_n_, &L_, &slots));
    check(nx_context_key_info(h_, &alpha_, nullptr, nullptr));
    moduli_.resize(L_ + alpha_);
    check(nx_context_moduli(h_, moduli_.data()));
}

CKKSContext::~CKKSContext()
{
    if (owns_) nx_context_destroy(h_);
}

uint32_t CKKSContext::galois_elt(int step) const
{
    uint32_t e = 0;
    check(nx_galois_elt_from_step(h_, step, &e));
    return e;
}

// ---------------------------------------------------------------- ciphertext
Ciphertext::Ciphertext(const CKKSContext &ctx) { bind(ctx); }

void Ciphertext::bind(const CKKSContext &ctx)
{
    if (h_) return;
    ctx_ = &ctx;
    check(nx_ciphertext_create(ctx.get(), &h_));
}

Ciphertext Ciphertext::copy_of(const CKKSContext &ctx, const nx_ciphertext *h)
{
    Ciphertext c(ctx);
    check(nx_ciphertext_copy(c.h_, h));
    return c;
}

Ciphertext::Ciphertext(const Ciphertext &o)
{
    if (o.h_) {
        bind(*o.ctx_);
        check(nx_ciphertext_copy(h_, o.h_));
    }
}

Ciphertext::Ciphertext(Ciphertext &&o) noexcept : ctx_(o.ctx_), h_(o.h_) { o.h_ = nullptr; }

Ciphertext &Ciphertext::operator=(const Ciphertext &o)
{
    if (this == &o) return *this;
    if (!o.h_) {
        nx_ciphertext_destroy(h_);
        h_ = nullptr;
        return *this;
    }
    bind(*o.ctx_);
    check(nx_ciphertext_copy(h_, o.h_));
    return *this;
}

Ciphertext &Ciphertext::operator=(Ciphertext &&o) noexcept
{
    if (this != &o) {
        nx_ciphertext_destroy(h_);
        ctx_ = o.ctx_;
        h_ = o.h_;
        o.h_ = nullptr;
    }
    return *this;
}

Ciphertext::~Ciphertext() { nx_ciphertext_destroy(h_); }

size_t Ciphertext::size() const
{
    size_t s = 0;
    check(nx_ciphertext_info(h_, &s, nullptr, nullptr, nullptr, nullptr));
    return s;
}

size_t Ciphertext::params_id() const
{
    size_t c = 0;
    check(nx_ciphertext_info(h_, nullptr, &c, nullptr, nullptr, nullptr));
    return c;
}

size_t Ciphertext::coeff_modulus_size() const
{
    size_t c = 0;
    check(nx_ciphertext_info(h_, nullptr, nullptr, &c, nullptr, nullptr));
    return c;
}

double Ciphertext::scale() const
{
    double s = 0;
    check(nx_ciphertext_info(h_, nullptr, nullptr, nullptr, &s, nullptr));
    return s;
}

void Ciphertext::set_scale(double s) { check(nx_ciphertext_set_scale(h_, s)); }

// ---------------------------------------------------------------- plaintext
Plaintext::Plaintext(const CKKSContext &ctx) { bind(ctx); }

void Plaintext::bind(const CKKSContext &ctx)
{
    if (h_) return;
    ctx_ = &ctx;
    check(nx_plaintext_create(ctx.get(), &h_));
}

Plaintext::Plaintext(const Plaintext &o)
{
    if (o.h_) {
        bind(*o.ctx_);
        check(nx_plaintext_copy(h_, o.h_));
    }
}

Plaintext::Plaintext(Plaintext &&o) noexcept : ctx_(o.ctx_), h_(o.h_) { o.h_ = nullptr; }

Plaintext &Plaintext::operator=(const Plaintext &o)
{
    if (this == &o) return *this;
    if (!o.h_) {
        nx_plaintext_destroy(h_);
        h_ = nullptr;
        return *this;
    }
    bind(*o.ctx_);
    check(nx_plaintext_copy(h_, o.h_));
    return *this;
}

Plaintext &Plaintext::operator=(Plaintext &&o) noexcept
{
    if (this != &o) {
        nx_plaintext_destroy(h_);
        ctx_ = o.ctx_;
        h_ = o.h_;
        o.h_ = nullptr;
    }
    return *this;
}

Plaintext::~Plaintext() { nx_plaintext_destroy(h_); }

size_t Plaintext::params_id() const
{
    size_t c = 0;
    check(nx_plaintext_info(h_, &c, nullptr, nullptr));
    return c;
}

double Plaintext::scale() const
{
    double s = 0;
    check(nx_plaintext_info(h_, nullptr, nullptr, &s));
    return s;
}

void Plaintext::set_scale(double s) { check(nx_plaintext_set_scale(h_, s)); }

// ---------------------------------------------------------------- keys
SecretKey::SecretKey(const CKKSContext &ctx, uint64_t seed) { check(nx_secret_key_generate(ctx.get(), seed, &h_)); }
SecretKey::~SecretKey()
{
    if (owns_) nx_secret_key_destroy(h_);
}

KSwitchKeys::KSwitchKeys(const CKKSContext &ctx) { check(nx_kswitch_keys_create(ctx.get(), &h_)); }
KSwitchKeys::~KSwitchKeys()
{
    if (owns_) nx_kswitch_keys_destroy(h_);
}

bool KSwitchKeys::has(uint32_t elt) const
{
    int p = 0;
    check(nx_kswitch_keys_has(h_, elt, &p));
    return p != 0;
}

// ---------------------------------------------------------------- encoder
void Encoder::encode(const std::vector<double> &values, size_t chain_index, double scale, Plaintext &plain) const
{
    plain.bind(*ctx_);
    check(nx_encode(ctx_->get(), values.data(), nullptr, values.size(), chain_index, scale, plain.get()));
}

void Encoder::encode(const std::vector<double> &values, double scale, Plaintext &plain) const
{
    encode(values, 1, scale, plain);
}

void Encoder::encode(const std::vector<std::complex<double>> &values, size_t chain_index, double scale,
                     Plaintext &plain) const
{
    std::vector<double> re(values.size()), im(values.size());
    for (size_t i = 0; i < values.size(); i++) {
        re[i] = values[i].real();
        im[i] = values[i].imag();
    }
    plain.bind(*ctx_);
    check(nx_encode(ctx_->get(), re.data(), im.data(), values.size(), chain_index, scale, plain.get()));
}

void Encoder::encode(const std::vector<std::complex<double>> &values, double scale, Plaintext &plain) const
{
    encode(values, 1, scale, plain);
}

void Encoder::encode(double value, size_t chain_index, double scale, Plaintext &plain) const
{
    plain.bind(*ctx_);
    check(nx_encode_constant(ctx_->get(), value, chain_index, scale, plain.get()));
}

void Encoder::encode(double value, double scale, Plaintext &plain) const { encode(value, 1, scale, plain); }

void Encoder::decode(const Plaintext &plain, std::vector<double> &values) const
{
    values.assign(slot_count(), 0.0);
    check(nx_decode(ctx_->get(), plain.get(), values.data(), nullptr, values.size()));
}

void Encoder::decode(const Plaintext &plain, std::vector<std::complex<double>> &values) const
{
    std::vector<double> re(slot_count()), im(slot_count());
    check(nx_decode(ctx_->get(), plain.get(), re.data(), im.data(), re.size()));
    values.resize(re.size());
    for (size_t i = 0; i < re.size(); i++) values[i] = {re[i], im[i]};
}

// ---------------------------------------------------------------- crypto
void Encryptor::encrypt(const Plaintext &plain, Ciphertext &ct)
{
    ct.bind(*ctx_);
    check(nx_encrypt_symmetric(ctx_->get(), sk_->get(), plain.get(), counter_++, ct.get()));
}

void Decryptor::decrypt(const Ciphertext &ct, Plaintext &plain) const
{
    plain.bind(*ctx_);
    check(nx_decrypt(ctx_->get(), sk_->get(), ct.get(), plain.get()));
}

void Decryptor::create_galois_keys_from_steps(const std::vector<int> &steps, GaloisKeys &keys) const
{
    std::vector<uint32_t> elts;
    for (int s : steps) elts.push_back(ctx_->galois_elt(s));
    create_galois_keys_from_elts(elts, keys);
}

void Decryptor::create_galois_keys_from_elts(const std::vector<uint32_t> &elts, GaloisKeys &keys) const
{
    check(nx_galois_keys_generate(ctx_->get(), sk_->get(), elts.data(), elts.size(), 0x67616c6f, keys.get()));
}

void Decryptor::create_relin_keys(RelinKeys &keys) const
{
    check(nx_relin_key_generate(ctx_->get(), sk_->get(), 0x72656c69, keys.get()));
}

// ---------------------------------------------------------------- evaluator
#define OUT(d) (d).bind(*ctx_)

void Evaluator::mod_switch_to_next_inplace(Ciphertext &ct) const { check(nx_mod_switch_to_next(ct.get(), ct.get())); }
void Evaluator::mod_switch_to_inplace(Ciphertext &ct, size_t chain_index) const
{
    check(nx_mod_switch_to(ct.get(), chain_index, ct.get()));
}
void Evaluator::mod_switch_to(const Ciphertext &ct, size_t chain_index, Ciphertext &dest) const
{
    Ciphertext out(*ctx_);
    check(nx_mod_switch_to(ct.get(), chain_index, out.get()));
    dest = std::move(out);
}
void Evaluator::mod_switch_to_inplace(Plaintext &pt, size_t chain_index) const
{
    check(nx_plain_mod_switch_to(pt.get(), chain_index, pt.get()));
}
void Evaluator::rescale_to_next_inplace(Ciphertext &ct) const { check(nx_rescale_to_next(ct.get(), ct.get())); }
void Evaluator::relinearize_inplace(Ciphertext &ct, const RelinKeys &rk) const
{
    check(nx_relinearize(ct.get(), rk.get(), ct.get()));
}
void Evaluator::square(const Ciphertext &ct, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_square(ct.get(), dest.get()));
}
void Evaluator::square_inplace(Ciphertext &ct) const { check(nx_square(ct.get(), ct.get())); }
void Evaluator::multiply(const Ciphertext &a, const Ciphertext &b, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_multiply(a.get(), b.get(), dest.get()));
}
void Evaluator::multiply_inplace(Ciphertext &a, const Ciphertext &b) const
{
    check(nx_multiply(a.get(), b.get(), a.get()));
}
void Evaluator::multiply_plain(const Ciphertext &ct, const Plaintext &p, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_multiply_plain(ct.get(), p.get(), dest.get()));
}
void Evaluator::multiply_plain_inplace(Ciphertext &ct, const Plaintext &p) const
{
    check(nx_multiply_plain(ct.get(), p.get(), ct.get()));
}
void Evaluator::add_plain(const Ciphertext &ct, const Plaintext &p, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_add_plain(ct.get(), p.get(), dest.get()));
}
void Evaluator::add_plain_inplace(Ciphertext &ct, const Plaintext &p) const
{
    check(nx_add_plain(ct.get(), p.get(), ct.get()));
}
void Evaluator::add(const Ciphertext &a, const Ciphertext &b, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_add(a.get(), b.get(), dest.get()));
}
void Evaluator::add_inplace(Ciphertext &a, const Ciphertext &b) const { check(nx_add(a.get(), b.get(), a.get())); }
void Evaluator::add_many(const std::vector<Ciphertext> &cts, Ciphertext &dest) const
{
    if (cts.size() < 2) throw std::invalid_argument("add_many requires at least 2 ciphertexts");
    std::vector<const nx_ciphertext *> hs;
    for (auto &c : cts) hs.push_back(c.get());
    OUT(dest);
    check(nx_add_many(hs.data(), hs.size(), dest.get()));
}
void Evaluator::sub_plain(const Ciphertext &ct, const Plaintext &p, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_sub_plain(ct.get(), p.get(), dest.get()));
}
void Evaluator::sub_plain_inplace(Ciphertext &ct, const Plaintext &p) const
{
    check(nx_sub_plain(ct.get(), p.get(), ct.get()));
}
void Evaluator::sub(const Ciphertext &a, const Ciphertext &b, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_sub(a.get(), b.get(), dest.get()));
}
void Evaluator::sub_inplace(Ciphertext &a, const Ciphertext &b) const { check(nx_sub(a.get(), b.get(), a.get())); }
void Evaluator::negate(const Ciphertext &ct, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_negate(ct.get(), dest.get()));
}
void Evaluator::negate_inplace(Ciphertext &ct) const { check(nx_negate(ct.get(), ct.get())); }
void Evaluator::rotate_vector(const Ciphertext &ct, int steps, const GaloisKeys &k, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_rotate_vector(ct.get(), steps, k.get(), dest.get()));
}
void Evaluator::rotate_vector_inplace(Ciphertext &ct, int steps, const GaloisKeys &k) const
{
    check(nx_rotate_vector(ct.get(), steps, k.get(), ct.get()));
}
void Evaluator::apply_galois(const Ciphertext &ct, uint32_t elt, const GaloisKeys &k, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_apply_galois(ct.get(), elt, k.get(), dest.get()));
}
void Evaluator::apply_galois_inplace(Ciphertext &ct, uint32_t elt, const GaloisKeys &k) const
{
    check(nx_apply_galois(ct.get(), elt, k.get(), ct.get()));
}
void Evaluator::rotate_vector_hoisted(const Ciphertext &ct, const std::vector<int> &steps, const GaloisKeys &k,
                                      std::vector<Ciphertext> &dest) const
{
    if (per_op_) {
        std::vector<Ciphertext> out(steps.size());
        for (size_t i = 0; i < steps.size(); i++) {
            if (steps[i] == 0)
                out[i] = ct;
            else
                rotate_vector(ct, steps[i], k, out[i]);
        }
        dest = std::move(out);
        return;
    }
    std::vector<Ciphertext> out(steps.size());
    std::vector<nx_ciphertext *> h(steps.size());
    for (size_t i = 0; i < steps.size(); i++) {
        out[i].bind(*ctx_);
        h[i] = out[i].get();
    }
    check(nx_rotate_hoisted(ct.get(), steps.data(), steps.size(), k.get(), h.data()));
    dest = std::move(out);
}
void Evaluator::multiply_plain_accumulate(const std::vector<const Ciphertext *> &cts,
                                          const std::vector<const Plaintext *> &pts, Ciphertext &dest) const
{
    if (per_op_) {
        if (cts.empty() || cts.size() != pts.size())
            throw std::invalid_argument("multiply_plain_accumulate: one plaintext per ciphertext");
        Ciphertext acc, t;
        for (size_t r = 0; r < cts.size(); r++) {
            multiply_plain(*cts[r], *pts[r], r ? t : acc);
            if (r) add_inplace(acc, t);
        }
        dest = std::move(acc);
        return;
    }
    std::vector<const nx_ciphertext *> c;
    std::vector<const nx_plaintext *> p;
    for (const Ciphertext *x : cts) c.push_back(x->get());
    for (const Plaintext *x : pts) p.push_back(x->get());
    Ciphertext out(*ctx_);
    check(nx_multiply_plain_accumulate(c.data(), p.data(), c.size(), out.get()));
    dest = std::move(out);
}

void Evaluator::linear_combination(const std::vector<const Ciphertext *> &cts, const std::vector<double> &values,
                                   const std::vector<double> &scales, size_t chain_index, Ciphertext &dest) const
{
    if (cts.size() != values.size() || cts.size() != scales.size())
        throw std::invalid_argument("linear_combination: term counts differ");
    if (per_op_) {
        // mod_switch_to, multiply_plain by the encoded constant, add: per term
        if (cts.empty()) throw std::invalid_argument("linear_combination: no terms");
        Ciphertext acc;
        for (size_t k = 0; k < cts.size(); k++) {
            Ciphertext t;
            mod_switch_to(*cts[k], chain_index, t);
            Plaintext p(*ctx_);
            check(nx_encode_constant(ctx_->get(), values[k], chain_index, scales[k], p.get()));
            multiply_plain_inplace(t, p);
            if (k)
                add_inplace(acc, t);
            else
                acc = std::move(t);
        }
        dest = std::move(acc);
        return;
    }
    std::vector<const nx_ciphertext *> c;
    for (const Ciphertext *x : cts) c.push_back(x->get());
    Ciphertext res(*ctx_);
    check(nx_linear_combination(c.data(), values.data(), scales.data(), c.size(), chain_index, res.get()));
    dest = std::move(res);
}

void Evaluator::multiply_plain_accumulate_multi(const std::vector<const Ciphertext *> &cts,
                                                const std::vector<const Plaintext *> &pts,
                                                std::vector<Ciphertext> &dest) const
{
    if (cts.empty() || pts.size() % cts.size())
        throw std::invalid_argument("multiply_plain_accumulate_multi: plaintexts must be rows x ciphertexts");
    const size_t rows = pts.size() / cts.size();
    if (per_op_) {
        const size_t K = cts.size();
        std::vector<Ciphertext> res(rows);
        for (size_t g = 0; g < rows; g++) {
            std::vector<const Ciphertext *> c;
            std::vector<const Plaintext *> p;
            for (size_t r = 0; r < K; r++)
                if (pts[g * K + r]) {
                    c.push_back(cts[r]);
                    p.push_back(pts[g * K + r]);
                }
            if (c.empty()) throw std::invalid_argument("multiply_plain_accumulate_multi: every row needs a term");
            multiply_plain_accumulate(c, p, res[g]);
        }
        dest = std::move(res);
        return;
    }
    std::vector<const nx_ciphertext *> c;
    std::vector<const nx_plaintext *> p;
    for (const Ciphertext *x : cts) c.push_back(x->get());
    for (const Plaintext *x : pts) p.push_back(x ? x->get() : nullptr);
    std::vector<Ciphertext> res;
    res.reserve(rows);
    std::vector<nx_ciphertext *> o;
    for (size_t g = 0; g < rows; g++) {
        res.emplace_back(*ctx_);
        o.push_back(res.back().get());
    }
    check(nx_multiply_plain_accumulate_multi(c.data(), c.size(), p.data(), rows, o.data()));
    dest = std::move(res);
}
void Evaluator::rescale_to_next_many(const std::vector<const Ciphertext *> &cts, std::vector<Ciphertext> &dest) const
{
    if (per_op_) {
        std::vector<Ciphertext> res(cts.size());
        for (size_t i = 0; i < cts.size(); i++) {
            res[i] = *cts[i];
            rescale_to_next_inplace(res[i]);
        }
        dest = std::move(res);
        return;
    }
    std::vector<const nx_ciphertext *> in;
    for (const Ciphertext *x : cts) in.push_back(x->get());
    std::vector<Ciphertext> res;
    res.reserve(cts.size());
    std::vector<nx_ciphertext *> o;
    for (size_t i = 0; i < cts.size(); i++) {
        res.emplace_back(*ctx_);
        o.push_back(res.back().get());
    }
    check(nx_rescale_to_next_many(in.data(), in.size(), o.data()));
    dest = std::move(res);
}

void Evaluator::multiply_sum(const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                             Ciphertext &dest) const
{
    if (a.empty() || a.size() != b.size()) throw std::invalid_argument("multiply_sum: operand counts differ");
    if (per_op_) {
        Ciphertext acc, t;
        for (size_t i = 0; i < a.size(); i++) {
            multiply(*a[i], *b[i], i ? t : acc);
            if (i) add_inplace(acc, t);
        }
        dest = std::move(acc);
        return;
    }
    std::vector<const nx_ciphertext *> x, y;
    for (size_t i = 0; i < a.size(); i++) {
        x.push_back(a[i]->get());
        y.push_back(b[i]->get());
    }
    Ciphertext out(*ctx_);
    check(nx_multiply_sum(x.data(), y.data(), x.size(), out.get()));
    dest = std::move(out);
}

void Evaluator::multiply_relin_rescale_many(const std::vector<const Ciphertext *> &a,
                                            const std::vector<const Ciphertext *> &b, const RelinKeys &rk,
                                            std::vector<Ciphertext> &dest, bool fused) const
{
    if (a.size() != b.size()) throw std::invalid_argument("multiply_relin_rescale_many: operand counts differ");
    if (per_op_) {
        // multiply, relinearize, rescale_to_next per pair (the fused single
        // rounding is a batched form; per op it is the unfused sequence)
        std::vector<Ciphertext> res(a.size());
        for (size_t i = 0; i < a.size(); i++) {
            multiply(*a[i], *b[i], res[i]);
            relinearize_inplace(res[i], rk);
            rescale_to_next_inplace(res[i]);
        }
        dest = std::move(res);
        return;
    }
    std::vector<const nx_ciphertext *> x, y;
    for (size_t i = 0; i < a.size(); i++) {
        x.push_back(a[i]->get());
        y.push_back(b[i]->get());
    }
    std::vector<Ciphertext> res;
    res.reserve(a.size());
    std::vector<nx_ciphertext *> o;
    for (size_t i = 0; i < a.size(); i++) {
        res.emplace_back(*ctx_);
        o.push_back(res.back().get());
    }
    if (fused)
        check(nx_multiply_relin_rescale_fused_many(x.data(), y.data(), x.size(), rk.get(), o.data()));
    else
        check(nx_multiply_relin_rescale_many(x.data(), y.data(), x.size(), rk.get(), o.data()));
    dest = std::move(res);
}

void Evaluator::multiply_periodic_mix(const std::vector<const Ciphertext *> &cts, const PeriodicPlaintexts &w,
                                      const std::vector<int32_t> &index, std::vector<Ciphertext> &out) const
{
    const size_t K = cts.size();
    if (!K || index.size() % K) throw std::invalid_argument("multiply_periodic_mix: index must be M x K");
    const size_t M = index.size() / K;
    if (per_op_) {
        // the masks as full plaintexts, one multiply_plain and add per term
        std::vector<Ciphertext> res(M);
        for (size_t m = 0; m < M; m++) {
            bool any = false;
            Ciphertext t;
            for (size_t k = 0; k < K; k++) {
                const int32_t e = index[m * K + k];
                if (e < 0) continue;
                multiply_plain(*cts[k], w.expanded((size_t)e), any ? t : res[m]);
                if (any) add_inplace(res[m], t);
                any = true;
            }
            if (!any) {
                // an output without terms is zero at the product's scale
                multiply_plain(*cts[0], w.expanded(0), res[m]);
                sub_inplace(res[m], res[m]);
            }
        }
        out = std::move(res);
        return;
    }
    std::vector<const nx_ciphertext *> c;
    for (const Ciphertext *x : cts) c.push_back(x->get());
    std::vector<Ciphertext> res;
    res.reserve(M);
    std::vector<nx_ciphertext *> o;
    for (size_t m = 0; m < M; m++) {
        res.emplace_back(*ctx_);
        o.push_back(res.back().get());
    }
    check(nx_multiply_periodic_mix(c.data(), K, w.get(), index.data(), M, o.data()));
    out = std::move(res);
}

PeriodicPlaintexts::PeriodicPlaintexts(const CKKSContext &ctx, const std::vector<double> &values, size_t period,
                                       size_t chain_index, double scale)
    : ctx_(&ctx)
{
    if (!period || values.size() % period) throw std::invalid_argument("PeriodicPlaintexts: values must be count x period");
    count_ = values.size() / period;
    check(nx_periodic_plain_create(ctx.get(), values.data(), period, count_, chain_index, scale, &h_));
}

PeriodicPlaintexts::~PeriodicPlaintexts() { nx_periodic_plain_destroy(h_); }

void PeriodicPlaintexts::expand(size_t index, Plaintext &out) const
{
    Plaintext r(*ctx_);
    check(nx_periodic_plain_expand(h_, index, r.get()));
    out = std::move(r);
}

const Plaintext &PeriodicPlaintexts::expanded(size_t index) const
{
    std::lock_guard<std::mutex> g(mu_);
    auto it = expanded_.find(index);
    if (it != expanded_.end()) return it->second;
    Plaintext r;
    expand(index, r);
    return expanded_.emplace(index, std::move(r)).first->second;
}

void Evaluator::rotate_vector_hoisted_many(const std::vector<const Ciphertext *> &cts, const std::vector<int> &steps,
                                           const GaloisKeys &k, std::vector<std::vector<Ciphertext>> &dest) const
{
    std::vector<std::vector<Ciphertext>> out(cts.size());
    if (per_op_) {
        for (size_t b = 0; b < cts.size(); b++) rotate_vector_hoisted(*cts[b], steps, k, out[b]);
        dest = std::move(out);
        return;
    }
    std::vector<const nx_ciphertext *> in;
    std::vector<nx_ciphertext *> h;
    for (size_t b = 0; b < cts.size(); b++) {
        in.push_back(cts[b]->get());
        out[b].resize(steps.size());
        for (auto &o : out[b]) {
            o.bind(*ctx_);
            h.push_back(o.get());
        }
    }
    check(nx_rotate_hoisted_many(in.data(), in.size(), steps.data(), steps.size(), k.get(), h.data()));
    dest = std::move(out);
}

void Evaluator::rotate_sum_many(const std::vector<std::vector<const Ciphertext *>> &groups,
                                const std::vector<std::vector<int>> &steps, const GaloisKeys &k,
                                std::vector<Ciphertext> &dest) const
{
    if (groups.size() != steps.size()) throw std::invalid_argument("rotate_sum_many: groups and steps must match");
    std::vector<Ciphertext> out(groups.size());
    if (per_op_) {
        for (size_t g = 0; g < groups.size(); g++) rotate_sum(groups[g], steps[g], k, out[g]);
        dest = std::move(out);
        return;
    }
    std::vector<const nx_ciphertext *> in;
    std::vector<int> st;
    std::vector<size_t> sizes;
    std::vector<nx_ciphertext *> h;
    for (size_t g = 0; g < groups.size(); g++) {
        if (groups[g].size() != steps[g].size()) throw std::invalid_argument("rotate_sum_many: inputs and steps");
        for (size_t i = 0; i < groups[g].size(); i++) {
            in.push_back(groups[g][i]->get());
            st.push_back(steps[g][i]);
        }
        sizes.push_back(groups[g].size());
        out[g].bind(*ctx_);
        h.push_back(out[g].get());
    }
    check(nx_rotate_sum_many(in.data(), st.data(), sizes.data(), sizes.size(), k.get(), h.data()));
    dest = std::move(out);
}

void Evaluator::hoisting_inplace(Ciphertext &ct, const GaloisKeys &k, const std::vector<int> &steps) const
{
    if (per_op_) {
        std::vector<const Ciphertext *> c(steps.size(), &ct);
        Ciphertext sum;
        rotate_sum(c, steps, k, sum);
        ct = std::move(sum);
        return;
    }
    check(nx_hoisting(ct.get(), steps.data(), steps.size(), k.get(), ct.get()));
}
void Evaluator::rotate_sum(const std::vector<const Ciphertext *> &cts, const std::vector<int> &steps,
                           const GaloisKeys &k, Ciphertext &dest) const
{
    if (cts.empty() || cts.size() != steps.size()) throw std::invalid_argument("rotate_sum: inputs and steps must match");
    if (per_op_) {
        Ciphertext acc, t;
        for (size_t i = 0; i < cts.size(); i++) {
            Ciphertext &dst = i ? t : acc;
            if (steps[i] == 0)
                dst = *cts[i];
            else
                rotate_vector(*cts[i], steps[i], k, dst);
            if (i) add_inplace(acc, t);
        }
        dest = std::move(acc);
        return;
    }
    std::vector<const nx_ciphertext *> h;
    for (const Ciphertext *c : cts) h.push_back(c->get());
    Ciphertext out(*ctx_);
    check(nx_rotate_sum(h.data(), steps.data(), h.size(), k.get(), out.get()));
    dest = std::move(out);
}
void Evaluator::complex_conjugate(const Ciphertext &ct, const GaloisKeys &k, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_complex_conjugate(ct.get(), k.get(), dest.get()));
}
void Evaluator::complex_conjugate_inplace(Ciphertext &ct, const GaloisKeys &k) const
{
    check(nx_complex_conjugate(ct.get(), k.get(), ct.get()));
}
void Evaluator::transform_to_ntt(const Ciphertext &ct, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_transform_to_ntt(ct.get(), dest.get()));
}
void Evaluator::transform_to_ntt_inplace(Ciphertext &ct) const { check(nx_transform_to_ntt(ct.get(), ct.get())); }
void Evaluator::transform_from_ntt(const Ciphertext &ct, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_transform_from_ntt(ct.get(), dest.get()));
}
void Evaluator::transform_from_ntt_inplace(Ciphertext &ct) const
{
    check(nx_transform_from_ntt(ct.get(), ct.get()));
}
void Evaluator::multiply_power_of_x(const Ciphertext &ct, int64_t e, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_multiply_power_of_x(ct.get(), e, dest.get()));
}
void Evaluator::multiply_const(const Ciphertext &ct, double v, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_multiply_const(ct.get(), v, dest.get()));
}
void Evaluator::multiply_const_inplace(Ciphertext &ct, double v) const
{
    check(nx_multiply_const(ct.get(), v, ct.get()));
}
void Evaluator::add_const(const Ciphertext &ct, double v, Ciphertext &dest) const
{
    OUT(dest);
    check(nx_add_const(ct.get(), v, dest.get()));
}
void Evaluator::add_const_inplace(Ciphertext &ct, double v) const { check(nx_add_const(ct.get(), v, ct.get())); }
void Evaluator::double_inplace(Ciphertext &ct) const { check(nx_double(ct.get(), ct.get())); }

// ---- reduced-error operations (SEAL-4.1-bs evaluator.cpp:420-594), through
// the C ABI; the _inplace forms write their first operand
void Evaluator::add_inplace_reduced_error(Ciphertext &a, const Ciphertext &b) const
{
    check(nx_add_reduced_error(a.get(), b.get(), a.get()));
}
void Evaluator::add_reduced_error(const Ciphertext &a, const Ciphertext &b, Ciphertext &dest) const
{
    // cuda/src/ckks_evaluator.cuh:338-345: dest = a, then dest += b (with dest
    // aliasing b, b's operand is the one adjusted)
    Ciphertext out(*ctx_);
    if (&b == &dest)
        check(nx_add_reduced_error(b.get(), a.get(), out.get()));
    else
        check(nx_add_reduced_error(a.get(), b.get(), out.get()));
    dest = std::move(out);
}
void Evaluator::sub_inplace_reduced_error(Ciphertext &a, const Ciphertext &b) const
{
    check(nx_sub_reduced_error(a.get(), b.get(), a.get()));
}
void Evaluator::sub_reduced_error(const Ciphertext &a, const Ciphertext &b, Ciphertext &dest) const
{
    Ciphertext out(*ctx_);
    check(nx_sub_reduced_error(a.get(), b.get(), out.get()));
    dest = std::move(out);
}
void Evaluator::multiply_inplace_reduced_error(Ciphertext &a, const Ciphertext &b, const RelinKeys &rk) const
{
    check(nx_multiply_reduced_error(a.get(), b.get(), rk.get(), a.get()));
}
void Evaluator::multiply_reduced_error(const Ciphertext &a, const Ciphertext &b, const RelinKeys &rk,
                                       Ciphertext &dest) const
{
    Ciphertext out(*ctx_);
    if (&b == &dest)
        check(nx_multiply_reduced_error(b.get(), a.get(), rk.get(), out.get()));
    else
        check(nx_multiply_reduced_error(a.get(), b.get(), rk.get(), out.get()));
    dest = std::move(out);
}

void Evaluator::multiply_vector_reduced_error(const Ciphertext &ct, const std::vector<double> &values,
                                              Ciphertext &dest) const
{
    Ciphertext out(*ctx_);
    check(nx_multiply_vector_reduced_error(ct.get(), values.data(), nullptr, values.size(), out.get()));
    dest = std::move(out);
}

void Evaluator::multiply_vector_reduced_error(const Ciphertext &ct, const std::vector<std::complex<double>> &values,
                                              Ciphertext &dest) const
{
    std::vector<double> re(values.size()), im(values.size());
    for (size_t i = 0; i < values.size(); i++) {
        re[i] = values[i].real();
        im[i] = values[i].imag();
    }
    Ciphertext out(*ctx_);
    check(nx_multiply_vector_reduced_error(ct.get(), re.data(), im.data(), values.size(), out.get()));
    dest = std::move(out);
}

#undef OUT

}  // namespace nexus
