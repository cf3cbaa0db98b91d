// C ABI over the evaluator programs (include/nexus_eval.h).
#include <memory>

#include "../../../include/nexus/ckks.hpp"
#include "../../../include/nexus_eval.h"

struct nx_evaluator {
    std::unique_ptr<nexus::CKKSContext> ctx;
    std::unique_ptr<nexus::SecretKey> sk;
    std::unique_ptr<nexus::KSwitchKeys> rk, gk;
    std::unique_ptr<nexus::CKKSEvaluator> ckks;
};

namespace nx {
void set_last_error(const char *msg);  // capi.cpp
}

namespace {

template <class F>
int run(F &&f)
{
    try {
        f();
        return NX_OK;
    } catch (const std::invalid_argument &e) {
        nx::set_last_error(e.what());
        return NX_ERR_INVALID_ARGUMENT;
    } catch (const std::out_of_range &e) {
        nx::set_last_error(e.what());
        return NX_ERR_OUT_OF_RANGE;
    } catch (const std::logic_error &e) {
        nx::set_last_error(e.what());
        return NX_ERR_LOGIC;
    } catch (const std::exception &e) {
        nx::set_last_error(e.what());
        return NX_ERR_UNKNOWN;
    }
}

void store(const nexus::Ciphertext &res, nx_ciphertext *out) { nexus::check(nx_ciphertext_copy(out, res.get())); }

void need(const void *p)
{
    if (!p) throw std::invalid_argument("null argument");
}

}  // namespace

extern "C" {

int nx_evaluator_create(nx_context *ctx, nx_secret_key *sk, nx_kswitch_keys *relin_keys,
                        nx_kswitch_keys *galois_keys, double scale, nx_evaluator **out)
{
    return run([&] {
        need(ctx);
        need(out);
        auto ev = std::make_unique<nx_evaluator>();
        ev->ctx = std::make_unique<nexus::CKKSContext>(ctx);
        if (sk) ev->sk = std::make_unique<nexus::SecretKey>(sk);
        if (relin_keys) ev->rk = std::make_unique<nexus::KSwitchKeys>(relin_keys);
        if (galois_keys) ev->gk = std::make_unique<nexus::KSwitchKeys>(galois_keys);
        ev->ckks = std::make_unique<nexus::CKKSEvaluator>(ev->ctx.get(), ev->sk.get(), ev->rk.get(), ev->gk.get(),
                                                          scale);
        *out = ev.release();
    });
}

int nx_evaluator_destroy(nx_evaluator *ev)
{
    return run([&] { delete ev; });
}

#define NX_IN(x) nexus::Ciphertext::copy_of(*ev->ctx, x)

int nx_eval_sgn(nx_evaluator *ev, const nx_ciphertext *x, int d_g, int d_f, double sgn_factor, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        store(ev->ckks->sgn_eval(NX_IN(x), d_g, d_f, sgn_factor), out);
    });
}

int nx_eval_invert_sqrt(nx_evaluator *ev, const nx_ciphertext *x, int d_newt, int d_gold, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        store(ev->ckks->invert_sqrt(NX_IN(x), d_newt, d_gold), out);
    });
}

int nx_eval_exp(nx_evaluator *ev, const nx_ciphertext *x, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        store(ev->ckks->exp(NX_IN(x)), out);
    });
}

int nx_eval_inverse(nx_evaluator *ev, const nx_ciphertext *x, int iter, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        store(ev->ckks->inverse(NX_IN(x), iter), out);
    });
}

int nx_eval_gelu(nx_evaluator *ev, const nx_ciphertext *x, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        nexus::GELUEvaluator g(*ev->ckks);
        nexus::Ciphertext res;
        g.gelu(NX_IN(x), res);
        store(res, out);
    });
}

int nx_eval_layer_norm(nx_evaluator *ev, const nx_ciphertext *x, int len, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        nexus::LNEvaluator g(*ev->ckks);
        nexus::Ciphertext res;
        g.layer_norm(NX_IN(x), res, len);
        store(res, out);
    });
}

int nx_eval_softmax(nx_evaluator *ev, const nx_ciphertext *x, int len, nx_ciphertext *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        nexus::SoftmaxEvaluator g(*ev->ckks);
        nexus::Ciphertext res;
        g.softmax(NX_IN(x), res, len);
        store(res, out);
    });
}

#undef NX_IN

}  // extern "C"
