// C ABI over the evaluator programs (include/nexus_eval.h).
#include "eval_handles.h"

using nexus_capi::need;
using nexus_capi::run;
using nexus_capi::store;

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

int nx_evaluator_set_per_op(nx_evaluator *ev, int on)
{
    return run([&] {
        need(ev);
        ev->ckks->evaluator.set_per_op(on != 0);
    });
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

int nx_eval_gelu_many(nx_evaluator *ev, const nx_ciphertext *const *x, size_t count, nx_ciphertext *const *out)
{
    return run([&] {
        need(ev);
        need(x);
        need(out);
        std::vector<nexus::Ciphertext> in;
        in.reserve(count);
        for (size_t i = 0; i < count; i++) {
            need(x[i]);
            need(out[i]);
            in.push_back(NX_IN(x[i]));
        }
        std::vector<const nexus::Ciphertext *> p;
        for (auto &c : in) p.push_back(&c);
        nexus::GELUEvaluator g(*ev->ckks);
        std::vector<nexus::Ciphertext> res;
        g.gelu_many(p, res);
        for (size_t i = 0; i < count; i++) store(res[i], out[i]);
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
