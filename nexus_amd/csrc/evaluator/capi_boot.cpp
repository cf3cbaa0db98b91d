// C ABI of the bootstrapper (include/nexus_eval.h, bootstrapping section).
#include <cstring>

#include "../../../include/nexus/bootstrapper.hpp"
#include "eval_handles.h"

using nexus_capi::need;
using nexus_capi::run;
using nexus_capi::store;

namespace {

const nexus::LinearLevel &plan_level(const nexus::BootPlan &p, int which, int level)
{
    if (which != 0 && which != 1) throw std::invalid_argument("which must be 0 (CoeffToSlot) or 1 (SlotToCoeff)");
    const auto &v = which == 0 ? p.cts : p.stc;
    if (level < 0 || (size_t)level >= v.size()) throw std::out_of_range("level out of range");
    return v[(size_t)level];
}

void copy_steps(const std::vector<int> &s, int *steps, size_t capacity, size_t *count)
{
    need(count);
    *count = s.size();
    if (steps) std::memcpy(steps, s.data(), std::min(capacity, s.size()) * sizeof(int));
}

nexus::Ciphertext input(nx_bootstrapper *bt, const nx_ciphertext *x)
{
    need(x);
    return nexus::Ciphertext::copy_of(*bt->ev->ctx, x);
}

}  // namespace

namespace nexus_capi {
const nexus::Bootstrapper::Params *bootstrapper_params(const nx_bootstrapper *bt) { return &bt->bt->params(); }
}  // namespace nexus_capi

extern "C" {

int nx_bootstrapper_create(nx_evaluator *ev, int log_slots, int K, int degree, int double_angles, int log_width,
                           double final_scale, nx_bootstrapper **out)
{
    return run([&] {
        need(ev);
        need(out);
        if (!ev->rk || !ev->gk) throw std::invalid_argument("the evaluator needs relinearisation and Galois keys");
        nexus::Bootstrapper::Params p;
        p.logn = log_slots;
        p.K = K;
        p.cheb_degree = degree;
        p.double_angles = double_angles;
        p.log_width = log_width;
        p.final_scale = final_scale;
        auto b = std::make_unique<nx_bootstrapper>();
        b->ev = ev;
        b->bt = std::make_unique<nexus::Bootstrapper>(*ev->ckks, p);
        *out = b.release();
    });
}

int nx_bootstrapper_destroy(nx_bootstrapper *bt)
{
    return run([&] { delete bt; });
}

int nx_bootstrapper_rotation_steps(const nx_bootstrapper *bt, int *steps, size_t capacity, size_t *count)
{
    return run([&] {
        need(bt);
        copy_steps(bt->bt->rotation_steps(), steps, capacity, count);
    });
}

int nx_bootstrapper_depth(const nx_bootstrapper *bt, int *depth)
{
    return run([&] {
        need(bt);
        need(depth);
        *depth = bt->bt->depth();
    });
}

int nx_bootstrap(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext res;
        bt->bt->bootstrap(input(bt, in), res);
        store(res, out);
    });
}

int nx_bootstrap_scaled(nx_bootstrapper *bt, const nx_ciphertext *in, double factor, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext res;
        bt->bt->bootstrap(input(bt, in), res, factor);
        store(res, out);
    });
}

int nx_bootstrap_many(nx_bootstrapper *bt, const nx_ciphertext *const *in, size_t count, const double *factors,
                      nx_ciphertext *const *out)
{
    return run([&] {
        need(bt);
        need(in);
        need(out);
        std::vector<nexus::Ciphertext> x;
        x.reserve(count);
        for (size_t i = 0; i < count; i++) {
            need(out[i]);
            x.push_back(input(bt, in[i]));
        }
        std::vector<const nexus::Ciphertext *> p;
        for (auto &c : x) p.push_back(&c);
        std::vector<double> f(count, 1.0);
        if (factors) f.assign(factors, factors + count);
        std::vector<nexus::Ciphertext> res;
        bt->bt->bootstrap_many(p, res, f);
        for (size_t i = 0; i < count; i++) store(res[i], out[i]);
    });
}

int nx_bootstrap_many_mixed(nx_bootstrapper *const *bts, const nx_ciphertext *const *in, size_t count,
                            const double *factors, nx_ciphertext *const *out)
{
    return run([&] {
        need(bts);
        need(in);
        need(out);
        std::vector<nexus::Ciphertext> x;
        std::vector<nexus::Bootstrapper *> b;
        x.reserve(count);
        for (size_t i = 0; i < count; i++) {
            need(bts[i]);
            need(out[i]);
            if (bts[i]->ev->ctx->get() != bts[0]->ev->ctx->get())
                throw std::invalid_argument("bootstrappers on different contexts");
            x.push_back(input(bts[i], in[i]));
            b.push_back(bts[i]->bt.get());
        }
        std::vector<const nexus::Ciphertext *> p;
        for (auto &c : x) p.push_back(&c);
        std::vector<double> f(count, 1.0);
        if (factors) f.assign(factors, factors + count);
        std::vector<nexus::Ciphertext> res;
        nexus::Bootstrapper::bootstrap_mixed(b, p, res, f);
        for (size_t i = 0; i < count; i++) store(res[i], out[i]);
    });
}

int nx_boot_mod_raise(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext res;
        bt->bt->mod_raise(input(bt, in), res);
        store(res, out);
    });
}

int nx_boot_sub_sum(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext x = input(bt, in);
        bt->bt->sub_sum(x);
        store(x, out);
    });
}

int nx_boot_coeff_to_slot(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *u_re, nx_ciphertext *u_im)
{
    return run([&] {
        need(bt);
        need(u_re);
        nexus::Ciphertext re, im;
        bt->bt->coeff_to_slot(input(bt, in), re, im);
        store(re, u_re);
        if (im.get()) {
            need(u_im);
            store(im, u_im);
        }
    });
}

int nx_boot_eval_mod(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext res;
        bt->bt->eval_mod(input(bt, in), res);
        store(res, out);
    });
}

int nx_boot_slot_to_coeff(nx_bootstrapper *bt, const nx_ciphertext *f_re, const nx_ciphertext *f_im,
                          double input_scale, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext re = input(bt, f_re), im;
        if (f_im) im = input(bt, f_im);
        nexus::Ciphertext res;
        bt->bt->slot_to_coeff(re, f_im ? &im : nullptr, input_scale, res);
        store(res, out);
    });
}

int nx_eval_argmax(nx_bootstrapper *bt, const nx_ciphertext *x, int len, nx_ciphertext *out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::Ciphertext in = input(bt, x), res;
        nexus::ArgmaxEvaluator am(*bt->ev->ckks, *bt->bt);
        am.argmax(in, res, len);
        store(res, out);
    });
}

int nx_boot_plan_info(int log_half_n, int log_slots, double K, int which, int level, int *levels, int *stride, int *g,
                      size_t *count)
{
    return run([&] {
        const nexus::BootPlan p = nexus::make_boot_plan(log_half_n, log_slots, K);
        if (levels) *levels = (int)(which == 0 ? p.cts.size() : p.stc.size());
        const nexus::LinearLevel &lv = plan_level(p, which, level);
        if (stride) *stride = lv.stride;
        if (g) *g = lv.g;
        if (count) *count = lv.diag.size();
    });
}

int nx_boot_plan_diagonals(int log_half_n, int log_slots, double K, int which, int level, int *js, int *is,
                           double *re, double *im)
{
    return run([&] {
        need(js);
        need(is);
        need(re);
        need(im);
        const nexus::BootPlan p = nexus::make_boot_plan(log_half_n, log_slots, K);
        const nexus::LinearLevel &lv = plan_level(p, which, level);
        const size_t Nh = size_t(1) << log_half_n;
        size_t k = 0;
        for (const auto &d : lv.diag) {
            js[k] = d.first.first;
            is[k] = d.first.second;
            for (size_t s = 0; s < Nh; s++) {
                re[k * Nh + s] = d.second[s].real();
                im[k * Nh + s] = d.second[s].imag();
            }
            k++;
        }
    });
}

int nx_boot_eval_mod_coeffs(int K, int degree, int double_angles, int log_width, double *coeffs, int *depth)
{
    return run([&] {
        const std::vector<double> c = nexus::eval_mod_coeffs(K, degree, double_angles, log_width);
        if (coeffs) std::memcpy(coeffs, c.data(), c.size() * sizeof(double));
        if (depth) *depth = nexus::chebyshev_depth(c);
    });
}

int nx_boot_plan_steps(int log_half_n, int log_slots, double K, int *steps, size_t capacity, size_t *count)
{
    return run([&] {
        const nexus::BootPlan p = nexus::make_boot_plan(log_half_n, log_slots, K);
        copy_steps(p.rotation_steps, steps, capacity, count);
    });
}

}  // extern "C"
