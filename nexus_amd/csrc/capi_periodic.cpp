// C ABI of periodic plaintexts and the block mix (include/nexus_ckks.h).
#include "handles.h"

extern "C" {

int nx_periodic_plain_create(nx_context *ctx, const double *values, size_t period, size_t count, size_t chain_index,
                             double scale, nx_periodic_plain **out)
{
    return nx::guard_call([&] {
        nx::need_ptr(ctx, "ctx");
        nx::need_ptr(values, "values");
        nx::need_ptr(out, "out");
        nx::Context &c = ctx->impl;
        auto p = std::make_unique<nx_periodic_plain>();
        nx::periodic_create(c, values, period, count, c.limbs_of(chain_index), scale, p->impl);
        *out = p.release();
    });
}

int nx_periodic_plain_destroy(nx_periodic_plain *p)
{
    delete p;
    return NX_OK;
}

int nx_periodic_plain_expand(const nx_periodic_plain *p, size_t index, nx_plaintext *out)
{
    return nx::guard_call([&] {
        nx::need_ptr(p, "p");
        nx::need_ptr(out, "out");
        nx::periodic_expand_entry(p->impl, index, out->impl);
    });
}

int nx_multiply_periodic_mix(const nx_ciphertext *const *in, size_t K, const nx_periodic_plain *w,
                             const int32_t *index, size_t M, nx_ciphertext *const *out)
{
    return nx::guard_call([&] {
        nx::need_ptr(in, "in");
        nx::need_ptr(w, "w");
        nx::need_ptr(index, "index");
        nx::need_ptr(out, "out");
        std::vector<const nx::Ciphertext *> a;
        for (size_t k = 0; k < K; k++) {
            nx::need_ptr(in[k], "in[k]");
            a.push_back(&in[k]->impl);
        }
        std::vector<nx::Ciphertext *> o;
        for (size_t m = 0; m < M; m++) {
            nx::need_ptr(out[m], "out[m]");
            o.push_back(&out[m]->impl);
        }
        // the table and term cache are internal state: the handle stays logically const
        nx::op_periodic_mix(a, const_cast<nx_periodic_plain *>(w)->impl, index, M, o);
    });
}

}  // extern "C"
