// C ABI of the hoisted rotations and the plaintext multiply-accumulate
// (include/nexus_ckks.h).
#include "../../include/nexus_ckks.h"

#include "handles.h"

extern "C" {

int nx_rotate_hoisted(const nx_ciphertext *a, const int *steps, size_t count, const nx_kswitch_keys *galois_keys,
                      nx_ciphertext *const *dests)
{
    return nx::guard_call([&] {
        nx::need_ptr(a, "a");
        nx::need_ptr(galois_keys, "galois_keys");
        nx::need_ptr(dests, "dests");
        if (count) nx::need_ptr(steps, "steps");
        for (size_t k = 0; k < count; k++) nx::need_ptr(dests[k], "dests[k]");
        std::vector<int> st(steps, steps + count);
        std::vector<nx::Ciphertext> out;
        nx::op_rotate_hoisted(a->impl, st, galois_keys->impl, out);
        for (size_t k = 0; k < count; k++) dests[k]->impl = std::move(out[k]);
    });
}

int nx_hoisting(const nx_ciphertext *a, const int *steps, size_t count, const nx_kswitch_keys *galois_keys,
                nx_ciphertext *dest)
{
    return nx::guard_call([&] {
        nx::need_ptr(a, "a");
        nx::need_ptr(steps, "steps");
        nx::need_ptr(galois_keys, "galois_keys");
        nx::need_ptr(dest, "dest");
        std::vector<int> st(steps, steps + count);
        nx::op_hoisting(a->impl, st, galois_keys->impl, dest->impl);
    });
}

int nx_rotate_sum(const nx_ciphertext *const *in, const int *steps, size_t count,
                  const nx_kswitch_keys *galois_keys, nx_ciphertext *dest)
{
    return nx::guard_call([&] {
        nx::need_ptr(in, "in");
        nx::need_ptr(steps, "steps");
        nx::need_ptr(galois_keys, "galois_keys");
        nx::need_ptr(dest, "dest");
        std::vector<const nx::Ciphertext *> v;
        for (size_t i = 0; i < count; i++) {
            nx::need_ptr(in[i], "in[i]");
            v.push_back(&in[i]->impl);
        }
        std::vector<int> st(steps, steps + count);
        nx::Ciphertext out;
        nx::op_rotate_sum(v, st, galois_keys->impl, out);
        dest->impl = std::move(out);
    });
}

int nx_rotate_hoisted_many(const nx_ciphertext *const *in, size_t count, const int *steps, size_t nsteps,
                           const nx_kswitch_keys *galois_keys, nx_ciphertext *const *dests)
{
    return nx::guard_call([&] {
        nx::need_ptr(in, "in");
        nx::need_ptr(galois_keys, "galois_keys");
        nx::need_ptr(dests, "dests");
        if (nsteps) nx::need_ptr(steps, "steps");
        std::vector<const nx::Ciphertext *> v;
        for (size_t b = 0; b < count; b++) {
            nx::need_ptr(in[b], "in[b]");
            v.push_back(&in[b]->impl);
        }
        for (size_t k = 0; k < count * nsteps; k++) nx::need_ptr(dests[k], "dests[k]");
        std::vector<std::vector<nx::Ciphertext>> out;
        nx::op_rotate_hoisted_many(v, std::vector<int>(steps, steps + nsteps), galois_keys->impl, out);
        for (size_t b = 0; b < count; b++)
            for (size_t k = 0; k < nsteps; k++) dests[b * nsteps + k]->impl = std::move(out[b][k]);
    });
}

int nx_rotate_sum_many(const nx_ciphertext *const *in, const int *steps, const size_t *group_sizes, size_t groups,
                       const nx_kswitch_keys *galois_keys, nx_ciphertext *const *dests)
{
    return nx::guard_call([&] {
        nx::need_ptr(in, "in");
        nx::need_ptr(steps, "steps");
        nx::need_ptr(group_sizes, "group_sizes");
        nx::need_ptr(galois_keys, "galois_keys");
        nx::need_ptr(dests, "dests");
        std::vector<std::vector<const nx::Ciphertext *>> g(groups);
        std::vector<std::vector<int>> st(groups);
        std::vector<nx::Ciphertext> res(groups);
        std::vector<nx::Ciphertext *> outs;
        size_t k = 0;
        for (size_t i = 0; i < groups; i++) {
            nx::need_ptr(dests[i], "dests[i]");
            for (size_t j = 0; j < group_sizes[i]; j++, k++) {
                nx::need_ptr(in[k], "in[k]");
                g[i].push_back(&in[k]->impl);
                st[i].push_back(steps[k]);
            }
            outs.push_back(&res[i]);
        }
        nx::op_rotate_sum_many(g, st, galois_keys->impl, outs);
        for (size_t i = 0; i < groups; i++) dests[i]->impl = std::move(res[i]);
    });
}

int nx_multiply_plain_accumulate(const nx_ciphertext *const *cts, const nx_plaintext *const *pts, size_t count,
                                 nx_ciphertext *dest)
{
    return nx::guard_call([&] {
        nx::need_ptr(cts, "cts");
        nx::need_ptr(pts, "pts");
        nx::need_ptr(dest, "dest");
        std::vector<const nx::Ciphertext *> c;
        std::vector<const nx::Plaintext *> p;
        for (size_t r = 0; r < count; r++) {
            nx::need_ptr(cts[r], "cts[r]");
            nx::need_ptr(pts[r], "pts[r]");
            c.push_back(&cts[r]->impl);
            p.push_back(&pts[r]->impl);
        }
        nx::op_dot_plain(c, p, dest->impl);
    });
}

int nx_linear_combination(const nx_ciphertext *const *cts, const double *values, const double *scales, size_t count,
                          size_t chain_index, nx_ciphertext *dest)
{
    return nx::guard_call([&] {
        nx::need_ptr(cts, "cts");
        nx::need_ptr(values, "values");
        nx::need_ptr(scales, "scales");
        nx::need_ptr(dest, "dest");
        nx::require(count >= 1, "count must be positive");
        std::vector<const nx::Ciphertext *> c;
        for (size_t r = 0; r < count; r++) {
            nx::need_ptr(cts[r], "cts[r]");
            c.push_back(&cts[r]->impl);
        }
        nx::require(c[0]->ctx != nullptr, "ciphertext is not initialised");
        nx::Context &ctx = *c[0]->ctx;
        nx::require(chain_index >= 1 && chain_index <= ctx.L, "chain_index out of range");
        nx::Ciphertext out;
        nx::op_linear_combination(c, std::vector<double>(values, values + count),
                                  std::vector<double>(scales, scales + count), ctx.limbs_of(chain_index), out);
        dest->impl = std::move(out);
    });
}

int nx_multiply_plain_accumulate_multi(const nx_ciphertext *const *cts, size_t count, const nx_plaintext *const *pts,
                                       size_t rows, nx_ciphertext *const *dests)
{
    return nx::guard_call([&] {
        nx::need_ptr(cts, "cts");
        nx::need_ptr(pts, "pts");
        nx::need_ptr(dests, "dests");
        std::vector<const nx::Ciphertext *> c;
        std::vector<const nx::Plaintext *> p;
        std::vector<nx::Ciphertext *> o;
        for (size_t r = 0; r < count; r++) {
            nx::need_ptr(cts[r], "cts[r]");
            c.push_back(&cts[r]->impl);
        }
        for (size_t i = 0; i < rows * count; i++) p.push_back(pts[i] ? &pts[i]->impl : nullptr);
        for (size_t g = 0; g < rows; g++) {
            nx::need_ptr(dests[g], "dests[g]");
            o.push_back(&dests[g]->impl);
        }
        nx::op_dot_plain_multi(c, p, o);
    });
}

}  // extern "C"
