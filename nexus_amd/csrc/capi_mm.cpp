// C ABI of the MatrixMul pipeline (include/nexus_mm.h).
#include "../../include/nexus_mm.h"

#include "handles.h"

namespace nx {
void mm_compress(Context &c, const SecretKey &sk, const double *val, size_t count, double scale, uint64_t seed,
                 Ciphertext &out);
void mm_expand(Context &c, const Ciphertext &ct, uint32_t m, const uint32_t *elts, const KSwitchKeys &gk,
               CtBatch &out);
void mm_accumulate_cts(Context &c, const std::vector<const Plaintext *> &a, const std::vector<const CtBatch *> &E,
                       size_t cols, std::vector<Ciphertext> &out);
}  // namespace nx

using nx::guard_call;
using nx::need_ptr;

extern "C" {

int nx_mm_compress(nx_context *ctx, const nx_secret_key *sk, const double *values, size_t count, double scale,
                   uint64_t seed, nx_ciphertext *out)
{
    return guard_call([&] {
        need_ptr(ctx, "ctx");
        need_ptr(sk, "sk");
        need_ptr(values, "values");
        need_ptr(out, "out");
        nx::mm_compress(ctx->impl, sk->impl, values, count, scale, seed, out->impl);
    });
}

int nx_mm_expand(const nx_ciphertext *ct, uint32_t m, const uint32_t *elts, const nx_kswitch_keys *galois_keys,
                 nx_ct_batch **out)
{
    return guard_call([&] {
        need_ptr(ct, "ct");
        need_ptr(elts, "elts");
        need_ptr(galois_keys, "galois_keys");
        need_ptr(out, "out");
        nx::require(ct->impl.ctx != nullptr, "ciphertext is not initialised");
        auto *b = new nx_ct_batch();
        try {
            nx::mm_expand(*ct->impl.ctx, ct->impl, m, elts, galois_keys->impl, b->impl);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int nx_ct_batch_count(const nx_ct_batch *batch, size_t *count)
{
    return guard_call([&] {
        need_ptr(batch, "batch");
        need_ptr(count, "count");
        *count = batch->impl.B;
    });
}

int nx_ct_batch_extract(const nx_ct_batch *batch, size_t index, nx_ciphertext *out)
{
    return guard_call([&] {
        need_ptr(batch, "batch");
        need_ptr(out, "out");
        const nx::CtBatch &b = batch->impl;
        nx::require(index < b.B, "index out of range", nx::kOutOfRange);
        nx::Context &c = *b.ctx;
        nx::Ciphertext r;
        r.ctx = &c;
        r.size = 2;
        r.limbs = b.limbs;
        r.scale = b.scale;
        r.ntt = true;
        r.data.reset(&c, 2 * b.limbs * c.n);
        NX_HIP(hipMemcpyAsync(r.data.get(), b.ct(index), 2 * b.limbs * c.n * 8, hipMemcpyDeviceToDevice, c.stream));
        out->impl = std::move(r);
    });
}

int nx_ct_batch_destroy(nx_ct_batch *batch)
{
    return guard_call([&] { delete batch; });
}

int nx_mm_accumulate(const nx_plaintext *const *a_pts, size_t inner, const nx_ct_batch *const *batches,
                     size_t n_batches, size_t cols, nx_ciphertext *const *outs)
{
    return guard_call([&] {
        need_ptr(a_pts, "a_pts");
        need_ptr(batches, "batches");
        need_ptr(outs, "outs");
        std::vector<const nx::Plaintext *> a;
        for (size_t j = 0; j < inner; j++) {
            need_ptr(a_pts[j], "a_pts[j]");
            a.push_back(&a_pts[j]->impl);
        }
        std::vector<const nx::CtBatch *> e;
        for (size_t k = 0; k < n_batches; k++) {
            need_ptr(batches[k], "batches[k]");
            e.push_back(&batches[k]->impl);
        }
        nx::require(!e.empty() && e[0]->ctx != nullptr, "no expanded ciphertexts");
        std::vector<nx::Ciphertext> res;
        nx::mm_accumulate_cts(*e[0]->ctx, a, e, cols, res);
        for (size_t i = 0; i < cols; i++) {
            need_ptr(outs[i], "outs[i]");
            outs[i]->impl = std::move(res[i]);
        }
    });
}

}  // extern "C"

extern "C" int nx_mod_raise(const nx_ciphertext *a, nx_ciphertext *dest)
{
    return nx::guard_call([&] {
        nx::need_ptr(a, "a");
        nx::need_ptr(dest, "dest");
        nx::op_mod_raise(a->impl, dest->impl);
    });
}
