// C ABI (include/nexus_ckks.h): converts C++ exceptions to status codes and
// keeps a thread-local error message.
#include "../../include/nexus_ckks.h"

#include <cstring>
#include <string>

#include "handles.h"
#include "numth.h"

namespace {
thread_local std::string g_err;
}

namespace nx {
void set_last_error(const char *msg) { g_err = msg ? msg : ""; }
}  // namespace nx

namespace {

template <class F>
int guard(F &&f)
{
    try {
        f();
        return NX_OK;
    } catch (const nx::Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::invalid_argument &e) {
        g_err = e.what();
        return NX_ERR_INVALID_ARGUMENT;
    } catch (const std::out_of_range &e) {
        g_err = e.what();
        return NX_ERR_OUT_OF_RANGE;
    } catch (const std::logic_error &e) {
        g_err = e.what();
        return NX_ERR_LOGIC;
    } catch (const std::bad_alloc &e) {
        g_err = e.what();
        return NX_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        g_err = e.what();
        return NX_ERR_UNKNOWN;
    }
}

void need(const void *p, const char *what)
{
    if (!p) nx::fail(nx::kInvalidArgument, std::string(what) + " is null");
}

nx::Context &ctx_of(const nx::Ciphertext &c)
{
    nx::require(c.ctx != nullptr, "ciphertext is not initialised");
    return *c.ctx;
}

}  // namespace

extern "C" {

const char *nx_last_error(void) { return g_err.c_str(); }

int nx_device_count(int *count)
{
    return guard([&] {
        need(count, "count");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) {
            (void)hipGetLastError();
            c = 0;
        }
        *count = c;
    });
}

int nx_device_synchronize(void)
{
    return guard([&] { NX_HIP(hipDeviceSynchronize()); });
}

int nx_context_create(int log_n, const int *bit_sizes, size_t count, uint32_t hamming_weight, size_t sparse_slots,
                      nx_context **out)
{
    return guard([&] {
        need(bit_sizes, "bit_sizes");
        need(out, "out");
        std::vector<int> bits(bit_sizes, bit_sizes + count);
        *out = new nx_context(log_n, bits, hamming_weight, sparse_slots);
    });
}

int nx_context_create_hybrid(int log_n, const int *bit_sizes, size_t count, size_t special_primes,
                             uint32_t hamming_weight, size_t sparse_slots, nx_context **out)
{
    return guard([&] {
        need(bit_sizes, "bit_sizes");
        need(out, "out");
        std::vector<int> bits(bit_sizes, bit_sizes + count);
        *out = new nx_context(log_n, bits, hamming_weight, sparse_slots, special_primes);
    });
}

int nx_coeff_modulus_create(size_t n, const int *bit_sizes, size_t count, uint64_t *moduli)
{
    return guard([&] {
        need(bit_sizes, "bit_sizes");
        need(moduli, "moduli");
        const std::vector<uint64_t> q = nx::coeff_modulus_create(n, std::vector<int>(bit_sizes, bit_sizes + count));
        for (size_t i = 0; i < q.size(); i++) moduli[i] = q[i];
    });
}

int nx_context_bind_lane(nx_context *ctx, size_t lane)
{
    return guard([&] {
        need(ctx, "ctx");
        ctx->impl.bind_lane(lane);
    });
}

int nx_context_key_info(const nx_context *ctx, size_t *special_primes, size_t *digits, size_t *key_words)
{
    return guard([&] {
        need(ctx, "ctx");
        if (special_primes) *special_primes = ctx->impl.alpha;
        if (digits) *digits = ctx->impl.dnum;
        if (key_words) *key_words = ctx->impl.key_words();
    });
}

int nx_context_destroy(nx_context *ctx)
{
    return guard([&] { delete ctx; });
}

int nx_context_info(const nx_context *ctx, int *log_n, size_t *data_primes, size_t *slot_count)
{
    return guard([&] {
        need(ctx, "ctx");
        if (log_n) *log_n = ctx->impl.logn;
        if (data_primes) *data_primes = ctx->impl.L;
        if (slot_count) *slot_count = ctx->impl.n / 2;
    });
}

int nx_context_moduli(const nx_context *ctx, uint64_t *moduli)
{
    return guard([&] {
        need(ctx, "ctx");
        need(moduli, "moduli");
        std::memcpy(moduli, ctx->impl.moduli.data(), ctx->impl.moduli.size() * 8);
    });
}

int nx_context_synchronize(nx_context *ctx)
{
    return guard([&] {
        need(ctx, "ctx");
        NX_HIP(hipStreamSynchronize(ctx->impl.stream));
    });
}

int nx_autotune(nx_context *ctx, int reps, char *report, size_t report_len, size_t *needed)
{
    return guard([&] {
        need(ctx, "ctx");
        const std::string r = nx::autotune(ctx->impl, reps);
        if (needed) *needed = r.size() + 1;
        if (report && report_len) {
            std::strncpy(report, r.c_str(), report_len - 1);
            report[report_len - 1] = 0;
        }
    });
}

int nx_galois_elt_from_step(const nx_context *ctx, int step, uint32_t *elt)
{
    return guard([&] {
        need(ctx, "ctx");
        need(elt, "elt");
        *elt = ctx->impl.elt_from_step(step);
    });
}

int nx_context_memory(const nx_context *ctx, size_t *in_use, size_t *cached)
{
    return guard([&] {
        need(ctx, "ctx");
        if (in_use) *in_use = ctx->impl.cache->bytes_in_use();
        if (cached) *cached = ctx->impl.cache->bytes_cached();
    });
}

int nx_context_trim(nx_context *ctx)
{
    return guard([&] {
        need(ctx, "ctx");
        ctx->impl.trim_caches();
    });
}

// ---------------------------------------------------------------- ciphertexts
int nx_ciphertext_create(nx_context *ctx, nx_ciphertext **out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        auto *c = new nx_ciphertext();
        c->impl.ctx = &ctx->impl;
        *out = c;
    });
}

int nx_ciphertext_destroy(nx_ciphertext *ct)
{
    return guard([&] { delete ct; });
}

int nx_ciphertext_copy(nx_ciphertext *dst, const nx_ciphertext *src)
{
    return guard([&] {
        need(dst, "dst");
        need(src, "src");
        nx::ct_copy(dst->impl, src->impl);
    });
}

int nx_ciphertext_upload(nx_ciphertext *ct, const uint64_t *data, size_t size, size_t chain_index, double scale,
                         int ntt_form)
{
    return guard([&] {
        need(ct, "ct");
        need(data, "data");
        nx::Context &c = ctx_of(ct->impl);
        nx::require(size >= 1 && size <= 3, "ciphertext size must be 1..3");
        const size_t limbs = c.limbs_of(chain_index);
        nx::ct_resize(ct->impl, &c, size, limbs);
        ct->impl.scale = scale;
        ct->impl.ntt = ntt_form != 0;
        NX_HIP(hipMemcpyAsync(ct->impl.data.get(), data, size * limbs * c.n * 8, hipMemcpyHostToDevice, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
    });
}

int nx_ciphertext_download(const nx_ciphertext *ct, uint64_t *data)
{
    return guard([&] {
        need(ct, "ct");
        need(data, "data");
        nx::Context &c = ctx_of(ct->impl);
        NX_HIP(hipMemcpyAsync(data, ct->impl.data.get(), ct->impl.size * ct->impl.limbs * c.n * 8,
                              hipMemcpyDeviceToHost, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
    });
}

int nx_ciphertext_info(const nx_ciphertext *ct, size_t *size, size_t *chain_index, size_t *coeff_modulus_size,
                       double *scale, int *ntt_form)
{
    return guard([&] {
        need(ct, "ct");
        const nx::Ciphertext &c = ct->impl;
        if (size) *size = c.size;
        if (chain_index) *chain_index = c.limbs ? c.ctx->chain_index_of(c.limbs) : 0;
        if (coeff_modulus_size) *coeff_modulus_size = c.limbs;
        if (scale) *scale = c.scale;
        if (ntt_form) *ntt_form = c.ntt ? 1 : 0;
    });
}

int nx_ciphertext_device_data(const nx_ciphertext *ct, uint64_t **data, size_t *words)
{
    return guard([&] {
        need(ct, "ct");
        need(data, "data");
        const nx::Ciphertext &c = ct->impl;
        *data = c.data.get();
        if (words) *words = c.size * c.limbs * (c.ctx ? c.ctx->n : 0);
    });
}

int nx_ciphertext_set_scale(nx_ciphertext *ct, double scale)
{
    return guard([&] {
        need(ct, "ct");
        ct->impl.scale = scale;
    });
}

// ---------------------------------------------------------------- plaintexts
int nx_plaintext_create(nx_context *ctx, nx_plaintext **out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        auto *p = new nx_plaintext();
        p->impl.ctx = &ctx->impl;
        *out = p;
    });
}

int nx_plaintext_destroy(nx_plaintext *pt)
{
    return guard([&] { delete pt; });
}

int nx_plaintext_copy(nx_plaintext *dst, const nx_plaintext *src)
{
    return guard([&] {
        need(dst, "dst");
        need(src, "src");
        nx::pt_copy(dst->impl, src->impl);
    });
}

int nx_plaintext_upload(nx_plaintext *pt, const uint64_t *data, size_t chain_index, double scale, int ntt_form)
{
    return guard([&] {
        need(pt, "pt");
        need(data, "data");
        nx::require(pt->impl.ctx != nullptr, "plaintext is not initialised");
        nx::Context &c = *pt->impl.ctx;
        const size_t limbs = c.limbs_of(chain_index);
        pt->impl.limbs = limbs;
        pt->impl.scale = scale;
        pt->impl.ntt = ntt_form != 0;
        pt->impl.data.reset(&c, limbs * c.n);
        NX_HIP(hipMemcpyAsync(pt->impl.data.get(), data, limbs * c.n * 8, hipMemcpyHostToDevice, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
    });
}

int nx_plaintext_download(const nx_plaintext *pt, uint64_t *data)
{
    return guard([&] {
        need(pt, "pt");
        need(data, "data");
        nx::Context &c = *pt->impl.ctx;
        NX_HIP(hipMemcpyAsync(data, pt->impl.data.get(), pt->impl.limbs * c.n * 8, hipMemcpyDeviceToHost, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
    });
}

int nx_plaintext_info(const nx_plaintext *pt, size_t *chain_index, size_t *coeff_modulus_size, double *scale)
{
    return guard([&] {
        need(pt, "pt");
        const nx::Plaintext &p = pt->impl;
        if (chain_index) *chain_index = p.limbs ? p.ctx->chain_index_of(p.limbs) : 0;
        if (coeff_modulus_size) *coeff_modulus_size = p.limbs;
        if (scale) *scale = p.scale;
    });
}

int nx_plaintext_set_scale(nx_plaintext *pt, double scale)
{
    return guard([&] {
        need(pt, "pt");
        pt->impl.scale = scale;
    });
}

// ---------------------------------------------------------------- keys
int nx_secret_key_generate(nx_context *ctx, uint64_t seed, nx_secret_key **out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        auto *k = new nx_secret_key();
        try {
            nx::secret_key_generate(ctx->impl, seed, k->impl);
        } catch (...) {
            delete k;
            throw;
        }
        *out = k;
    });
}

int nx_secret_key_from_coeffs(nx_context *ctx, const int8_t *coeffs, nx_secret_key **out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(coeffs, "coeffs");
        need(out, "out");
        auto *k = new nx_secret_key();
        try {
            nx::secret_key_from_coeffs(ctx->impl, coeffs, k->impl);
        } catch (...) {
            delete k;
            throw;
        }
        *out = k;
    });
}

int nx_secret_key_coeffs(const nx_secret_key *sk, int8_t *coeffs)
{
    return guard([&] {
        need(sk, "sk");
        need(coeffs, "coeffs");
        std::memcpy(coeffs, sk->impl.coeff.data(), sk->impl.coeff.size());
    });
}

int nx_secret_key_destroy(nx_secret_key *sk)
{
    return guard([&] { delete sk; });
}

int nx_kswitch_keys_create(nx_context *ctx, nx_kswitch_keys **out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        auto *k = new nx_kswitch_keys();
        k->impl.ctx = &ctx->impl;
        *out = k;
    });
}

int nx_kswitch_keys_destroy(nx_kswitch_keys *keys)
{
    return guard([&] { delete keys; });
}

int nx_relin_key_generate(nx_context *ctx, const nx_secret_key *sk, uint64_t seed, nx_kswitch_keys *keys)
{
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(keys, "keys");
        nx::relin_key_generate(ctx->impl, sk->impl, seed, keys->impl);
    });
}

int nx_galois_keys_generate(nx_context *ctx, const nx_secret_key *sk, const uint32_t *elts, size_t count,
                            uint64_t seed, nx_kswitch_keys *keys)
{
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(keys, "keys");
        need(elts, "elts");
        nx::galois_keys_generate(ctx->impl, sk->impl, elts, count, seed, keys->impl);
    });
}

int nx_kswitch_key_upload(nx_kswitch_keys *keys, uint32_t elt, const uint64_t *data)
{
    return guard([&] {
        need(keys, "keys");
        need(data, "data");
        nx::Context &c = *keys->impl.ctx;
        const size_t words = c.key_words();
        nx::DevBuf &b = keys->impl.keys[elt];
        keys->impl.seeds.erase(elt);  // arbitrary words: the uniform half is read, not regenerated
        b.reset(&c, words);
        NX_HIP(hipMemcpyAsync(b.get(), data, words * 8, hipMemcpyHostToDevice, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
    });
}

int nx_kswitch_key_download(const nx_kswitch_keys *keys, uint32_t elt, uint64_t *data)
{
    return guard([&] {
        need(keys, "keys");
        need(data, "data");
        auto it = keys->impl.keys.find(elt);
        nx::require(it != keys->impl.keys.end(), "key not present");
        nx::Context &c = *keys->impl.ctx;
        NX_HIP(hipMemcpyAsync(data, it->second.get(), it->second.words() * 8, hipMemcpyDeviceToHost, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
    });
}

int nx_kswitch_keys_has(const nx_kswitch_keys *keys, uint32_t elt, int *present)
{
    return guard([&] {
        need(keys, "keys");
        need(present, "present");
        *present = keys->impl.keys.count(elt) ? 1 : 0;
    });
}

int nx_encrypt_symmetric(nx_context *ctx, const nx_secret_key *sk, const nx_plaintext *pt, uint64_t seed,
                         nx_ciphertext *out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(pt, "pt");
        need(out, "out");
        nx::encrypt_symmetric(ctx->impl, sk->impl, pt->impl, seed, out->impl);
    });
}

int nx_decrypt(nx_context *ctx, const nx_secret_key *sk, const nx_ciphertext *ct, nx_plaintext *out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(sk, "sk");
        need(ct, "ct");
        need(out, "out");
        nx::decrypt(ctx->impl, sk->impl, ct->impl, out->impl);
    });
}

// ---------------------------------------------------------------- encoder
int nx_encode(nx_context *ctx, const double *re, const double *im, size_t count, size_t chain_index, double scale,
              nx_plaintext *out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        nx::require(re != nullptr || count == 0, "values cannot be null");
        nx::encode_vector(ctx->impl, re, im, count, ctx->impl.limbs_of(chain_index), scale, out->impl);
    });
}

int nx_encode_constant(nx_context *ctx, double value, size_t chain_index, double scale, nx_plaintext *out)
{
    return guard([&] {
        need(ctx, "ctx");
        need(out, "out");
        nx::encode_constant(ctx->impl, value, ctx->impl.limbs_of(chain_index), scale, out->impl);
    });
}

int nx_decode(nx_context *ctx, const nx_plaintext *pt, double *re, double *im, size_t count)
{
    return guard([&] {
        need(ctx, "ctx");
        need(pt, "pt");
        need(re, "re");
        nx::decode(ctx->impl, pt->impl, re, im, count);
    });
}

// ---------------------------------------------------------------- evaluator
#define NX_BIN(name, fn)                                                                \
    int name(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest)      \
    {                                                                                   \
        return guard([&] {                                                              \
            need(a, "a");                                                               \
            need(b, "b");                                                               \
            need(dest, "dest");                                                         \
            nx::fn(a->impl, b->impl, dest->impl);                                       \
        });                                                                             \
    }
NX_BIN(nx_add, op_add)
NX_BIN(nx_sub, op_sub)
NX_BIN(nx_multiply, op_multiply)
#undef NX_BIN

#define NX_PLAIN(name, fn)                                                              \
    int name(const nx_ciphertext *a, const nx_plaintext *p, nx_ciphertext *dest)       \
    {                                                                                   \
        return guard([&] {                                                              \
            need(a, "a");                                                               \
            need(p, "p");                                                               \
            need(dest, "dest");                                                         \
            nx::fn(a->impl, p->impl, dest->impl);                                       \
        });                                                                             \
    }
NX_PLAIN(nx_add_plain, op_add_plain)
NX_PLAIN(nx_sub_plain, op_sub_plain)
NX_PLAIN(nx_multiply_plain, op_multiply_plain)
#undef NX_PLAIN

#define NX_UN(name, fn)                                                                 \
    int name(const nx_ciphertext *a, nx_ciphertext *dest)                               \
    {                                                                                   \
        return guard([&] {                                                              \
            need(a, "a");                                                               \
            need(dest, "dest");                                                         \
            nx::fn(a->impl, dest->impl);                                                \
        });                                                                             \
    }
NX_UN(nx_negate, op_negate)
NX_UN(nx_square, op_square)
NX_UN(nx_rescale_to_next, op_rescale_to_next)
NX_UN(nx_mod_switch_to_next, op_mod_switch_to_next)
NX_UN(nx_transform_to_ntt, op_to_ntt)
NX_UN(nx_transform_from_ntt, op_from_ntt)
#undef NX_UN

int nx_rescale_to_next_many(const nx_ciphertext *const *cts, size_t count, nx_ciphertext *const *dests)
{
    return guard([&] {
        need(cts, "cts");
        need(dests, "dests");
        std::vector<const nx::Ciphertext *> in;
        std::vector<nx::Ciphertext *> out;
        for (size_t i = 0; i < count; i++) {
            need(cts[i], "cts[i]");
            need(dests[i], "dests[i]");
            in.push_back(&cts[i]->impl);
            out.push_back(&dests[i]->impl);
        }
        nx::op_rescale_many(in, out);
    });
}

static int multiply_relin_rescale_many(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count,
                                       const nx_kswitch_keys *relin_keys, nx_ciphertext *const *dests, bool fused)
{
    return guard([&] {
        need(a, "a");
        need(b, "b");
        need(relin_keys, "relin_keys");
        need(dests, "dests");
        std::vector<const nx::Ciphertext *> x, y;
        std::vector<nx::Ciphertext *> out;
        for (size_t i = 0; i < count; i++) {
            need(a[i], "a[i]");
            need(b[i], "b[i]");
            need(dests[i], "dests[i]");
            x.push_back(&a[i]->impl);
            y.push_back(&b[i]->impl);
            out.push_back(&dests[i]->impl);
        }
        if (fused)
            nx::op_multiply_relin_rescale_fused_many(x, y, relin_keys->impl, out);
        else
            nx::op_multiply_relin_rescale_many(x, y, relin_keys->impl, out);
    });
}

int nx_multiply_relin_rescale_many(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count,
                                   const nx_kswitch_keys *relin_keys, nx_ciphertext *const *dests)
{
    return multiply_relin_rescale_many(a, b, count, relin_keys, dests, false);
}

int nx_multiply_relin_rescale_fused_many(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count,
                                         const nx_kswitch_keys *relin_keys, nx_ciphertext *const *dests)
{
    return multiply_relin_rescale_many(a, b, count, relin_keys, dests, true);
}

int nx_multiply_sum(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(b, "b");
        need(dest, "dest");
        std::vector<const nx::Ciphertext *> x, y;
        for (size_t i = 0; i < count; i++) {
            need(a[i], "a[i]");
            need(b[i], "b[i]");
            x.push_back(&a[i]->impl);
            y.push_back(&b[i]->impl);
        }
        nx::op_multiply_sum(x, y, dest->impl);
    });
}

int nx_relinearize(const nx_ciphertext *a, const nx_kswitch_keys *relin_keys, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(relin_keys, "relin_keys");
        need(dest, "dest");
        nx::op_relinearize(a->impl, relin_keys->impl, dest->impl);
    });
}

int nx_mod_switch_to(const nx_ciphertext *a, size_t chain_index, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(dest, "dest");
        nx::op_mod_switch_to(a->impl, chain_index, dest->impl);
    });
}

int nx_plain_mod_switch_to(const nx_plaintext *a, size_t chain_index, nx_plaintext *dest)
{
    return guard([&] {
        need(a, "a");
        need(dest, "dest");
        nx::op_plain_mod_switch_to(a->impl, chain_index, dest->impl);
    });
}

int nx_apply_galois(const nx_ciphertext *a, uint32_t elt, const nx_kswitch_keys *galois_keys, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(galois_keys, "galois_keys");
        need(dest, "dest");
        nx::op_apply_galois(a->impl, elt, galois_keys->impl, dest->impl);
    });
}

int nx_rotate_vector(const nx_ciphertext *a, int steps, const nx_kswitch_keys *galois_keys, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(galois_keys, "galois_keys");
        need(dest, "dest");
        nx::op_rotate(a->impl, steps, galois_keys->impl, dest->impl);
    });
}

int nx_complex_conjugate(const nx_ciphertext *a, const nx_kswitch_keys *galois_keys, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(galois_keys, "galois_keys");
        need(dest, "dest");
        nx::op_conjugate(a->impl, galois_keys->impl, dest->impl);
    });
}

int nx_add_many(const nx_ciphertext *const *cts, size_t count, nx_ciphertext *dest)
{
    return guard([&] {
        need(cts, "cts");
        need(dest, "dest");
        std::vector<const nx::Ciphertext *> v;
        for (size_t i = 0; i < count; i++) {
            need(cts[i], "cts[i]");
            v.push_back(&cts[i]->impl);
        }
        nx::op_add_many(v, dest->impl);
    });
}

int nx_multiply_const(const nx_ciphertext *a, double value, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(dest, "dest");
        nx::Context &c = ctx_of(a->impl);
        nx::Plaintext p;
        nx::encode_constant(c, value, a->impl.limbs, a->impl.scale, p);
        nx::op_multiply_plain(a->impl, p, dest->impl);
    });
}

int nx_add_const(const nx_ciphertext *a, double value, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(dest, "dest");
        nx::Context &c = ctx_of(a->impl);
        nx::Plaintext p;
        nx::encode_constant(c, value, a->impl.limbs, a->impl.scale, p);
        nx::op_add_plain(a->impl, p, dest->impl);
    });
}

static int reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, nx::ReducedOp op,
                         const nx_kswitch_keys *rk, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "encrypted1");
        need(b, "encrypted2");
        need(dest, "destination");
        if (op == nx::ReducedOp::Mul) need(rk, "relin_keys");
        nx::op_reduced_error(a->impl, b->impl, op, rk ? &rk->impl : nullptr, dest->impl);
    });
}

int nx_add_reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest)
{
    return reduced_error(a, b, nx::ReducedOp::Add, nullptr, dest);
}

int nx_sub_reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest)
{
    return reduced_error(a, b, nx::ReducedOp::Sub, nullptr, dest);
}

int nx_multiply_reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, const nx_kswitch_keys *relin_keys,
                              nx_ciphertext *dest)
{
    return reduced_error(a, b, nx::ReducedOp::Mul, relin_keys, dest);
}

int nx_multiply_vector_reduced_error(const nx_ciphertext *a, const double *re, const double *im, size_t count,
                                     nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "encrypted");
        need(re, "values");
        need(dest, "destination");
        nx::Context &c = ctx_of(a->impl);
        nx::Plaintext p;
        nx::encode_vector(c, re, im, count, a->impl.limbs, a->impl.scale, p);
        nx::op_multiply_plain(a->impl, p, dest->impl);
    });
}

int nx_double(const nx_ciphertext *a, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "encrypted");
        need(dest, "destination");
        nx::op_add(a->impl, a->impl, dest->impl);
    });
}

int nx_rescale_to_next_noalloc(nx_context *ctx, const uint64_t *in, size_t polys, size_t limbs, uint64_t *out)
{
    return guard([&] {
        need(ctx, "ctx");
        nx::rescale_noalloc(ctx->impl, in, polys, limbs, out);
    });
}

int nx_multiply_power_of_x(const nx_ciphertext *a, int64_t exponent, nx_ciphertext *dest)
{
    return guard([&] {
        need(a, "a");
        need(dest, "dest");
        nx::op_multiply_power_of_x(a->impl, exponent, dest->impl);
    });
}

}  // extern "C"
