// C ABI of SEAL-format serialization (include/nexus_ckks.h).
#include <cstring>

#include "../../include/nexus_ckks.h"
#include "handles.h"

namespace {

// copy out with the size query convention: buf == NULL or cap too small -> size only
void emit(const std::vector<uint8_t> &b, uint8_t *buf, size_t cap, size_t *size)
{
    nx::need_ptr(size, "size");
    *size = b.size();
    if (buf) {
        nx::require(cap >= b.size(), "buffer too small", nx::kOutOfRange);
        std::memcpy(buf, b.data(), b.size());
    }
}

}  // namespace

extern "C" {

int nx_hash_blake2b(const uint8_t *data, size_t size, uint8_t *digest, size_t digest_size)
{
    return nx::guard_call([&] {
        if (size) nx::need_ptr(data, "data");
        nx::need_ptr(digest, "digest");
        nx::blake2b(digest, digest_size, data, size);
    });
}

int nx_parms_id(const nx_context *ctx, size_t chain_index, uint64_t *id)
{
    return nx::guard_call([&] {
        nx::need_ptr(ctx, "ctx");
        nx::need_ptr(id, "id");
        const nx::Context &c = ctx->impl;
        nx::parms_id(c, chain_index == 0 ? 0 : c.limbs_of(chain_index), id);
    });
}

int nx_ciphertext_save(const nx_ciphertext *ct, uint8_t *buf, size_t capacity, size_t *size)
{
    return nx::guard_call([&] {
        nx::need_ptr(ct, "ct");
        emit(nx::save_ciphertext(ct->impl), buf, capacity, size);
    });
}

int nx_ciphertext_load(nx_context *ctx, const uint8_t *buf, size_t size, nx_ciphertext *ct)
{
    return nx::guard_call([&] {
        nx::need_ptr(ctx, "ctx");
        nx::need_ptr(buf, "buf");
        nx::need_ptr(ct, "ct");
        nx::load_ciphertext(ctx->impl, buf, size, ct->impl);
    });
}

int nx_plaintext_save(const nx_plaintext *pt, uint8_t *buf, size_t capacity, size_t *size)
{
    return nx::guard_call([&] {
        nx::need_ptr(pt, "pt");
        emit(nx::save_plaintext(pt->impl), buf, capacity, size);
    });
}

int nx_plaintext_load(nx_context *ctx, const uint8_t *buf, size_t size, nx_plaintext *pt)
{
    return nx::guard_call([&] {
        nx::need_ptr(ctx, "ctx");
        nx::need_ptr(buf, "buf");
        nx::need_ptr(pt, "pt");
        nx::load_plaintext(ctx->impl, buf, size, pt->impl);
    });
}

int nx_kswitch_keys_save(const nx_kswitch_keys *keys, int galois, uint8_t *buf, size_t capacity, size_t *size)
{
    return nx::guard_call([&] {
        nx::need_ptr(keys, "keys");
        emit(nx::save_kswitch_keys(keys->impl, galois != 0), buf, capacity, size);
    });
}

int nx_kswitch_keys_load(nx_context *ctx, const uint8_t *buf, size_t size, int galois, nx_kswitch_keys *keys)
{
    return nx::guard_call([&] {
        nx::need_ptr(ctx, "ctx");
        nx::need_ptr(buf, "buf");
        nx::need_ptr(keys, "keys");
        nx::load_kswitch_keys(ctx->impl, buf, size, galois != 0, keys->impl);
    });
}

}  // extern "C"
