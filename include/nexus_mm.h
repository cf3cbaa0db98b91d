/*
 * nexus_mm.h -- C ABI of the ciphertext-plaintext MatrixMul (reference
 * MMEvaluator, src/matrix_mul.{h,cpp}; CUDA mirror cuda/src/matrix_mul.cu).
 */
#ifndef NEXUS_MM_H
#define NEXUS_MM_H

#include "nexus_ckks.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A device-resident batch of ciphertexts of one shape (the expanded operand). */
typedef struct nx_ct_batch nx_ct_batch;

/* MMEvaluator::enc_compress_ciphertext (src/matrix_mul.cpp:58-99): coefficient i
 * of the plaintext is round(values[i] * scale); encrypted at the top level. */
int nx_mm_compress(nx_context *ctx, const nx_secret_key *sk, const double *values, size_t count, double scale,
                   uint64_t seed, nx_ciphertext *out);
/* MMEvaluator::expand_ciphertext (src/matrix_mul.cpp:101-129): m ciphertexts, the
 * k-th encrypting m * coefficient_k; elts[i] (i < log2 m) are the Galois elements
 * of the rounds (the reference uses n / 2^i + 1). */
int nx_mm_expand(const nx_ciphertext *ct, uint32_t m, const uint32_t *elts, const nx_kswitch_keys *galois_keys,
                 nx_ct_batch **out);
int nx_ct_batch_count(const nx_ct_batch *batch, size_t *count);
int nx_ct_batch_extract(const nx_ct_batch *batch, size_t index, nx_ciphertext *out);
int nx_ct_batch_destroy(nx_ct_batch *batch);
/* MMEvaluator::matrix_mul's product (src/matrix_mul.cpp:213-236): outs[i] =
 * sum_{j < inner} expanded[i * inner + j] * a_pts[j], i < cols, the expanded
 * ciphertexts taken in order across `batches`. */
int nx_mm_accumulate(const nx_plaintext *const *a_pts, size_t inner, const nx_ct_batch *const *batches,
                     size_t n_batches, size_t cols, nx_ciphertext *const *outs);

#ifdef __cplusplus
}
#endif

#endif /* NEXUS_MM_H */
