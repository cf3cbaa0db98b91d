/*
 * nexus_eval.h -- C ABI of the NEXUS evaluator programs built on the engine
 * (reference src/ckks_evaluator.cpp, src/gelu.cpp, src/layer_norm.cpp,
 * src/softmax.cpp; CUDA mirrors cuda/src/{ckks_evaluator,gelu,layer_norm,
 * softmax}.cu).  An nx_evaluator bundles what NEXUS's CKKSEvaluator holds by
 * pointer: context, relinearisation and Galois keys, the working scale.
 */
#ifndef NEXUS_EVAL_H
#define NEXUS_EVAL_H

#include "nexus_ckks.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nx_evaluator nx_evaluator;

/* CKKSEvaluator(context, encryptor, decryptor, encoder, evaluator, scale,
 * relin_keys, galois_keys) -- src/ckks_evaluator.h:53-79.  sk may be null
 * (then decrypt-based helpers are unavailable). */
int nx_evaluator_create(nx_context *ctx, nx_secret_key *sk, nx_kswitch_keys *relin_keys,
                        nx_kswitch_keys *galois_keys, double scale, nx_evaluator **out);
int nx_evaluator_destroy(nx_evaluator *ev);

/* CKKSEvaluator::sgn_eval (src/ckks_evaluator.cpp:181-212) */
int nx_eval_sgn(nx_evaluator *ev, const nx_ciphertext *x, int d_g, int d_f, double sgn_factor, nx_ciphertext *out);
/* CKKSEvaluator::invert_sqrt (src/ckks_evaluator.cpp:480-485) */
int nx_eval_invert_sqrt(nx_evaluator *ev, const nx_ciphertext *x, int d_newt, int d_gold, nx_ciphertext *out);
/* CKKSEvaluator::exp (src/ckks_evaluator.cpp:534-548) */
int nx_eval_exp(nx_evaluator *ev, const nx_ciphertext *x, nx_ciphertext *out);
/* CKKSEvaluator::inverse (src/ckks_evaluator.cpp:487-509) */
int nx_eval_inverse(nx_evaluator *ev, const nx_ciphertext *x, int iter, nx_ciphertext *out);
/* GeLUEvaluator::gelu (src/gelu.cpp:5-111) */
int nx_eval_gelu(nx_evaluator *ev, const nx_ciphertext *x, nx_ciphertext *out);
/* LNEvaluator::layer_norm (src/layer_norm.cpp:6-33) */
int nx_eval_layer_norm(nx_evaluator *ev, const nx_ciphertext *x, int len, nx_ciphertext *out);
/* SoftmaxEvaluator::softmax (src/softmax.cpp:6-42) */
int nx_eval_softmax(nx_evaluator *ev, const nx_ciphertext *x, int len, nx_ciphertext *out);

#ifdef __cplusplus
}
#endif

#endif /* NEXUS_EVAL_H */
