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
/* Per-op mode (on = 1): the programs over this evaluator -- the evaluators,
 * the bootstrapper, the BERT layer and the server lanes created while it is on
 * -- issue only single-ciphertext operations, one library call per term, as
 * the reference's evaluators call Phantom (e.g. cuda/src/gelu.cu:8-117,
 * cuda/src/bootstrapping/Bootstrapper.cu:1855-1919: rotate_vector,
 * multiply_plain, add, multiply, relinearize, rescale_to_next): no hoisting,
 * no sums of rotations with one ModDown, no batched products, multiply-
 * accumulates, linear combinations or block mixes, no lockstep GELU or
 * EvalMod.  The results decrypt like the default (batched) mode's.  No
 * counterpart in the reference: it is the plain per-call drop-in the batched
 * engine is measured against. */
int nx_evaluator_set_per_op(nx_evaluator *ev, int on);

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
/* gelu of count inputs of one level and scale in lockstep (each
 * multiplication, key switch and rescale one batched launch); out[i] equals
 * nx_eval_gelu(x[i]) word for word.  No counterpart in the reference. */
int nx_eval_gelu_many(nx_evaluator *ev, const nx_ciphertext *const *x, size_t count, nx_ciphertext *const *out);
/* LNEvaluator::layer_norm (src/layer_norm.cpp:6-33) */
int nx_eval_layer_norm(nx_evaluator *ev, const nx_ciphertext *x, int len, nx_ciphertext *out);
/* SoftmaxEvaluator::softmax (src/softmax.cpp:6-42) */
int nx_eval_softmax(nx_evaluator *ev, const nx_ciphertext *x, int len, nx_ciphertext *out);

/* ---- bootstrapping (reference Bootstrapper, cuda/src/bootstrapping/Bootstrapper.cu;
 * src/bootstrapping/Bootstrapper.cpp).  nx_bootstrapper_create plays the
 * constructor + prepare_mod_polynomial + generate_LT_coefficient_3
 * (cuda/src/bootstrapping.cu:78-115): log_slots (reference logn; 0 = N/2
 * slots), K (boundary_K), degree of the EvalMod interpolant (sin_cos_deg),
 * double_angles (scale_factor), final_scale (0 = the input's scale).  The
 * evaluator must hold the relinearisation key and Galois keys for every step
 * of nx_bootstrapper_rotation_steps. */
typedef struct nx_bootstrapper nx_bootstrapper;
int nx_bootstrapper_create(nx_evaluator *ev, int log_slots, int K, int degree, int double_angles, int log_width,
                           double final_scale, nx_bootstrapper **out);
int nx_bootstrapper_destroy(nx_bootstrapper *bt);
/* addLeftRotKeys_Linear_to_vector_3 (Bootstrapper.cu:82-178): rotation steps
 * (0 = conjugation); *count receives the number, at most `capacity` written */
int nx_bootstrapper_rotation_steps(const nx_bootstrapper *bt, int *steps, size_t capacity, size_t *count);
/* levels consumed from the top data level */
int nx_bootstrapper_depth(const nx_bootstrapper *bt, int *depth);
/* bootstrap_3 (Bootstrapper.cu:3358-3370): a ciphertext at any level (it is
 * switched to the lowest first) -> fresh ciphertext `depth` levels below the top */
int nx_bootstrap(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out);
/* the same with the output holding factor x the input message (the factor is
 * folded into the last SlotToCoeff level: no extra level; the bootstrapping
 * error is scaled by it).  No counterpart in the reference. */
int nx_bootstrap_scaled(nx_bootstrapper *bt, const nx_ciphertext *in, double factor, nx_ciphertext *out);
/* out[i] = bootstrap of in[i] times factors[i] (factors may be NULL: all 1)
 * for i < count: every CoeffToSlot / SlotToCoeff level batched over the
 * inputs, one EvalMod over every input in lockstep (each multiplication, key
 * switch and rescale one batched launch): nx_bootstrap_many_mixed with bt for
 * every input.  Decrypts like count nx_bootstrap_scaled calls.  No
 * counterpart in the reference, which bootstraps one at a time. */
int nx_bootstrap_many(nx_bootstrapper *bt, const nx_ciphertext *const *in, size_t count, const double *factors,
                      nx_ciphertext *const *out);
/* Bootstrapping of a batch with one bootstrapper per input (of one context
 * and EvalMod parameters, any slot counts: dense and sparse packings mixed,
 * inputs at any level): out[i] = factors[i] (NULL: 1) x in[i]'s message
 * through bts[i]'s transforms (bootstrap_full_3 / bootstrap_sparse_3,
 * cuda/src/bootstrapping/Bootstrapper.cu:2987-3113).  Every CoeffToSlot and
 * SlotToCoeff level runs once over the inputs of each bootstrapper (baby
 * rotations, ModDowns and rescales batched), one EvalMod over every input's
 * halves.  Word for word each input's nx_bootstrap_scaled with one special
 * prime.  No counterpart in the reference, which bootstraps one at a time. */
int nx_bootstrap_many_mixed(nx_bootstrapper *const *bts, const nx_ciphertext *const *in, size_t count,
                            const double *factors, nx_ciphertext *const *out);
/* stages of bootstrap_sparse_3 / bootstrap_full_3 (Bootstrapper.cu:2987-3113) */
int nx_boot_mod_raise(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out);
int nx_boot_sub_sum(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out);
/* u_im is written for dense packing only (may be null for sparse) */
int nx_boot_coeff_to_slot(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *u_re, nx_ciphertext *u_im);
int nx_boot_eval_mod(nx_bootstrapper *bt, const nx_ciphertext *in, nx_ciphertext *out);
/* f_im: dense packing only (null for sparse); input_scale = scale of the bootstrapped input */
int nx_boot_slot_to_coeff(nx_bootstrapper *bt, const nx_ciphertext *f_re, const nx_ciphertext *f_im,
                          double input_scale, nx_ciphertext *out);

/* ArgmaxEvaluator::argmax (src/argmax.cpp:9-64): one-hot indicator of the
 * maximum of the first len (power of two) slots, bootstrapping once per
 * round; x is left unchanged */
int nx_eval_argmax(nx_bootstrapper *bt, const nx_ciphertext *x, int len, nx_ciphertext *out);

/* ---- serving: a request loop over the evaluator programs.  Workers, one per
 * lane of the context (nx_context_bind_lane), take requests in arrival order as
 * soon as they are free and the request's footprint (6 x its input ciphertext
 * bytes) fits in memory_budget (0 = no limit); requests on different lanes
 * overlap on the device and share the evaluator's keys.  boot (optional) gives
 * the parameters of bootstrap jobs.  No counterpart in the reference (its
 * drivers run one call at a time, src/main.cpp). */
typedef struct nx_server nx_server;
enum { NX_JOB_GELU = 0, NX_JOB_SOFTMAX = 1, NX_JOB_LAYER_NORM = 2, NX_JOB_BOOTSTRAP = 3, NX_JOB_BERT = 4 };
int nx_server_create(nx_evaluator *ev, int lanes, size_t memory_budget, const nx_bootstrapper *boot,
                     nx_server **out);
/* BERT layer calls in flight at once, 1..lanes (default 1): each call's work
 * overlaps on the device with the others'.  No counterpart in the reference. */
int nx_server_set_bert_concurrency(nx_server *s, int n);
int nx_server_destroy(nx_server *s);
/* a server that also runs BERT-layer requests (nx_server_submit_bert) with
 * the layer's weights shared by every lane.  pack != 0: continuous batching
 * -- a lane that takes a BERT request packs the oldest queued BERT requests
 * whose rows still fit into the same set of ciphertexts (token rotations),
 * runs one layer call with attention kept inside each sequence and returns
 * each request's rows moved back to [0, len) with the others cleared (one
 * level); a request that runs alone returns the layer output as is.
 * bert_footprint: admission bytes of one layer call (0 = 8 GiB). */
typedef struct nx_bert_layer nx_bert_layer;
int nx_server_create_bert(nx_evaluator *ev, int lanes, size_t memory_budget, const nx_bootstrapper *boot,
                          const nx_bert_layer *layer, int pack, size_t bert_footprint, nx_server **out);
/* one request = count jobs (kind, row length for softmax / LayerNorm, input);
 * the inputs are copied */
int nx_server_submit(nx_server *s, const int *kinds, const int *lens, const nx_ciphertext *const *inputs,
                     size_t count, uint64_t *request_id);
/* one BERT-layer request: a sequence encrypted at rows [0, seq_len) of n =
 * hidden / 256 ciphertexts (the inputs are copied); its wait yields n outputs */
int nx_server_submit_bert(nx_server *s, const nx_ciphertext *const *x, size_t n, int seq_len, uint64_t *request_id);
/* one BERT-layer request over a set the caller packed: `segments` sequences at
 * rows [starts[k], starts[k] + lens[k]) of the n ciphertexts (attention kept
 * inside each); one layer call, never packed with other requests; its wait
 * yields the n layer outputs.  No counterpart in the reference. */
int nx_server_submit_bert_set(nx_server *s, const nx_ciphertext *const *x, size_t n, const int *starts,
                              const int *lens, size_t segments, uint64_t *request_id);
/* blocks until the request is done; count outputs, one per job */
int nx_server_wait(nx_server *s, uint64_t request_id, nx_ciphertext *const *outputs, size_t count);
int nx_server_drain(nx_server *s);
/* paused != 0: admit nothing until resumed (requests submitted meanwhile are
 * batched together afterwards) */
int nx_server_pause(nx_server *s, int paused);
/* latency from submission to completion on the device */
int nx_server_stats(const nx_server *s, size_t *completed, double *p50_ms, double *p99_ms, double *mean_ms,
                    size_t *peak_concurrent);
/* layer calls that carried more than one BERT request, and those requests */
int nx_server_batching_stats(const nx_server *s, size_t *packed_calls, size_t *packed_requests, size_t *peak_footprint);

/* ---- encrypted BERT-base encoder layer (include/nexus/bert.hpp).  NEXUS ships
 * the pieces (MatrixMul with plaintext weights, GELU src/gelu.cpp, softmax
 * src/softmax.cpp, LayerNorm src/layer_norm.cpp, Bootstrapper) but no
 * composed layer.  Token matrices are packed in token groups: token i owns
 * slots [256 i, 256 i + 256) of each of the features / 256 ciphertexts, and
 * ciphertext c holds features [256 c, 256 c + 256).  Weights are row-major
 * [in][out] float64, applied as x @ W: wq, wk, wv, wo hidden x hidden, w1
 * hidden x ffn, w2 ffn x hidden; hidden = 64 heads; no biases.  The chain needs
 * 19 compute levels below the bootstrapper's. */
enum { NX_BERT_WQ = 0, NX_BERT_WK, NX_BERT_WV, NX_BERT_WO, NX_BERT_W1, NX_BERT_W2 };
int nx_bert_layer_create(nx_bootstrapper *bt, int hidden, int heads, int ffn, const double *wq, const double *wk,
                         const double *wv, const double *wo, const double *w1, const double *w2, int baby,
                         nx_bert_layer **out);
int nx_bert_layer_destroy(nx_bert_layer *layer);
/* rotation steps the layer needs besides the bootstrapper's (host only) */
int nx_bert_layer_rotation_steps(size_t slots, int hidden, int heads, int ffn, int baby, int *steps, size_t capacity,
                                 size_t *count);
/* x: hidden / 256 ciphertexts (>= 2 limbs, the evaluator's scale, rows >=
 * seq_len zero) -> the layer output, hidden / 256 ciphertexts at 2 limbs */
int nx_bert_layer_forward(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t n, int seq_len,
                          nx_ciphertext *const *out);
/* several sequences packed in one set of ciphertexts: segment k is rows
 * [starts[k], starts[k] + lens[k]); attention stays within each segment, so
 * each sequence's rows equal what it would get alone; other rows must be zero */
int nx_bert_layer_forward_packed(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t n, const int *starts,
                                 const int *lens, size_t segments, nx_ciphertext *const *out);
/* `sets` packed sets in lockstep (x and out: sets * n handles, set s at
 * [s n, s n + n)); set s has seg_counts[s] segments, their starts and lengths
 * listed set after set.  Every stage runs over all sets, and each
 * bootstrapping step refreshes every set's ciphertexts together (real
 * messages paired across sets).  No counterpart in the reference. */
int nx_bert_layer_forward_sets(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t sets, size_t n,
                               const int *seg_counts, const int *starts, const int *lens, nx_ciphertext *const *out);
/* stages: factor * x W (one limb out); attention of bootstrapped q (already
 * scaled by 1/8), k, v; x * invert_sqrt(mean(x^2)) per token */
int nx_bert_layer_linear(nx_bert_layer *layer, int matrix, const nx_ciphertext *const *x, size_t n_in, double factor,
                         nx_ciphertext *const *out, size_t n_out);
int nx_bert_layer_attention(nx_bert_layer *layer, const nx_ciphertext *const *q, const nx_ciphertext *const *k,
                            const nx_ciphertext *const *v, size_t n, int seq_len, nx_ciphertext *const *out);
int nx_bert_layer_layer_norm(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t n,
                             nx_ciphertext *const *out);
/* stage timing: when on, forward() synchronises after each stage; the last
 * forward's stages are read back as comma-separated names and milliseconds */
int nx_bert_layer_set_stage_timing(nx_bert_layer *layer, int on);
int nx_bert_layer_stage_ms(const nx_bert_layer *layer, char *names, size_t names_len, double *ms, size_t capacity,
                           size_t *count);

/* Host-only view of the linear-transform plan (genfftcoeff_3 / geninvfftcoeff_3,
 * Bootstrapper.cu:1137,1545); no device use.  which = 0: CoeffToSlot, 1:
 * SlotToCoeff; level in application order.  A level computes
 * y = sum_(j,i) rot(D_(j,i) (.) rot(x, i*stride), g*j*stride) over N/2 slots. */
int nx_boot_plan_info(int log_half_n, int log_slots, double K, int which, int level, int *levels, int *stride, int *g,
                      size_t *count);
/* diagonals: js[count], is[count], re/im[count][N/2] */
int nx_boot_plan_diagonals(int log_half_n, int log_slots, double K, int which, int level, int *js, int *is,
                           double *re, double *im);
/* EvalMod polynomial (Chebyshev basis on [-1, 1], degree+1 coefficients) and
 * the levels its evaluation consumes */
int nx_boot_eval_mod_coeffs(int K, int degree, int double_angles, int log_width, double *coeffs, int *depth);
/* rotation steps of a plan (0 = conjugation) */
int nx_boot_plan_steps(int log_half_n, int log_slots, double K, int *steps, size_t capacity, size_t *count);

#ifdef __cplusplus
}
#endif

#endif /* NEXUS_EVAL_H */
