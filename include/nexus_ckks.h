/*
 * nexus_ckks.h -- C ABI of the MI355X-native CKKS evaluation engine.
 *
 * Drop-in boundary: these entry points replace the operations NEXUS's
 * CKKSEvaluator facade reaches into its HE library for (reference
 * cuda/src/ckks_evaluator.cuh:11-499 over Phantom's include/evaluate.cuh, and
 * src/ckks_evaluator.h:12-103 over SEAL's seal::Evaluator).  Plain pointers and
 * sizes only; every call is stream-ordered on the context's HIP stream and
 * returns NX_OK or a negative status (message in nx_last_error()).
 *
 * Levels are Phantom-style chain indices (reference
 * cuda/thirdparty/phantom-fhe/include/ciphertext.h:193-195): 0 = key level
 * (all primes incl. the special prime), 1 = top data level (L primes), L = one
 * prime left.  Residue arrays are [poly][limb][N] uint64 in limb order q_0..
 * (reference SEAL PolyIter / Phantom ciphertext.h:41,222-228).
 */
#ifndef NEXUS_CKKS_H
#define NEXUS_CKKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nx_context nx_context;
typedef struct nx_ciphertext nx_ciphertext;
typedef struct nx_plaintext nx_plaintext;
typedef struct nx_periodic_plain nx_periodic_plain;
typedef struct nx_secret_key nx_secret_key;
typedef struct nx_kswitch_keys nx_kswitch_keys;

enum {
    NX_OK = 0,
    NX_ERR_INVALID_ARGUMENT = -1, /* std::invalid_argument in the reference */
    NX_ERR_LOGIC = -2,            /* std::logic_error */
    NX_ERR_OUT_OF_RANGE = -3,     /* std::out_of_range */
    NX_ERR_DEVICE = -4,           /* HIP runtime error */
    NX_ERR_NO_MEMORY = -5,
    NX_ERR_UNKNOWN = -9
};

/* Thread-local message of the last failing call on this thread. */
const char *nx_last_error(void);
/* Number of visible GPUs (0 without a GPU). */
int nx_device_count(int *count);
int nx_device_synchronize(void);

/* ---- context: replaces EncryptionParameters + CoeffModulus::Create + SEALContext
 * (src/main.cpp:44-49) / PhantomContext (cuda/src/main.cu:264-270,
 * phantom-fhe/src/context.cu:125-238).  bit_sizes lists data primes then the
 * special prime; hamming_weight 0 = uniform ternary secret
 * (SEAL-4.1-bs encryptionparams.h:188-231); sparse_slots 0 = N/2. */
int nx_context_create(int log_n, const int *bit_sizes, size_t count, uint32_t hamming_weight, size_t sparse_slots,
                      nx_context **out);
/* Hybrid key switching: the last special_primes of bit_sizes are special
 * primes; key-switching digits are special_primes consecutive data primes
 * (Phantom EncryptionParameters::set_special_modulus_size,
 * cuda/thirdparty/phantom-fhe/include/context.cuh; ModUp/ModDown
 * rns_bconv.cu:455-634,776-828).  special_primes = 1 is nx_context_create. */
int nx_context_create_hybrid(int log_n, const int *bit_sizes, size_t count, size_t special_primes,
                             uint32_t hamming_weight, size_t sparse_slots, nx_context **out);
/* special primes, key-switching digits at the top level, words of one key */
int nx_context_key_info(const nx_context *ctx, size_t *special_primes, size_t *digits, size_t *key_words);
int nx_context_destroy(nx_context *ctx);
/* CoeffModulus::Create (SEAL modulus.cpp:143-184, util/numth.cpp:278-310):
 * for each bit size (2..60) the largest primes = 1 mod 2n below 2^bits, the
 * first occurrence of a size taking the largest.  Host only (no GPU needed). */
int nx_coeff_modulus_create(size_t n, const int *bit_sizes, size_t count, uint64_t *moduli);

int nx_context_info(const nx_context *ctx, int *log_n, size_t *data_primes, size_t *slot_count);
/* moduli: L + special words (q_0..q_{L-1}, p_0..) */
int nx_context_moduli(const nx_context *ctx, uint64_t *moduli);
int nx_context_synchronize(nx_context *ctx);
/* Launch-shape autotuner (no counterpart in the reference, which fixes its
 * launch shapes): times the NTT tile widths per pass, direction and batch
 * bucket, and the multiply-accumulate thread layout per limb bucket, at this
 * context's ring size, `reps` launches each.  A candidate is chosen only if its
 * output equals the default shape's word for word.  The choices are
 * process-wide and apply to later launches at this ring size.  report (may be
 * null) receives a JSON summary; *needed its size including the terminator. */
int nx_autotune(nx_context *ctx, int reps, char *report, size_t report_len, size_t *needed);
/* Lanes: bind the calling thread to lane `lane` (1..64) of the context, a HIP
 * stream with its own memory cache; 0 returns the thread to the default stream.
 * Work of threads on different lanes overlaps on the device and shares the
 * context's keys.  A ciphertext written on one lane and read on another must be
 * synchronised first (nx_context_synchronize on the writing thread).  No
 * counterpart in the reference, whose Phantom path runs one global stream
 * (cuda/thirdparty/phantom-fhe/include/util/globals.h). */
int nx_context_bind_lane(nx_context *ctx, size_t lane);
/* Galois element of a rotation step (SEAL util/galois.cpp:53-90, generator 5). */
int nx_galois_elt_from_step(const nx_context *ctx, int step, uint32_t *elt);
/* Device memory held by the context's cache: in use / cached, bytes. */
int nx_context_memory(const nx_context *ctx, size_t *in_use, size_t *cached);
/* Release the cached (free) device blocks of every lane of the context, after
 * a device synchronisation: between workload phases whose working sets differ,
 * so that the next phase does not meet a device filled by the last one's
 * cache.  No counterpart in the reference (its allocator is
 * cuda/thirdparty/phantom-fhe/include/util/cuda_wrapper.cuh). */
int nx_context_trim(nx_context *ctx);

/* ---- ciphertexts / plaintexts (value types with deep copy; reference
 * PhantomCiphertext / seal::Ciphertext) */
int nx_ciphertext_create(nx_context *ctx, nx_ciphertext **out);
int nx_ciphertext_destroy(nx_ciphertext *ct);
int nx_ciphertext_copy(nx_ciphertext *dst, const nx_ciphertext *src);
int nx_ciphertext_upload(nx_ciphertext *ct, const uint64_t *data, size_t size, size_t chain_index, double scale,
                         int ntt_form);
int nx_ciphertext_download(const nx_ciphertext *ct, uint64_t *data);
int nx_ciphertext_info(const nx_ciphertext *ct, size_t *size, size_t *chain_index, size_t *coeff_modulus_size,
                       double *scale, int *ntt_form);
int nx_ciphertext_set_scale(nx_ciphertext *ct, double scale);
/* Device address of the residues ([size][limbs][N] words; *words may be NULL),
 * valid until the ciphertext is written or destroyed (reference
 * PhantomCiphertext::data(), cuda/thirdparty/phantom-fhe/include/ciphertext.h:222). */
int nx_ciphertext_device_data(const nx_ciphertext *ct, uint64_t **data, size_t *words);

int nx_plaintext_create(nx_context *ctx, nx_plaintext **out);
int nx_plaintext_destroy(nx_plaintext *pt);
int nx_plaintext_copy(nx_plaintext *dst, const nx_plaintext *src);
int nx_plaintext_upload(nx_plaintext *pt, const uint64_t *data, size_t chain_index, double scale, int ntt_form);
int nx_plaintext_download(const nx_plaintext *pt, uint64_t *data);
int nx_plaintext_info(const nx_plaintext *pt, size_t *chain_index, size_t *coeff_modulus_size, double *scale);
int nx_plaintext_set_scale(nx_plaintext *pt, double scale);

/* ---- keys (reference KeyGenerator, src/main.cpp:51-58; PhantomSecretKey
 * cuda/src/main.cu:272-278).  Key-switching keys are device resident;
 * elt 0 addresses the relinearisation key of a set. */
int nx_secret_key_generate(nx_context *ctx, uint64_t seed, nx_secret_key **out);
int nx_secret_key_from_coeffs(nx_context *ctx, const int8_t *coeffs, nx_secret_key **out);
int nx_secret_key_coeffs(const nx_secret_key *sk, int8_t *coeffs);
int nx_secret_key_destroy(nx_secret_key *sk);
int nx_kswitch_keys_create(nx_context *ctx, nx_kswitch_keys **out);
int nx_kswitch_keys_destroy(nx_kswitch_keys *keys);
int nx_relin_key_generate(nx_context *ctx, const nx_secret_key *sk, uint64_t seed, nx_kswitch_keys *keys);
int nx_galois_keys_generate(nx_context *ctx, const nx_secret_key *sk, const uint32_t *elts, size_t count,
                            uint64_t seed, nx_kswitch_keys *keys);
/* key material [digits][2][L + special][N] words (nx_context_key_info) */
int nx_kswitch_key_upload(nx_kswitch_keys *keys, uint32_t elt, const uint64_t *data);
int nx_kswitch_key_download(const nx_kswitch_keys *keys, uint32_t elt, uint64_t *data);
int nx_kswitch_keys_has(const nx_kswitch_keys *keys, uint32_t elt, int *present);

/* ---- encryptor / decryptor (reference cuda/src/ckks_evaluator.cuh:78-94,396-419) */
int nx_encrypt_symmetric(nx_context *ctx, const nx_secret_key *sk, const nx_plaintext *pt, uint64_t seed,
                         nx_ciphertext *out);
int nx_decrypt(nx_context *ctx, const nx_secret_key *sk, const nx_ciphertext *ct, nx_plaintext *out);

/* ---- encoder (reference Encoder, cuda/src/ckks_evaluator.cuh:11-76;
 * SEAL ckks.h:457-760) */
int nx_encode(nx_context *ctx, const double *re, const double *im, size_t count, size_t chain_index, double scale,
              nx_plaintext *out);
int nx_encode_constant(nx_context *ctx, double value, size_t chain_index, double scale, nx_plaintext *out);
int nx_decode(nx_context *ctx, const nx_plaintext *pt, double *re, double *im, size_t count);

/* ---- evaluator (reference Evaluator, cuda/src/ckks_evaluator.cuh:115-394;
 * SEAL evaluator.cpp).  dest may alias any input. */
int nx_add(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest);
int nx_sub(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest);
int nx_negate(const nx_ciphertext *a, nx_ciphertext *dest);
int nx_add_plain(const nx_ciphertext *a, const nx_plaintext *p, nx_ciphertext *dest);
int nx_sub_plain(const nx_ciphertext *a, const nx_plaintext *p, nx_ciphertext *dest);
int nx_multiply_plain(const nx_ciphertext *a, const nx_plaintext *p, nx_ciphertext *dest);
int nx_multiply(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest);
int nx_square(const nx_ciphertext *a, nx_ciphertext *dest);
int nx_relinearize(const nx_ciphertext *a, const nx_kswitch_keys *relin_keys, nx_ciphertext *dest);
int nx_rescale_to_next(const nx_ciphertext *a, nx_ciphertext *dest);
/* Batched forms (no counterpart in the reference, which runs these per
 * ciphertext): dests[i] = rescale_to_next(cts[i]), and dests[i] =
 * rescale_to_next(relinearize(a[i] * b[i])), word for word the per-ciphertext
 * results; every stage runs as one launch over the batch.  Inputs of one
 * level (and size); dests may alias inputs. */
int nx_rescale_to_next_many(const nx_ciphertext *const *cts, size_t count, nx_ciphertext *const *dests);
int nx_multiply_relin_rescale_many(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count,
                                   const nx_kswitch_keys *relin_keys, nx_ciphertext *const *dests);
/* dest = sum_i a[i] * b[i] as one size-3 ciphertext: word for word multiply and
 * add of every pair (the reference's loop of evaluator.multiply + add,
 * e.g. cuda/src/matrix_mul.cu:202-275), with no tensor written out.  Pairs of
 * one level and product scale, size 2, primes of at most 60 bits. */
int nx_multiply_sum(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count, nx_ciphertext *dest);
/* The same products with relinearisation and rescale in one rounding (no
 * counterpart in the reference): with hybrid key switching the key-switch
 * accumulator plus P (c0, c1) is divided by P q_(l-1) in one centred
 * base conversion, saving one NTT and two passes per product.  Each result
 * coefficient is within 1 of nx_multiply_relin_rescale_many's (the same
 * integer on every limb); the two differ only where the exact quotient lies
 * within 1/(2 q_(l-1)) of a rounding boundary.  Without hybrid key switching
 * it is nx_multiply_relin_rescale_many. */
int nx_multiply_relin_rescale_fused_many(const nx_ciphertext *const *a, const nx_ciphertext *const *b, size_t count,
                                         const nx_kswitch_keys *relin_keys, nx_ciphertext *const *dests);
int nx_mod_switch_to_next(const nx_ciphertext *a, nx_ciphertext *dest);
int nx_mod_switch_to(const nx_ciphertext *a, size_t chain_index, nx_ciphertext *dest);
int nx_plain_mod_switch_to(const nx_plaintext *a, size_t chain_index, nx_plaintext *dest);
int nx_apply_galois(const nx_ciphertext *a, uint32_t elt, const nx_kswitch_keys *galois_keys, nx_ciphertext *dest);
int nx_rotate_vector(const nx_ciphertext *a, int steps, const nx_kswitch_keys *galois_keys, nx_ciphertext *dest);
int nx_complex_conjugate(const nx_ciphertext *a, const nx_kswitch_keys *galois_keys, nx_ciphertext *dest);
int nx_transform_to_ntt(const nx_ciphertext *a, nx_ciphertext *dest);
int nx_transform_from_ntt(const nx_ciphertext *a, nx_ciphertext *dest);
int nx_add_many(const nx_ciphertext *const *cts, size_t count, nx_ciphertext *dest);
/* SEAL-4.1-bs Evaluator extensions (evaluator.cpp:395-418): the constant is
 * encoded at the ciphertext's scale */
int nx_multiply_const(const nx_ciphertext *a, double value, nx_ciphertext *dest);
int nx_add_const(const nx_ciphertext *a, double value, nx_ciphertext *dest);
/* SEAL-4.1-bs reduced-error operations (evaluator.cpp:420-594; CUDA mirror
 * cuda/src/ckks_evaluator.cu:465-589, cuda/src/ckks_evaluator.cuh:338-393):
 * dest = a + b / a - b / relinearize(a * b) after bringing the operand with
 * more limbs down by multiply_const(s_other q_last / s^2), a scale override
 * to s_other q_last, a rescale and a mod switch, so that both sides carry the
 * same scale exactly.  At equal levels a takes b's scale.  dest may alias an
 * input. */
int nx_add_reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest);
int nx_sub_reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, nx_ciphertext *dest);
int nx_multiply_reduced_error(const nx_ciphertext *a, const nx_ciphertext *b, const nx_kswitch_keys *relin_keys,
                              nx_ciphertext *dest);
/* multiply_vector_reduced_error (cuda/src/ckks_evaluator.cuh:371-393; SEAL
 * multiply_vector_inplace, evaluator.cpp:412-418): values (count <= slots,
 * zero padded; im may be NULL) encoded at the ciphertext's level and scale,
 * then multiply_plain. */
int nx_multiply_vector_reduced_error(const nx_ciphertext *a, const double *re, const double *im, size_t count,
                                     nx_ciphertext *dest);
/* double_inplace (cuda/src/ckks_evaluator.cuh:367-369): dest = a + a */
int nx_double(const nx_ciphertext *a, nx_ciphertext *dest);
/* Allocation-free rescale: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The residues of nx_rescale_to_next
 * (SEAL util/rns.cpp:830-905, q_last / 2 rounding) for `polys` (1..3)
 * polynomials of `limbs` (2..L) NTT-form limbs at device pointer `in`
 * ([polys][limbs][N] words, limb i on prime q_i) into device memory `out`
 * ([polys][limbs - 1][N], not overlapping `in`); the scale is the caller's to
 * divide by q_(limbs-1).  Calls no hipMalloc or hipFree: its scratch is a
 * 32 MiB static device array (polys x limbs x N <= 4 Mi words, larger inputs
 * are rejected), one per GPU, held until the process ends.  Stream-ordered on
 * the calling thread's stream of ctx; calls that overlap on one GPU share the
 * array and race. */
int nx_rescale_to_next_noalloc(nx_context *ctx, const uint64_t *in, size_t polys, size_t limbs, uint64_t *out);
/* multiply by x^exponent (coefficient shift; reference matrix_mul.cu:84-127) */
int nx_multiply_power_of_x(const nx_ciphertext *a, int64_t exponent, nx_ciphertext *dest);
/* ModRaise of a lowest-level ciphertext to the top data level (reference
 * src/bootstrapping/Bootstrapper.cpp:2937-2991, cuda/src/bootstrapping/Bootstrapper.cu:2813-2860);
 * the scale is left unchanged. */
int nx_mod_raise(const nx_ciphertext *a, nx_ciphertext *dest);
/* Hoisted rotations: one ModUp of c1 shared by every step (reference Phantom
 * hoisting_inplace, cuda/thirdparty/phantom-fhe/src/evaluate.cu:1727-1920).
 * nx_rotate_hoisted writes rot(a, steps[k]) to dests[k] (step 0 = copy);
 * nx_hoisting writes sum_k rot(a, steps[k]) with a single ModDown.  Results
 * decrypt like nx_rotate_vector's; residues differ within key-switching noise. */
int nx_rotate_hoisted(const nx_ciphertext *a, const int *steps, size_t count, const nx_kswitch_keys *galois_keys,
                      nx_ciphertext *const *dests);
int nx_hoisting(const nx_ciphertext *a, const int *steps, size_t count, const nx_kswitch_keys *galois_keys,
                nx_ciphertext *dest);
/* dest = sum_i rot(in[i], steps[i]) over distinct inputs of one level and scale
 * (step 0 = the input itself) with a single ModDown: each input's ModUp feeds
 * one accumulator in the extended basis (the BSGS giant-step sum of the
 * reference's linear transforms, Bootstrapper.cu:1855-1919, without a ModDown
 * per term).  Decrypts like the sum of nx_rotate_vector results. */
int nx_rotate_sum(const nx_ciphertext *const *in, const int *steps, size_t count,
                  const nx_kswitch_keys *galois_keys, nx_ciphertext *dest);
/* Batched forms of the two above (no counterpart in the reference): every
 * input of one level rotated by every step (dests[b nsteps + k] =
 * rot(in[b], steps[k]), word for word nx_rotate_hoisted of each input, the
 * ModUps, inner products and ModDowns batched across inputs); and several
 * sums of rotations at once (group i: the next group_sizes[i] entries of in /
 * steps into dests[i], word for word nx_rotate_sum of each group, the ModDowns
 * of all groups one batch).  The BSGS levels of a batch of bootstraps use them. */
int nx_rotate_hoisted_many(const nx_ciphertext *const *in, size_t count, const int *steps, size_t nsteps,
                           const nx_kswitch_keys *galois_keys, nx_ciphertext *const *dests);
int nx_rotate_sum_many(const nx_ciphertext *const *in, const int *steps, const size_t *group_sizes, size_t groups,
                       const nx_kswitch_keys *galois_keys, nx_ciphertext *const *dests);
/* dest = sum_r cts[r] * pts[r] (size-2 ciphertexts of one level, NTT form) in one
 * pass: the per-giant-step accumulation of the reference's BSGS transforms
 * (cuda/src/bootstrapping/Bootstrapper.cu:1855-1919, multiply_vector_reduced_error
 * + add per term) and of the MatrixMul row sums (cuda/src/matrix_mul.cu:202-275). */
int nx_multiply_plain_accumulate(const nx_ciphertext *const *cts, const nx_plaintext *const *pts, size_t count,
                                 nx_ciphertext *dest);
/* dests[g] = sum_r pts[g count + r] * cts[r] for g < rows (pts entries may be
 * NULL: no term; every row needs one): all giant steps of one BSGS output at
 * once (no counterpart in the reference, which runs one multiply-accumulate per
 * giant step, cuda/src/bootstrapping/Bootstrapper.cu:1855-1919).  Each
 * ciphertext is read once for every 4 rows instead of once per row; each row
 * equals nx_multiply_plain_accumulate over its terms word for word. */
/* dest = sum_r cts[r] * encode_constant(values[r], scales[r]) at chain_index
 * (each input read at its first limbs, no copy): word for word mod_switch_to,
 * multiply_plain by the encoded constant and add per term (no counterpart in
 * the reference, which runs a multiply_const and an add per term, e.g. the
 * leaves of its Chebyshev evaluation, cuda/src/bootstrapping/common/
 * Polynomial.cu:429-447).  The
 * result carries scale cts[0].scale * scales[0]. */
int nx_linear_combination(const nx_ciphertext *const *cts, const double *values, const double *scales, size_t count,
                          size_t chain_index, nx_ciphertext *dest);
int nx_multiply_plain_accumulate_multi(const nx_ciphertext *const *cts, size_t count, const nx_plaintext *const *pts,
                                       size_t rows, nx_ciphertext *const *dests);

/* ---- Periodic plaintexts and the block mix.  No counterpart in the
 * reference, which multiplies such masks as full plaintexts term by term
 * (multiply_plain + add per mask).  A real slot vector repeating every
 * `period` slots (a power of two <= slots) is a polynomial in X^(N / (2 period));
 * encoded exactly (coefficients round(scale c_k) of the real inverse
 * embedding, not the FFT path of nx_encode) its NTT form is constant over
 * blocks of N / (2 period) positions, so `count` of them are held as
 * 2 period words per limb.  values: [count][period]. */
int nx_periodic_plain_create(nx_context *ctx, const double *values, size_t period, size_t count, size_t chain_index,
                             double scale, nx_periodic_plain **out);
int nx_periodic_plain_destroy(nx_periodic_plain *p);
/* The full NTT-form plaintext of entry `index` (decodes to its slot vector). */
int nx_periodic_plain_expand(const nx_periodic_plain *p, size_t index, nx_plaintext *out);
/* out[m] = sum_k W[index[m K + k]] (.) in[k] for m < M (index -1: no term):
 * K <= 128 size-2 inputs of one level and scale, the level of W, each read
 * once for all outputs; M <= 128 outputs at that level with scale
 * in.scale * W.scale.  Equals nx_multiply_plain_accumulate over the expanded
 * plaintexts word for word.  Outputs may alias inputs. */
int nx_multiply_periodic_mix(const nx_ciphertext *const *in, size_t K, const nx_periodic_plain *w,
                             const int32_t *index, size_t M, nx_ciphertext *const *out);

/* ---- SEAL 4.1 serialization, compression mode "none" (reference
 * thirdparty/SEAL-4.1-bs/native/src/seal/serialization.h:76-91 header,
 * ciphertext.cpp / plaintext.cpp / kswitchkeys.cpp save_members and load).
 * Save: buf == NULL returns the size in *size only.  Load validates the header,
 * the parms_id (which level of this context), sizes and residue ranges.
 * Key sets: galois = 0 saves the relinearisation key (one entry), 1 the Galois
 * keys (N entries indexed by (elt - 1) / 2, empty where absent). */
int nx_ciphertext_save(const nx_ciphertext *ct, uint8_t *buf, size_t capacity, size_t *size);
int nx_ciphertext_load(nx_context *ctx, const uint8_t *buf, size_t size, nx_ciphertext *ct);
int nx_plaintext_save(const nx_plaintext *pt, uint8_t *buf, size_t capacity, size_t *size);
int nx_plaintext_load(nx_context *ctx, const uint8_t *buf, size_t size, nx_plaintext *pt);
int nx_kswitch_keys_save(const nx_kswitch_keys *keys, int galois, uint8_t *buf, size_t capacity, size_t *size);
int nx_kswitch_keys_load(nx_context *ctx, const uint8_t *buf, size_t size, int galois, nx_kswitch_keys *keys);
/* SEAL's parms_id of a level (EncryptionParameters::compute_parms_id,
 * encryptionparams.cpp): BLAKE2b-256 of {2 (CKKS), N, moduli of the level, 0};
 * chain_index 0 = key level.  id: 4 words. */
int nx_parms_id(const nx_context *ctx, size_t chain_index, uint64_t *id);
/* BLAKE2b (RFC 7693, unkeyed), digest_size 1..64 (SEAL util/hash.h HashFunction) */
int nx_hash_blake2b(const uint8_t *data, size_t size, uint8_t *digest, size_t digest_size);

#ifdef __cplusplus
}
#endif

#endif /* NEXUS_CKKS_H */
