// Ciphertext-level CKKS operations: SEAL-4.1-bs semantics (scale bookkeeping,
// level rules, error messages) over the gfx950 kernels.  Every result is built
// in a fresh buffer from the context's cache and then moved into the
// destination, so any aliasing of destination and inputs is safe.
#include <algorithm>
#include <cmath>

#include "context.h"
#include "prof.h"

namespace nx {

namespace {

size_t poly_words(const Context &c, size_t limbs) { return limbs * c.n; }

void d2d(Context &c, uint64_t *dst, const uint64_t *src, size_t words)
{
    if (words) NX_HIP(hipMemcpyAsync(dst, src, words * 8, hipMemcpyDeviceToDevice, c.stream));
}

}  // namespace

// Copy `limbs` limbs of each of `polys` polys between differently strided buffers.
void copy_limbs(Context &c, uint64_t *dst, size_t dst_stride, const uint64_t *src, size_t src_stride, size_t polys,
                size_t limbs)
{
    if (!polys || !limbs) return;
    poly_copy_limbs(c.tables, dst, dst_stride, src, src_stride, (uint32_t)polys, (uint32_t)limbs, c.stream);
}

namespace {

bool scale_within_bounds(const Context &c, double scale, size_t limbs)
{
    return !(scale <= 0 || (int)std::log2(scale) >= c.total_bits(limbs));
}

void same_params(const Ciphertext &a, const Ciphertext &b)
{
    require(a.ctx && a.ctx == b.ctx, "encrypted1 and encrypted2 parameter mismatch");
    require(a.limbs == b.limbs, "encrypted1 and encrypted2 parameter mismatch");
    require(a.ntt == b.ntt, "NTT form mismatch");
}

void valid(const Ciphertext &a)
{
    require(a.ctx != nullptr && a.size >= 2 && a.limbs >= 1, "encrypted is not valid for encryption parameters");
}

}  // namespace

void ct_resize(Ciphertext &ct, Context *c, size_t size, size_t limbs)
{
    ct.ctx = c;
    ct.size = size;
    ct.limbs = limbs;
    ct.data.reset(c, size * limbs * c->n);
}

void ct_copy(Ciphertext &dst, const Ciphertext &src)
{
    if (&dst == &src) return;
    Context &c = *src.ctx;
    DevBuf buf(&c, src.size * poly_words(c, src.limbs));
    d2d(c, buf.get(), src.data.get(), buf.words());
    dst.ctx = src.ctx;
    dst.size = src.size;
    dst.limbs = src.limbs;
    dst.scale = src.scale;
    dst.ntt = src.ntt;
    dst.data = std::move(buf);
}

void pt_copy(Plaintext &dst, const Plaintext &src)
{
    if (&dst == &src) return;
    Context &c = *src.ctx;
    DevBuf buf(&c, poly_words(c, src.limbs));
    d2d(c, buf.get(), src.data.get(), buf.words());
    dst.ctx = src.ctx;
    dst.limbs = src.limbs;
    dst.scale = src.scale;
    dst.ntt = src.ntt;
    dst.data = std::move(buf);
}

static void add_sub(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst, bool sub)
{
    valid(a);
    valid(b);
    same_params(a, b);
    tally("add", a.limbs);
    require(are_close(a.scale, b.scale), "scale mismatch");
    Context &c = *a.ctx;
    const size_t pw = poly_words(c, a.limbs);
    const size_t sz = std::max(a.size, b.size), mn = std::min(a.size, b.size);
    DevBuf buf(&c, sz * pw);
    if (sub)
        poly_sub(c.tables, buf.get(), a.data.get(), b.data.get(), (uint32_t)mn, (uint32_t)a.limbs, c.stream);
    else
        poly_add(c.tables, buf.get(), a.data.get(), b.data.get(), (uint32_t)mn, (uint32_t)a.limbs, c.stream);
    if (a.size > mn) d2d(c, buf.get() + mn * pw, a.data.get() + mn * pw, (a.size - mn) * pw);
    if (b.size > mn) {
        if (sub)
            poly_negate(c.tables, buf.get() + mn * pw, b.data.get() + mn * pw, (uint32_t)(b.size - mn),
                        (uint32_t)b.limbs, c.stream);
        else
            d2d(c, buf.get() + mn * pw, b.data.get() + mn * pw, (b.size - mn) * pw);
    }
    const double scale = a.scale;
    const bool ntt = a.ntt;
    const size_t limbs = a.limbs;
    dst.ctx = &c;
    dst.size = sz;
    dst.limbs = limbs;
    dst.scale = scale;
    dst.ntt = ntt;
    dst.data = std::move(buf);
}

void op_add(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst) { add_sub(a, b, dst, false); }
void op_sub(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst) { add_sub(a, b, dst, true); }

void op_negate(const Ciphertext &a, Ciphertext &dst)
{
    valid(a);
    tally("add", a.limbs);
    Context &c = *a.ctx;
    DevBuf buf(&c, a.size * poly_words(c, a.limbs));
    poly_negate(c.tables, buf.get(), a.data.get(), (uint32_t)a.size, (uint32_t)a.limbs, c.stream);
    Ciphertext tmp;
    tmp.ctx = &c;
    tmp.size = a.size;
    tmp.limbs = a.limbs;
    tmp.scale = a.scale;
    tmp.ntt = a.ntt;
    tmp.data = std::move(buf);
    dst = std::move(tmp);
}

static void plain_check(const Ciphertext &a, const Plaintext &p)
{
    valid(a);
    require(p.ctx == a.ctx && p.limbs == a.limbs, "encrypted and plain parameter mismatch");
    require(p.ntt == a.ntt, "NTT form mismatch");
}

static void add_sub_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst, bool sub)
{
    plain_check(a, p);
    tally("add", a.limbs);
    require(are_close(a.scale, p.scale), "scale mismatch");
    Context &c = *a.ctx;
    const size_t pw = poly_words(c, a.limbs);
    DevBuf buf(&c, a.size * pw);
    poly_add_plain_first(c.tables, buf.get(), a.data.get(), p.data.get(), (uint32_t)a.size, (uint32_t)a.limbs, sub,
                         c.stream);
    Ciphertext tmp;
    tmp.ctx = &c;
    tmp.size = a.size;
    tmp.limbs = a.limbs;
    tmp.scale = a.scale;
    tmp.ntt = a.ntt;
    tmp.data = std::move(buf);
    dst = std::move(tmp);
}

void op_add_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst) { add_sub_plain(a, p, dst, false); }
void op_sub_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst) { add_sub_plain(a, p, dst, true); }

void op_multiply_plain(const Ciphertext &a, const Plaintext &p, Ciphertext &dst)
{
    plain_check(a, p);
    tally("mul_plain", a.limbs);
    require(a.ntt, "encrypted must be in NTT form");
    Context &c = *a.ctx;
    const double scale = a.scale * p.scale;
    require(scale_within_bounds(c, scale, a.limbs), "scale out of bounds");
    DevBuf buf(&c, a.size * poly_words(c, a.limbs));
    poly_mul_plain(c.tables, buf.get(), a.data.get(), p.data.get(), (uint32_t)a.size, (uint32_t)a.limbs, c.stream);
    Ciphertext tmp;
    tmp.ctx = &c;
    tmp.size = a.size;
    tmp.limbs = a.limbs;
    tmp.scale = scale;
    tmp.ntt = true;
    tmp.data = std::move(buf);
    dst = std::move(tmp);
}

void op_multiply(const Ciphertext &a, const Ciphertext &b, Ciphertext &dst)
{
    valid(a);
    valid(b);
    same_params(a, b);
    tally("tensor", a.limbs);
    require(a.ntt && b.ntt, "encrypted1 or encrypted2 must be in NTT form");
    require(a.size == 2 && b.size == 2, "multiply expects two size-2 ciphertexts (relinearize first)");
    Context &c = *a.ctx;
    const double scale = a.scale * b.scale;
    require(scale_within_bounds(c, scale, a.limbs), "scale out of bounds");
    DevBuf buf(&c, 3 * poly_words(c, a.limbs));
    poly_tensor(c.tables, buf.get(), a.data.get(), b.data.get(), (uint32_t)a.limbs, c.stream);
    Ciphertext tmp;
    tmp.ctx = &c;
    tmp.size = 3;
    tmp.limbs = a.limbs;
    tmp.scale = scale;
    tmp.ntt = true;
    tmp.data = std::move(buf);
    dst = std::move(tmp);
}

void op_square(const Ciphertext &a, Ciphertext &dst) { op_multiply(a, a, dst); }

// sum_r a_r b_r as one size-3 ciphertext: multiply and add of every pair, each
// tensor never written out (word for word the multiply / add sequence)
void op_multiply_sum(const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                     Ciphertext &dst)
{
    require(!a.empty() && a.size() == b.size(), "operand counts must match");
    const Ciphertext &a0 = *a[0];
    valid(a0);
    Context &c = *a0.ctx;
    const size_t l = a0.limbs;
    const double scale = a0.scale * b[0]->scale;
    for (size_t r = 0; r < a.size(); r++) {
        valid(*a[r]);
        valid(*b[r]);
        same_params(*a[r], *b[r]);
        require(a[r]->ctx == &c && a[r]->limbs == l, "operands must share the level");
        require(a[r]->ntt && b[r]->ntt, "encrypted1 or encrypted2 must be in NTT form");
        require(a[r]->size == 2 && b[r]->size == 2, "multiply expects two size-2 ciphertexts (relinearize first)");
        require(are_close(a[r]->scale * b[r]->scale, scale), "scale mismatch");
    }
    require(scale_within_bounds(c, scale, l), "scale out of bounds");
    tally("tensor", l, (double)a.size());
    tally("add", l, (double)(a.size() - 1));
    DevBuf buf(&c, 3 * poly_words(c, l));
    for (size_t base = 0; base < a.size(); base += (size_t)kPtrMax) {
        const size_t cnt = std::min<size_t>(kPtrMax, a.size() - base);
        std::vector<const uint64_t *> pa(cnt), pb(cnt);
        for (size_t r = 0; r < cnt; r++) {
            pa[r] = a[base + r]->data.get();
            pb[r] = b[base + r]->data.get();
        }
        poly_tensor_sum(c.tables, buf.get(), pa.data(), pb.data(), (uint32_t)cnt, (uint32_t)l, base > 0, c.stream);
    }
    Ciphertext tmp;
    tmp.ctx = &c;
    tmp.size = 3;
    tmp.limbs = l;
    tmp.scale = scale;
    tmp.ntt = true;
    tmp.data = std::move(buf);
    dst = std::move(tmp);
}

// ---------------------------------------------------------------- key switching
// Adds key_switch(target) into (ct0, ct1); evaluator.cpp:2724-3021.
// zero_c1: write instead of accumulate into ct1 (apply_galois clears c1 first).
// Key switch of B targets ([B][l][n], NTT) added into B ciphertexts
// ([B][2][l][n]); every stage runs as one launch over the whole batch.
void ks_out_primes(const Context &c, size_t l, uint8_t *outp)
{
    for (size_t i = 0; i < l; i++) outp[i] = (uint8_t)i;
    for (size_t k = 0; k < c.alpha; k++) outp[l + k] = (uint8_t)(c.L + k);
}

static DevBuf ks_modup_impl(Context &c, const uint64_t *targets, const uint64_t *const *target_list, size_t B,
                            size_t l)
{
    // ModUp: INTT the l limbs; each digit (one prime, or alpha consecutive
    // primes in hybrid mode) extended to every output prime; NTT.
    const size_t n = c.n, nout = l + c.alpha, digits = c.digits(l);
    uint8_t outp[kMaxPrimes];
    ks_out_primes(c, l, outp);
    DevBuf coef(&c, B * l * n);
    if (target_list) {
        NttBatch b;
        b.data = coef.get();
        b.poly_stride = l * n;
        b.polys = (uint32_t)B;
        b.limbs = (uint32_t)l;
        b.prime_by_poly = 0;
        for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)i;
        b.src_list = target_list;
        ntt_inverse(c.tables, b, c.stream);
    } else {
        ntt_inv_from(c, coef.get(), targets, B, l, l * n, l * n);
    }
    DevBuf T(&c, B * nout * digits * n);
    if (c.alpha > 1 && ks_modup_col_ok(c.tables, (uint32_t)c.alpha) && ks_modup_col_get((uint32_t)B) >= 0) {
        // the conversion feeds the forward column pass directly; the digits'
        // own limbs are copied in NTT form; only the row pass remains
        ks_modup_copy_own(c.tables, T.get(), targets, target_list, (uint32_t)B, (uint32_t)l, (uint32_t)c.alpha,
                          (uint32_t)digits, (uint32_t)nout, c.stream);
        ks_modup_convert_col(c.tables, T.get(), coef.get(), c.d_modup_w, (uint32_t)B, (uint32_t)l,
                             (uint32_t)c.alpha, (uint32_t)c.key_limbs(), outp, (uint32_t)nout, c.stream);
        NttBatch b;
        b.skip_l = (uint32_t)l;
        b.skip_alpha = (uint32_t)c.alpha;
        b.rows_only = 1;
        b.data = T.get();
        b.poly_stride = digits * n;
        b.polys = (uint32_t)(B * nout);
        b.limbs = (uint32_t)digits;
        b.prime_by_poly = 1;
        b.poly_period = (uint32_t)nout;
        for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = i < (int)nout ? outp[i] : 0;
        ntt_forward(c.tables, b, c.stream);
        return T;
    }
    if (c.alpha == 1)
        ks_modup_expand(c.tables, T.get(), coef.get(), (uint32_t)B, (uint32_t)l, outp, (uint32_t)nout, c.stream);
    else if (ks_conv_fused_ok((uint32_t)c.alpha))
        ks_modup_convert_fused(c.tables, T.get(), coef.get(), c.d_modup_w, c.d_modup_conv, (uint32_t)B, (uint32_t)l,
                               (uint32_t)c.alpha, (uint32_t)c.key_limbs(), outp, (uint32_t)nout, c.stream);
    else
        ks_modup_convert(c.tables, T.get(), coef.get(), c.d_modup_w, c.d_modup_conv, (uint32_t)B, (uint32_t)l,
                         (uint32_t)c.alpha, (uint32_t)c.key_limbs(), outp, (uint32_t)nout, c.stream);
    NttBatch b;
    if (c.alpha > 1) {
        // digit j on its own primes is its input: copied in NTT form, not transformed
        ks_modup_copy_own(c.tables, T.get(), targets, target_list, (uint32_t)B, (uint32_t)l, (uint32_t)c.alpha,
                          (uint32_t)digits, (uint32_t)nout, c.stream);
        b.skip_l = (uint32_t)l;
        b.skip_alpha = (uint32_t)c.alpha;
    }
    b.data = T.get();
    b.poly_stride = digits * n;
    b.polys = (uint32_t)(B * nout);
    b.limbs = (uint32_t)digits;
    b.prime_by_poly = 1;
    b.poly_period = (uint32_t)nout;
    for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = i < (int)nout ? outp[i] : 0;
    ntt_forward(c.tables, b, c.stream);
    return T;
}

DevBuf ks_modup(Context &c, const uint64_t *targets, size_t B, size_t l)
{
    return ks_modup_impl(c, targets, nullptr, B, l);
}

DevBuf ks_modup_list(Context &c, const uint64_t *const *targets, size_t B, size_t l)
{
    require(B >= 1 && B <= (size_t)kNttSrcMax, "too many ModUp targets in one batch");
    return ks_modup_impl(c, nullptr, targets, B, l);
}

void key_switch_batch(Context &c, uint64_t *cts, const uint64_t *targets, size_t B, size_t l, const DevBuf &key,
                      bool zero_c1, const uint64_t *base, const uint64_t *aseed)
{
    tally("ks", l, (double)B);
    const size_t n = c.n, nout = l + c.alpha;
    uint8_t outp[kMaxPrimes];
    ks_out_primes(c, l, outp);
    DevBuf T = ks_modup(c, targets, B, l);
    // inner product with the device-resident key
    DevBuf acc(&c, B * 2 * nout * n);
    ks_inner_product(c.tables, acc.get(), T.get(), key.get(), (uint32_t)B, (uint32_t)c.digits(l),
                     (uint32_t)c.key_limbs(), outp, (uint32_t)nout, c.stream, aseed);
    ks_moddown(c, cts, acc.get(), B, l, zero_c1, base);
}

DevBuf ks_moddown_prepare(Context &c, uint64_t *acc, size_t B, size_t l, bool transform, bool *col_done)
{
    const size_t n = c.n, L = c.L, alpha = c.alpha;
    const size_t nout = l + alpha;
    // The special-prime limbs of acc are transformed in place: INTT of the
    // special limbs of every poly.
    {
        NttBatch b;
        b.data = acc + l * n;
        b.poly_stride = nout * n;
        b.polys = (uint32_t)(2 * B);
        b.limbs = (uint32_t)alpha;
        b.prime_by_poly = 0;
        for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)(L + (i < (int)alpha ? i : 0));
        ntt_inverse(c.tables, b, c.stream);
    }
    DevBuf tmp(&c, B * 2 * l * n);
    if (col_done) *col_done = false;
    if (alpha > 1 && (col_done || transform) && c.d_moddown_conv_d && ks_moddown_col_ok(c.tables, (uint32_t)alpha) &&
        ks_moddown_col_get((uint32_t)(2 * B)) >= 0) {
        // the centred conversion feeds the forward column pass directly
        uint8_t inp[kMaxPrimes];
        for (size_t k = 0; k < alpha; k++) inp[k] = (uint8_t)(L + k);
        ks_moddown_convert_col(c.tables, tmp.get(), acc, nout * n, c.d_moddown_w, c.d_moddown_conv_d, inp,
                               (uint32_t)alpha, (uint32_t)(2 * B), (uint32_t)l, (uint32_t)L, c.stream);
        if (transform) ntt_fwd_epi(c, tmp.get(), 2 * B, l, l * n, NttEpilogue(), true);
        if (col_done) *col_done = !transform;
        return tmp;
    }
    if (alpha == 1 && (col_done || transform) && center_col_ok(c.tables)) {
        // SEAL's centring inside the forward column pass
        rns_center_last_col(c.tables, tmp.get(), acc + l * n, nout * n, (uint32_t)L, (uint32_t)(2 * B), (uint32_t)l,
                            c.stream);
        if (transform) ntt_fwd_epi(c, tmp.get(), 2 * B, l, l * n, NttEpilogue(), true);
        if (col_done) *col_done = !transform;
        return tmp;
    }
    if (alpha == 1)  // SEAL: centred (rounded) division by p
        rns_center_last(c.tables, tmp.get(), acc + l * n, nout * n, (uint32_t)L, (uint32_t)(2 * B), (uint32_t)l,
                        c.stream);
    else if (ks_conv_fused_ok((uint32_t)alpha))  // hybrid: centred conversion of the P part
        ks_moddown_convert_fused(c.tables, tmp.get(), acc, nout * n, c.d_moddown_w, c.d_moddown_conv,
                                 c.d_moddown_conv_d, (uint32_t)(2 * B), (uint32_t)l, (uint32_t)alpha, (uint32_t)L,
                                 c.stream);
    else
        ks_moddown_convert(c.tables, tmp.get(), acc, nout * n, c.d_moddown_w, c.d_moddown_conv, (uint32_t)(2 * B),
                           (uint32_t)l, (uint32_t)alpha, (uint32_t)L, c.stream);
    if (transform) ntt_fwd(c, tmp.get(), 2 * B, l, l * n);
    return tmp;
}

void ks_moddown(Context &c, uint64_t *cts, uint64_t *acc, size_t B, size_t l, bool zero_c1, const uint64_t *base)
{
    // cts = (acc - NTT(conversion)) P^-1 (+ base or cts): the combination is the
    // forward transform's last pass
    const size_t n = c.n, nout = l + c.alpha;
    bool col_done = false;
    DevBuf tmp = ks_moddown_prepare(c, acc, B, l, false, &col_done);
    NttEpilogue e;
    e.out = cts;
    e.out_stride = l * n;
    e.x = acc;
    e.x_stride = nout * n;
    e.inv = c.d_p_inv;
    e.accumulate = zero_c1 ? 2 : 1;
    e.base = base;
    e.base_stride = l * n;
    ntt_fwd_epi(c, tmp.get(), 2 * B, l, l * n, e, col_done);
}

static void key_switch_impl(Context &c, Ciphertext &ct, const uint64_t *target, const DevBuf &key, bool zero_c1,
                            const uint64_t *aseed)
{
    key_switch_batch(c, ct.data.get(), target, 1, ct.limbs, key, zero_c1, nullptr, aseed);
}

void key_switch_into(Context &c, Ciphertext &ct, const uint64_t *target, const DevBuf &key)
{
    key_switch_impl(c, ct, target, key, false, nullptr);
}

static const DevBuf &find_key(const KSwitchKeys &k, uint32_t elt, const char *msg)
{
    auto it = k.keys.find(elt);
    require(it != k.keys.end(), msg);
    return it->second;
}

void op_relinearize(const Ciphertext &a, const KSwitchKeys &rk, Ciphertext &dst)
{
    valid(a);
    require(a.size == 3, "encrypted size must be 3 to relinearize");
    require(rk.ctx == a.ctx, "relin_keys is not valid for encryption parameters");
    require(a.ntt, "CKKS encrypted must be in NTT form");
    Context &c = *a.ctx;
    const DevBuf &key = find_key(rk, 0, "relinearization key not present");
    const size_t pw = poly_words(c, a.limbs);
    Ciphertext out;
    out.ctx = &c;
    out.size = 2;
    out.limbs = a.limbs;
    out.scale = a.scale;
    out.ntt = true;
    out.data.reset(&c, 2 * pw);
    // out = (c0, c1) + key_switch(c2): c0, c1 are read where they lie
    key_switch_batch(c, out.data.get(), a.data.get() + 2 * pw, 1, a.limbs, key, false, a.data.get(),
                     rk.seeds_of(0));
    dst = std::move(out);
}

void op_apply_galois(const Ciphertext &a, uint32_t elt, const KSwitchKeys &gk, Ciphertext &dst)
{
    valid(a);
    require(a.size == 2, "encrypted size must be 2");
    require(gk.ctx == a.ctx, "galois_keys is not valid for encryption parameters");
    Context &c = *a.ctx;
    require((elt & 1) && elt < 2 * c.n, "Galois element is not valid");
    const DevBuf &key = find_key(gk, elt, "Galois key not present");
    const uint32_t *perm = c.galois_table(elt);
    const size_t pw = poly_words(c, a.limbs);
    Ciphertext out;
    out.ctx = &c;
    out.size = 2;
    out.limbs = a.limbs;
    out.scale = a.scale;
    out.ntt = true;
    out.data.reset(&c, 2 * pw);
    DevBuf t(&c, pw);
    poly_permute(out.data.get(), a.data.get(), perm, 1, (uint32_t)a.limbs, c.logn, c.stream);
    poly_permute(t.get(), a.data.get() + pw, perm, 1, (uint32_t)a.limbs, c.logn, c.stream);
    key_switch_impl(c, out, t.get(), key, true, gk.seeds_of(elt));
    dst = std::move(out);
}

static std::vector<int> naf(int value)
{
    // util/numth.h:22-42
    std::vector<int> res;
    const bool sign = value < 0;
    value = std::abs(value);
    for (int i = 0; value; i++) {
        int zi = (value & 1) ? 2 - (value & 3) : 0;
        value = (value - zi) >> 1;
        if (zi) res.push_back((sign ? -zi : zi) * (1 << i));
    }
    return res;
}

void op_rotate(const Ciphertext &a, int steps, const KSwitchKeys &gk, Ciphertext &dst)
{
    // evaluator.cpp:2667-2722
    Context &c = *a.ctx;
    if (steps == 0) {
        ct_copy(dst, a);
        return;
    }
    const uint32_t elt = c.elt_from_step(steps);
    if (gk.keys.count(elt)) {
        op_apply_galois(a, elt, gk, dst);
        return;
    }
    std::vector<int> terms = naf(steps);
    require(terms.size() > 1, "Galois key not present");
    Ciphertext cur;
    ct_copy(cur, a);
    for (int t : terms) {
        if ((size_t)std::abs(t) == (c.n >> 1)) continue;
        op_rotate(cur, t, gk, cur);
    }
    dst = std::move(cur);
}

void op_conjugate(const Ciphertext &a, const KSwitchKeys &gk, Ciphertext &dst)
{
    op_apply_galois(a, (uint32_t)(2 * a.ctx->n - 1), gk, dst);
}

// ---------------------------------------------------------------- rescale & level
void op_rescale_to_next(const Ciphertext &a, Ciphertext &dst)
{
    // evaluator.cpp:1402-1481 + util/rns.cpp:830-905
    valid(a);
    tally("rescale", a.limbs, (double)a.size / 2);
    require(a.ntt, "CKKS encrypted must be in NTT form");
    require(a.limbs >= 2, "end of modulus switching chain reached");
    Context &c = *a.ctx;
    const size_t l = a.limbs, n = c.n, P = a.size;
    DevBuf last(&c, P * n);
    {
        // INTT of the last limb of every poly, read in place
        NttBatch b;
        b.data = last.get();
        b.poly_stride = n;
        b.polys = (uint32_t)P;
        b.limbs = 1;
        b.prime_by_poly = 0;
        for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)(l - 1);
        b.src = a.data.get() + (l - 1) * n;
        b.src_poly_stride = l * n;
        ntt_inverse(c.tables, b, c.stream);
    }
    DevBuf tmp(&c, P * (l - 1) * n);
    const bool col = center_col_ok(c.tables);  // centring inside the forward column pass
    if (col)
        rns_center_last_col(c.tables, tmp.get(), last.get(), n, (uint32_t)(l - 1), (uint32_t)P, (uint32_t)(l - 1),
                            c.stream);
    else
        rns_center_last(c.tables, tmp.get(), last.get(), n, (uint32_t)(l - 1), (uint32_t)P, (uint32_t)(l - 1),
                        c.stream);
    // out = (a - NTT(centred last limb)) q_(l-1)^-1, stored by the transform's last pass
    DevBuf out(&c, P * (l - 1) * n);
    NttEpilogue e;
    e.out = out.get();
    e.out_stride = (l - 1) * n;
    e.x = a.data.get();
    e.x_stride = l * n;
    e.inv = c.rescale_inv(l);
    ntt_fwd_epi(c, tmp.get(), P, l - 1, (l - 1) * n, e, col);
    const double scale = a.scale / (double)c.moduli[l - 1];
    Ciphertext r;
    r.ctx = &c;
    r.size = P;
    r.limbs = l - 1;
    r.scale = scale;
    r.ntt = true;
    r.data = std::move(out);
    dst = std::move(r);
}

// Rescale tail shared by the batched forms: `count` ciphertexts of `polys`
// polys at l limbs, given by pointer, into fresh results at l - 1 limbs.  One
// INTT of all last limbs, one centring, one NTT, one combination.
static void rescale_batch(Context &c, const std::vector<const uint64_t *> &x, size_t polys, size_t l,
                          const std::vector<uint64_t *> &outs)
{
    tally("rescale", l, (double)x.size() * (double)polys / 2);
    const size_t n = c.n, B = x.size(), P = B * polys;
    require(P <= (size_t)kNttSrcMax, "rescale batch too large");
    std::vector<const uint64_t *> lasts;
    for (size_t r = 0; r < B; r++)
        for (size_t p = 0; p < polys; p++) lasts.push_back(x[r] + (p * l + l - 1) * n);
    DevBuf last(&c, P * n);
    {
        NttBatch b;
        b.data = last.get();
        b.poly_stride = n;
        b.polys = (uint32_t)P;
        b.limbs = 1;
        b.prime_by_poly = 0;
        for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)(l - 1);
        b.src_list = lasts.data();
        ntt_inverse(c.tables, b, c.stream);
    }
    DevBuf tmp(&c, P * (l - 1) * n);
    const bool col = center_col_ok(c.tables);  // centring inside the forward column pass
    if (col)
        rns_center_last_col(c.tables, tmp.get(), last.get(), n, (uint32_t)(l - 1), (uint32_t)P, (uint32_t)(l - 1),
                            c.stream);
    else
        rns_center_last(c.tables, tmp.get(), last.get(), n, (uint32_t)(l - 1), (uint32_t)P, (uint32_t)(l - 1),
                        c.stream);
    NttEpilogue e;  // outs[r] = (x[r] - NTT(centred last limbs)) q_(l-1)^-1
    for (size_t r = 0; r < B; r++) {
        e.outs[r] = outs[r];
        e.xs[r] = x[r];
    }
    e.x_limbs = (uint32_t)l;
    e.polys_per = (uint32_t)polys;
    e.inv = c.rescale_inv(l);
    ntt_fwd_epi(c, tmp.get(), P, l - 1, (l - 1) * n, e, col);
}

void op_rescale_many(const std::vector<const Ciphertext *> &in, const std::vector<Ciphertext *> &out)
{
    require(!in.empty() && in.size() == out.size(), "inputs and outputs must match");
    const Ciphertext &a0 = *in[0];
    valid(a0);
    Context &c = *a0.ctx;
    const size_t l = a0.limbs, P = a0.size, n = c.n;
    for (const Ciphertext *a : in) {
        valid(*a);
        require(a->ctx == &c && a->size == P && a->limbs == l, "rescale batch needs one shape");
        require(a->ntt, "CKKS encrypted must be in NTT form");
    }
    require(l >= 2, "end of modulus switching chain reached");
    std::vector<Ciphertext> res(in.size());
    const size_t chunk = (size_t)kNttSrcMax / P;
    for (size_t base = 0; base < in.size(); base += chunk) {
        const size_t cnt = std::min(chunk, in.size() - base);
        std::vector<const uint64_t *> x;
        std::vector<uint64_t *> o;
        for (size_t r = 0; r < cnt; r++) {
            const Ciphertext &a = *in[base + r];
            Ciphertext &t = res[base + r];
            t.ctx = &c;
            t.size = P;
            t.limbs = l - 1;
            t.scale = a.scale / (double)c.moduli[l - 1];
            t.ntt = true;
            t.data.reset(&c, P * (l - 1) * n);
            x.push_back(a.data.get());
            o.push_back(t.data.get());
        }
        rescale_batch(c, x, P, l, o);
    }
    for (size_t r = 0; r < in.size(); r++) *out[r] = std::move(res[r]);
}

void op_multiply_relin_rescale_many(const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                                    const KSwitchKeys &rk, const std::vector<Ciphertext *> &out)
{
    // multiply (evaluator.cpp:1003-1040), relinearize (2724-3021) and
    // rescale_to_next (1402-1481) of every pair, each stage one batched launch
    require(!a.empty() && a.size() == b.size() && a.size() == out.size(), "operand counts must match");
    const Ciphertext &a0 = *a[0];
    valid(a0);
    Context &c = *a0.ctx;
    const size_t l = a0.limbs, n = c.n, pw = l * n;
    require(rk.ctx == &c, "relin_keys is not valid for encryption parameters");
    const DevBuf &key = find_key(rk, 0, "relinearization key not present");
    require(l >= 2, "end of modulus switching chain reached");
    for (size_t r = 0; r < a.size(); r++) {
        valid(*a[r]);
        valid(*b[r]);
        same_params(*a[r], *b[r]);
        require(a[r]->ctx == &c && a[r]->limbs == l, "batch operands must share the level");
        require(a[r]->ntt && b[r]->ntt, "encrypted1 or encrypted2 must be in NTT form");
        require(a[r]->size == 2 && b[r]->size == 2, "multiply expects two size-2 ciphertexts (relinearize first)");
        require(scale_within_bounds(c, a[r]->scale * b[r]->scale, l), "scale out of bounds");
    }
    std::vector<Ciphertext> res(a.size());
    const size_t chunk = (size_t)kNttSrcMax / 2;
    for (size_t base = 0; base < a.size(); base += chunk) {
        const size_t cnt = std::min(chunk, a.size() - base);
        std::vector<const uint64_t *> pa, pb;
        for (size_t r = 0; r < cnt; r++) {
            pa.push_back(a[base + r]->data.get());
            pb.push_back(b[base + r]->data.get());
        }
        DevBuf c01(&c, cnt * 2 * pw), c2(&c, cnt * pw), ks(&c, cnt * 2 * pw);
        poly_tensor_many(c.tables, c01.get(), c2.get(), pa.data(), pb.data(), (uint32_t)cnt, (uint32_t)l, c.stream);
        key_switch_batch(c, ks.get(), c2.get(), cnt, l, key, false, c01.get(), rk.seeds_of(0));
        std::vector<const uint64_t *> x;
        std::vector<uint64_t *> o;
        for (size_t r = 0; r < cnt; r++) {
            Ciphertext &t = res[base + r];
            t.ctx = &c;
            t.size = 2;
            t.limbs = l - 1;
            t.scale = a[base + r]->scale * b[base + r]->scale / (double)c.moduli[l - 1];
            t.ntt = true;
            t.data.reset(&c, 2 * (l - 1) * n);
            x.push_back(ks.get() + r * 2 * pw);
            o.push_back(t.data.get());
        }
        rescale_batch(c, x, 2, l, o);
    }
    for (size_t r = 0; r < a.size(); r++) *out[r] = std::move(res[r]);
}

void op_multiply_relin_rescale_fused_many(const std::vector<const Ciphertext *> &a,
                                          const std::vector<const Ciphertext *> &b, const KSwitchKeys &rk,
                                          const std::vector<Ciphertext *> &out)
{
    // The relinearised product (c0, c1) + ModDown_P(acc) is rescaled by q_(l-1)
    // in the same rounding: Y = acc + P (c0, c1) over Q_l P is divided by
    // P' = P q_(l-1) with one centred conversion of its limbs
    // (q_(l-1), p_0..p_(alpha-1)).  out_i = (acc_i - [Y]_(P'),i) P'^-1
    // + d_i q_(l-1)^-1, one rounding instead of two: each coefficient is within
    // 1 of op_multiply_relin_rescale_many's and the same integer on every limb.
    require(!a.empty() && a.size() == b.size() && a.size() == out.size(), "operand counts must match");
    const Ciphertext &a0 = *a[0];
    valid(a0);
    Context &c = *a0.ctx;
    if (c.alpha == 1 || !ks_moddown_rescale_ok(c.tables, (uint32_t)c.alpha)) {
        op_multiply_relin_rescale_many(a, b, rk, out);
        return;
    }
    const size_t l = a0.limbs, n = c.n, pw = l * n, alpha = c.alpha, nout = l + alpha;
    require(rk.ctx == &c, "relin_keys is not valid for encryption parameters");
    const DevBuf &key = find_key(rk, 0, "relinearization key not present");
    require(l >= 2, "end of modulus switching chain reached");
    for (size_t r = 0; r < a.size(); r++) {
        valid(*a[r]);
        valid(*b[r]);
        same_params(*a[r], *b[r]);
        require(a[r]->ctx == &c && a[r]->limbs == l, "batch operands must share the level");
        require(a[r]->ntt && b[r]->ntt, "encrypted1 or encrypted2 must be in NTT form");
        require(a[r]->size == 2 && b[r]->size == 2, "multiply expects two size-2 ciphertexts (relinearize first)");
        require(scale_within_bounds(c, a[r]->scale * b[r]->scale, l), "scale out of bounds");
    }
    uint8_t outp[kMaxPrimes];
    ks_out_primes(c, l, outp);
    const size_t A = alpha + 1, lo = l - 1;
    std::vector<Ciphertext> res(a.size());
    // every buffer of a chunk is contiguous: up to one epilogue output per product
    const size_t chunk = (size_t)std::min(kEpiMax, kPtrMax);
    for (size_t base = 0; base < a.size(); base += chunk) {
        const size_t cnt = std::min(chunk, a.size() - base);
        std::vector<const uint64_t *> pa, pb;
        for (size_t r = 0; r < cnt; r++) {
            pa.push_back(a[base + r]->data.get());
            pb.push_back(b[base + r]->data.get());
        }
        DevBuf c01(&c, cnt * 2 * pw), c2(&c, cnt * pw);
        poly_tensor_many(c.tables, c01.get(), c2.get(), pa.data(), pb.data(), (uint32_t)cnt, (uint32_t)l, c.stream);
        DevBuf acc(&c, cnt * 2 * nout * n);
        {
            DevBuf T = ks_modup(c, c2.get(), cnt, l);
            ks_inner_product(c.tables, acc.get(), T.get(), key.get(), (uint32_t)cnt, (uint32_t)c.digits(l),
                             (uint32_t)c.key_limbs(), outp, (uint32_t)nout, c.stream, rk.seeds_of(0));
        }
        // limb q_(l-1) of Y = acc + P d, then (q_(l-1), p_0..p_(alpha-1)) to coefficients
        rns_add_limb_scaled(c.tables, acc.get(), nout * n, c01.get(), pw, (uint32_t)lo, c.mdr_pmod[l],
                            (uint32_t)(2 * cnt), c.stream);
        {
            NttBatch nb;
            nb.data = acc.get() + lo * n;
            nb.poly_stride = nout * n;
            nb.polys = (uint32_t)(2 * cnt);
            nb.limbs = (uint32_t)A;
            nb.prime_by_poly = 0;
            for (int i = 0; i < kMaxPrimes; i++)
                nb.prime_of[i] = (uint8_t)(i == 0 ? lo : c.L + (i - 1 < (int)alpha ? i - 1 : 0));
            ntt_inverse(c.tables, nb, c.stream);
        }
        DevBuf tmp(&c, cnt * 2 * lo * n);  // conversion, then the transform's first pass
        bool col_done = false;
        if (ks_moddown_col_ok(c.tables, (uint32_t)A) && ks_moddown_col_get((uint32_t)(2 * cnt)) >= 0) {
            // the conversion feeds the forward column pass directly
            uint8_t inp[kMaxPrimes];
            for (size_t k = 0; k < A; k++) inp[k] = (uint8_t)(k == 0 ? lo : c.L + k - 1);
            ks_moddown_convert_col(c.tables, tmp.get(), acc.get(), nout * n, c.d_mdr_w + l * A,
                                   c.d_mdr_conv_d + l * (A + 1) * c.L, inp, (uint32_t)A, (uint32_t)(2 * cnt),
                                   (uint32_t)lo, (uint32_t)c.L, c.stream);
            col_done = true;
        } else {
            ks_moddown_rescale_convert(c.tables, tmp.get(), acc.get(), nout * n, c.d_mdr_w + l * A,
                                       c.d_mdr_conv + l * (A + 1) * c.L, c.d_mdr_conv_d + l * (A + 1) * c.L,
                                       (uint32_t)(2 * cnt), (uint32_t)lo, (uint32_t)alpha, (uint32_t)c.L, c.stream);
        }
        NttEpilogue e;  // out_i = (acc_i - NTT(tmp)_i) P'^-1 + d_i q_(l-1)^-1
        for (size_t r = 0; r < cnt; r++) {
            Ciphertext &t = res[base + r];
            t.ctx = &c;
            t.size = 2;
            t.limbs = lo;
            t.scale = a[base + r]->scale * b[base + r]->scale / (double)c.moduli[l - 1];
            t.ntt = true;
            t.data.reset(&c, 2 * lo * n);
            e.outs[r] = t.data.get();
        }
        e.x = acc.get();
        e.x_stride = nout * n;
        e.inv = c.d_mdr_inv + l * c.L;
        e.d = c01.get();
        e.d_stride = pw;
        e.inv2 = c.rescale_inv(l);
        ntt_fwd_epi(c, tmp.get(), 2 * cnt, lo, lo * n, e, col_done);
    }
    for (size_t r = 0; r < a.size(); r++) *out[r] = std::move(res[r]);
}

static void drop_to(const Ciphertext &a, size_t limbs, Ciphertext &dst)
{
    Context &c = *a.ctx;
    if (limbs == a.limbs) {
        ct_copy(dst, a);
        return;
    }
    DevBuf out(&c, a.size * limbs * c.n);
    copy_limbs(c, out.get(), limbs * c.n, a.data.get(), a.limbs * c.n, a.size, limbs);
    Ciphertext r;
    r.ctx = &c;
    r.size = a.size;
    r.limbs = limbs;
    r.scale = a.scale;
    r.ntt = a.ntt;
    r.data = std::move(out);
    dst = std::move(r);
}

void op_mod_switch_to_next(const Ciphertext &a, Ciphertext &dst)
{
    // evaluator.cpp:1483-1580 (CKKS: drop the last limb)
    valid(a);
    require(a.limbs >= 2, "end of modulus switching chain reached");
    require(a.ntt, "CKKS encrypted must be in NTT form");
    require(scale_within_bounds(*a.ctx, a.scale, a.limbs - 1), "scale out of bounds");
    drop_to(a, a.limbs - 1, dst);
}

void op_mod_switch_to(const Ciphertext &a, size_t chain_index, Ciphertext &dst)
{
    // evaluator.cpp:1630-1680
    valid(a);
    Context &c = *a.ctx;
    const size_t target = c.limbs_of(chain_index);
    require(target <= a.limbs, "cannot switch to higher level modulus");
    require(scale_within_bounds(c, a.scale, target), "scale out of bounds");
    drop_to(a, target, dst);
}

void op_plain_mod_switch_to(const Plaintext &a, size_t chain_index, Plaintext &dst)
{
    require(a.ctx != nullptr, "plain is not valid for encryption parameters");
    require(a.ntt, "plain is not in NTT form");
    Context &c = *a.ctx;
    const size_t target = c.limbs_of(chain_index);
    require(target <= a.limbs, "cannot switch to higher level modulus");
    if (target == a.limbs) {
        pt_copy(dst, a);
        return;
    }
    DevBuf out(&c, target * c.n);
    d2d(c, out.get(), a.data.get(), target * c.n);
    Plaintext r;
    r.ctx = &c;
    r.limbs = target;
    r.scale = a.scale;
    r.ntt = true;
    r.data = std::move(out);
    dst = std::move(r);
}

// ---------------------------------------------------------------- reduced-error alignment
void op_reduced_error(const Ciphertext &a, const Ciphertext &b, ReducedOp op, const KSwitchKeys *rk, Ciphertext &dst)
{
    // SEAL-4.1-bs evaluator.cpp:420-594 (CUDA mirror cuda/src/ckks_evaluator.cu:465-589):
    // the operand with more limbs is multiplied by the constant
    // s_other q_last / s^2 (encoded at its own scale), given the scale
    // s_other q_last, rescaled and dropped to the other's level, so both
    // operands carry the same scale exactly; the scale formulas and their
    // double-precision order are SEAL's.  Equal levels take b's scale.
    valid(a);
    valid(b);
    require(a.ctx == b.ctx, "encrypted1 and encrypted2 parameter mismatch");
    require(op != ReducedOp::Mul || rk != nullptr, "relin_keys is null");
    Context &c = *a.ctx;
    auto apply = [&](const Ciphertext &x, const Ciphertext &y, Ciphertext &out) {
        if (op == ReducedOp::Add)
            op_add(x, y, out);
        else if (op == ReducedOp::Sub)
            op_sub(x, y, out);
        else
            op_multiply(x, y, out);
    };
    Ciphertext res;
    if (a.limbs == b.limbs) {
        Ciphertext e1;
        ct_copy(e1, a);
        e1.scale = b.scale;
        apply(e1, b, res);
    } else if (a.limbs < b.limbs) {
        const double qlast = (double)c.moduli[b.limbs - 1];
        const double adjust = a.scale * qlast / (b.scale * b.scale);
        Plaintext k;
        encode_constant(c, adjust, b.limbs, b.scale, k);
        Ciphertext e2;
        op_multiply_plain(b, k, e2);
        e2.scale = a.scale * qlast;
        op_rescale_to_next(e2, e2);
        op_mod_switch_to(e2, c.chain_index_of(a.limbs), e2);
        Ciphertext e1;
        ct_copy(e1, a);
        e1.scale = e2.scale;
        apply(e1, e2, res);
    } else {
        const double qlast = (double)c.moduli[a.limbs - 1];
        const double adjust = b.scale * qlast / (a.scale * a.scale);
        Plaintext k;
        encode_constant(c, adjust, a.limbs, a.scale, k);
        Ciphertext e1;
        op_multiply_plain(a, k, e1);
        e1.scale = b.scale * qlast;
        op_rescale_to_next(e1, e1);
        op_mod_switch_to(e1, c.chain_index_of(b.limbs), e1);
        e1.scale = b.scale;
        apply(e1, b, res);
    }
    if (op == ReducedOp::Mul) op_relinearize(res, *rk, res);
    dst = std::move(res);
}

void op_to_ntt(const Ciphertext &a, Ciphertext &dst)
{
    valid(a);
    require(!a.ntt, "encrypted is already in NTT form");
    Context &c = *a.ctx;
    const size_t pw = poly_words(c, a.limbs);
    Ciphertext r;
    r.ctx = &c;
    r.size = a.size;
    r.limbs = a.limbs;
    r.scale = a.scale;
    r.data.reset(&c, a.size * pw);
    ntt_fwd_from(c, r.data.get(), a.data.get(), a.size, a.limbs, pw, pw);
    r.ntt = true;
    dst = std::move(r);
}

void op_from_ntt(const Ciphertext &a, Ciphertext &dst)
{
    valid(a);
    require(a.ntt, "encrypted is not in NTT form");
    Context &c = *a.ctx;
    const size_t pw = poly_words(c, a.limbs);
    Ciphertext r;
    r.ctx = &c;
    r.size = a.size;
    r.limbs = a.limbs;
    r.scale = a.scale;
    r.data.reset(&c, a.size * pw);
    ntt_inv_from(c, r.data.get(), a.data.get(), a.size, a.limbs, pw, pw);
    r.ntt = false;
    dst = std::move(r);
}

void op_multiply_power_of_x(const Ciphertext &a, int64_t exponent, Ciphertext &dst)
{
    // x^(2n) = 1 in Z[x]/(x^n + 1): reduce the exponent into [0, 2n)
    valid(a);
    Context &c = *a.ctx;
    const int64_t m = 2 * (int64_t)c.n;
    const uint32_t k = (uint32_t)(((exponent % m) + m) % m);
    const size_t pw = poly_words(c, a.limbs);
    DevBuf coef(&c, a.size * pw);
    if (a.ntt)
        ntt_inv_from(c, coef.get(), a.data.get(), a.size, a.limbs, pw, pw);
    else
        d2d(c, coef.get(), a.data.get(), a.size * pw);
    DevBuf out(&c, a.size * pw);
    poly_negacyclic_shift(c.tables, out.get(), coef.get(), k, (uint32_t)a.size, (uint32_t)a.limbs, c.stream);
    if (a.ntt) ntt_fwd(c, out.get(), a.size, a.limbs, pw);
    Ciphertext r;
    r.ctx = &c;
    r.size = a.size;
    r.limbs = a.limbs;
    r.scale = a.scale;
    r.ntt = a.ntt;
    r.data = std::move(out);
    dst = std::move(r);
}

void op_mod_raise(const Ciphertext &a, Ciphertext &dst)
{
    // src/bootstrapping/Bootstrapper.cpp:2937-2991: lowest-level ciphertext -> top data level
    valid(a);
    require(a.size == 2, "Ciphertexts of size 2 are supported only!");
    require(a.limbs == 1, "Ciphertexts in the lowest level are supported only!");
    Context &c = *a.ctx;
    const size_t n = c.n, L = c.L;
    DevBuf low(&c, 2 * n);
    if (a.ntt)
        ntt_inv_from(c, low.get(), a.data.get(), 2, 1, n, n);
    else
        d2d(c, low.get(), a.data.get(), 2 * n);
    DevBuf out(&c, 2 * L * n);
    mod_raise(c.tables, out.get(), low.get(), 2, (uint32_t)L, c.stream);
    ntt_fwd(c, out.get(), 2, L, L * n);
    Ciphertext r;
    r.ctx = &c;
    r.size = 2;
    r.limbs = L;
    r.scale = a.scale;
    r.ntt = true;
    r.data = std::move(out);
    dst = std::move(r);
}

void op_dot_plain(const std::vector<const Ciphertext *> &cts, const std::vector<const Plaintext *> &pts,
                  Ciphertext &dst)
{
    require(!cts.empty() && cts.size() == pts.size(), "ciphertext and plaintext counts must match");
    const Ciphertext &a0 = *cts[0];
    valid(a0);
    require(a0.size == 2 && a0.ntt, "encrypted must be size 2 in NTT form");
    Context &c = *a0.ctx;
    const double scale = a0.scale * pts[0]->scale;
    require(scale_within_bounds(c, scale, a0.limbs), "scale out of bounds");
    std::vector<const uint64_t *> cp, pp;
    for (size_t r = 0; r < cts.size(); r++) {
        const Ciphertext &a = *cts[r];
        const Plaintext &p = *pts[r];
        valid(a);
        require(a.ctx == &c && a.size == 2 && a.limbs == a0.limbs && a.ntt, "encrypted parameter mismatch");
        plain_check(a, p);
        require(are_close(a.scale * p.scale, scale), "scale mismatch");
        cp.push_back(a.data.get());
        pp.push_back(p.data.get());
    }
    DevBuf acc(&c, 2 * a0.limbs * c.n);
    DevBuf part;
    for (size_t base = 0; base < cp.size(); base += kMaxDot) {
        const size_t cnt = std::min<size_t>(kMaxDot, cp.size() - base);
        uint64_t *dst_words = base ? (part.reset(&c, acc.words()), part.get()) : acc.get();
        poly_dot_plain(c.tables, dst_words, cp.data() + base, pp.data() + base, (uint32_t)cnt, (uint32_t)a0.limbs,
                       c.stream);
        if (base) poly_add(c.tables, acc.get(), acc.get(), part.get(), 2, (uint32_t)a0.limbs, c.stream);
    }
    Ciphertext out;
    out.ctx = &c;
    out.size = 2;
    out.limbs = a0.limbs;
    out.scale = scale;
    out.ntt = true;
    out.data = std::move(acc);
    dst = std::move(out);
}

void op_linear_combination(const std::vector<const Ciphertext *> &xs, const std::vector<double> &values,
                           const std::vector<double> &scales, size_t limbs, Ciphertext &dst)
{
    // sum_k x_k * encode_constant(values[k], scales[k]) over the first `limbs`
    // limbs of each x_k: word for word mod_switch + multiply_plain + add of each
    // term, with the sum at the first term's scale
    require(!xs.empty() && xs.size() == values.size() && xs.size() == scales.size(), "term counts must match");
    const Ciphertext &a0 = *xs[0];
    valid(a0);
    Context &c = *a0.ctx;
    require(limbs >= 1 && limbs <= a0.limbs, "level out of range");
    for (const Ciphertext *x : xs) {
        valid(*x);
        require(x->ctx == &c && x->size == 2 && x->ntt && x->limbs >= limbs, "encrypted parameter mismatch");
    }
    const double scale = a0.scale * scales[0];
    require(scale_within_bounds(c, scale, limbs), "scale out of bounds");
    // each term is an add of x_k * constant at scale x_k.scale * scales[k]:
    // SEAL's add requires it to equal the sum's scale
    for (size_t k = 0; k < xs.size(); k++) require(are_close(xs[k]->scale * scales[k], scale), "scale mismatch");
    const size_t K = xs.size();
    std::vector<std::vector<uint64_t>> res(K, std::vector<uint64_t>(kMaxPrimes, 0));
    for (size_t k = 0; k < K; k++) constant_residues(c, values[k], limbs, scales[k], res[k].data());
    Ciphertext out;
    out.ctx = &c;
    out.size = 2;
    out.limbs = limbs;
    out.scale = scale;
    out.ntt = true;
    out.data.reset(&c, 2 * limbs * c.n);
    for (size_t k0 = 0; k0 < K; k0 += kLinMax) {
        const size_t cnt = std::min<size_t>(kLinMax, K - k0);
        const uint64_t *x[kLinMax], *w[kLinMax];
        uint64_t stride[kLinMax];
        for (size_t k = 0; k < cnt; k++) {
            x[k] = xs[k0 + k]->data.get();
            stride[k] = xs[k0 + k]->limbs * c.n;
            w[k] = res[k0 + k].data();
        }
        poly_lincomb(c.tables, out.data.get(), x, stride, w, (uint32_t)cnt, (uint32_t)limbs, k0 > 0, c.stream);
    }
    dst = std::move(out);
}

void op_dot_plain_multi(const std::vector<const Ciphertext *> &cts, const std::vector<const Plaintext *> &pts,
                        const std::vector<Ciphertext *> &outs)
{
    // outs[g] = sum_r pts[g K + r] * cts[r] (null plaintexts: no term), the
    // giant steps of one BSGS output at once; each row equals op_dot_plain over
    // its terms word for word
    const size_t K = cts.size(), G = outs.size();
    require(K >= 1 && G >= 1 && pts.size() == G * K, "need rows x ciphertexts plaintexts");
    const Ciphertext &a0 = *cts[0];
    valid(a0);
    require(a0.size == 2 && a0.ntt, "encrypted must be size 2 in NTT form");
    Context &c = *a0.ctx;
    const size_t l = a0.limbs;
    for (const Ciphertext *a : cts) {
        valid(*a);
        require(a->ctx == &c && a->size == 2 && a->limbs == l && a->ntt, "encrypted parameter mismatch");
    }
    std::vector<double> scale(G, 0.0);
    for (size_t g = 0; g < G; g++) {
        for (size_t r = 0; r < K; r++) {
            const Plaintext *p = pts[g * K + r];
            if (!p) continue;
            plain_check(*cts[r], *p);
            const double sc = cts[r]->scale * p->scale;
            if (scale[g] == 0.0) scale[g] = sc;
            require(are_close(sc, scale[g]), "scale mismatch");
        }
        require(scale[g] != 0.0, "every output needs a term");
        require(scale_within_bounds(c, scale[g], l), "scale out of bounds");
    }
    std::vector<Ciphertext> res(G);
    std::vector<uint64_t *> op(G);
    for (size_t g = 0; g < G; g++) {
        Ciphertext &t = res[g];
        t.ctx = &c;
        t.size = 2;
        t.limbs = l;
        t.scale = scale[g];
        t.ntt = true;
        t.data.reset(&c, 2 * l * c.n);
        op[g] = t.data.get();
    }
    std::vector<const uint64_t *> cp(K), pp;
    for (size_t r = 0; r < K; r++) cp[r] = cts[r]->data.get();
    for (size_t g0 = 0; g0 < G; g0 += kMultiDotG) {
        const size_t rows = std::min<size_t>(kMultiDotG, G - g0);
        for (size_t r0 = 0; r0 < K; r0 += kMultiDotK) {
            const size_t cnt = std::min<size_t>(kMultiDotK, K - r0);
            pp.assign(rows * cnt, nullptr);
            for (size_t g = 0; g < rows; g++)
                for (size_t r = 0; r < cnt; r++) {
                    const Plaintext *p = pts[(g0 + g) * K + r0 + r];
                    pp[g * cnt + r] = p ? p->data.get() : nullptr;
                }
            poly_dot_plain_multi(c.tables, op.data() + g0, cp.data() + r0, pp.data(), (uint32_t)rows, (uint32_t)cnt,
                                 (uint32_t)l, r0 > 0, c.stream);
        }
    }
    for (size_t g = 0; g < G; g++) *outs[g] = std::move(res[g]);
}

void op_add_many(const std::vector<const Ciphertext *> &in, Ciphertext &dst)
{
    require(!in.empty(), "encrypteds cannot be empty");
    Ciphertext acc;
    if (in.size() == 1) {
        ct_copy(acc, *in[0]);
    } else {
        op_add(*in[0], *in[1], acc);
        for (size_t i = 2; i < in.size(); i++) op_add(acc, *in[i], acc);
    }
    dst = std::move(acc);
}

}  // namespace nx
