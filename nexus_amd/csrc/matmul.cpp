// Ciphertext-plaintext MatrixMul of NEXUS (reference src/matrix_mul.cpp:58-246,
// cuda/src/matrix_mul.cu:12-275), batched for the GPU:
//  * enc_compress_ciphertext: one operand is coefficient-packed (value * 1e10
//    rounded into the polynomial coefficients) and encrypted;
//  * expand_ciphertext: log2(m) rounds of "x + tau_i(x)" / "x^-2^i (x - tau_i(x))"
//    turn one packed ciphertext into m ciphertexts, the k-th encrypting
//    m * coefficient_k as a constant.  Each round is one batched automorphism
//    (key switch over all 2^i ciphertexts at once) and one fused
//    multiply-by-x^k / add kernel; multiplying by x^k is a pointwise product
//    with NTT(x^k), the same residues as SEAL's INTT -> shift -> NTT;
//  * the product: out_i = sum_j expanded[i * inner + j] (x) a_j, one
//    accumulation kernel for all output columns (multiply_plain + add_many).
#include <cmath>

#include "context.h"

namespace nx {

void encode_vector(Context &c, const double *re, const double *im, size_t count, size_t limbs, double scale,
                   Plaintext &out);

namespace {

void d2d(Context &c, uint64_t *dst, const uint64_t *src, size_t words)
{
    if (words) NX_HIP(hipMemcpyAsync(dst, src, words * 8, hipMemcpyDeviceToDevice, c.stream));
}

}  // namespace

// NTT-domain image of x^k (0 <= k < 2n) over limbs 0..limbs-1, cached per context.
const uint64_t *power_of_x_ntt(Context &c, uint32_t k, size_t limbs)
{
    std::lock_guard<std::mutex> g(c.xk_mu);
    auto key = std::make_pair(k, limbs);
    auto it = c.xk_cache.find(key);
    if (it != c.xk_cache.end()) return it->second.get();
    const size_t n = c.n;
    std::vector<uint64_t> h(limbs * n, 0);
    const uint32_t pos = k % (uint32_t)n;
    const bool neg = (k / (uint32_t)n) & 1;
    for (size_t i = 0; i < limbs; i++) h[i * n + pos] = neg ? c.moduli[i] - 1 : 1;
    DevBuf b(&c, limbs * n);
    NX_HIP(hipMemcpyAsync(b.get(), h.data(), limbs * n * 8, hipMemcpyHostToDevice, c.stream));
    ntt_fwd(c, b.get(), 1, limbs, limbs * n);
    NX_HIP(hipStreamSynchronize(c.stream));
    return c.xk_cache.emplace(key, std::move(b)).first->second.get();
}

// src/matrix_mul.cpp:58-99: coefficient i of the plaintext = round(val[i] * scale)
// reduced into every limb, then NTT; the ciphertext is Enc(0) + plaintext at `scale`.
void mm_compress(Context &c, const SecretKey &sk, const double *val, size_t count, double scale, uint64_t seed,
                 Ciphertext &out)
{
    const size_t n = c.n, limbs = c.L;
    require(count <= n, "too many values for one packed ciphertext");
    std::vector<uint64_t> h(limbs * n, 0);
    for (size_t i = 0; i < count; i++) {
        const double cd = std::round(val[i] * scale);
        const bool neg = std::signbit(cd);
        const uint64_t cu = (uint64_t)std::fabs(cd);
        for (size_t j = 0; j < limbs; j++) {
            const uint64_t q = c.moduli[j];
            const uint64_t r = cu % q;
            h[j * n + i] = neg ? (r ? q - r : 0) : r;
        }
    }
    Plaintext p;
    p.ctx = &c;
    p.limbs = limbs;
    p.scale = scale;
    p.ntt = true;
    p.data.reset(&c, limbs * n);
    NX_HIP(hipMemcpyAsync(p.data.get(), h.data(), limbs * n * 8, hipMemcpyHostToDevice, c.stream));
    ntt_fwd(c, p.data.get(), 1, limbs, limbs * n);
    NX_HIP(hipStreamSynchronize(c.stream));
    // Enc(encode(zeros)): the zero vector encodes to the zero polynomial
    Plaintext zero;
    zero.ctx = &c;
    zero.limbs = limbs;
    zero.scale = scale;
    zero.ntt = true;
    zero.data.reset(&c, limbs * n);
    uint64_t zeros[kMaxPrimes] = {0};
    poly_fill_limbs(zero.data.get(), zeros, 1, (uint32_t)limbs, c.logn, c.stream);
    Ciphertext z;
    encrypt_symmetric(c, sk, zero, seed, z);
    z.scale = scale;
    op_add_plain(z, p, out);
}

// src/matrix_mul.cpp:101-129: oblivious expansion with Galois elements elts[i]
void mm_expand(Context &c, const Ciphertext &ct, uint32_t m, const uint32_t *elts, const KSwitchKeys &gk,
               CtBatch &out)
{
    require(ct.size == 2 && ct.ntt, "expand needs a size-2 NTT-form ciphertext");
    require(m >= 1 && (m & (m - 1)) == 0 && m <= c.n, "m must be a power of two <= n");
    const size_t n = c.n, l = ct.limbs, ctw = 2 * l * n;
    uint32_t logm = 0;
    while ((1u << logm) < m) logm++;
    CtBatch cur;
    cur.ctx = &c;
    cur.B = 1;
    cur.limbs = l;
    cur.scale = ct.scale;
    cur.data.reset(&c, ctw);
    d2d(c, cur.data.get(), ct.data.get(), ctw);
    for (uint32_t i = 0; i < logm; i++) {
        const size_t S = cur.B;
        const uint32_t elt = elts[i];
        auto it = gk.keys.find(elt);
        require(it != gk.keys.end(), "Galois key not present");
        // rot = tau_elt(cur): permute both polys, key-switch the permuted c1
        const uint32_t *perm = c.galois_table(elt);
        DevBuf rot(&c, S * ctw), tgt(&c, S * l * n);
        // c0 of every ciphertext -> rot, c1 -> tgt: permute a [S][2][l] view
        // row by row (rows of one ciphertext are contiguous)
        DevBuf perm_all(&c, S * ctw);
        poly_permute(perm_all.get(), cur.data.get(), perm, (uint32_t)(2 * S), (uint32_t)l, c.logn, c.stream);
        poly_copy_limbs(c.tables, rot.get(), ctw, perm_all.get(), ctw, (uint32_t)S, (uint32_t)l, c.stream);
        poly_copy_limbs(c.tables, tgt.get(), l * n, perm_all.get() + l * n, ctw, (uint32_t)S, (uint32_t)l, c.stream);
        key_switch_batch(c, rot.get(), tgt.get(), S, l, it->second, true, nullptr, gk.seeds_of(elt));
        // next round: [cur + rot | cur * x^(2n - 2^i) + rot * x^((2n - 2^i) * elt mod 2n)]
        CtBatch nxt;
        nxt.ctx = &c;
        nxt.B = 2 * S;
        nxt.limbs = l;
        nxt.scale = ct.scale;
        nxt.data.reset(&c, 2 * S * ctw);
        poly_add(c.tables, nxt.data.get(), cur.data.get(), rot.get(), (uint32_t)(2 * S), (uint32_t)l, c.stream);
        const uint64_t two_n = 2 * (uint64_t)n;
        const uint32_t k_raw = (uint32_t)(two_n - (1ull << i));
        const uint32_t k_idx = (uint32_t)(((uint64_t)k_raw * elt) % two_n);
        const uint64_t *P = power_of_x_ntt(c, k_raw, l), *Q = power_of_x_ntt(c, k_idx, l);
        poly_mul2_add(c.tables, nxt.data.get() + S * ctw, cur.data.get(), P, rot.get(), Q, (uint32_t)(2 * S),
                      (uint32_t)l, c.stream);
        cur = std::move(nxt);
    }
    out = std::move(cur);
}

// src/matrix_mul.cpp:213-236: out[i] = sum_j E[i * inner + j] (x) a[j]
void mm_accumulate_cts(Context &c, const std::vector<const Plaintext *> &a, const std::vector<const CtBatch *> &E,
                       size_t cols, std::vector<Ciphertext> &out)
{
    const size_t inner = a.size(), n = c.n;
    require(inner > 0 && cols > 0, "empty matrix product");
    const size_t l = E[0]->limbs, ctw = 2 * l * n;
    size_t total = 0;
    for (auto *e : E) {
        require(e->limbs == l, "expanded ciphertexts must share a level");
        total += e->B;
    }
    require(total >= cols * inner, "not enough expanded ciphertexts");
    for (auto *p : a) require(p->limbs == l && p->ntt, "plaintext level mismatch");
    // gather: expanded ciphertexts and plaintexts into contiguous operands
    DevBuf Ebuf, Abuf(&c, inner * l * n);
    const uint64_t *Eptr;
    if (E.size() == 1) {
        Eptr = E[0]->data.get();
    } else {
        Ebuf.reset(&c, cols * inner * ctw);
        size_t off = 0;
        for (auto *e : E) {
            const size_t take = std::min(e->B, cols * inner - off);
            if (!take) break;
            d2d(c, Ebuf.get() + off * ctw, e->data.get(), take * ctw);
            off += take;
        }
        Eptr = Ebuf.get();
    }
    for (size_t j = 0; j < inner; j++) d2d(c, Abuf.get() + j * l * n, a[j]->data.get(), l * n);
    DevBuf res(&c, cols * ctw);
    mm_accumulate(c.tables, res.get(), Eptr, Abuf.get(), (uint32_t)cols, (uint32_t)inner, (uint32_t)l, c.stream);
    out.resize(cols);
    const double scale = E[0]->scale * a[0]->scale;
    for (size_t i = 0; i < cols; i++) {
        Ciphertext r;
        r.ctx = &c;
        r.size = 2;
        r.limbs = l;
        r.scale = scale;
        r.ntt = true;
        r.data.reset(&c, ctw);
        d2d(c, r.data.get(), res.get() + i * ctw, ctw);
        out[i] = std::move(r);
    }
}

}  // namespace nx
