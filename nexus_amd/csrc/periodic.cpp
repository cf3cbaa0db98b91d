// Periodic plaintexts and the block mix (kernels_mix.hip).
//
// Encoding.  With slot j holding m(psi^(5^j)) (SEAL's slot order, psi =
// exp(i pi / n)), a real vector v of period P evaluates a polynomial
// m(X) = sum_{k < B} c_k X^(k n / B), B = 2 P, at zeta^(5^j mod 2B) with
// zeta = exp(2 pi i / 2B).  Inverting over the B primitive 2B-th roots
// (5^j and -5^j for j < P) gives
//     c_k = (1 / P) sum_{j < P} v_j cos(pi (5^j mod 2B) k / B),
// evaluated here in long double and rounded half away from zero at `scale`.
// No other coefficient is non-zero, so the NTT form is exactly constant over
// blocks of n / B positions; creation checks that on the device.
//
// The sums are one length-2B transform per entry: with z_g = v_j at g = 5^j
// mod 2B (zero elsewhere), sum_j v_j cos(pi g_j k / B) = Re sum_g z_g
// exp(2 pi i g k / 2B), an unnormalised inverse DFT (radix-2, long double),
// O(B log B) per entry instead of O(P B).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

#include "context.h"
#include "numth.h"

namespace nx {

namespace {

typedef std::complex<long double> cld;

// a <- (sum_g a_g exp(2 pi i g k / M))_k for M = a.size() a power of two; w[t] = exp(2 pi i t / M), t < M/2
void inverse_dft(std::vector<cld> &a, const std::vector<cld> &w, int logm)
{
    const size_t M = a.size();
    for (size_t i = 0; i < M; i++) {
        const size_t j = reverse_bits((uint32_t)i, logm);
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= M; len <<= 1) {
        const size_t half = len / 2, step = M / len;
        for (size_t s0 = 0; s0 < M; s0 += len)
            for (size_t j = 0; j < half; j++) {
                const cld t = a[s0 + j + half] * w[j * step];
                a[s0 + j + half] = a[s0 + j] - t;
                a[s0 + j] += t;
            }
    }
}

}  // namespace

void periodic_create(Context &c, const double *values, size_t period, size_t count, size_t limbs, double scale,
                     PeriodicPlain &out)
{
    require(values != nullptr, "values is null");
    require(period >= 1 && (period & (period - 1)) == 0 && period <= c.n / 2,
            "period must be a power of two between 1 and the slot count");
    require(count >= 1 && count <= ((size_t)1 << 20), "count must be in 1..2^20");
    require(limbs >= 1 && limbs <= c.L, "chain index out of range");
    require(std::isfinite(scale) && scale > 0 && scale < std::ldexp(1.0, 62), "scale out of bounds");
    const size_t B = 2 * period, n = c.n;
    // 5^j mod 2B
    std::vector<uint64_t> g(period);
    uint64_t x = 1;
    for (size_t j = 0; j < period; j++) {
        g[j] = x;
        x = (x * 5) % (2 * B);
    }
    std::vector<uint64_t> res(count * limbs * B);
    const long double pi = 3.141592653589793238462643383279502884L;
    const size_t M = 2 * B;
    const int logm = __builtin_ctzll((unsigned long long)M);
    std::vector<cld> w(M / 2), z(M);
    for (size_t t = 0; t < M / 2; t++) {
        const long double ang = 2 * pi * (long double)t / (long double)M;
        w[t] = cld(cosl(ang), sinl(ang));
    }
    for (size_t e = 0; e < count; e++) {
        const double *v = values + e * period;
        for (size_t j = 0; j < period; j++) require(std::isfinite(v[j]), "values must be finite");
        std::fill(z.begin(), z.end(), cld(0));
        for (size_t j = 0; j < period; j++) z[g[j]] = (long double)v[j];
        inverse_dft(z, w, logm);
        for (size_t k = 0; k < B; k++) {
            long double ck = z[k].real();
            ck = ck / (long double)period * (long double)scale;
            require(fabsl(ck) < 4.611686018427387904e18L, "encoded value too large for the scale");
            const long long r = llroundl(ck);
            for (size_t i = 0; i < limbs; i++) {
                const uint64_t q = c.moduli[i];
                const uint64_t a = (uint64_t)(r < 0 ? -(unsigned long long)r : (unsigned long long)r) % q;
                res[(e * limbs + i) * B + k] = (r < 0 && a) ? q - a : a;
            }
        }
    }
    const size_t rows = count * limbs;
    DevBuf coef(&c, res.size()), work(&c, rows * n), table(&c, rows * B);
    DevBuf flags(&c, (rows * B + 7) / 8);
    NX_HIP(hipMemcpyAsync(coef.get(), res.data(), res.size() * 8, hipMemcpyHostToDevice, c.stream));
    uint8_t *fl = reinterpret_cast<uint8_t *>(flags.get());
    periodic_scatter(c.tables, work.get(), fl, coef.get(), (uint32_t)rows, (uint32_t)B, c.stream);
    NttBatch b;
    b.data = work.get();
    b.poly_stride = limbs * n;
    b.polys = (uint32_t)count;
    b.limbs = (uint32_t)limbs;
    b.prime_by_poly = 0;
    for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)i;
    ntt_forward(c.tables, b, c.stream);
    periodic_gather(c.tables, table.get(), fl, work.get(), (uint32_t)rows, (uint32_t)B, c.stream);
    std::vector<uint8_t> hf(rows * B);
    NX_HIP(hipMemcpyAsync(hf.data(), fl, hf.size(), hipMemcpyDeviceToHost, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));
    for (uint8_t f : hf) require(f == 0, "periodic plaintext is not constant over its NTT blocks", kLogicError);
    std::lock_guard<std::mutex> lk(out.mu);
    out.ctx = &c;
    out.count = count;
    out.period = period;
    out.blocks = B;
    out.limbs = limbs;
    out.scale = scale;
    out.table = std::move(table);
    out.terms.clear();
}

void periodic_expand_entry(const PeriodicPlain &p, size_t index, Plaintext &out)
{
    require(p.ctx != nullptr, "periodic plaintext is not initialised");
    require(index < p.count, "entry index out of range");
    Context &c = *p.ctx;
    Plaintext r;
    r.ctx = &c;
    r.limbs = p.limbs;
    r.scale = p.scale;
    r.ntt = true;
    r.data.reset(&c, p.limbs * c.n);
    periodic_expand(c.tables, r.data.get(), p.table.get() + index * p.limbs * p.blocks, (uint32_t)p.limbs,
                    (uint32_t)p.blocks, c.stream);
    out = std::move(r);
}

void op_periodic_mix(const std::vector<const Ciphertext *> &in, PeriodicPlain &w, const int32_t *index, size_t M,
                     const std::vector<Ciphertext *> &out)
{
    require(w.ctx != nullptr, "periodic plaintext is not initialised");
    Context &c = *w.ctx;
    const size_t K = in.size();
    require(K >= 1 && K <= (size_t)kMixMax, "block mix needs 1..128 inputs");
    require(M >= 1 && M <= (size_t)kMixMax && out.size() == M, "block mix needs 1..128 outputs");
    require(index != nullptr, "index is null");
    require(c.n >= 64, "block mix needs a ring of at least 64 coefficients");
    require(block_mix_lds_bytes((uint32_t)K, (uint32_t)w.count, (uint32_t)w.blocks, c.logn) <= 65536,
            "block mix tile exceeds 64 KiB of LDS (too many inputs for this period)");
    const Ciphertext &a0 = *in[0];
    require(a0.ctx == &c && a0.size == 2 && a0.ntt, "block mix inputs must be size-2 NTT-form ciphertexts");
    require(a0.limbs == w.limbs, "encrypted and plain parameter mismatch");
    for (const Ciphertext *a : in) {
        require(a != nullptr, "input is null");
        require(a->ctx == &c && a->size == 2 && a->ntt && a->limbs == a0.limbs, "encrypted parameter mismatch");
        require(are_close(a->scale, a0.scale), "scale mismatch");
    }
    for (Ciphertext *o : out) require(o != nullptr, "output is null");
    const double scale = a0.scale * w.scale;
    require(!(scale <= 0 || (int)std::log2(scale) >= c.total_bits(a0.limbs)), "scale out of bounds");
    std::vector<int32_t> key(index, index + M * K);
    for (int32_t e : key) require(e >= -1 && e < (int32_t)w.count, "entry index out of range");
    key.push_back((int32_t)M);
    key.push_back((int32_t)K);
    const PeriodicPlain::Terms *terms;
    {
        std::lock_guard<std::mutex> lk(w.mu);
        auto it = w.terms.find(key);
        if (it == w.terms.end()) {
            // first[M + 1] (u32, padded to a multiple of four) then runs (k0,
            // entry0, len, dentry) as int32 quads: consecutive terms of an output
            // with k + 1 and a constant entry step share one run
            const size_t head = (M + 1 + 3) & ~(size_t)3;
            std::vector<int32_t> h(head, 0);
            std::vector<int32_t> runs;
            for (size_t m = 0; m < M; m++) {
                h[m] = (int32_t)(runs.size() / 4);
                size_t k = 0;
                while (k < K) {
                    const int32_t e = index[m * K + k];
                    if (e < 0) {
                        k++;
                        continue;
                    }
                    int32_t len = 1, de = 0;
                    if (k + 1 < K && index[m * K + k + 1] >= 0) {
                        de = index[m * K + k + 1] - e;
                        while (k + len < K && index[m * K + k + len] >= 0 &&
                               index[m * K + k + len] == e + (int32_t)len * de)
                            len++;
                    }
                    runs.push_back((int32_t)k);
                    runs.push_back(e);
                    runs.push_back(len);
                    runs.push_back(de);
                    k += (size_t)len;
                }
            }
            h[M] = (int32_t)(runs.size() / 4);
            const uint32_t nterms = (uint32_t)(runs.size() / 4);
            h.insert(h.end(), runs.begin(), runs.end());
            PeriodicPlain::Terms t;
            t.words.reset(&c, (h.size() + 1) / 2);
            // stream-ordered: the buffer may come from work still queued on the stream
            NX_HIP(hipMemcpyAsync(t.words.get(), h.data(), h.size() * 4, hipMemcpyHostToDevice, c.stream));
            NX_HIP(hipStreamSynchronize(c.stream));
            t.M = (uint32_t)M;
            t.nterms = nterms;
            it = w.terms.emplace(std::move(key), std::move(t)).first;
        }
        terms = &it->second;
    }
    const size_t n = c.n, l = a0.limbs;
    std::vector<Ciphertext> res(M);
    MixPtrs p;
    std::memset(&p, 0, sizeof p);
    for (size_t k = 0; k < K; k++) p.in[k] = in[k]->data.get();
    for (size_t m = 0; m < M; m++) {
        res[m].ctx = &c;
        res[m].size = 2;
        res[m].limbs = l;
        res[m].scale = scale;
        res[m].ntt = true;
        res[m].data.reset(&c, 2 * l * n);
        p.out[m] = res[m].data.get();
    }
    const uint32_t *first = reinterpret_cast<const uint32_t *>(terms->words.get());
    const int32_t *runs = reinterpret_cast<const int32_t *>(first + ((M + 1 + 3) & ~(size_t)3));
    block_mix(c.tables, p, w.table.get(), runs, first, (uint32_t)K, (uint32_t)M, terms->nterms, (uint32_t)l,
              (uint32_t)w.count, (uint32_t)w.blocks, c.stream);
    for (size_t m = 0; m < M; m++) *out[m] = std::move(res[m]);
}

}  // namespace nx
