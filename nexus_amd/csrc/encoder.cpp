// CKKS encoder / decoder (SEAL-4.1-bs ckks.cpp:13-71, ckks.h:457-760,
// util/croots.cpp, util/dwthandler.h).  Encoding runs on the device: the
// complex FFT, rounding and RNS decomposition in kernels_encode.hip reproduce
// SEAL's floating-point order bit for bit, then the NTT.  Decoding composes
// the CRT value and runs the inverse FFT on the host in SEAL's order.
// Compiled with -ffp-contract=off: no fused multiply-adds may change the
// rounding of the double-precision butterflies.
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>

#include "context.h"
#include "numth.h"
#include "prof.h"

namespace nx {

namespace {

typedef std::complex<double> cd;

struct EncoderTables {
    int logn = 0;
    std::vector<size_t> index_map;  // matrix_reps_index_map_
    std::vector<cd> roots, iroots;  // root_powers_, inv_root_powers_
};

std::mutex g_enc_mu;
std::map<int, std::unique_ptr<EncoderTables>> g_enc;

const double kPi = 3.1415926535897932384626433832795028842;

cd complex_root(const std::vector<cd> &base, size_t m, size_t index)
{
    // 8-fold symmetry of the m-th roots of unity
    index &= m - 1;
    if (index <= m / 8) return base[index];
    if (index <= m / 4) {
        cd a = base[m / 4 - index];
        return cd(a.imag(), a.real());
    }
    if (index <= m / 2) return -std::conj(complex_root(base, m, m / 2 - index));
    if (index <= 3 * m / 4) return -complex_root(base, m, index - m / 2);
    return std::conj(complex_root(base, m, m - index));
}

const EncoderTables &encoder_tables(int logn)
{
    std::lock_guard<std::mutex> g(g_enc_mu);
    auto it = g_enc.find(logn);
    if (it != g_enc.end()) return *it->second;
    auto t = std::make_unique<EncoderTables>();
    t->logn = logn;
    const size_t n = size_t(1) << logn, slots = n / 2, m = 2 * n;
    t->index_map.resize(n);
    uint64_t pos = 1;
    for (size_t i = 0; i < slots; i++) {
        uint64_t i1 = (pos - 1) >> 1, i2 = (m - pos - 1) >> 1;
        t->index_map[i] = reverse_bits((uint32_t)i1, logn);
        t->index_map[slots | i] = reverse_bits((uint32_t)i2, logn);
        pos = (pos * 5) & (m - 1);
    }
    t->roots.assign(n, cd(0, 0));
    t->iroots.assign(n, cd(0, 0));
    if (m >= 8) {
        std::vector<cd> base(m / 8 + 1);
        for (size_t i = 0; i <= m / 8; i++) base[i] = std::polar<double>(1.0, 2 * kPi * (double)i / (double)m);
        for (size_t i = 1; i < n; i++) {
            t->roots[i] = complex_root(base, m, reverse_bits((uint32_t)i, logn));
            t->iroots[i] = std::conj(complex_root(base, m, reverse_bits((uint32_t)(i - 1), logn) + 1));
        }
    } else if (m == 4) {
        t->roots[1] = cd(0, 1);
        t->iroots[1] = cd(0, -1);
    }
    auto &ref = *t;
    g_enc[logn] = std::move(t);
    return ref;
}

// Cooley-Tukey over doubles, natural in, bit-reversed out (transform_to_rev).
void fft_to_rev(cd *v, int logn, const cd *roots)
{
    const size_t n = size_t(1) << logn;
    size_t gap = n >> 1, m = 1;
    const cd *r = roots;
    for (; m < (n >> 1); m <<= 1) {
        size_t off = 0;
        for (size_t i = 0; i < m; i++) {
            const cd w = *++r;
            cd *x = v + off, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                const cd u = x[j], t = y[j] * w;
                x[j] = u + t;
                y[j] = u - t;
            }
            off += gap << 1;
        }
        gap >>= 1;
    }
    for (size_t i = 0; i < m; i++) {
        const cd w = *++r;
        const cd u = v[0], t = v[1] * w;
        v[0] = u + t;
        v[1] = u - t;
        v += 2;
    }
}

// |coeffd| (already rounded, non-negative) -> residue mod q.  Handles the
// 64-bit, 128-bit and multi-word decompositions (ckks.h:523-610).
struct CoeffDecomposer {
    int bits;  // bit bound of the largest coefficient
    double two64 = std::pow(2.0, 64);
    std::vector<uint64_t> words;
    void load(double cd_abs)
    {
        words.clear();
        if (bits <= 64) {
            words.push_back((uint64_t)cd_abs);
        } else if (bits <= 128) {
            words.push_back((uint64_t)std::fmod(cd_abs, two64));
            words.push_back((uint64_t)(cd_abs / two64));
        } else {
            while (cd_abs >= 1) {
                words.push_back((uint64_t)std::fmod(cd_abs, two64));
                cd_abs /= two64;
            }
        }
    }
    uint64_t mod(uint64_t q) const
    {
        u128 r = 0;
        for (size_t i = words.size(); i-- > 0;) r = ((r << 64) | words[i]) % q;
        return (uint64_t)r;
    }
};

}  // namespace

void upload_encoder_tables(Context &c)
{
    const EncoderTables &t = encoder_tables(c.logn);
    std::vector<uint32_t> idx(c.n);
    for (size_t i = 0; i < c.n; i++) idx[i] = (uint32_t)t.index_map[i];
    static_assert(sizeof(cd) == sizeof(double2), "complex<double> is two doubles");
    NX_HIP(hipMalloc(&c.d_enc_index, c.n * sizeof(uint32_t)));
    NX_HIP(hipMalloc(&c.d_enc_roots, c.n * sizeof(double2)));
    NX_HIP(hipMemcpy(c.d_enc_index, idx.data(), c.n * sizeof(uint32_t), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(c.d_enc_roots, t.iroots.data(), c.n * sizeof(double2), hipMemcpyHostToDevice));
}

// ckks.h:457-620 encode_internal: values -> index_map scatter with conjugates
// -> transform_from_rev with scale / n -> round -> residues -> NTT.  The
// largest |coefficient| comes back to the host for SEAL's size check.
void encode_vector(Context &c, const double *re, const double *im, size_t count, size_t limbs, double scale,
                   Plaintext &out)
{
    const size_t n = c.n, slots = n / 2;
    require(count <= slots, "values_size is too large");
    require(limbs >= 1 && limbs <= c.L, "parms_id is not valid for encryption parameters");
    tally("encode", limbs);
    const int total_bits = c.total_bits(limbs);
    require(!(scale <= 0 || ((int)std::log2(scale) + 1 >= total_bits)), "scale out of bounds");
    const uint32_t blocks = encode_blocks(c.logn);
    // work: [n] double2 | re[count] | im[count] | blockmax[blocks]
    DevBuf work(&c, 2 * n + 2 * count + blocks);
    double2 *v = reinterpret_cast<double2 *>(work.get());
    double *d_re = reinterpret_cast<double *>(work.get() + 2 * n);
    double *d_im = d_re + count;
    uint64_t *blockmax = work.get() + 2 * n + 2 * count;
    if (count) {
        NX_HIP(hipMemcpyAsync(d_re, re, count * 8, hipMemcpyHostToDevice, c.stream));
        if (im) NX_HIP(hipMemcpyAsync(d_im, im, count * 8, hipMemcpyHostToDevice, c.stream));
    }
    Plaintext p;
    p.ctx = &c;
    p.limbs = limbs;
    p.scale = scale;
    p.ntt = true;
    p.data.reset(&c, limbs * n);
    encode_fft(c.tables, p.data.get(), v, d_re, im ? d_im : nullptr, (uint32_t)count, c.d_enc_index, c.d_enc_roots,
               scale / (double)n, (uint32_t)limbs, blockmax, c.stream);
    std::vector<uint64_t> maxima(blocks);
    NX_HIP(hipMemcpyAsync(maxima.data(), blockmax, blocks * 8, hipMemcpyDeviceToHost, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));
    uint64_t mb = 0;
    for (uint64_t x : maxima) mb = std::max(mb, x);
    double max_coeff;
    std::memcpy(&max_coeff, &mb, 8);
    require(std::isfinite(max_coeff), "encoded values are too large");
    const int bits = (int)std::ceil(std::log2(std::max(max_coeff, 1.0))) + 1;
    require(bits < total_bits, "encoded values are too large");
    ntt_fwd(c, p.data.get(), 1, limbs, limbs * n);
    out = std::move(p);
}

// Residues of round(value * scale) for the first `limbs` primes
// (ckks.cpp:73-200, encode_internal(double)).
void constant_residues(Context &c, double value, size_t limbs, double scale, uint64_t *res)
{
    require(limbs >= 1 && limbs <= c.L, "parms_id is not valid for encryption parameters");
    const int total_bits = c.total_bits(limbs);
    require(!(scale <= 0 || ((int)std::log2(scale) >= total_bits)), "scale out of bounds");
    value *= scale;
    const int bits = (int)std::log2(std::fabs(value)) + 2;
    require(bits < total_bits, "encoded value is too large");
    double x = std::round(value);
    const bool neg = std::signbit(x);
    CoeffDecomposer dec;
    dec.bits = bits;
    dec.load(std::fabs(x));
    for (size_t j = 0; j < limbs; j++) {
        const uint64_t q = c.moduli[j];
        const uint64_t r = dec.mod(q);
        res[j] = neg ? (r ? q - r : 0) : r;
    }
}

void encode_constant(Context &c, double value, size_t limbs, double scale, Plaintext &out)
{
    uint64_t res[kMaxPrimes];
    constant_residues(c, value, limbs, scale, res);
    Plaintext p;
    p.ctx = &c;
    p.limbs = limbs;
    p.scale = scale;
    p.ntt = true;
    p.data.reset(&c, limbs * c.n);
    // a constant polynomial is the same constant in every NTT slot
    poly_fill_limbs(p.data.get(), res, 1, (uint32_t)limbs, c.logn, c.stream);
    out = std::move(p);
}

void decode(Context &c, const Plaintext &pt, double *re, double *im, size_t count)
{
    require(pt.ctx == &c, "plain is not valid for encryption parameters");
    require(pt.ntt, "plain is not in NTT form");
    const size_t n = c.n, l = pt.limbs;
    require(!(pt.scale <= 0 || (int)std::log2(pt.scale) >= c.total_bits(l)), "scale out of bounds");
    const size_t sparse = c.sparse_slots;
    require(count <= sparse, "too many values requested");
    DevBuf tmp(&c, l * n);
    NX_HIP(hipMemcpyAsync(tmp.get(), pt.data.get(), l * n * 8, hipMemcpyDeviceToDevice, c.stream));
    ntt_inv(c, tmp.get(), 1, l, l * n);
    std::vector<uint64_t> coef(l * n);
    NX_HIP(hipMemcpyAsync(coef.data(), tmp.get(), l * n * 8, hipMemcpyDeviceToHost, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));

    // CRT composition into [0, Q): x = sum_i [x_i * (Q/q_i)^-1]_{q_i} * (Q/q_i) mod Q
    const size_t W = l;  // words
    std::vector<uint64_t> Q(W + 1, 0), punct(l * (W + 1), 0), inv(l);
    auto mul_word = [](const uint64_t *a, size_t w, uint64_t b, uint64_t *o) {
        u128 carry = 0;
        for (size_t i = 0; i < w; i++) {
            u128 t = (u128)a[i] * b + carry;
            o[i] = (uint64_t)t;
            carry = t >> 64;
        }
        o[w] = (uint64_t)carry;
    };
    std::vector<uint64_t> tw(W + 2);
    Q[0] = 1;
    for (size_t i = 0; i < l; i++) {
        mul_word(Q.data(), W, c.moduli[i], tw.data());
        std::copy(tw.begin(), tw.begin() + W, Q.begin());
    }
    for (size_t i = 0; i < l; i++) {
        uint64_t *p = punct.data() + i * (W + 1);
        p[0] = 1;
        uint64_t r = 1;
        for (size_t j = 0; j < l; j++) {
            if (j == i) continue;
            mul_word(p, W, c.moduli[j], tw.data());
            std::copy(tw.begin(), tw.begin() + W, p);
            r = host_mul_mod(r, c.moduli[j] % c.moduli[i], c.moduli[i]);
        }
        inv[i] = host_inv_mod(r, c.moduli[i]);
    }
    // upper_half_threshold = (Q + 1) / 2
    std::vector<uint64_t> thr(W, 0);
    {
        u128 carry = 1;
        for (size_t w = 0; w < W; w++) {
            u128 s = (u128)Q[w] + carry;
            thr[w] = (uint64_t)s;
            carry = s >> 64;
        }
        for (size_t w = 0; w < W; w++) thr[w] = (thr[w] >> 1) | (w + 1 < W ? (thr[w + 1] << 63) : 0);
    }
    auto geq = [&](const uint64_t *a, const uint64_t *b) {
        for (size_t w = W; w-- > 0;)
            if (a[w] != b[w]) return a[w] > b[w];
        return true;
    };
    const EncoderTables &t = encoder_tables(c.logn);
    std::vector<cd> res(n);
    std::vector<uint64_t> acc(W + 1), prod(W + 2);
    const double inv_scale = 1.0 / pt.scale, two64 = std::pow(2.0, 64);
    const size_t slots = n / 2;
    const size_t sparsity = slots / sparse;
    for (size_t i = 0; i < n; i++) {
        if (sparsity > 1 && (((i - 1) & (sparsity - 1)) != sparsity - 1)) {
            res[i] = 0.0;
            continue;
        }
        std::fill(acc.begin(), acc.end(), 0);
        for (size_t k = 0; k < l; k++) {
            const uint64_t v = host_mul_mod(coef[k * n + i], inv[k], c.moduli[k]);
            mul_word(punct.data() + k * (W + 1), W, v, prod.data());
            u128 carry = 0;
            for (size_t w = 0; w <= W; w++) {
                u128 s = (u128)acc[w] + prod[w] + carry;
                acc[w] = (uint64_t)s;
                carry = s >> 64;
            }
            // acc < 2Q: subtract Q once if acc >= Q
            bool ge = acc[W] != 0 || geq(acc.data(), Q.data());
            if (ge) {
                uint64_t borrow = 0;
                for (size_t w = 0; w <= W; w++) {
                    uint64_t qw = w < W ? Q[w] : 0;
                    uint64_t sub = qw + borrow;
                    uint64_t nb = (sub < borrow) || (acc[w] < sub);
                    acc[w] -= sub;
                    borrow = nb;
                }
            }
        }
        double r = 0.0, s64 = inv_scale;
        if (geq(acc.data(), thr.data())) {
            for (size_t j = 0; j < W; j++, s64 *= two64) {
                if (acc[j] > Q[j]) {
                    uint64_t d = acc[j] - Q[j];
                    r += d ? (double)d * s64 : 0.0;
                } else {
                    uint64_t d = Q[j] - acc[j];
                    r -= d ? (double)d * s64 : 0.0;
                }
            }
        } else {
            for (size_t j = 0; j < W; j++, s64 *= two64) r += acc[j] ? (double)acc[j] * s64 : 0.0;
        }
        res[i] = r;
    }
    fft_to_rev(res.data(), c.logn, t.roots.data());
    for (size_t i = 0; i < count; i++) {
        const cd v = res[t.index_map[i]];
        re[i] = v.real();
        if (im) im[i] = v.imag();
    }
}

}  // namespace nx
