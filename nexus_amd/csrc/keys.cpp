// Secret / key-switching key generation, encryption and decryption.
//
// The key structure is SEAL's (keygenerator.cpp:56-336, rlwe.cpp): the secret
// is ternary (optionally with a fixed Hamming weight, the bs fork's
// set_secret_key_hamming_weight), and key-switching key digit j is
//   ( -a_j s + e_j + [i in digit j] (P mod q_i) s' ,  a_j )   in every key-level limb i,
// so the residues that key switching produces follow from the same algebra.
// The randomness is this engine's own counter-based generator, not SEAL's
// Blake2/SHAKE stream, so keys are not byte-identical to SEAL's keys.
#include <algorithm>
#include <random>

#include "context.h"

namespace nx {

namespace {

uint64_t mix(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

uint64_t subseed(uint64_t seed, uint64_t a, uint64_t b = 0) { return mix(mix(seed ^ mix(a)) ^ mix(b + 0x51ed)); }

void key_primes(const Context &c, size_t count, uint8_t *p)
{
    for (size_t i = 0; i < count; i++) p[i] = (uint8_t)i;
}

// Error polynomial (centred binomial) in NTT form over limbs 0..count-1.
void sample_error_ntt(Context &c, uint64_t *out, size_t count, uint64_t seed)
{
    uint8_t p[kMaxPrimes];
    key_primes(c, count, p);
    sample_cbd(c.tables, out, (uint32_t)count, p, seed, c.stream);
    ntt_fwd(c, out, 1, count, count * c.n);
}

// (b, a) = (-a s + e, a) at `count` limbs, written to out[0..1][count][n].
void encrypt_zero_sk(Context &c, const SecretKey &sk, uint64_t *out, size_t count, size_t stride, uint64_t seed)
{
    uint8_t p[kMaxPrimes];
    key_primes(c, count, p);
    uint64_t *b = out, *a = out + stride;
    sample_uniform(c.tables, a, 1, (uint32_t)count, p, subseed(seed, 1), c.stream);
    DevBuf e(&c, count * c.n);
    sample_error_ntt(c, e.get(), count, subseed(seed, 2));
    poly_mul_add(c.tables, b, a, sk.ntt.get(), e.get(), (uint32_t)count, p, 1, c.stream);
}

// One key-switching key for new_key (NTT form, key-level limbs) ->
// keygenerator.cpp:303-336.  Digit j adds (P mod q_i) new_key to the limbs i of
// its primes (one prime with one special prime; alpha consecutive primes in
// hybrid mode), P the product of the special primes.
// aseeds receives, per digit, the seed of the digit's uniform half (poly 1):
// its residues are sample_uniform's words under that seed.
void kswitch_key(Context &c, const SecretKey &sk, const uint64_t *new_key, uint64_t seed, DevBuf &out,
                 DevBuf &aseeds)
{
    const size_t L = c.L, K = c.key_limbs(), n = c.n;
    out.reset(&c, c.key_words());
    std::vector<uint64_t> hs(c.dnum);
    for (size_t j = 0; j < c.dnum; j++) {
        uint64_t *dj = out.get() + j * 2 * K * n;
        hs[j] = subseed(subseed(seed, 100 + j), 1);  // the uniform sample's seed inside encrypt_zero_sk
        encrypt_zero_sk(c, sk, dj, K, K * n, subseed(seed, 100 + j));
        for (size_t i = j * c.alpha; i < std::min(L, (j + 1) * c.alpha); i++) {
            const uint64_t qi = c.moduli[i];
            uint64_t factor = 1;
            for (size_t k = 0; k < c.alpha; k++) factor = host_mul_mod(factor, c.moduli[L + k] % qi, qi);
            limb_add_scaled(c.tables, dj + i * n, new_key + i * n,
                            make_ulonglong2(factor, shoup_quotient(factor, qi)), (uint32_t)i, c.stream);
        }
    }
    aseeds.reset(&c, c.dnum);
    NX_HIP(hipMemcpyAsync(aseeds.get(), hs.data(), hs.size() * 8, hipMemcpyHostToDevice, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));  // hs is a local
}

}  // namespace

void secret_key_from_coeffs(Context &c, const int8_t *coeff, SecretKey &sk)
{
    const size_t K = c.key_limbs(), n = c.n;
    sk.ctx = &c;
    sk.coeff.assign(coeff, coeff + n);
    for (int8_t v : sk.coeff) require(v >= -1 && v <= 1, "secret key coefficients must be ternary");
    sk.ntt.reset(&c, K * n);
    DevBuf d8(&c, (n + 7) / 8);
    NX_HIP(hipMemcpyAsync(d8.get(), sk.coeff.data(), n, hipMemcpyHostToDevice, c.stream));
    uint8_t p[kMaxPrimes];
    key_primes(c, K, p);
    small_to_rns(c.tables, sk.ntt.get(), (const int8_t *)d8.get(), (uint32_t)K, p, c.stream);
    ntt_fwd(c, sk.ntt.get(), 1, K, K * n);
    NX_HIP(hipStreamSynchronize(c.stream));
}

void secret_key_generate(Context &c, uint64_t seed, SecretKey &sk)
{
    const size_t n = c.n;
    std::vector<int8_t> s(n, 0);
    std::mt19937_64 rng(seed);
    if (c.hamming_weight) {
        // sparse ternary: `h` distinct positions with random signs (rlwe.cpp sample_poly_sparse_ternary)
        require(c.hamming_weight <= n, "hamming weight too large");
        std::vector<uint32_t> idx(n);
        for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
        for (size_t i = 0; i < c.hamming_weight; i++) {
            size_t j = i + rng() % (n - i);
            std::swap(idx[i], idx[j]);
            s[idx[i]] = (rng() & 1) ? 1 : -1;
        }
    } else {
        for (size_t i = 0; i < n; i++) s[i] = (int8_t)((int)(rng() % 3) - 1);
    }
    secret_key_from_coeffs(c, s.data(), sk);
}

void relin_key_generate(Context &c, const SecretKey &sk, uint64_t seed, KSwitchKeys &out)
{
    const size_t K = c.key_limbs(), n = c.n;
    uint8_t p[kMaxPrimes];
    key_primes(c, K, p);
    DevBuf s2(&c, K * n);
    poly_mul_add(c.tables, s2.get(), sk.ntt.get(), sk.ntt.get(), nullptr, (uint32_t)K, p, 0, c.stream);
    out.ctx = &c;
    kswitch_key(c, sk, s2.get(), seed, out.keys[0], out.seeds[0]);
    NX_HIP(hipStreamSynchronize(c.stream));
}

void galois_keys_generate(Context &c, const SecretKey &sk, const uint32_t *elts, size_t count, uint64_t seed,
                          KSwitchKeys &out)
{
    const size_t K = c.key_limbs(), n = c.n;
    out.ctx = &c;
    DevBuf rot(&c, K * n);
    for (size_t i = 0; i < count; i++) {
        const uint32_t elt = elts[i];
        require((elt & 1) && elt < 2 * n, "Galois element is not valid");
        if (out.keys.count(elt)) continue;
        const uint32_t *perm = c.galois_table(elt);
        poly_permute(rot.get(), sk.ntt.get(), perm, 1, (uint32_t)K, c.logn, c.stream);
        kswitch_key(c, sk, rot.get(), subseed(seed, elt, 7), out.keys[elt], out.seeds[elt]);
    }
    NX_HIP(hipStreamSynchronize(c.stream));
}

void encrypt_symmetric(Context &c, const SecretKey &sk, const Plaintext &pt, uint64_t seed, Ciphertext &out)
{
    require(pt.ctx == &c && pt.ntt, "plain is not valid for encryption parameters");
    const size_t l = pt.limbs, n = c.n;
    Ciphertext ct;
    ct.ctx = &c;
    ct.size = 2;
    ct.limbs = l;
    ct.scale = pt.scale;
    ct.ntt = true;
    ct.data.reset(&c, 2 * l * n);
    encrypt_zero_sk(c, sk, ct.data.get(), l, l * n, seed);
    poly_add(c.tables, ct.data.get(), ct.data.get(), pt.data.get(), 1, (uint32_t)l, c.stream);
    out = std::move(ct);
}

void decrypt(Context &c, const SecretKey &sk, const Ciphertext &ct, Plaintext &out)
{
    require(ct.ctx == &c && ct.ntt, "encrypted is not valid for encryption parameters");
    require(ct.size == 2 || ct.size == 3, "encrypted size must be 2 or 3");
    const size_t l = ct.limbs, n = c.n;
    uint8_t p[kMaxPrimes];
    key_primes(c, l, p);
    Plaintext r;
    r.ctx = &c;
    r.limbs = l;
    r.scale = ct.scale;
    r.ntt = true;
    r.data.reset(&c, l * n);
    const uint64_t *c0 = ct.data.get(), *c1 = c0 + l * n;
    poly_mul_add(c.tables, r.data.get(), c1, sk.ntt.get(), c0, (uint32_t)l, p, 0, c.stream);
    if (ct.size == 3) {
        DevBuf t(&c, l * n);
        poly_mul_add(c.tables, t.get(), c0 + 2 * l * n, sk.ntt.get(), nullptr, (uint32_t)l, p, 0, c.stream);
        poly_mul_add(c.tables, r.data.get(), t.get(), sk.ntt.get(), r.data.get(), (uint32_t)l, p, 0, c.stream);
    }
    out = std::move(r);
}

}  // namespace nx
