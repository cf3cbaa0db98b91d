#include "numth.h"

#include <stdexcept>
#include <unordered_map>

#include "modarith.h"

namespace nx {

bool is_prime(uint64_t n)
{
    // Deterministic Miller-Rabin: these bases decide every n < 2^64.
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (n < 2) return false;
    for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
        if (n == p) return true;
        if (n % p == 0) return false;
    }
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) {
        d >>= 1;
        r++;
    }
    for (uint64_t b : bases) {
        uint64_t a = b % n;
        if (a == 0) continue;
        uint64_t x = host_pow_mod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int i = 1; i < r; i++) {
            x = host_mul_mod(x, x, n);
            if (x == n - 1) {
                composite = false;
                break;
            }
        }
        if (composite) return false;
    }
    return true;
}

std::vector<uint64_t> get_primes(uint64_t factor, int bit_size, size_t count)
{
    std::vector<uint64_t> out;
    uint64_t value = ((uint64_t(1) << bit_size) - 1) / factor * factor + 1;
    uint64_t lower = uint64_t(1) << (bit_size - 1);
    while (out.size() < count && value > lower) {
        if (is_prime(value)) out.push_back(value);
        value -= factor;
    }
    if (out.size() < count) throw std::logic_error("failed to find enough qualifying primes");
    return out;
}

std::vector<uint64_t> coeff_modulus_create(size_t n, const std::vector<int> &bit_sizes)
{
    if (n < 2 || (n & (n - 1))) throw std::invalid_argument("poly_modulus_degree is invalid");
    std::unordered_map<int, size_t> count;
    for (int b : bit_sizes) {
        // SEAL_USER_MOD_BIT_COUNT_MAX = 60 (util/defines.h:40): the lazy 128-bit
        // accumulators sum up to 128 products of two residues below 2^60
        if (b < 2 || b > 60) throw std::invalid_argument("bit_sizes is invalid");
        count[b]++;
    }
    std::unordered_map<int, std::vector<uint64_t>> pool;
    for (auto &kv : count) pool[kv.first] = get_primes(2 * (uint64_t)n, kv.first, kv.second);
    std::vector<uint64_t> out;
    for (int b : bit_sizes) {
        out.push_back(pool[b].back());
        pool[b].pop_back();
    }
    return out;
}

uint64_t minimal_primitive_root(uint64_t degree, uint64_t q)
{
    if ((q - 1) % degree) throw std::invalid_argument("invalid modulus");
    uint64_t cofactor = (q - 1) / degree;
    uint64_t root = 0;
    for (uint64_t g = 2; g < q && !root; g++) {
        uint64_t c = host_pow_mod(g, cofactor, q);
        if (host_pow_mod(c, degree >> 1, q) == q - 1) root = c;
    }
    if (!root) throw std::invalid_argument("invalid modulus");
    // The primitive degree-th roots are the odd powers of `root`.
    uint64_t sq = host_mul_mod(root, root, q), cur = root, best = root;
    for (uint64_t i = 0; i < degree; i += 2) {
        if (cur < best) best = cur;
        cur = host_mul_mod(cur, sq, q);
    }
    return best;
}

uint32_t reverse_bits(uint32_t x, int bits)
{
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

void ntt_root_tables(int logn, uint64_t q, std::vector<uint64_t> &fwd, std::vector<uint64_t> &inv)
{
    size_t n = size_t(1) << logn;
    uint64_t psi = minimal_primitive_root(2 * n, q);
    uint64_t ipsi = host_inv_mod(psi, q);
    fwd.assign(n, 0);
    inv.assign(n, 0);
    uint64_t p = psi;
    for (size_t i = 1; i < n; i++) {
        fwd[reverse_bits((uint32_t)i, logn)] = p;
        p = host_mul_mod(p, psi, q);
    }
    fwd[0] = 1;
    p = ipsi;
    for (size_t i = 1; i < n; i++) {
        inv[reverse_bits((uint32_t)(i - 1), logn) + 1] = p;
        p = host_mul_mod(p, ipsi, q);
    }
    inv[0] = 1;
}

}  // namespace nx
