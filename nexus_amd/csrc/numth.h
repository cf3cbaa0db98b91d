// Host-side number theory for context set-up: NTT-friendly prime selection,
// minimal primitive roots and the SEAL-ordered root tables.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace nx {

bool is_prime(uint64_t n);

// Primes = 1 mod factor below 2^bit_size, in descending order
// (thirdparty/SEAL-4.1-bs/native/src/seal/util/numth.cpp:278-310).
std::vector<uint64_t> get_primes(uint64_t factor, int bit_size, size_t count);

// CoeffModulus::Create(poly_modulus_degree, bit_sizes)
// (thirdparty/SEAL-4.1-bs/native/src/seal/modulus.cpp:143-184): the first
// occurrence of a bit size receives the smallest of the primes selected for it.
std::vector<uint64_t> coeff_modulus_create(size_t n, const std::vector<int> &bit_sizes);

// Smallest primitive `degree`-th root of unity mod q (numth.cpp:386-412).
uint64_t minimal_primitive_root(uint64_t degree, uint64_t q);

uint32_t reverse_bits(uint32_t x, int bits);

// SEAL NTT tables (util/ntt.cpp:241-300):
//   fwd[brv(i)] = psi^i for i >= 1, fwd[0] = 1
//   inv[brv(i-1)+1] = psi^-i for i >= 1, inv[0] = 1
void ntt_root_tables(int logn, uint64_t q, std::vector<uint64_t> &fwd, std::vector<uint64_t> &inv);

}  // namespace nx
