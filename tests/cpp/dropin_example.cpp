// A compiled C++ caller of the drop-in (include/nexus/ckks.hpp), written the
// way a NEXUS evaluator calls its CKKSEvaluator (reference
// cuda/src/main.cu:264-278, cuda/src/ckks_evaluator.cuh:421-499): keys from the
// decryptor, encode / encrypt, multiply + relinearize + rescale, rotations,
// add_many, the reduced-error add, decrypt / decode.  Exit status 0 when every
// decrypted result is within 2^-16 of the plaintext arithmetic (scale 2^40,
// one or two key switches), 1 otherwise; the failing check is printed.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "nexus/ckks.hpp"

namespace {

double max_err(const std::vector<double> &a, const std::vector<double> &b)
{
    double m = 0;
    for (size_t i = 0; i < a.size() && i < b.size(); i++) m = std::max(m, std::fabs(a[i] - b[i]));
    return m;
}

}  // namespace

int main()
{
    try {
        nexus::CKKSContext ctx(13, {60, 40, 40, 60});
        nexus::SecretKey sk(ctx, 1);
        nexus::RelinKeys rk(ctx);
        nexus::GaloisKeys gk(ctx);
        nexus::Decryptor dec(&ctx, &sk);
        dec.create_relin_keys(rk);
        dec.create_galois_keys_from_steps({1, 4, -1}, gk);
        nexus::CKKSEvaluator ckks(&ctx, &sk, &rk, &gk, std::pow(2.0, 40));
        const size_t n = ckks.slot_count;
        std::mt19937_64 rng(7);
        std::uniform_real_distribution<double> u(-1, 1);
        std::vector<std::vector<double>> v(4, std::vector<double>(n));
        std::vector<nexus::Ciphertext> c(4);
        for (int k = 0; k < 4; k++) {
            for (auto &x : v[k]) x = u(rng);
            nexus::Plaintext p(ctx);
            ckks.encoder.encode(v[k], ckks.scale, p);
            ckks.encryptor.encrypt(p, c[k]);
        }
        auto decrypt = [&](const nexus::Ciphertext &ct) {
            nexus::Plaintext p(ctx);
            ckks.decryptor.decrypt(ct, p);
            std::vector<double> r;
            ckks.encoder.decode(p, r);
            return r;
        };
        const double tol = std::pow(2.0, -16);
        int failures = 0;
        auto expect = [&](const char *what, const std::vector<double> &got, const std::vector<double> &want) {
            const double e = max_err(got, want);
            std::printf("%-28s max error %.3g\n", what, e);
            if (!(e <= tol)) failures++;
        };
        // product, relinearised and rescaled
        nexus::Ciphertext prod;
        ckks.evaluator.multiply(c[0], c[1], prod);
        ckks.evaluator.relinearize_inplace(prod, rk);
        ckks.evaluator.rescale_to_next_inplace(prod);
        std::vector<double> want(n);
        for (size_t i = 0; i < n; i++) want[i] = v[0][i] * v[1][i];
        expect("multiply/relin/rescale", decrypt(prod), want);
        // rotations: one with its own key, two composed from keys by their NAF
        // (5 = 4 + 1, 3 = 4 - 1; evaluator.cpp:2667-2722)
        for (int step : {1, 5, 3}) {
            nexus::Ciphertext r;
            ckks.evaluator.rotate_vector(c[2], step, gk, r);
            for (size_t i = 0; i < n; i++) want[i] = v[2][((long)i + step + (long)n) % (long)n];
            char what[32];
            std::snprintf(what, sizeof what, "rotate_vector(%d)", step);
            expect(what, decrypt(r), want);
        }
        // add_many over four ciphertexts
        nexus::Ciphertext sum;
        ckks.evaluator.add_many(c, sum);
        for (size_t i = 0; i < n; i++) want[i] = v[0][i] + v[1][i] + v[2][i] + v[3][i];
        expect("add_many", decrypt(sum), want);
        // reduced-error add of a rescaled product and a fresh ciphertext (levels differ)
        nexus::Ciphertext mix;
        ckks.evaluator.add_reduced_error(prod, c[3], mix);
        for (size_t i = 0; i < n; i++) want[i] = v[0][i] * v[1][i] + v[3][i];
        expect("add_reduced_error", decrypt(mix), want);
        // the reference's exception types survive the C boundary
        bool threw = false;
        try {
            nexus::Ciphertext bad;
            ckks.evaluator.add(prod, c[0], bad);  // different levels
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        std::printf("%-28s %s\n", "level mismatch throws", threw ? "yes" : "no");
        if (!threw) failures++;
        std::printf(failures ? "FAILED %d\n" : "ok\n", failures);
        return failures ? 1 : 0;
    } catch (const std::exception &e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
}
