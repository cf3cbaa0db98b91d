// Autotuner of launch shapes at a context's ring size: the tile widths and the
// radix (8 or 16) of the two-pass NTT (per pass, direction and power-of-two
// batch bucket), the thread layout of the fused multiply-accumulate (per limb
// bucket), the base conversions fused into the NTT column pass or not (and
// their outputs per workgroup, tile width and threads), and the arithmetic of
// the ModDown epilogue.  Every
// candidate's output is compared word for word with the default shape's on
// the same data before it may be chosen; the choices are process-wide.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

namespace nx {

namespace {

struct Events {
    hipEvent_t a, b;
    Events()
    {
        NX_HIP(hipEventCreate(&a));
        NX_HIP(hipEventCreate(&b));
    }
    ~Events()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

// residues of `count` limbs, limb i below the prime (i mod period)
void fill_random(Context &c, std::vector<uint64_t> &h, size_t count, uint64_t seed, size_t period)
{
    h.resize(count * c.n);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    for (size_t i = 0; i < count; i++) {
        const uint64_t q = c.moduli[i % period];
        for (size_t k = 0; k < c.n; k++) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            h[i * c.n + k] = x % q;
        }
    }
}

std::vector<uint64_t> download(Context &c, const uint64_t *d, size_t words)
{
    std::vector<uint64_t> h(words);
    NX_HIP(hipMemcpyAsync(h.data(), d, words * 8, hipMemcpyDeviceToHost, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));
    return h;
}

}  // namespace

std::string autotune(Context &c, int reps)
{
    require(reps >= 1 && reps <= 1000, "reps must be in 1..1000");
    const size_t n = c.n, P = c.moduli.size();
    Events ev;
    std::string rep = "{\"logn\": " + std::to_string(c.logn) + ", \"ntt\": [";
    bool first_entry = true;
    auto time_ms = [&](auto &&fn) {
        fn();  // warm
        NX_HIP(hipEventRecord(ev.a, c.stream));
        for (int r = 0; r < reps; r++) fn();
        NX_HIP(hipEventRecord(ev.b, c.stream));
        NX_HIP(hipEventSynchronize(ev.b));
        float ms = 0;
        NX_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        return (double)ms / reps;
    };
    if (c.logn >= 14) {
        // up to 512 limbs (256 MB at n = 2^16): the largest bucket is timed on a
        // working set beyond the last-level cache, like the batched ModUp NTTs
        const uint32_t max_batch = 512;
        // limbs per poly of a tuning batch: a power of two <= min(primes, 16)
        uint32_t lp = 1;
        while (lp * 2 <= std::min<size_t>(P, 16)) lp *= 2;
        std::vector<uint64_t> h;
        fill_random(c, h, max_batch, 7, lp);
        DevBuf src(&c, max_batch * n), work(&c, max_batch * n);
        NX_HIP(hipMemcpyAsync(src.get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, c.stream));
        NX_HIP(hipStreamSynchronize(c.stream));
        const int cols[3] = {4, 8, 16}, rows[2] = {4, 8};
        for (int fwd = 1; fwd >= 0; fwd--)
            for (uint32_t batch = 1; batch <= max_batch; batch <<= 1) {
                // batch = polys x limbs with limb i on prime i mod P
                const uint32_t limbs = std::min(batch, lp);
                NttBatch b;
                b.data = work.get();
                b.polys = batch / limbs;
                b.limbs = limbs;
                b.poly_stride = (uint64_t)limbs * n;
                b.prime_by_poly = 0;
                for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)(i % lp);
                auto run = [&] {
                    if (fwd)
                        ntt_forward(c.tables, b, c.stream);
                    else
                        ntt_inverse(c.tables, b, c.stream);
                };
                auto reset = [&] {
                    NX_HIP(hipMemcpyAsync(work.get(), src.get(), (size_t)batch * n * 8, hipMemcpyDeviceToDevice,
                                          c.stream));
                };
                // reference output of the default shapes
                ntt_set_tile(c.logn, 0, fwd != 0, batch, 0);
                ntt_set_tile(c.logn, 1, fwd != 0, batch, 0);
                ntt_set_radix16(c.logn, 0, fwd != 0, batch, false);
                ntt_set_radix16(c.logn, 1, fwd != 0, batch, false);
                reset();
                run();
                const std::vector<uint64_t> ref = download(c, work.get(), (size_t)batch * n);
                const double t_def = time_ms(run);
                double best = t_def;
                int best_c = 0, best_r = 0;
                for (int cw : cols)
                    for (int rw : rows) {
                        ntt_set_tile(c.logn, 0, fwd != 0, batch, cw);
                        ntt_set_tile(c.logn, 1, fwd != 0, batch, rw);
                        reset();
                        run();
                        if (download(c, work.get(), (size_t)batch * n) != ref) continue;  // never chosen
                        const double t = time_ms(run);
                        if (t < best * 0.98) {
                            best = t;
                            best_c = cw;
                            best_r = rw;
                        }
                    }
                ntt_set_tile(c.logn, 0, fwd != 0, batch, best_c);
                ntt_set_tile(c.logn, 1, fwd != 0, batch, best_r);
                // radix-16 steps per pass (double-precision path) with the chosen tiles
                int r16[2] = {0, 0}, matched[2] = {0, 0};
                for (int pass = 0; pass < 2; pass++) {
                    ntt_set_radix16(c.logn, pass, fwd != 0, batch, true);
                    reset();
                    run();
                    const bool same = download(c, work.get(), (size_t)batch * n) == ref;
                    matched[pass] = same ? 1 : 0;
                    const double t = same ? time_ms(run) : 1e30;
                    if (t < best * 0.98) {
                        best = t;
                        r16[pass] = 1;
                    } else {
                        ntt_set_radix16(c.logn, pass, fwd != 0, batch, false);
                    }
                }
                char buf[280];
                std::snprintf(buf, sizeof buf,
                              "%s{\"dir\": \"%s\", \"batch\": %u, \"col\": %d, \"row\": %d, \"radix16\": [%d, %d], "
                              "\"radix16_matched\": [%d, %d], \"us\": %.2f, \"default_us\": %.2f}",
                              first_entry ? "" : ", ", fwd ? "fwd" : "inv", batch, best_c, best_r, r16[0], r16[1],
                              matched[0], matched[1], best * 1e3, t_def * 1e3);
                rep += buf;
                first_entry = false;
            }
    }
    rep += "], \"dot_plain\": [";
    first_entry = true;
    {
        const uint32_t terms = 16;
        for (uint32_t limbs : {2u, 4u, 8u, 16u, 20u}) {
            if (limbs > P) break;
            std::vector<uint64_t> h;
            std::vector<DevBuf> cts(terms), pts(terms);
            std::vector<const uint64_t *> cp(terms), pp(terms);
            for (uint32_t r = 0; r < terms; r++) {
                cts[r].reset(&c, 2 * limbs * n);
                pts[r].reset(&c, limbs * n);
                fill_random(c, h, limbs, 100 + r, limbs);
                // both polynomials of a ciphertext use the same limb primes
                NX_HIP(hipMemcpyAsync(cts[r].get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, c.stream));
                NX_HIP(hipMemcpyAsync(cts[r].get() + limbs * n, h.data(), h.size() * 8, hipMemcpyHostToDevice,
                                      c.stream));
                fill_random(c, h, limbs, 200 + r, limbs);
                NX_HIP(hipMemcpyAsync(pts[r].get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, c.stream));
                NX_HIP(hipStreamSynchronize(c.stream));
                cp[r] = cts[r].get();
                pp[r] = pts[r].get();
            }
            DevBuf out(&c, 2 * limbs * n);
            auto run = [&] { poly_dot_plain(c.tables, out.get(), cp.data(), pp.data(), terms, limbs, c.stream); };
            dot_plain_set_layout(limbs, 0);
            run();
            const std::vector<uint64_t> ref = download(c, out.get(), 2 * limbs * n);
            const double t0 = time_ms(run);
            dot_plain_set_layout(limbs, 1);
            run();
            const bool same = download(c, out.get(), 2 * limbs * n) == ref;
            const double t1 = same ? time_ms(run) : 1e30;
            const int choice = (same && t1 < t0 * 0.98) ? 1 : 0;
            dot_plain_set_layout(limbs, choice);
            char buf[200];
            std::snprintf(buf, sizeof buf, "%s{\"limbs\": %u, \"layout\": %d, \"us\": [%.2f, %.2f]}",
                          first_entry ? "" : ", ", limbs, choice, t0 * 1e3, same ? t1 * 1e3 : -1.0);
            rep += buf;
            first_entry = false;
        }
    }
    rep += "], \"modup\": [";
    first_entry = true;
    if (c.alpha > 1 && ks_modup_col_ok(c.tables, (uint32_t)c.alpha)) {
        // hybrid ModUp at the top level: the conversion fused into the forward
        // column pass (outputs per workgroup: default rule, 4, 8, 12, all of a
        // digit; 8- or 4-column tiles (+256), 256 or 512 threads (+512)) against
        // separate conversion and column-pass launches
        const size_t l = c.L;
        for (uint32_t B : {1u, 2u, 4u, 8u, 16u, 32u}) {
            std::vector<uint64_t> h;
            fill_random(c, h, B * l, 300 + B, l);
            DevBuf tg(&c, B * l * n);
            NX_HIP(hipMemcpyAsync(tg.get(), h.data(), h.size() * 8, hipMemcpyHostToDevice, c.stream));
            NX_HIP(hipStreamSynchronize(c.stream));
            const size_t words = B * (l + c.alpha) * c.digits(l) * n;
            auto run = [&] { (void)ks_modup(c, tg.get(), B, l); };
            ks_modup_col_set(B, -1);
            std::vector<uint64_t> ref;
            {
                DevBuf T = ks_modup(c, tg.get(), B, l);
                ref = download(c, T.get(), words);
            }
            const double t_off = time_ms(run);
            double best = t_off;
            int best_m = -1;
            std::string times;
            char buf[64];
            for (int m : {0, 4, 8, 12, 64, 256, 256 + 8, 256 + 12, 512, 512 + 8, 512 + 12}) {
                ks_modup_col_set(B, m);
                DevBuf T = ks_modup(c, tg.get(), B, l);
                const bool same = download(c, T.get(), words) == ref;
                const double t = same ? time_ms(run) : 1e30;
                std::snprintf(buf, sizeof buf, ", %.2f", same ? t * 1e3 : -1.0);
                times += buf;
                if (t < best * 0.98) {
                    best = t;
                    best_m = m;
                }
            }
            ks_modup_col_set(B, best_m);
            char line[300];
            std::snprintf(line, sizeof line,
                          "%s{\"batch\": %u, \"limbs\": %zu, \"mode\": %d, \"us\": %.2f, \"unfused_us\": %.2f, "
                          "\"fused_us\": [%s]}",
                          first_entry ? "" : ", ", B, l, best_m, best * 1e3, t_off * 1e3, times.c_str() + 2);
            rep += line;
            first_entry = false;
        }
    }
    rep += "], \"moddown\": [";
    first_entry = true;
    bool epi_int = false;
    std::string epi_rep = "null";
    if (c.alpha > 1 && c.d_moddown_conv_d && ks_moddown_col_ok(c.tables, (uint32_t)c.alpha)) {
        // hybrid ModDown at the top level (B ciphertexts, 2B polys): the centred
        // conversion fused into the forward column pass against separate launches
        const size_t l = c.L, nout = l + c.alpha;
        std::vector<uint64_t> pr(nout);
        for (size_t i = 0; i < nout; i++) pr[i] = i < l ? i : c.L + (i - l);
        for (uint32_t B : {1u, 2u, 4u, 8u, 16u}) {
            const size_t aw = (size_t)B * 2 * nout * n, ow = (size_t)B * 2 * l * n;
            std::vector<uint64_t> h(aw);
            uint64_t x = 0x2545F4914F6CDD1Dull + B;
            for (size_t p = 0; p < (size_t)B * 2; p++)
                for (size_t i = 0; i < nout; i++) {
                    const uint64_t q = c.moduli[pr[i]];
                    for (size_t k = 0; k < n; k++) {
                        x ^= x << 13;
                        x ^= x >> 7;
                        x ^= x << 17;
                        h[(p * nout + i) * n + k] = x % q;
                    }
                }
            DevBuf src(&c, aw), acc(&c, aw), out(&c, ow);
            NX_HIP(hipMemcpyAsync(src.get(), h.data(), aw * 8, hipMemcpyHostToDevice, c.stream));
            auto run = [&] {
                NX_HIP(hipMemcpyAsync(acc.get(), src.get(), aw * 8, hipMemcpyDeviceToDevice, c.stream));
                NX_HIP(hipMemsetAsync(out.get(), 0, ow * 8, c.stream));
                ks_moddown(c, out.get(), acc.get(), B, l, false, nullptr);
            };
            ks_moddown_col_set(2 * B, -1);
            run();
            const std::vector<uint64_t> ref = download(c, out.get(), ow);
            const double t_off = time_ms(run);
            double best = t_off;
            int best_m = -1;
            std::string times;
            char buf[64];
            for (int m : {0, 4, 8, 12, 64, 256, 256 + 8, 256 + 12, 512, 512 + 8, 512 + 12}) {
                ks_moddown_col_set(2 * B, m);
                run();
                const bool same = download(c, out.get(), ow) == ref;
                const double t = same ? time_ms(run) : 1e30;
                std::snprintf(buf, sizeof buf, ", %.2f", same ? t * 1e3 : -1.0);
                times += buf;
                if (t < best * 0.98) {
                    best = t;
                    best_m = m;
                }
            }
            ks_moddown_col_set(2 * B, best_m);
            if (B == 16) {
                // the ModDown's combining epilogue in double-precision or integer products
                ntt_set_epilogue_int(true);
                run();
                const bool same = download(c, out.get(), ow) == ref;
                const double t_int = same ? time_ms(run) : 1e30;
                ntt_set_epilogue_int(false);
                const double t_fp = time_ms(run);
                epi_int = same && t_int < t_fp * 0.98;
                ntt_set_epilogue_int(epi_int);
                char eb[120];
                std::snprintf(eb, sizeof eb, "{\"integer\": %d, \"fp_us\": %.2f, \"int_us\": %.2f}", epi_int ? 1 : 0,
                              t_fp * 1e3, same ? t_int * 1e3 : -1.0);
                epi_rep = eb;
            }
            char line[300];
            std::snprintf(line, sizeof line,
                          "%s{\"batch\": %u, \"limbs\": %zu, \"mode\": %d, \"us\": %.2f, \"unfused_us\": %.2f, "
                          "\"fused_us\": [%s]}",
                          first_entry ? "" : ", ", B, l, best_m, best * 1e3, t_off * 1e3, times.c_str() + 2);
            rep += line;
            first_entry = false;
        }
    }
    rep += "], \"epilogue\": " + epi_rep + "}";
    NX_HIP(hipStreamSynchronize(c.stream));
    return rep;
}

}  // namespace nx
