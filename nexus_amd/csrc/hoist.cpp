// Hoisted rotations (reference Phantom hoisting_inplace,
// cuda/thirdparty/phantom-fhe/src/evaluate.cu:1727-1920): the ModUp of c1 --
// INTT, lift of every digit to every output prime, NTT, i.e. l(l+1) of the
// l^2+3l+2 NTTs of a key switch -- is done once and shared by all steps.
//
// The automorphism is applied to the lifted digits (a permutation in the NTT
// domain), so digit j of a hoisted rotation is sigma(lift(c1 mod q_j)) rather
// than lift(sigma(c1) mod q_j): both are decompositions of sigma(c1), the
// results decrypt to the same values, and the residues differ from the
// one-at-a-time rotation by key-switching noise.
#include "context.h"
#include "prof.h"

namespace nx {

namespace {

struct Prepared {
    std::vector<const uint64_t *> keys;
    std::vector<const uint64_t *> seeds;  // per key: its uniform half's digit seeds, or null
    std::vector<const uint32_t *> perms;
    std::vector<size_t> index;  // position of each non-identity step in `steps`
};

Prepared prepare(Context &c, const std::vector<int> &steps, const KSwitchKeys &gk)
{
    Prepared p;
    for (size_t k = 0; k < steps.size(); k++) {
        if (steps[k] == 0) continue;
        const uint32_t elt = c.elt_from_step(steps[k]);
        auto it = gk.keys.find(elt);
        require(it != gk.keys.end(), "Galois key not present in hoisting");
        p.keys.push_back(it->second.get());
        p.seeds.push_back(gk.seeds_of(elt));
        p.perms.push_back(c.galois_table(elt));
        p.index.push_back(k);
    }
    return p;
}

void check_input(const Ciphertext &a, const KSwitchKeys &gk)
{
    require(a.ctx != nullptr && a.size == 2, "ciphertext size must be 2");
    require(a.ntt, "CKKS encrypted must be in NTT form");
    require(gk.ctx == a.ctx, "galois_keys is not valid for encryption parameters");
}

}  // namespace

void op_rotate_hoisted(const Ciphertext &a, const std::vector<int> &steps, const KSwitchKeys &gk,
                       std::vector<Ciphertext> &dst)
{
    check_input(a, gk);
    Context &c = *a.ctx;
    const size_t n = c.n, l = a.limbs, nout = l + c.alpha, pw = l * n;
    Prepared p = prepare(c, steps, gk);
    tally("ks_hoisted", l, (double)p.keys.size());
    std::vector<Ciphertext> out(steps.size());
    for (size_t k = 0; k < steps.size(); k++)
        if (steps[k] == 0) ct_copy(out[k], a);
    if (!p.keys.empty()) {
        uint8_t outp[kMaxPrimes];
        ks_out_primes(c, l, outp);
        DevBuf T = ks_modup(c, a.data.get() + pw, 1, l);
        for (size_t base = 0; base < p.keys.size(); base += kMaxHoist) {
            const size_t cnt = std::min<size_t>(kMaxHoist, p.keys.size() - base);
            DevBuf acc(&c, cnt * 2 * nout * n);
            ks_inner_product_rot(c.tables, acc.get(), T.get(), p.keys.data() + base, p.perms.data() + base,
                                 (uint32_t)cnt, (uint32_t)c.digits(l), (uint32_t)c.key_limbs(), outp, (uint32_t)nout,
                                 c.stream, nullptr, p.seeds.data() + base);
            bool col_done = false;
            DevBuf tmp = ks_moddown_prepare(c, acc.get(), cnt, l, false, &col_done);
            // each result written in place by the transform's last pass:
            // c0 = sigma(c0) + ks0, c1 = ks1
            NttEpilogue e;
            for (size_t r = 0; r < cnt; r++) {
                Ciphertext &o = out[p.index[base + r]];
                o.ctx = &c;
                o.size = 2;
                o.limbs = l;
                o.scale = a.scale;
                o.ntt = true;
                o.data.reset(&c, 2 * pw);
                e.outs[r] = o.data.get();
                e.perm[r] = p.perms[base + r];
            }
            e.x = acc.get();
            e.x_stride = nout * n;
            e.inv = c.d_p_inv;
            e.a0 = a.data.get();
            ntt_fwd_epi(c, tmp.get(), 2 * cnt, l, l * n, e, col_done);
        }
    }
    dst = std::move(out);
}

void op_rotate_hoisted_many(const std::vector<const Ciphertext *> &in, const std::vector<int> &steps,
                            const KSwitchKeys &gk, std::vector<std::vector<Ciphertext>> &dst)
{
    // every input rotated by every step, the steps of each input hoisted: the
    // ModUps of all inputs in one batch, the (input, rotation) inner products
    // up to 32 per launch, their ModDowns batched likewise -- word for word
    // op_rotate_hoisted of each input
    require(!in.empty(), "no inputs");
    const Ciphertext &a0 = *in[0];
    check_input(a0, gk);
    Context &c = *a0.ctx;
    const size_t n = c.n, l = a0.limbs, nout = l + c.alpha, pw = l * n, digits = c.digits(l);
    for (const Ciphertext *a : in) {
        check_input(*a, gk);
        require(a->ctx == &c && a->limbs == l, "inputs must share the level");
    }
    Prepared p = prepare(c, steps, gk);
    tally("ks_hoisted", l, (double)(p.keys.size() * in.size()));
    const size_t B = in.size(), R = p.keys.size();
    std::vector<std::vector<Ciphertext>> out(B);
    for (auto &o : out) o.resize(steps.size());
    for (size_t b = 0; b < B; b++)
        for (size_t k = 0; k < steps.size(); k++)
            if (steps[k] == 0) ct_copy(out[b][k], *in[b]);
    if (R) {
        uint8_t outp[kMaxPrimes];
        ks_out_primes(c, l, outp);
        const size_t tstride = nout * digits * n;
        // ModUp of every input, up to kNttSrcMax per batch
        std::vector<DevBuf> Ts;
        std::vector<const uint64_t *> tp(B);
        for (size_t b0 = 0; b0 < B; b0 += (size_t)kNttSrcMax) {
            const size_t cnt = std::min<size_t>((size_t)kNttSrcMax, B - b0);
            std::vector<const uint64_t *> c1s(cnt);
            for (size_t b = 0; b < cnt; b++) c1s[b] = in[b0 + b]->data.get() + pw;
            Ts.push_back(ks_modup_list(c, c1s.data(), cnt, l));
            for (size_t b = 0; b < cnt; b++) tp[b0 + b] = Ts.back().get() + b * tstride;
        }
        // (input, rotation) pairs in chunks of kMaxHoist
        const size_t P = B * R;
        for (size_t base = 0; base < P; base += (size_t)kMaxHoist) {
            const size_t cnt = std::min<size_t>((size_t)kMaxHoist, P - base);
            std::vector<const uint64_t *> keys(cnt), tps(cnt), seeds(cnt);
            std::vector<const uint32_t *> perms(cnt);
            for (size_t r = 0; r < cnt; r++) {
                const size_t b = (base + r) / R, k = (base + r) % R;
                keys[r] = p.keys[k];
                seeds[r] = p.seeds[k];
                perms[r] = p.perms[k];
                tps[r] = tp[b];
            }
            DevBuf acc(&c, cnt * 2 * nout * n);
            ks_inner_product_rot(c.tables, acc.get(), nullptr, keys.data(), perms.data(), (uint32_t)cnt,
                                 (uint32_t)digits, (uint32_t)c.key_limbs(), outp, (uint32_t)nout, c.stream,
                                 tps.data(), seeds.data());
            bool col_done = false;
            DevBuf tmp = ks_moddown_prepare(c, acc.get(), cnt, l, false, &col_done);
            NttEpilogue e;
            for (size_t r = 0; r < cnt; r++) {
                const size_t b = (base + r) / R, k = (base + r) % R;
                Ciphertext &o = out[b][p.index[k]];
                o.ctx = &c;
                o.size = 2;
                o.limbs = l;
                o.scale = in[b]->scale;
                o.ntt = true;
                o.data.reset(&c, 2 * pw);
                e.outs[r] = o.data.get();
                e.perm[r] = perms[r];
                e.a0s[r] = in[b]->data.get();
            }
            e.x = acc.get();
            e.x_stride = nout * n;
            e.inv = c.d_p_inv;
            ntt_fwd_epi(c, tmp.get(), 2 * cnt, l, l * n, e, col_done);
        }
    }
    dst = std::move(out);
}

void op_rotate_sum_many(const std::vector<std::vector<const Ciphertext *>> &groups,
                        const std::vector<std::vector<int>> &steps, const KSwitchKeys &gk,
                        const std::vector<Ciphertext *> &dst)
{
    // several independent sums of rotations (one output per group), each as
    // op_rotate_sum, with the ModDowns of all groups as one batch
    require(!groups.empty() && groups.size() == steps.size() && dst.size() == groups.size(),
            "groups, steps and outputs must match");
    const Ciphertext &a0 = *groups[0].at(0);
    check_input(a0, gk);
    Context &c = *a0.ctx;
    const size_t n = c.n, l = a0.limbs, nout = l + c.alpha, pw = l * n, digits = c.digits(l);
    const size_t G = groups.size();
    DevBuf outs(&c, G * 2 * pw), acc(&c, G * 2 * nout * n);
    uint8_t outp[kMaxPrimes];
    ks_out_primes(c, l, outp);
    // the switched inputs of every group, in group order: their ModUps run in
    // chunks of up to kMaxHoist inputs across group boundaries, and each group's
    // inner products read its inputs' digits from the chunks that hold them
    struct Sw {
        size_t g;
        const uint64_t *key, *seed, *c1;
        const uint32_t *perm;
    };
    std::vector<Sw> sw;
    for (size_t g = 0; g < G; g++) {
        const auto &in = groups[g];
        require(!in.empty() && in.size() == steps[g].size(), "inputs and steps must match");
        std::vector<const uint64_t *> ins;
        std::vector<const uint32_t *> perms;
        for (size_t i = 0; i < in.size(); i++) {
            check_input(*in[i], gk);
            require(in[i]->ctx == &c && in[i]->limbs == l, "inputs must share the level");
            require(are_close(in[i]->scale, in[0]->scale), "inputs of a sum must share the scale");
            ins.push_back(in[i]->data.get());
            if (steps[g][i] == 0) {
                perms.push_back(nullptr);
                continue;
            }
            Prepared p = prepare(c, std::vector<int>{steps[g][i]}, gk);
            tally("ks_hoisted", l);
            perms.push_back(p.perms[0]);
            sw.push_back({g, p.keys[0], p.seeds[0], in[i]->data.get() + pw, p.perms[0]});
        }
        uint64_t *og = outs.get() + g * 2 * pw;
        DevBuf part;
        for (size_t base = 0; base < ins.size(); base += kMaxHoist) {
            const size_t cnt = std::min<size_t>(kMaxHoist, ins.size() - base);
            uint64_t *o = base ? (part.reset(&c, 2 * pw), part.get()) : og;
            rot_base(c.tables, o, ins.data() + base, perms.data() + base, (uint32_t)cnt, (uint32_t)l, c.stream);
            if (base) poly_add(c.tables, og, og, o, 2, (uint32_t)l, c.stream);
        }
    }
    std::vector<char> started(G, 0);
    for (size_t base = 0; base < sw.size(); base += kMaxHoist) {
        const size_t cnt = std::min<size_t>(kMaxHoist, sw.size() - base);
        std::vector<const uint64_t *> c1s(cnt);
        for (size_t r = 0; r < cnt; r++) c1s[r] = sw[base + r].c1;
        DevBuf T = ks_modup_list(c, c1s.data(), cnt, l);
        for (size_t r0 = 0; r0 < cnt;) {
            const size_t g = sw[base + r0].g;
            size_t r1 = r0;
            while (r1 < cnt && sw[base + r1].g == g) r1++;
            std::vector<const uint64_t *> keys, seeds;
            std::vector<const uint32_t *> rperms;
            for (size_t r = r0; r < r1; r++) {
                keys.push_back(sw[base + r].key);
                seeds.push_back(sw[base + r].seed);
                rperms.push_back(sw[base + r].perm);
            }
            ks_inner_product_rot_sum(c.tables, acc.get() + g * 2 * nout * n, T.get() + r0 * nout * digits * n,
                                     keys.data(), rperms.data(), (uint32_t)(r1 - r0), (uint32_t)digits,
                                     (uint32_t)c.key_limbs(), outp, (uint32_t)nout, started[g] != 0, c.stream,
                                     nout * digits * n, seeds.data());
            started[g] = 1;
            r0 = r1;
        }
    }
    for (size_t g = 0; g < G; g++)
        if (!started[g]) {
            // no switched term: a zero accumulator leaves the base unchanged
            std::vector<uint64_t> zeros(nout, 0);
            poly_fill_limbs(acc.get() + g * 2 * nout * n, zeros.data(), 2, (uint32_t)nout, c.logn, c.stream);
        }
    ks_moddown(c, outs.get(), acc.get(), G, l, false);
    for (size_t g = 0; g < G; g++) {
        Ciphertext r;
        r.ctx = &c;
        r.size = 2;
        r.limbs = l;
        r.scale = groups[g][0]->scale;
        r.ntt = true;
        r.data.reset(&c, 2 * pw);
        copy_limbs(c, r.data.get(), 2 * pw, outs.get() + g * 2 * pw, 2 * pw, 1, 2 * l);
        *dst[g] = std::move(r);
    }
}

void op_hoisting(const Ciphertext &a, const std::vector<int> &steps, const KSwitchKeys &gk, Ciphertext &dst)
{
    // hoisting_inplace: sum_k rot(a, steps[k]) with one ModUp and one ModDown
    check_input(a, gk);
    require(!steps.empty(), "steps cannot be empty");
    Context &c = *a.ctx;
    const size_t n = c.n, l = a.limbs, nout = l + c.alpha, pw = l * n;
    Prepared p = prepare(c, steps, gk);
    tally("ks_hoisted", l, (double)p.keys.size());
    Ciphertext out;
    out.ctx = &c;
    out.size = 2;
    out.limbs = l;
    out.scale = a.scale;
    out.ntt = true;
    out.data.reset(&c, 2 * pw);
    // c0 = sum_k sigma_k(c0) and c1 = 0, plus (c0, c1) once per identity step;
    // the ModDown then accumulates the switched sum into both
    size_t identity = steps.size() - p.keys.size();
    {
        std::vector<uint64_t> zeros(l, 0);
        poly_fill_limbs(out.data.get(), zeros.data(), 2, (uint32_t)l, c.logn, c.stream);
    }
    DevBuf t(&c, pw);
    for (size_t r = 0; r < p.perms.size(); r++) {
        poly_permute(t.get(), a.data.get(), p.perms[r], 1, (uint32_t)l, c.logn, c.stream);
        poly_add(c.tables, out.data.get(), out.data.get(), t.get(), 1, (uint32_t)l, c.stream);
    }
    for (size_t r = 0; r < identity; r++) {
        poly_add(c.tables, out.data.get(), out.data.get(), a.data.get(), 2, (uint32_t)l, c.stream);
    }
    if (!p.keys.empty()) {
        uint8_t outp[kMaxPrimes];
        ks_out_primes(c, l, outp);
        DevBuf T = ks_modup(c, a.data.get() + pw, 1, l);
        DevBuf acc(&c, 2 * nout * n);
        for (size_t base = 0; base < p.keys.size(); base += kMaxHoist) {
            const size_t cnt = std::min<size_t>(kMaxHoist, p.keys.size() - base);
            ks_inner_product_rot_sum(c.tables, acc.get(), T.get(), p.keys.data() + base, p.perms.data() + base,
                                     (uint32_t)cnt, (uint32_t)c.digits(l), (uint32_t)c.key_limbs(), outp,
                                     (uint32_t)nout, base > 0, c.stream, 0, p.seeds.data() + base);
        }
        ks_moddown(c, out.data.get(), acc.get(), 1, l, false);
    }
    dst = std::move(out);
}

void op_rotate_sum(const std::vector<const Ciphertext *> &in, const std::vector<int> &steps, const KSwitchKeys &gk,
                   Ciphertext &dst)
{
    // sum_i rot(in_i, steps_i) with one ModDown: the ModUps of all rotated
    // inputs run as one batch, one inner product reads each input's ModUp
    // with its rotation key into one accumulator in the extended basis (the
    // ModDown is linear), one kernel sums the sigma_i(c0_i) in the data basis,
    // and the accumulator is brought down once.  The result decrypts like the
    // sum of the separate rotations with one rounding instead of one per term.
    require(!in.empty() && in.size() == steps.size(), "inputs and steps must match");
    const Ciphertext &a0 = *in[0];
    check_input(a0, gk);
    Context &c = *a0.ctx;
    const size_t n = c.n, l = a0.limbs, nout = l + c.alpha, pw = l * n, digits = c.digits(l);
    for (const Ciphertext *a : in) {
        check_input(*a, gk);
        require(a->ctx == &c && a->limbs == l, "inputs must share the level");
        require(are_close(a->scale, a0.scale), "inputs must share the scale");
    }
    Ciphertext out;
    out.ctx = &c;
    out.size = 2;
    out.limbs = l;
    out.scale = a0.scale;
    out.ntt = true;
    out.data.reset(&c, 2 * pw);
    // data-basis part: sigma_i(c0_i), and (c0, c1) of the identity terms
    std::vector<const uint64_t *> ins, c1s, keys, seeds;
    std::vector<const uint32_t *> perms, rperms;
    for (size_t i = 0; i < in.size(); i++) {
        ins.push_back(in[i]->data.get());
        if (steps[i] == 0) {
            perms.push_back(nullptr);
            continue;
        }
        Prepared p = prepare(c, std::vector<int>{steps[i]}, gk);
        tally("ks_hoisted", l);
        perms.push_back(p.perms[0]);
        rperms.push_back(p.perms[0]);
        keys.push_back(p.keys[0]);
        seeds.push_back(p.seeds[0]);
        c1s.push_back(in[i]->data.get() + pw);
    }
    {
        DevBuf part;
        for (size_t base = 0; base < ins.size(); base += kMaxHoist) {
            const size_t cnt = std::min<size_t>(kMaxHoist, ins.size() - base);
            uint64_t *o = base ? (part.reset(&c, 2 * pw), part.get()) : out.data.get();
            rot_base(c.tables, o, ins.data() + base, perms.data() + base, (uint32_t)cnt, (uint32_t)l, c.stream);
            if (base) poly_add(c.tables, out.data.get(), out.data.get(), o, 2, (uint32_t)l, c.stream);
        }
    }
    if (!keys.empty()) {
        uint8_t outp[kMaxPrimes];
        ks_out_primes(c, l, outp);
        DevBuf acc(&c, 2 * nout * n);
        for (size_t base = 0; base < keys.size(); base += kMaxHoist) {
            const size_t cnt = std::min<size_t>(kMaxHoist, keys.size() - base);
            DevBuf T = ks_modup_list(c, c1s.data() + base, cnt, l);
            ks_inner_product_rot_sum(c.tables, acc.get(), T.get(), keys.data() + base, rperms.data() + base,
                                     (uint32_t)cnt, (uint32_t)digits, (uint32_t)c.key_limbs(), outp, (uint32_t)nout,
                                     base > 0, c.stream, nout * digits * n, seeds.data() + base);
        }
        ks_moddown(c, out.data.get(), acc.get(), 1, l, false);
    }
    dst = std::move(out);
}

}  // namespace nx
