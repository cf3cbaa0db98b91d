// Encrypted BERT-base layer program (include/nexus/bert.hpp).
#include "../../../include/nexus/bert.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace nexus {

namespace {

int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// smallest power of two >= sqrt(n)
int sqrt_pow2(size_t n)
{
    int b = 1;
    while ((size_t)b * (size_t)b < n) b <<= 1;
    return b;
}

constexpr int kOffBaby = 8;  // baby steps of the within-head offset rotations

}  // namespace

BertLayer::BertLayer(CKKSEvaluator &ck, Bootstrapper &bt, const BertShape &shape, const std::vector<const double *> &w,
                     const Options &opt)
    : ck_(ck), bt_(bt), shape_(shape), opt_(opt)
{
    slots_ = ck.context->poly_degree() / 2;
    if (shape.hidden <= 0 || shape.ffn <= 0 || shape.hidden % kGroup || shape.ffn % kGroup)
        throw std::invalid_argument("BertLayer: hidden and ffn must be positive multiples of 256");
    if (shape.heads * kHeadDim != shape.hidden) throw std::invalid_argument("BertLayer: head dimension must be 64");
    if (slots_ < (size_t)kGroup) throw std::invalid_argument("BertLayer: fewer slots than one token group");
    if (opt.baby < 2 || opt.baby > kGroup || (opt.baby & (opt.baby - 1)))
        throw std::invalid_argument("BertLayer: baby steps must be a power of two in [2, 256]");
    if (w.size() != (size_t)kMatrices) throw std::invalid_argument("BertLayer: six weight matrices expected");
    T_ = slots_ / kGroup;
    const int D = shape.hidden, F = shape.ffn;
    const int dims[kMatrices][2] = {{D, D}, {D, D}, {D, D}, {D, D}, {D, F}, {F, D}};
    for (int m = 0; m < kMatrices; m++) {
        if (!w[m]) throw std::invalid_argument("BertLayer: null weight matrix");
        Diagonals &d = diag_[m];
        d.n_in = dims[m][0] / kGroup;
        d.n_out = dims[m][1] / kGroup;
        d.a_min = floor_div(-(kGroup - 1), opt.baby);
        d.a_max = floor_div(kGroup - 1, opt.baby);
        d.w.assign(w[m], w[m] + (size_t)dims[m][0] * dims[m][1]);
    }
}

std::vector<int> BertLayer::rotation_steps(size_t slots, const BertShape &shape, const Options &opt)
{
    (void)shape;
    std::vector<int> s;
    const int T = (int)(slots / kGroup);
    // linear maps: babies 1..B-1, giants B a
    for (int b = 1; b < opt.baby; b++) s.push_back(b);
    for (int a = floor_div(-(kGroup - 1), opt.baby); a <= floor_div(kGroup - 1, opt.baby); a++)
        if (a) s.push_back(opt.baby * a);
    // token rotations: 256 b and 256 tb a
    const int tb = sqrt_pow2((size_t)T);
    for (int b = 1; b < tb && b < T; b++) s.push_back(kGroup * b);
    for (int a = 1; a * tb < T; a++) s.push_back(kGroup * tb * a);
    // within-head offsets d in [-63, 63]: babies 1..7, giants 8 a
    for (int b = 1; b < kOffBaby; b++) s.push_back(b);
    for (int a = floor_div(-(kHeadDim - 1), kOffBaby); a <= floor_div(kHeadDim - 1, kOffBaby); a++)
        if (a) s.push_back(kOffBaby * a);
    // sums and broadcasts within 64-slot blocks and 256-slot groups
    for (int st = 1; st < kGroup; st <<= 1) {
        s.push_back(st);
        s.push_back(-st);
    }
    const int half = (int)slots;
    for (int &x : s) {
        x %= half;
        if (x < 0) x += half;
        if (x > half / 2) x -= half;  // keep the nearer representative
    }
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    s.erase(std::remove(s.begin(), s.end(), 0), s.end());
    return s;
}

void BertLayer::encode_diagonals(Matrix m, double factor, size_t chain_index, double scale)
{
    Diagonals &d = diag_[m];
    const int B = opt_.baby, n_a = d.a_max - d.a_min + 1;
    const size_t per = (size_t)d.n_in * d.n_out * n_a * B;
    if (d.pt.size() == per && d.factor == factor && d.chain == chain_index && d.scale == scale) return;
    d.pt.assign(per, Plaintext());
    d.present.assign(per, 0);
    d.factor = factor;
    d.chain = chain_index;
    d.scale = scale;
    const int out_dim = d.n_out * kGroup;
    std::vector<double> pattern(kGroup), v(slots_);
    for (int o = 0; o < d.n_out; o++)
        for (int c = 0; c < d.n_in; c++)
            for (int a = d.a_min; a <= d.a_max; a++)
                for (int b = 0; b < B; b++) {
                    const int e = B * a + b;
                    if (e <= -kGroup || e >= kGroup) continue;
                    // D_e[256 i + j] = W[256 c + j + e][256 o + j] for 0 <= j + e < 256
                    bool any = false;
                    for (int j = 0; j < kGroup; j++) {
                        const int f = j + e;
                        double x = 0;
                        if (f >= 0 && f < kGroup)
                            x = d.w[(size_t)(kGroup * c + f) * out_dim + (size_t)(kGroup * o + j)] * factor;
                        pattern[j] = x;
                        any |= x != 0;
                    }
                    if (!any) continue;
                    // pre-rotated by the giant step: D'[y] = D_e[y - B a]
                    const int shift = ((B * a) % kGroup + kGroup) % kGroup;
                    for (size_t y = 0; y < slots_; y++) v[y] = pattern[((int)(y % kGroup) - shift + kGroup) % kGroup];
                    const size_t k = (((size_t)o * d.n_in + c) * n_a + (size_t)(a - d.a_min)) * B + b;
                    d.pt[k] = Plaintext(*ck_.context);
                    ck_.encoder.encode(v, chain_index, scale, d.pt[k]);
                    d.present[k] = 1;
                }
}

std::vector<Ciphertext> BertLayer::linear(Matrix m, const std::vector<Ciphertext> &x, double factor)
{
    if (m < 0 || m >= kMatrices) throw std::invalid_argument("BertLayer: unknown matrix");
    Diagonals &d = diag_[m];
    if (x.size() != (size_t)d.n_in) throw std::invalid_argument("BertLayer: wrong number of input ciphertexts");
    const Evaluator &ev = ck_.evaluator;
    const double delta = ck_.scale;
    const size_t two = chain_of(2);
    // inputs at two limbs and exactly the evaluator's scale
    std::vector<Ciphertext> in(x.size());
    for (size_t c = 0; c < x.size(); c++) {
        if (x[c].coeff_modulus_size() < 2) throw std::invalid_argument("BertLayer: linear input needs two limbs");
        if (std::fabs(x[c].scale() / delta - 1.0) > 1e-3)
            throw std::invalid_argument("BertLayer: linear input is not at the evaluator's scale");
        in[c] = x[c];
        if (in[c].coeff_modulus_size() > 2) ev.mod_switch_to_inplace(in[c], two);
        in[c].set_scale(delta);
    }
    // the plaintexts' scale is the prime the rescale drops: outputs land at delta
    const double sigma = (double)prime(2);
    encode_diagonals(m, factor, two, sigma);
    const int B = opt_.baby, n_a = d.a_max - d.a_min + 1;
    std::vector<int> bsteps;
    for (int b = 1; b < B; b++) bsteps.push_back(b);
    std::vector<std::vector<Ciphertext>> baby(x.size());
    for (size_t c = 0; c < x.size(); c++) {
        std::vector<Ciphertext> r;
        ev.rotate_vector_hoisted(in[c], bsteps, *ck_.galois_keys, r);
        baby[c].reserve(B);
        baby[c].push_back(in[c]);
        for (auto &t : r) baby[c].push_back(std::move(t));
    }
    std::vector<Ciphertext> out(d.n_out);
    for (int o = 0; o < d.n_out; o++) {
        bool have = false;
        for (int a = d.a_min; a <= d.a_max; a++) {
            std::vector<const Ciphertext *> terms;
            std::vector<const Plaintext *> weights;
            for (int c = 0; c < d.n_in; c++)
                for (int b = 0; b < B; b++) {
                    const size_t k = (((size_t)o * d.n_in + c) * n_a + (size_t)(a - d.a_min)) * B + b;
                    if (!d.present[k]) continue;
                    terms.push_back(&baby[c][b]);
                    weights.push_back(&d.pt[k]);
                }
            if (terms.empty()) continue;
            Ciphertext acc;
            ev.multiply_plain_accumulate(terms, weights, acc);
            if (a) ev.rotate_vector_inplace(acc, B * a, *ck_.galois_keys);
            if (!have) {
                out[o] = std::move(acc);
                have = true;
            } else {
                ev.add_inplace(out[o], acc);
            }
        }
        if (!have) throw std::logic_error("BertLayer: all-zero weight block");
        ev.rescale_to_next_inplace(out[o]);
    }
    return out;
}

std::vector<Ciphertext> BertLayer::token_rotations(const Ciphertext &x, int t0, int count)
{
    // rot(x, 256 t) for t in [t0, t0 + count), t = tb a + b: hoisted babies
    // 256 b, then per baby hoisted giants 256 tb a
    const Evaluator &ev = ck_.evaluator;
    const int tb = sqrt_pow2(T_);
    std::vector<Ciphertext> out(count);
    std::vector<int> need_b;
    for (int t = t0; t < t0 + count; t++)
        if (std::find(need_b.begin(), need_b.end(), t % tb) == need_b.end()) need_b.push_back(t % tb);
    std::sort(need_b.begin(), need_b.end());
    std::vector<int> bsteps;
    for (int b : need_b)
        if (b) bsteps.push_back(kGroup * b);
    std::vector<Ciphertext> br;
    if (!bsteps.empty()) ev.rotate_vector_hoisted(x, bsteps, *ck_.galois_keys, br);
    size_t bi = 0;
    for (int b : need_b) {
        const Ciphertext &base = b ? br[bi++] : x;
        std::vector<int> as, gsteps;
        for (int t = t0; t < t0 + count; t++)
            if (t % tb == b) as.push_back(t / tb);
        for (int a : as)
            if (a) gsteps.push_back(kGroup * tb * a);
        std::vector<Ciphertext> gr;
        if (!gsteps.empty()) ev.rotate_vector_hoisted(base, gsteps, *ck_.galois_keys, gr);
        size_t gi = 0;
        for (int a : as) out[tb * a + b - t0] = a ? std::move(gr[gi++]) : base;
    }
    return out;
}

std::vector<Ciphertext> BertLayer::offset_rotations(const Ciphertext &x, int dmin, int dmax)
{
    // rot(x, d) for d in [dmin, dmax], d = 8 a + b
    const Evaluator &ev = ck_.evaluator;
    std::vector<Ciphertext> out(dmax - dmin + 1);
    std::vector<int> bsteps;
    for (int b = 1; b < kOffBaby; b++) bsteps.push_back(b);
    std::vector<Ciphertext> br;
    ev.rotate_vector_hoisted(x, bsteps, *ck_.galois_keys, br);
    for (int b = 0; b < kOffBaby; b++) {
        const Ciphertext &base = b ? br[b - 1] : x;
        std::vector<int> as, gsteps;
        for (int d = dmin; d <= dmax; d++)
            if (((d % kOffBaby) + kOffBaby) % kOffBaby == b) as.push_back(floor_div(d, kOffBaby));
        for (int a : as)
            if (a) gsteps.push_back(kOffBaby * a);
        std::vector<Ciphertext> gr;
        if (!gsteps.empty()) ev.rotate_vector_hoisted(base, gsteps, *ck_.galois_keys, gr);
        size_t gi = 0;
        for (int a : as) out[kOffBaby * a + b - dmin] = a ? std::move(gr[gi++]) : base;
    }
    return out;
}

const Plaintext &BertLayer::offset_mask(int p, size_t limbs, double scale)
{
    auto key = std::make_tuple(p, limbs, scale);
    auto it = masks_.find(key);
    if (it != masks_.end()) return it->second;
    std::vector<double> v(slots_, 0.0);
    for (size_t y = (size_t)p; y < slots_; y += kHeadDim) v[y] = 1.0;
    Plaintext pt(*ck_.context);
    ck_.encoder.encode(v, chain_of(limbs), scale, pt);
    return masks_.emplace(key, std::move(pt)).first->second;
}

Ciphertext BertLayer::scaled(const Ciphertext &x, double value)
{
    const Evaluator &ev = ck_.evaluator;
    Ciphertext y = x;
    if (y.coeff_modulus_size() > 2) ev.mod_switch_to_inplace(y, chain_of(2));
    y.set_scale(ck_.scale);
    Plaintext c(*ck_.context);
    ck_.encoder.encode(value, y.params_id(), (double)prime(y.coeff_modulus_size()), c);
    ev.multiply_plain_inplace(y, c);
    ev.rescale_to_next_inplace(y);
    return y;
}

std::vector<Ciphertext> BertLayer::attention(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                                             const std::vector<Ciphertext> &v, int s)
{
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (q.size() != n || k.size() != n || v.size() != n)
        throw std::invalid_argument("BertLayer: attention needs hidden / 256 ciphertexts of q, k and v");
    if (s < 1 || (size_t)s > T_) throw std::invalid_argument("BertLayer: sequence length out of range");
    const Evaluator &ev = ck_.evaluator;
    const RelinKeys &rk = *ck_.relin_keys;
    const GaloisKeys &gk = *ck_.galois_keys;
    const int T = (int)T_;
    const int halves = (T + kHeadDim - 1) / kHeadDim;
    std::vector<Ciphertext> out(n);
    for (size_t c = 0; c < n; c++) {
        Ciphertext qc = q[c], kc = k[c];
        if (kc.coeff_modulus_size() > qc.coeff_modulus_size()) ev.mod_switch_to_inplace(kc, qc.params_id());
        if (qc.coeff_modulus_size() > kc.coeff_modulus_size()) ev.mod_switch_to_inplace(qc, kc.params_id());
        // ---- scores, compacted: S^eta[256 i + 64 h + t'] = q_i . k_(i + 64 eta + t') on head h
        std::vector<Ciphertext> S(halves);
        std::vector<int> tcount(halves);
        for (int eta = 0; eta < halves; eta++) {
            const int t0 = kHeadDim * eta, tc = std::min(kHeadDim, T - t0);
            tcount[eta] = tc;
            std::vector<Ciphertext> kr = token_rotations(kc, t0, tc);
            std::vector<Ciphertext> P(tc);
            for (int tau = 0; tau < tc; tau++) {
                ev.multiply(qc, kr[tau], P[tau]);
                ev.relinearize_inplace(P[tau], rk);
                ev.rescale_to_next_inplace(P[tau]);
            }
            kr.clear();
            const size_t lp = P[0].coeff_modulus_size();
            const double ms = (double)prime(lp);
            // S = sum_d rot(Z_d, d), Z_d = sum_tau A_(tau + d) (.) P_tau, d = 8 a + b
            const int dmin = -(tc - 1), dmax = kHeadDim - 1;
            bool have = false;
            for (int a = floor_div(dmin, kOffBaby); a <= floor_div(dmax, kOffBaby); a++) {
                Ciphertext Y;
                bool haveY = false;
                for (int b = 0; b < kOffBaby; b++) {
                    const int d = kOffBaby * a + b;
                    if (d < dmin || d > dmax) continue;
                    std::vector<const Ciphertext *> terms;
                    std::vector<const Plaintext *> weights;
                    for (int tau = std::max(0, -d); tau < tc && tau + d < kHeadDim; tau++) {
                        terms.push_back(&P[tau]);
                        weights.push_back(&offset_mask(tau + d, lp, ms));
                    }
                    if (terms.empty()) continue;
                    Ciphertext Z;
                    ev.multiply_plain_accumulate(terms, weights, Z);
                    if (b) ev.rotate_vector_inplace(Z, b, gk);
                    if (!haveY) {
                        Y = std::move(Z);
                        haveY = true;
                    } else {
                        ev.add_inplace(Y, Z);
                    }
                }
                if (!haveY) continue;
                if (a) ev.rotate_vector_inplace(Y, kOffBaby * a, gk);
                if (!have) {
                    S[eta] = std::move(Y);
                    have = true;
                } else {
                    ev.add_inplace(S[eta], Y);
                }
            }
            ev.rescale_to_next_inplace(S[eta]);
        }
        // ---- softmax over t for every (query, head): exp, key mask, sum, 1 / sum
        std::vector<Ciphertext> E(halves);
        for (int eta = 0; eta < halves; eta++) {
            E[eta] = ck_.exp(S[eta]);
            const size_t le = E[eta].coeff_modulus_size();
            std::vector<double> km(slots_, 0.0);
            for (int i = 0; i < s; i++)
                for (int h = 0; h < kGroup / kHeadDim; h++)
                    for (int tp = 0; tp < tcount[eta]; tp++)
                        if ((i + kHeadDim * eta + tp) % T < s) km[(size_t)kGroup * i + kHeadDim * h + tp] = 1.0;
            Plaintext kmp(*ck_.context);
            ck_.encoder.encode(km, E[eta].params_id(), (double)prime(le), kmp);
            ev.multiply_plain_inplace(E[eta], kmp);
            ev.rescale_to_next_inplace(E[eta]);
        }
        Ciphertext tot = E[0];
        for (int eta = 1; eta < halves; eta++) ev.add_inplace(tot, E[eta]);
        for (int st = 1; st < kHeadDim; st <<= 1) {
            Ciphertext r;
            ev.rotate_vector(tot, st, gk, r);
            ev.add_inplace(tot, r);
        }
        {
            const size_t lt = tot.coeff_modulus_size();
            std::vector<double> m0(slots_, 0.0);
            for (size_t y = 0; y < slots_; y += kHeadDim) m0[y] = 1.0 / s;
            Plaintext p0(*ck_.context);
            ck_.encoder.encode(m0, tot.params_id(), (double)prime(lt), p0);
            ev.multiply_plain_inplace(tot, p0);
            ev.rescale_to_next_inplace(tot);
        }
        for (int st = 1; st < kHeadDim; st <<= 1) {
            Ciphertext r;
            ev.rotate_vector(tot, -st, gk, r);
            ev.add_inplace(tot, r);
        }
        Ciphertext inv = ck_.inverse(tot);
        {
            Plaintext cs(*ck_.context);
            ck_.encoder.encode(1.0 / s, inv.params_id(), (double)prime(inv.coeff_modulus_size()), cs);
            ev.multiply_plain_inplace(inv, cs);
            ev.rescale_to_next_inplace(inv);
        }
        // ---- probabilities, refreshed; expansion over the 64 features; times v
        Ciphertext vc = v[c];
        Ciphertext acc;
        bool have_acc = false;
        for (int eta = 0; eta < halves; eta++) {
            const int t0 = kHeadDim * eta, tc = tcount[eta];
            Ciphertext Pe = E[eta];
            ev.mod_switch_to_inplace(Pe, inv.params_id());
            ev.multiply_inplace(Pe, inv);
            ev.relinearize_inplace(Pe, rk);
            ev.rescale_to_next_inplace(Pe);
            Ciphertext Pb;
            bt_.bootstrap(Pe, Pb);
            // Pb_tau[256 i + 64 h + f] = P[256 i + 64 h + tau] = sum_d A_(tau - d) (.) rot(P, d)
            std::vector<Ciphertext> R = offset_rotations(Pb, -(kHeadDim - 1), tc - 1);
            const size_t lb = Pb.coeff_modulus_size();
            const double ms = (double)prime(lb);
            Ciphertext vlow = vc;
            if (vlow.coeff_modulus_size() > lb - 1) ev.mod_switch_to_inplace(vlow, chain_of(lb - 1));
            std::vector<Ciphertext> vr = token_rotations(vlow, t0, tc);
            for (int tau = 0; tau < tc; tau++) {
                std::vector<const Ciphertext *> terms;
                std::vector<const Plaintext *> weights;
                for (int d = tau - (kHeadDim - 1); d <= tau; d++) {
                    terms.push_back(&R[d + (kHeadDim - 1)]);
                    weights.push_back(&offset_mask(tau - d, lb, ms));
                }
                Ciphertext pb;
                ev.multiply_plain_accumulate(terms, weights, pb);
                ev.rescale_to_next_inplace(pb);
                Ciphertext prod;
                ev.multiply(pb, vr[tau], prod);
                if (!have_acc) {
                    acc = std::move(prod);
                    have_acc = true;
                } else {
                    ev.add_inplace(acc, prod);
                }
            }
        }
        ev.relinearize_inplace(acc, rk);
        ev.rescale_to_next_inplace(acc);
        out[c] = std::move(acc);
    }
    return out;
}

std::vector<Ciphertext> BertLayer::layer_norm(const std::vector<Ciphertext> &r)
{
    // layer_norm.cpp:6-33 over the 768 features of a token spread across the
    // ciphertexts: sum of squares inside the 256-slot group, broadcast back,
    // invert_sqrt (4 Newton, 2 Goldschmidt steps), times the input
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (r.size() != n) throw std::invalid_argument("BertLayer: layer_norm needs hidden / 256 ciphertexts");
    const Evaluator &ev = ck_.evaluator;
    const GaloisKeys &gk = *ck_.galois_keys;
    Ciphertext tot;
    for (size_t c = 0; c < n; c++) {
        Ciphertext sq;
        ev.square(r[c], sq);
        ev.relinearize_inplace(sq, *ck_.relin_keys);
        ev.rescale_to_next_inplace(sq);
        if (c == 0)
            tot = std::move(sq);
        else
            ev.add_inplace(tot, sq);
    }
    for (int st = 1; st < kGroup; st <<= 1) {
        Ciphertext t;
        ev.rotate_vector(tot, st, gk, t);
        ev.add_inplace(tot, t);
    }
    {
        std::vector<double> m0(slots_, 0.0);
        for (size_t y = 0; y < slots_; y += kGroup) m0[y] = 1.0 / shape_.hidden;
        Plaintext p0(*ck_.context);
        ck_.encoder.encode(m0, tot.params_id(), tot.scale(), p0);
        ev.multiply_plain_inplace(tot, p0);
        ev.rescale_to_next_inplace(tot);
    }
    for (int st = 1; st < kGroup; st <<= 1) {
        Ciphertext t;
        ev.rotate_vector(tot, -st, gk, t);
        ev.add_inplace(tot, t);
    }
    Ciphertext y = ck_.invert_sqrt(tot, 4, 2);
    std::vector<Ciphertext> out(n);
    for (size_t c = 0; c < n; c++) {
        Ciphertext x = r[c];
        ev.mod_switch_to_inplace(x, y.params_id());
        ev.multiply(y, x, out[c]);
        ev.relinearize_inplace(out[c], *ck_.relin_keys);
        ev.rescale_to_next_inplace(out[c]);
    }
    return out;
}

std::vector<Ciphertext> BertLayer::forward(const std::vector<Ciphertext> &x, int s)
{
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (x.size() != n) throw std::invalid_argument("BertLayer: input needs hidden / 256 ciphertexts");
    if (s < 1 || (size_t)s > T_) throw std::invalid_argument("BertLayer: sequence length out of range");
    const Evaluator &ev = ck_.evaluator;
    const double pre = opt_.ln_prescale;
    auto boot_all = [&](std::vector<Ciphertext> &v, double factor) {
        for (auto &c : v) {
            Ciphertext o;
            bt_.bootstrap(c, o, factor);
            c = std::move(o);
        }
    };
    auto add_scaled = [&](std::vector<Ciphertext> &y, const std::vector<Ciphertext> &res) {
        for (size_t c = 0; c < y.size(); c++) {
            Ciphertext r8 = scaled(res[c], 1.0 / 8);
            r8.set_scale(y[c].scale());
            ev.add_inplace(y[c], r8);
        }
    };
    // attention block; the bootstrapping range factors are folded into the weights
    std::vector<Ciphertext> q = linear(Wq, x, 1.0 / 32), k = linear(Wk, x, 0.25), v = linear(Wv, x, 0.25);
    boot_all(q, 4);
    boot_all(k, 4);
    boot_all(v, 4);
    std::vector<Ciphertext> o = attention(q, k, v, s);
    q.clear();
    k.clear();
    v.clear();
    std::vector<Ciphertext> r1 = linear(Wo, o, 1.0 / 8);
    add_scaled(r1, x);
    boot_all(r1, 8 * pre);
    std::vector<Ciphertext> z = layer_norm(r1);
    for (auto &c : z) c.set_scale(ck_.scale);
    // feed-forward block
    std::vector<Ciphertext> h = linear(W1, z, 0.25);
    boot_all(h, 4);
    GELUEvaluator gelu(ck_);
    for (auto &c : h) {
        Ciphertext g;
        gelu.gelu(c, g);
        c = std::move(g);
    }
    std::vector<Ciphertext> r2 = linear(W2, h, 1.0 / 8);
    add_scaled(r2, z);
    boot_all(r2, 8 * pre);
    std::vector<Ciphertext> y = layer_norm(r2);
    for (auto &c : y) c.set_scale(ck_.scale);
    return y;
}

}  // namespace nexus
