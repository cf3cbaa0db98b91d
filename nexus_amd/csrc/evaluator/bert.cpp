// Encrypted BERT-base layer program (include/nexus/bert.hpp).
#include "../../../include/nexus/bert.hpp"

#include <algorithm>
#include <chrono>
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
    shared_ = std::make_shared<Shared>();
    const int D = shape.hidden, F = shape.ffn;
    const int dims[kMatrices][2] = {{D, D}, {D, D}, {D, D}, {D, D}, {D, F}, {F, D}};
    for (int m = 0; m < kMatrices; m++) {
        if (!w[m]) throw std::invalid_argument("BertLayer: null weight matrix");
        Weights &d = shared_->m[m];
        d.n_in = dims[m][0] / kGroup;
        d.n_out = dims[m][1] / kGroup;
        d.a_min = floor_div(-(kGroup - 1), opt.baby);
        d.a_max = floor_div(kGroup - 1, opt.baby);
        d.w.assign(w[m], w[m] + (size_t)dims[m][0] * dims[m][1]);
    }
}

BertLayer::BertLayer(const BertLayer &o, CKKSEvaluator &ck, Bootstrapper &bt)
    : ck_(ck), bt_(bt), shape_(o.shape_), opt_(o.opt_), slots_(o.slots_), T_(o.T_), shared_(o.shared_)
{
    if (ck.context->get() != o.ck_.context->get()) throw std::invalid_argument("BertLayer: view on another context");
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
    if (opt.direct_rotations) {
        // every token rotation 256 t and every within-head offset d in [-63, 63]
        for (int t = 1; t < T; t++) s.push_back(kGroup * t);
        for (int d = -(kHeadDim - 1); d < kHeadDim; d++)
            if (d) s.push_back(d);
    } else {
        // token rotations: 256 b and 256 tb a
        const int tb = sqrt_pow2((size_t)T);
        for (int b = 1; b < tb && b < T; b++) s.push_back(kGroup * b);
        for (int a = 1; a * tb < T; a++) s.push_back(kGroup * tb * a);
        // within-head offsets d in [-63, 63]: babies 1..7, giants 8 a
        for (int b = 1; b < kOffBaby; b++) s.push_back(b);
        for (int a = floor_div(-(kHeadDim - 1), kOffBaby); a <= floor_div(kHeadDim - 1, kOffBaby); a++)
            if (a) s.push_back(kOffBaby * a);
    }
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

std::shared_ptr<const BertLayer::Encoded> BertLayer::diagonals(Matrix m, double factor, size_t chain_index,
                                                                double scale)
{
    std::lock_guard<std::mutex> g(shared_->mu);
    Weights &d = shared_->m[m];
    const auto key = std::make_tuple(factor, chain_index, scale);
    for (auto &e : d.enc)
        if (e.first == key) return e.second;
    const int B = opt_.baby, n_a = d.a_max - d.a_min + 1;
    auto enc = std::make_shared<Encoded>();
    // inputs per group: the block mix stages a 64-position tile of every input
    // (512 bytes each) and the tile's weight words of every entry (one per
    // 512-slot-period block the tile touches) in 64 KiB of LDS
    const int stride = (int)(2 * slots_ / (2 * kGroup));  // NTT positions per block
    const int bpt = stride >= 64 ? 1 : 64 / stride;
    enc->group = std::max(1, std::min(d.n_in, 65536 / (B * (512 + 8 * n_a * bpt))));
    enc->groups = (d.n_in + enc->group - 1) / enc->group;
    enc->giant_any.assign((size_t)d.n_out * n_a, 0);
    const int out_dim = d.n_out * kGroup;
    std::vector<double> pattern(kGroup);
    for (int o = 0; o < d.n_out; o++)
        for (int gi = 0; gi < enc->groups; gi++) {
            const int c0 = gi * enc->group, c1 = std::min(d.n_in, c0 + enc->group);
            const int K = (c1 - c0) * B;
            const size_t count = (size_t)(c1 - c0) * n_a * B;
            std::vector<double> vals(count * kGroup, 0.0);
            std::vector<int32_t> index((size_t)n_a * K, -1);
            for (int c = c0; c < c1; c++)
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
                        // pre-rotated by the giant step: D'[y] = D_e[y - B a], still 256-periodic
                        const int shift = ((B * a) % kGroup + kGroup) % kGroup;
                        const size_t entry = ((size_t)(c - c0) * n_a + (size_t)(a - d.a_min)) * B + b;
                        for (int y = 0; y < kGroup; y++)
                            vals[entry * kGroup + y] = pattern[(y - shift + kGroup) % kGroup];
                        index[(size_t)(a - d.a_min) * K + (size_t)(c - c0) * B + b] = (int32_t)entry;
                        enc->giant_any[(size_t)o * n_a + (a - d.a_min)] = 1;
                    }
            enc->sets.push_back(
                std::make_unique<PeriodicPlaintexts>(*ck_.context, vals, (size_t)kGroup, chain_index, scale));
            enc->index.push_back(std::move(index));
        }
    // complete on this thread's stream before any lane reads them
    ck_.context->synchronize();
    if (d.enc.size() >= 2) d.enc.erase(d.enc.begin());
    d.enc.emplace_back(key, enc);
    return enc;
}

std::vector<Ciphertext> BertLayer::linear(Matrix m, const std::vector<Ciphertext> &x, double factor)
{
    if (m < 0 || m >= kMatrices) throw std::invalid_argument("BertLayer: unknown matrix");
    const Weights &d = shared_->m[m];
    if (x.size() != (size_t)d.n_in) throw std::invalid_argument("BertLayer: wrong number of input ciphertexts");
    const Evaluator &ev = ck_.evaluator;
    const double delta = ck_.scale;
    const size_t two = chain_of(2);
    // inputs at two limbs, all at one scale s_in; the diagonals are encoded
    // at delta q_1 / s_in so that the rescale (dropping q_1) lands exactly on
    // delta -- no scale is overwritten
    std::vector<Ciphertext> in(x.size());
    const double s_in = x[0].scale();
    if (std::fabs(s_in / delta - 1.0) > 0.05)
        throw std::invalid_argument("BertLayer: linear input is far from the evaluator's scale");
    for (size_t c = 0; c < x.size(); c++) {
        if (x[c].coeff_modulus_size() < 2) throw std::invalid_argument("BertLayer: linear input needs two limbs");
        if (x[c].scale() != s_in) throw std::invalid_argument("BertLayer: linear inputs at different scales");
        in[c] = x[c];
        if (in[c].coeff_modulus_size() > 2) ev.mod_switch_to_inplace(in[c], two);
    }
    const double sigma = delta * (double)prime(2) / s_in;
    const std::shared_ptr<const Encoded> enc = diagonals(m, factor, two, sigma);
    const int B = opt_.baby, n_a = d.a_max - d.a_min + 1;
    std::vector<int> bsteps;
    for (int b = 1; b < B; b++) bsteps.push_back(b);
    // the babies of every input in one batched hoisted call (one ModUp launch)
    std::vector<std::vector<Ciphertext>> baby(x.size());
    {
        std::vector<const Ciphertext *> ip;
        for (auto &c : in) ip.push_back(&c);
        std::vector<std::vector<Ciphertext>> r;
        ev.rotate_vector_hoisted_many(ip, bsteps, *ck_.galois_keys, r);
        for (size_t c = 0; c < x.size(); c++) {
            baby[c].reserve(B);
            baby[c].push_back(std::move(in[c]));
            for (auto &t : r[c]) baby[c].push_back(std::move(t));
        }
    }
    // every output's giant-step accumulators as block mixes over the babies of
    // each input group (the periodic diagonals are 512 words a limb); then the
    // giant rotations of every output summed in one call (one ModDown launch
    // for all outputs) and all outputs rescaled together
    std::vector<std::vector<Ciphertext>> accs(d.n_out);
    std::vector<std::vector<const Ciphertext *>> ap(d.n_out);
    std::vector<std::vector<int>> gsteps(d.n_out);
    for (int o = 0; o < d.n_out; o++) {
        for (int gi = 0; gi < enc->groups; gi++) {
            const int c0 = gi * enc->group, c1 = std::min(d.n_in, c0 + enc->group);
            std::vector<const Ciphertext *> terms;
            for (int c = c0; c < c1; c++)
                for (int b = 0; b < B; b++) terms.push_back(&baby[c][b]);
            std::vector<Ciphertext> part;
            ev.multiply_periodic_mix(terms, *enc->sets[(size_t)o * enc->groups + gi],
                                     enc->index[(size_t)o * enc->groups + gi], part);
            if (accs[o].empty())
                accs[o] = std::move(part);
            else
                for (int a = 0; a < n_a; a++) ev.add_inplace(accs[o][a], part[a]);
        }
        for (int a = d.a_min; a <= d.a_max; a++)
            if (enc->giant_any[(size_t)o * n_a + (a - d.a_min)]) {
                ap[o].push_back(&accs[o][a - d.a_min]);
                gsteps[o].push_back(B * a);
            }
        if (ap[o].empty()) throw std::logic_error("BertLayer: all-zero weight block");
    }
    std::vector<Ciphertext> sums;
    ev.rotate_sum_many(ap, gsteps, *ck_.galois_keys, sums);
    accs.clear();
    std::vector<const Ciphertext *> sp;
    for (auto &x : sums) sp.push_back(&x);
    std::vector<Ciphertext> out;
    ev.rescale_to_next_many(sp, out);
    return out;
}

std::vector<Ciphertext> BertLayer::token_rotations(const Ciphertext &x, int t0, int count)
{
    std::vector<int> ts;
    for (int t = t0; t < t0 + count; t++) ts.push_back(t);
    std::vector<std::vector<Ciphertext>> r = token_rotations_many({&x}, ts);
    return std::move(r[0]);
}

std::vector<std::vector<Ciphertext>> BertLayer::token_rotations_many(const std::vector<const Ciphertext *> &xs,
                                                                     const std::vector<int> &ts)
{
    // rot(x, 256 t) for every x and every t of ts (distinct, in [0, T)), t = tb a + b.
    // A range of 64 tokens has about 4 giants a and 16 babies b, so the giants
    // 256 tb a come first, hoisted from each x together with the babies of a = 0
    // (one ModUp of each x, all x in one batched call), then every giant's
    // babies 256 b in one more batched call: far fewer ModUps than one per
    // baby, the same number of key switches.
    const Evaluator &ev = ck_.evaluator;
    // direct: every rotation hoisted from one ModUp of x (tb = T: a = 0 only)
    const int tb = opt_.direct_rotations ? (int)T_ : sqrt_pow2(T_);
    const size_t X = xs.size();
    std::vector<std::vector<Ciphertext>> out(X, std::vector<Ciphertext>(ts.size()));
    std::vector<int> pos(T_, -1);  // t -> its index in ts
    for (size_t k = 0; k < ts.size(); k++) {
        if (ts[k] < 0 || (size_t)ts[k] >= T_ || pos[ts[k]] >= 0)
            throw std::invalid_argument("BertLayer: token offsets must be distinct and in [0, T)");
        pos[ts[k]] = (int)k;
    }
    std::vector<int> need_a;
    for (int t : ts)
        if (std::find(need_a.begin(), need_a.end(), t / tb) == need_a.end()) need_a.push_back(t / tb);
    std::sort(need_a.begin(), need_a.end());
    auto babies_of = [&](int a) {
        std::vector<int> bs;
        for (int t : ts)
            if (t / tb == a) bs.push_back(t % tb);
        return bs;
    };
    // from each x: every giant a != 0 and the babies of a = 0
    std::vector<int> xsteps;
    std::vector<std::pair<int, int>> xwhat;  // (a, b) of each x rotation
    for (int a : need_a) {
        if (a) {
            xsteps.push_back(kGroup * tb * a);
            xwhat.push_back({a, 0});
        } else {
            for (int b : babies_of(0))
                if (b) {
                    xsteps.push_back(kGroup * b);
                    xwhat.push_back({0, b});
                }
        }
    }
    std::vector<std::vector<Ciphertext>> xr;
    if (!xsteps.empty()) ev.rotate_vector_hoisted_many(xs, xsteps, *ck_.galois_keys, xr);
    const int amax = need_a.empty() ? 0 : need_a.back();
    std::vector<std::vector<Ciphertext>> giant(X, std::vector<Ciphertext>(amax + 1));
    for (size_t x = 0; x < X && !xsteps.empty(); x++)
        for (size_t i = 0; i < xsteps.size(); i++) {
            const int a = xwhat[i].first, b = xwhat[i].second;
            if (b)
                out[x][pos[tb * a + b]] = std::move(xr[x][i]);
            else
                giant[x][a] = std::move(xr[x][i]);
        }
    xr.clear();
    // every giant's babies in one batched hoisted call; the union of their baby
    // steps, the few outside a giant's offsets rotated and dropped
    std::vector<const Ciphertext *> gs;
    std::vector<std::pair<size_t, int>> gwhat;  // (x, a) of each giant
    std::vector<int> ub;
    for (size_t x = 0; x < X; x++)
        for (int a : need_a)
            if (a) {
                gs.push_back(&giant[x][a]);
                gwhat.push_back({x, a});
                for (int b : babies_of(a))
                    if (b && std::find(ub.begin(), ub.end(), b) == ub.end()) ub.push_back(b);
            }
    std::sort(ub.begin(), ub.end());
    std::vector<std::vector<Ciphertext>> gr;
    if (!gs.empty() && !ub.empty()) {
        std::vector<int> bsteps;
        for (int b : ub) bsteps.push_back(kGroup * b);
        ev.rotate_vector_hoisted_many(gs, bsteps, *ck_.galois_keys, gr);
    }
    for (size_t k = 0; k < gwhat.size() && !gr.empty(); k++) {
        const size_t x = gwhat[k].first;
        const int a = gwhat[k].second;
        for (int b : babies_of(a)) {
            if (!b) continue;
            const size_t j = std::find(ub.begin(), ub.end(), b) - ub.begin();
            out[x][pos[tb * a + b]] = std::move(gr[k][j]);
        }
    }
    for (size_t x = 0; x < X; x++)
        for (int a : need_a)
            for (int b : babies_of(a))
                if (!b) out[x][pos[tb * a]] = a ? std::move(giant[x][a]) : *xs[x];  // the giant itself
    return out;
}

Ciphertext BertLayer::rotate_tokens(const Ciphertext &x, int t)
{
    // rot(x, 256 t) with the token-rotation keys: 256 tb a then 256 b
    const Evaluator &ev = ck_.evaluator;
    const int T = (int)T_;
    t = ((t % T) + T) % T;
    const int tb = sqrt_pow2(T_);
    Ciphertext y = x;
    if (t / tb) ev.rotate_vector_inplace(y, kGroup * tb * (t / tb), *ck_.galois_keys);
    if (t % tb) ev.rotate_vector_inplace(y, kGroup * (t % tb), *ck_.galois_keys);
    return y;
}

std::vector<Ciphertext> BertLayer::shift_rows(const std::vector<Ciphertext> &x, int offset)
{
    if (offset < 0 || (size_t)offset >= T_) throw std::invalid_argument("BertLayer: row offset out of range");
    std::vector<Ciphertext> out;
    for (const auto &c : x) out.push_back(offset ? rotate_tokens(c, -offset) : c);
    return out;
}

std::vector<Ciphertext> BertLayer::extract_rows(const std::vector<Ciphertext> &x, int offset, int len)
{
    if (offset < 0 || len < 1 || (size_t)(offset + len) > T_)
        throw std::invalid_argument("BertLayer: rows out of range");
    const Evaluator &ev = ck_.evaluator;
    std::vector<Ciphertext> out;
    for (const auto &c : x) {
        Ciphertext y = offset ? rotate_tokens(c, offset) : c;
        const size_t l = y.coeff_modulus_size();
        if (l < 2) throw std::invalid_argument("BertLayer: extract_rows needs two limbs");
        auto key = std::make_tuple(-1000 - len, l, (double)prime(l), std::vector<int>());
        auto it = masks_.find(key);
        if (it == masks_.end()) {
            std::vector<double> v(slots_, 0.0);
            std::fill(v.begin(), v.begin() + (size_t)kGroup * len, 1.0);
            Plaintext pt(*ck_.context);
            ck_.encoder.encode(v, chain_of(l), (double)prime(l), pt);
            it = masks_.emplace(key, std::move(pt)).first;
        }
        ev.multiply_plain_inplace(y, it->second);
        ev.rescale_to_next_inplace(y);
        out.push_back(std::move(y));
    }
    return out;
}

std::vector<Ciphertext> BertLayer::offset_rotations(const Ciphertext &x, int dmin, int dmax)
{
    std::vector<std::vector<Ciphertext>> r = offset_rotations_many({&x}, dmin, dmax);
    return std::move(r[0]);
}

std::vector<std::vector<Ciphertext>> BertLayer::offset_rotations_many(const std::vector<const Ciphertext *> &xs,
                                                                      int dmin, int dmax)
{
    // rot(x, d) for every x and d in [dmin, dmax], d = 8 a + b
    const Evaluator &ev = ck_.evaluator;
    const size_t X = xs.size();
    std::vector<std::vector<Ciphertext>> out(X, std::vector<Ciphertext>(dmax - dmin + 1));
    if (opt_.direct_rotations) {
        // every offset hoisted from one ModUp of each x
        std::vector<int> steps;
        for (int d = dmin; d <= dmax; d++)
            if (d) steps.push_back(d);
        std::vector<std::vector<Ciphertext>> r;
        if (!steps.empty()) ev.rotate_vector_hoisted_many(xs, steps, *ck_.galois_keys, r);
        for (size_t x = 0; x < X; x++) {
            size_t i = 0;
            for (int d = dmin; d <= dmax; d++) out[x][d - dmin] = d ? std::move(r[x][i++]) : *xs[x];
        }
        return out;
    }
    // babies 1..7 of every x in one batched hoisted call, then the giants of
    // every x and every baby (8 bases each) in one more: the union of their
    // giant steps, the few outside a base's offsets rotated and dropped
    std::vector<int> bsteps;
    for (int b = 1; b < kOffBaby; b++) bsteps.push_back(b);
    std::vector<std::vector<Ciphertext>> br;
    ev.rotate_vector_hoisted_many(xs, bsteps, *ck_.galois_keys, br);
    auto giants_of = [&](int b) {
        std::vector<int> as;
        for (int d = dmin; d <= dmax; d++)
            if (((d % kOffBaby) + kOffBaby) % kOffBaby == b) as.push_back(floor_div(d, kOffBaby));
        return as;
    };
    std::vector<int> ua;
    for (int b = 0; b < kOffBaby; b++)
        for (int a : giants_of(b))
            if (a && std::find(ua.begin(), ua.end(), a) == ua.end()) ua.push_back(a);
    std::sort(ua.begin(), ua.end());
    std::vector<const Ciphertext *> bases;  // base (x, b) at x kOffBaby + b
    for (size_t x = 0; x < X; x++)
        for (int b = 0; b < kOffBaby; b++) bases.push_back(b ? &br[x][b - 1] : xs[x]);
    std::vector<std::vector<Ciphertext>> gr;
    if (!ua.empty()) {
        std::vector<int> gsteps;
        for (int a : ua) gsteps.push_back(kOffBaby * a);
        ev.rotate_vector_hoisted_many(bases, gsteps, *ck_.galois_keys, gr);
    }
    for (size_t x = 0; x < X; x++) {
        for (int b = 0; b < kOffBaby; b++)
            for (int a : giants_of(b)) {
                if (!a) continue;
                const size_t j = std::find(ua.begin(), ua.end(), a) - ua.begin();
                out[x][kOffBaby * a + b - dmin] = std::move(gr[x * kOffBaby + b][j]);
            }
        for (int b = 0; b < kOffBaby; b++) {
            const int d = b;  // giant a = 0: the base itself
            if (d >= dmin && d <= dmax) out[x][d - dmin] = b ? std::move(br[x][b - 1]) : *xs[x];
        }
    }
    return out;
}

const PeriodicPlaintexts &BertLayer::offset_masks(size_t limbs, double scale)
{
    const auto key = std::make_pair(limbs, scale);
    auto it = offset_masks_.find(key);
    if (it != offset_masks_.end()) return *it->second;
    std::vector<double> eye(kHeadDim * kHeadDim, 0.0);
    for (int p = 0; p < kHeadDim; p++) eye[p * kHeadDim + p] = 1.0;
    auto w = std::make_unique<PeriodicPlaintexts>(*ck_.context, eye, kHeadDim, chain_of(limbs), scale);
    return *offset_masks_.emplace(key, std::move(w)).first->second;
}

std::vector<int> BertLayer::segment_of(const std::vector<Segment> &segments) const
{
    std::vector<int> seg(T_, -1);
    for (size_t k = 0; k < segments.size(); k++) {
        const Segment &g = segments[k];
        if (g.len < 1 || g.start < 0 || (size_t)g.start + (size_t)g.len > T_)
            throw std::invalid_argument("BertLayer: segment out of range");
        for (int i = g.start; i < g.start + g.len; i++) {
            if (seg[i] >= 0) throw std::invalid_argument("BertLayer: overlapping segments");
            seg[i] = (int)k;
        }
    }
    if (segments.empty()) throw std::invalid_argument("BertLayer: no segment");
    return seg;
}

const Plaintext &BertLayer::key_mask(int eta, int tc, const std::vector<int> &seg, size_t limbs)
{
    // 1 where query i and key (i + 64 eta + t') mod T lie in one segment, t' < tc;
    // at the scale of the prime the next rescale drops
    const double scale = (double)prime(limbs);
    auto key = std::make_tuple(-1 - eta, limbs, scale, seg);
    auto it = masks_.find(key);
    if (it != masks_.end()) return it->second;
    std::vector<double> km(slots_, 0.0);
    const int T = (int)T_;
    for (int i = 0; i < T; i++) {
        if (seg[i] < 0) continue;
        for (int h = 0; h < kGroup / kHeadDim; h++)
            for (int tp = 0; tp < tc; tp++)
                if (seg[(i + kHeadDim * eta + tp) % T] == seg[i]) km[(size_t)kGroup * i + kHeadDim * h + tp] = 1.0;
    }
    Plaintext pt(*ck_.context);
    ck_.encoder.encode(km, chain_of(limbs), scale, pt);
    return masks_.emplace(key, std::move(pt)).first->second;
}

const Plaintext &BertLayer::row_mask(int tag, const std::vector<int> &seg, const std::vector<Segment> &segments,
                                     size_t limbs, double scale, int period)
{
    // 1 / (length of token i's segment) at slots 256 i + m period of token i's
    // group (0 outside the segments)
    auto key = std::make_tuple(-100 - tag, limbs, scale, seg);
    auto it = masks_.find(key);
    if (it != masks_.end()) return it->second;
    std::vector<double> v(slots_, 0.0);
    for (size_t i = 0; i < T_; i++) {
        if (seg[i] < 0) continue;
        const double w = 1.0 / segments[seg[i]].len;
        for (int y = 0; y < kGroup; y += period) v[kGroup * i + y] = w;
    }
    Plaintext pt(*ck_.context);
    ck_.encoder.encode(v, chain_of(limbs), scale, pt);
    return masks_.emplace(key, std::move(pt)).first->second;
}

void BertLayer::bootstrap_real(const std::vector<Ciphertext *> &cts, double factor)
{
    // real messages two at a time: y = a + i b (X^(N/2) multiplies every slot
    // by i), y' = bootstrap(y) factor / 2, then a' = y' + conj(y') and
    // b' = -i (y' - conj(y')) -- one bootstrap and one conjugation per pair;
    // all of them bootstrapped together (one EvalMod over the whole set)
    const Evaluator &ev = ck_.evaluator;
    const int64_t N = (int64_t)ck_.context->poly_degree();
    std::vector<std::pair<Ciphertext *, Ciphertext *>> items;
    std::vector<Ciphertext> ys;
    ys.reserve(cts.size());
    std::vector<const Ciphertext *> in;
    std::vector<double> factors;
    size_t k = 0;
    while (k < cts.size()) {
        Ciphertext &a = *cts[k];
        if (k + 1 < cts.size() && cts[k + 1]->scale() == a.scale() &&
            cts[k + 1]->coeff_modulus_size() == a.coeff_modulus_size()) {
            Ciphertext &b = *cts[k + 1];
            ys.emplace_back();
            ev.multiply_power_of_x(b, N / 2, ys.back());
            ev.add_inplace(ys.back(), a);
            in.push_back(&ys.back());
            factors.push_back(factor / 2);
            items.emplace_back(&a, &b);
            k += 2;
        } else {
            in.push_back(&a);
            factors.push_back(factor);
            items.emplace_back(&a, nullptr);
            k += 1;
        }
    }
    // at most kBootChunk bootstraps per batched call: a bootstrap costs the same
    // per input from about 5 inputs on (26.5 ms alone, 24.1 ms in 10, 23.8 ms in
    // 32), while the lockstep EvalMod's working set grows with every input
    constexpr size_t kBootChunk = 8;
    std::vector<Ciphertext> outs;
    for (size_t b0 = 0; b0 < in.size(); b0 += kBootChunk) {
        const size_t b1 = std::min(in.size(), b0 + kBootChunk);
        std::vector<const Ciphertext *> ci(in.begin() + b0, in.begin() + b1);
        std::vector<double> cf(factors.begin() + b0, factors.begin() + b1);
        std::vector<Ciphertext> co;
        bt_.bootstrap_many(ci, co, cf);
        for (auto &c : co) outs.push_back(std::move(c));
    }
    ys.clear();
    for (size_t i = 0; i < items.size(); i++) {
        Ciphertext &yb = outs[i];
        if (items[i].second) {
            Ciphertext yc, d;
            ev.complex_conjugate(yb, *ck_.galois_keys, yc);
            ev.add(yb, yc, *items[i].first);
            ev.sub(yb, yc, d);
            ev.multiply_power_of_x(d, 3 * N / 2, *items[i].second);
        } else {
            *items[i].first = std::move(yb);
        }
    }
}

Ciphertext BertLayer::scaled(const Ciphertext &x, double value)
{
    // x * value at two limbs -> one limb, landing exactly on the evaluator's scale
    const Evaluator &ev = ck_.evaluator;
    Ciphertext y = x;
    if (y.coeff_modulus_size() > 2) ev.mod_switch_to_inplace(y, chain_of(2));
    Plaintext c(*ck_.context);
    ck_.encoder.encode(value, y.params_id(), ck_.scale * (double)prime(y.coeff_modulus_size()) / y.scale(), c);
    ev.multiply_plain_inplace(y, c);
    ev.rescale_to_next_inplace(y);
    return y;
}

std::vector<Ciphertext> BertLayer::attention(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                                             const std::vector<Ciphertext> &v, int s)
{
    if (s < 1 || (size_t)s > T_) throw std::invalid_argument("BertLayer: sequence length out of range");
    return attention(q, k, v, std::vector<Segment>{{0, s}});
}

std::vector<Ciphertext> BertLayer::attention(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                                             const std::vector<Ciphertext> &v, const std::vector<Segment> &segments)
{
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (q.size() != n || k.size() != n || v.size() != n)
        throw std::invalid_argument("BertLayer: attention needs hidden / 256 ciphertexts of q, k and v");
    AttnState st = attention_probs(q, k, segments);
    {
        std::vector<Ciphertext *> all;
        for (auto &row : st.Pe)
            for (int eta : st.active) all.push_back(&row[eta]);
        bootstrap_real(all, 1.0);
    }
    mark("attention.bootstrap");
    return attention_values(st, v);
}

BertLayer::AttnState BertLayer::attention_probs(const std::vector<Ciphertext> &q, const std::vector<Ciphertext> &k,
                                                const std::vector<Segment> &segments)
{
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (q.size() != n || k.size() != n)
        throw std::invalid_argument("BertLayer: attention needs hidden / 256 ciphertexts of q and k");
    AttnState st;
    st.seg = segment_of(segments);
    const std::vector<int> &seg = st.seg;
    const Evaluator &ev = ck_.evaluator;
    const RelinKeys &rk = *ck_.relin_keys;
    const GaloisKeys &gk = *ck_.galois_keys;
    const int T = (int)T_;
    const int halves = (T + kHeadDim - 1) / kHeadDim;
    // token offsets t = 64 eta + t' that can pair two rows of one segment:
    // t < lmax or t > T - lmax (cyclically); half eta computes t' in [lo, hi)
    int lmax = 0;
    for (const Segment &g : segments) lmax = std::max(lmax, g.len);
    st.halves = halves;
    std::vector<int> &tlo = st.tlo, &thi = st.thi;
    tlo.assign(halves, 0);
    thi.assign(halves, 0);
    for (int eta = 0; eta < halves; eta++) {
        const int t0 = kHeadDim * eta, full = std::min(kHeadDim, T - t0);
        const int hi1 = std::max(0, std::min(full, lmax - t0));
        const int lo2 = std::max(0, std::min(full, T - lmax + 1 - t0));
        const bool a1 = hi1 > 0, a2 = lo2 < full;
        tlo[eta] = a1 ? 0 : (a2 ? lo2 : 0);
        thi[eta] = a2 ? full : (a1 ? hi1 : 0);
    }
    std::vector<int> &active = st.active;
    for (int eta = 0; eta < halves; eta++)
        if (thi[eta] > tlo[eta]) active.push_back(eta);
    // probabilities of every ciphertext and active half, refreshed together
    std::vector<std::vector<Ciphertext>> &Pe = st.Pe;
    Pe.assign(n, std::vector<Ciphertext>(halves));
    // q and k of every ciphertext at one level; the token rotations of every k
    // over every active half's offsets in one batched call
    std::vector<Ciphertext> qs(n), ks(n);
    for (size_t c = 0; c < n; c++) {
        qs[c] = q[c];
        ks[c] = k[c];
        if (ks[c].coeff_modulus_size() > qs[c].coeff_modulus_size()) ev.mod_switch_to_inplace(ks[c], qs[c].params_id());
        if (qs[c].coeff_modulus_size() > ks[c].coeff_modulus_size()) ev.mod_switch_to_inplace(qs[c], ks[c].params_id());
    }
    std::vector<int> ts;
    for (int eta : active)
        for (int t = kHeadDim * eta + tlo[eta]; t < kHeadDim * eta + thi[eta]; t++) ts.push_back(t);
    std::vector<std::vector<Ciphertext>> KR;
    {
        std::vector<const Ciphertext *> kp;
        for (auto &x : ks) kp.push_back(&x);
        KR = token_rotations_many(kp, ts);
    }
    ks.clear();
    mark("attention.k_rotations");
    // ---- scores, compacted: S^eta[256 i + 64 h + t'] = q_i . k_(i + 64 eta + t') on head h.
    // Every stage over every (ciphertext, active half) at once: the products in
    // one batched call, a block mix per (ciphertext, half), the partial rotation
    // sums Y_a of all of them in one call, their final sums in one more.
    std::vector<std::vector<Ciphertext>> P(n * halves);  // [c halves + eta][tau - lo]
    {
        std::vector<const Ciphertext *> qa, kb;
        for (size_t c = 0; c < n; c++) {
            size_t tpos = 0;
            for (int eta : active) {
                for (int i = 0; i < thi[eta] - tlo[eta]; i++) {
                    qa.push_back(&qs[c]);
                    kb.push_back(&KR[c][tpos + i]);
                }
                tpos += thi[eta] - tlo[eta];
            }
        }
        std::vector<Ciphertext> prod;
        ev.multiply_relin_rescale_many(qa, kb, rk, prod, true);
        KR.clear();
        size_t k = 0;
        for (size_t c = 0; c < n; c++)
            for (int eta : active)
                for (int i = 0; i < thi[eta] - tlo[eta]; i++) P[c * halves + eta].push_back(std::move(prod[k++]));
    }
    mark("attention.qk_products");
    // S = sum_d rot(Z_d, d), Z_d = sum_tau A_(tau + d) (.) P_tau, d = 8 a + b:
    // Y_a = sum_b rot(Z_(8a+b), b), S = sum_a rot(Y_a, 8 a), each sum of
    // rotations with a single ModDown
    std::vector<std::vector<Ciphertext>> Zall(n * halves);
    std::vector<std::vector<const Ciphertext *>> zgroups;
    std::vector<std::vector<int>> zsteps, ysteps(n * halves);
    std::vector<size_t> ygroup;  // (c, eta) of each Y group
    for (size_t c = 0; c < n; c++)
        for (int eta : active) {
            const int lo = tlo[eta], hi = thi[eta];
            std::vector<Ciphertext> &Pc = P[c * halves + eta];
            const size_t lp = Pc[0].coeff_modulus_size();
            const double ms = (double)prime(lp);
            const int dmin = -(hi - 1), dmax = kHeadDim - 1 - lo;
            // every Z_d in one block mix: each P_tau is read once for all offsets
            {
                std::vector<const Ciphertext *> in;
                for (auto &p : Pc) in.push_back(&p);
                const int K = hi - lo, M = dmax - dmin + 1;
                std::vector<int32_t> index((size_t)M * K, -1);
                for (int m = 0; m < M; m++)
                    for (int k = 0; k < K; k++) {
                        const int pidx = lo + k + dmin + m;  // tau + d
                        if (pidx >= 0 && pidx < kHeadDim) index[(size_t)m * K + k] = pidx;
                    }
                ev.multiply_periodic_mix(in, offset_masks(lp, ms), index, Zall[c * halves + eta]);
            }
            Pc.clear();
            for (int a = floor_div(dmin, kOffBaby); a <= floor_div(dmax, kOffBaby); a++) {
                std::vector<const Ciphertext *> zp;
                std::vector<int> zs;
                for (int b = 0; b < kOffBaby; b++) {
                    const int d = kOffBaby * a + b;
                    if (d < dmin || d > dmax) continue;
                    zp.push_back(&Zall[c * halves + eta][d - dmin]);
                    zs.push_back(b);
                }
                if (zp.empty()) continue;
                zgroups.push_back(std::move(zp));
                zsteps.push_back(std::move(zs));
                ysteps[c * halves + eta].push_back(kOffBaby * a);
                ygroup.push_back(c * halves + eta);
            }
        }
    std::vector<Ciphertext> Ys;
    ev.rotate_sum_many(zgroups, zsteps, gk, Ys);
    zgroups.clear();
    Zall.clear();
    std::vector<std::vector<Ciphertext>> S(n, std::vector<Ciphertext>(halves));
    {
        std::vector<std::vector<const Ciphertext *>> ygroups(n * halves);
        for (size_t g = 0; g < Ys.size(); g++) ygroups[ygroup[g]].push_back(&Ys[g]);
        std::vector<std::vector<const Ciphertext *>> gp;
        std::vector<std::vector<int>> gs;
        std::vector<size_t> which;
        for (size_t ce = 0; ce < n * halves; ce++)
            if (!ygroups[ce].empty()) {
                gp.push_back(ygroups[ce]);
                gs.push_back(ysteps[ce]);
                which.push_back(ce);
            }
        std::vector<Ciphertext> sums;
        ev.rotate_sum_many(gp, gs, gk, sums);
        std::vector<const Ciphertext *> sp;
        for (auto &x : sums) sp.push_back(&x);
        std::vector<Ciphertext> resc;
        ev.rescale_to_next_many(sp, resc);
        for (size_t k = 0; k < which.size(); k++) S[which[k] / halves][which[k] % halves] = std::move(resc[k]);
    }
    Ys.clear();
    mark("attention.compaction");
    // ---- softmax over t for every (query, head): exp, key mask, sum, 1 / sum,
    // every ciphertext in lockstep (each product, rescale and rotation one
    // batched call over the ciphertexts)
    std::vector<std::vector<Ciphertext>> E(n, std::vector<Ciphertext>(halves));
    {
        std::vector<const Ciphertext *> sp;
        for (size_t c = 0; c < n; c++)
            for (int eta : active) sp.push_back(&S[c][eta]);
        std::vector<Ciphertext> ex = ck_.exp_many(sp);
        S.clear();
        std::vector<const Ciphertext *> mp;
        size_t k = 0;
        for (size_t c = 0; c < n; c++)
            for (int eta : active) {
                E[c][eta] = std::move(ex[k++]);
                const int full = std::min(kHeadDim, T - kHeadDim * eta);
                ev.multiply_plain_inplace(E[c][eta], key_mask(eta, full, seg, E[c][eta].coeff_modulus_size()));
                mp.push_back(&E[c][eta]);
            }
        std::vector<Ciphertext> resc;
        ev.rescale_to_next_many(mp, resc);
        k = 0;
        for (size_t c = 0; c < n; c++)
            for (int eta : active) E[c][eta] = std::move(resc[k++]);
    }
    std::vector<Ciphertext> tot(n);
    for (size_t c = 0; c < n; c++) {
        bool have_tot = false;
        for (int eta : active) {
            if (!have_tot) {
                tot[c] = E[c][eta];
                have_tot = true;
            } else {
                ev.add_inplace(tot[c], E[c][eta]);
            }
        }
    }
    auto rotate_add_all = [&](int st) {
        std::vector<const Ciphertext *> tp;
        for (auto &x : tot) tp.push_back(&x);
        std::vector<std::vector<Ciphertext>> r;
        ev.rotate_vector_hoisted_many(tp, {st}, gk, r);
        for (size_t c = 0; c < n; c++) ev.add_inplace(tot[c], r[c][0]);
    };
    auto mask_rescale_all = [&](std::vector<Ciphertext> &v, int tag, int period) {
        std::vector<const Ciphertext *> vp;
        for (auto &x : v) {
            ev.multiply_plain_inplace(x, row_mask(tag, seg, segments, x.coeff_modulus_size(),
                                                  (double)prime(x.coeff_modulus_size()), period));
            vp.push_back(&x);
        }
        std::vector<Ciphertext> r;
        ev.rescale_to_next_many(vp, r);
        v = std::move(r);
    };
    for (int st = 1; st < kHeadDim; st <<= 1) rotate_add_all(st);
    // sum / len at the first slot of every 64-slot block
    mask_rescale_all(tot, 0, kHeadDim);
    for (int st = 1; st < kHeadDim; st <<= 1) rotate_add_all(-st);
    std::vector<Ciphertext> inv;
    {
        std::vector<const Ciphertext *> tp;
        for (auto &x : tot) tp.push_back(&x);
        inv = ck_.inverse_many(tp);
    }
    tot.clear();
    mask_rescale_all(inv, 1, 1);
    {
        std::vector<const Ciphertext *> pa, pb;
        for (size_t c = 0; c < n; c++)
            for (int eta : active) {
                ev.mod_switch_to_inplace(E[c][eta], inv[c].params_id());
                pa.push_back(&E[c][eta]);
                pb.push_back(&inv[c]);
            }
        std::vector<Ciphertext> pr;
        ev.multiply_relin_rescale_many(pa, pb, rk, pr);
        size_t k = 0;
        for (size_t c = 0; c < n; c++)
            for (int eta : active) Pe[c][eta] = std::move(pr[k++]);
    }
    mark("attention.softmax");
    return st;
}

std::vector<Ciphertext> BertLayer::attention_values(AttnState &st, const std::vector<Ciphertext> &v)
{
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (v.size() != n || st.Pe.size() != n)
        throw std::invalid_argument("BertLayer: attention needs hidden / 256 ciphertexts of v");
    const Evaluator &ev = ck_.evaluator;
    const RelinKeys &rk = *ck_.relin_keys;
    const std::vector<int> &tlo = st.tlo, &thi = st.thi, &active = st.active;
    std::vector<std::vector<Ciphertext>> &Pe = st.Pe;
    // ---- expansion over the 64 features of each head; times the rotated v
    // the token rotations of every v (at one level below the probabilities)
    // over every active half's offsets in one batched call
    std::vector<int> ts;
    for (int eta : active)
        for (int t = kHeadDim * eta + tlo[eta]; t < kHeadDim * eta + thi[eta]; t++) ts.push_back(t);
    std::vector<std::vector<Ciphertext>> VR;
    if (!active.empty()) {
        const size_t lb = Pe[0][active[0]].coeff_modulus_size();
        std::vector<Ciphertext> vs(n);
        for (size_t c = 0; c < n; c++) {
            if (Pe[c][active[0]].coeff_modulus_size() != lb)
                throw std::logic_error("BertLayer: probabilities at different levels");
            vs[c] = v[c];
            if (vs[c].coeff_modulus_size() > lb - 1) ev.mod_switch_to_inplace(vs[c], chain_of(lb - 1));
        }
        std::vector<const Ciphertext *> vp;
        for (auto &x : vs) vp.push_back(&x);
        VR = token_rotations_many(vp, ts);
    }
    mark("attention.v_rotations");
    // the offset rotations of every ciphertext's probabilities, per half in one batched call
    std::vector<std::vector<std::vector<Ciphertext>>> PR(st.halves);
    for (int eta : active) {
        std::vector<const Ciphertext *> pp;
        for (size_t c = 0; c < n; c++) pp.push_back(&Pe[c][eta]);
        PR[eta] = offset_rotations_many(pp, tlo[eta] - (kHeadDim - 1), thi[eta] - 1);
    }
    mark("attention.p_offsets");
    std::vector<Ciphertext> out(n);
    for (size_t c = 0; c < n; c++) {
        Ciphertext acc;
        bool have_acc = false;
        size_t tpos = 0;  // this half's first offset in ts
        for (int eta : active) {
            const int t0 = kHeadDim * eta, lo = tlo[eta], hi = thi[eta];
            const Ciphertext &Pb = Pe[c][eta];
            // Pb_tau[256 i + 64 h + f] = P[256 i + 64 h + tau] = sum_d A_(tau - d) (.) rot(P, d)
            const int dlo = lo - (kHeadDim - 1);
            std::vector<Ciphertext> R = std::move(PR[eta][c]);
            const size_t lb = Pb.coeff_modulus_size();
            const double ms = (double)prime(lb);
            std::vector<Ciphertext> vr(hi - lo);
            for (int i = 0; i < hi - lo; i++) vr[i] = std::move(VR[c][tpos + i]);
            tpos += hi - lo;
            (void)t0;
            // every Pb_tau in one block mix over the rotations R_d
            std::vector<Ciphertext> pbs;
            {
                std::vector<const Ciphertext *> in;
                for (auto &r : R) in.push_back(&r);
                const int K = (int)R.size(), M = hi - lo;
                std::vector<int32_t> index((size_t)M * K, -1);
                for (int m = 0; m < M; m++)
                    for (int k = 0; k < K; k++) {
                        const int pidx = (lo + m) - (dlo + k);  // tau - d
                        if (pidx >= 0 && pidx < kHeadDim) index[(size_t)m * K + k] = pidx;
                    }
                ev.multiply_periodic_mix(in, offset_masks(lb, ms), index, pbs);
            }
            R.clear();
            {
                std::vector<const Ciphertext *> pp;
                for (auto &x : pbs) pp.push_back(&x);
                ev.rescale_to_next_many(pp, pbs);
            }
            // sum_tau Pb_tau (x) rot(v, 256 tau) in one pass (no product written out)
            {
                std::vector<const Ciphertext *> pa, pv;
                for (int tau = lo; tau < hi; tau++) {
                    pa.push_back(&pbs[tau - lo]);
                    pv.push_back(&vr[tau - lo]);
                }
                Ciphertext part;
                ev.multiply_sum(pa, pv, part);
                if (!have_acc) {
                    acc = std::move(part);
                    have_acc = true;
                } else {
                    ev.add_inplace(acc, part);
                }
            }
            mark("attention.expansion");
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
    // the squares of every ciphertext in one batched multiply-relinearise-rescale
    Ciphertext tot;
    {
        std::vector<const Ciphertext *> rp;
        for (auto &x : r) rp.push_back(&x);
        std::vector<Ciphertext> sq;
        ev.multiply_relin_rescale_many(rp, rp, *ck_.relin_keys, sq);
        tot = std::move(sq[0]);
        for (size_t c = 1; c < n; c++) ev.add_inplace(tot, sq[c]);
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
    // every ciphertext times y in one batched multiply-relinearise-rescale
    std::vector<Ciphertext> xs(n);
    std::vector<const Ciphertext *> yp(n, &y), xp;
    for (size_t c = 0; c < n; c++) {
        ev.mod_switch_to(r[c], y.params_id(), xs[c]);
        xp.push_back(&xs[c]);
    }
    std::vector<Ciphertext> out;
    ev.multiply_relin_rescale_many(yp, xp, *ck_.relin_keys, out);
    return out;
}

void BertLayer::mark(const char *name)
{
    if (!opt_.stage_timing) return;
    ck_.context->synchronize();
    const auto now = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(now - t_mark_).count();
    t_mark_ = now;
    for (auto &e : stage_ms_)
        if (e.first == name) {
            e.second += ms;
            return;
        }
    stage_ms_.emplace_back(name, ms);
}

std::vector<Ciphertext> BertLayer::forward(const std::vector<Ciphertext> &x, int s)
{
    if (s < 1 || (size_t)s > T_) throw std::invalid_argument("BertLayer: sequence length out of range");
    return forward(x, std::vector<Segment>{{0, s}});
}

std::vector<Ciphertext> BertLayer::forward(const std::vector<Ciphertext> &x, const std::vector<Segment> &segments)
{
    const size_t n = (size_t)shape_.hidden / kGroup;
    if (x.size() != n) throw std::invalid_argument("BertLayer: input needs hidden / 256 ciphertexts");
    segment_of(segments);  // validates
    const Evaluator &ev = ck_.evaluator;
    const double pre = opt_.ln_prescale;
    auto boot_all = [&](std::vector<Ciphertext> &v, double factor) {
        std::vector<Ciphertext *> p;
        for (auto &c : v) p.push_back(&c);
        bootstrap_real(p, factor);
    };
    auto add_scaled = [&](std::vector<Ciphertext> &y, const std::vector<Ciphertext> &res) {
        for (size_t c = 0; c < y.size(); c++) {
            Ciphertext r8 = scaled(res[c], 1.0 / 8);
            r8.set_scale(y[c].scale());  // both are the evaluator's scale, by construction
            ev.add_inplace(y[c], r8);
        }
    };
    stage_ms_.clear();
    if (opt_.stage_timing) ck_.context->synchronize(), t_mark_ = std::chrono::steady_clock::now();
    // attention block; the bootstrapping range factors are folded into the weights
    std::vector<Ciphertext> q = linear(Wq, x, 1.0 / 32), k = linear(Wk, x, 0.25), v = linear(Wv, x, 0.25);
    mark("qkv");
    {
        std::vector<Ciphertext *> p;
        for (auto *set : {&q, &k, &v})
            for (auto &c : *set) p.push_back(&c);
        bootstrap_real(p, 4);
    }
    mark("bootstrap_qkv");
    std::vector<Ciphertext> o = attention(q, k, v, segments);
    q.clear();
    k.clear();
    v.clear();
    mark("attention");
    std::vector<Ciphertext> r1 = linear(Wo, o, 1.0 / 8);
    add_scaled(r1, x);
    mark("wo_residual");
    boot_all(r1, 8 * pre);
    mark("bootstrap_ln1");
    std::vector<Ciphertext> z = layer_norm(r1);
    mark("ln1");
    // feed-forward block
    std::vector<Ciphertext> h = linear(W1, z, 0.25);
    mark("ffn1");
    boot_all(h, 4);
    mark("bootstrap_ffn");
    {
        GELUEvaluator gelu(ck_);
        std::vector<const Ciphertext *> hp;
        for (auto &c : h) hp.push_back(&c);
        std::vector<Ciphertext> g;
        gelu.gelu_many(hp, g);
        h = std::move(g);
    }
    mark("gelu");
    std::vector<Ciphertext> r2 = linear(W2, h, 1.0 / 8);
    add_scaled(r2, z);
    mark("ffn2_residual");
    boot_all(r2, 8 * pre);
    mark("bootstrap_ln2");
    std::vector<Ciphertext> y = layer_norm(r2);
    mark("ln2");
    return y;
}

std::vector<std::vector<Ciphertext>> BertLayer::forward_sets(const std::vector<std::vector<Ciphertext>> &xs,
                                                             const std::vector<std::vector<Segment>> &segments)
{
    const size_t n = (size_t)shape_.hidden / kGroup, S = xs.size();
    if (segments.size() != S) throw std::invalid_argument("BertLayer: one segment list per set");
    if (!S) return {};
    for (size_t s = 0; s < S; s++) {
        if (xs[s].size() != n) throw std::invalid_argument("BertLayer: input needs hidden / 256 ciphertexts");
        segment_of(segments[s]);  // validates
    }
    const Evaluator &ev = ck_.evaluator;
    const double pre = opt_.ln_prescale;
    // every set's ciphertexts of one stage bootstrapped together (pairs across sets)
    auto boot_sets = [&](std::vector<std::vector<Ciphertext>> &v, double factor) {
        std::vector<Ciphertext *> p;
        for (auto &set : v)
            for (auto &c : set) p.push_back(&c);
        bootstrap_real(p, factor);
    };
    auto add_scaled = [&](std::vector<Ciphertext> &y, const std::vector<Ciphertext> &res) {
        for (size_t c = 0; c < y.size(); c++) {
            Ciphertext r8 = scaled(res[c], 1.0 / 8);
            r8.set_scale(y[c].scale());  // both are the evaluator's scale, by construction
            ev.add_inplace(y[c], r8);
        }
    };
    stage_ms_.clear();
    if (opt_.stage_timing) ck_.context->synchronize(), t_mark_ = std::chrono::steady_clock::now();
    std::vector<std::vector<Ciphertext>> q(S), k(S), v(S);
    for (size_t s = 0; s < S; s++) {
        q[s] = linear(Wq, xs[s], 1.0 / 32);
        k[s] = linear(Wk, xs[s], 0.25);
        v[s] = linear(Wv, xs[s], 0.25);
    }
    mark("qkv");
    {
        std::vector<Ciphertext *> p;
        for (size_t s = 0; s < S; s++)
            for (auto *set : {&q[s], &k[s], &v[s]})
                for (auto &c : *set) p.push_back(&c);
        bootstrap_real(p, 4);
    }
    mark("bootstrap_qkv");
    std::vector<AttnState> st(S);
    for (size_t s = 0; s < S; s++) {
        st[s] = attention_probs(q[s], k[s], segments[s]);
        q[s].clear();
        k[s].clear();
    }
    {
        std::vector<Ciphertext *> all;
        for (auto &a : st)
            for (auto &row : a.Pe)
                for (int eta : a.active) all.push_back(&row[eta]);
        bootstrap_real(all, 1.0);
    }
    mark("attention.bootstrap");
    std::vector<std::vector<Ciphertext>> r1(S);
    for (size_t s = 0; s < S; s++) {
        std::vector<Ciphertext> o = attention_values(st[s], v[s]);
        st[s] = AttnState();
        v[s].clear();
        r1[s] = linear(Wo, o, 1.0 / 8);
        add_scaled(r1[s], xs[s]);
    }
    mark("attention_wo_residual");
    boot_sets(r1, 8 * pre);
    mark("bootstrap_ln1");
    std::vector<std::vector<Ciphertext>> z(S), h(S);
    for (size_t s = 0; s < S; s++) {
        z[s] = layer_norm(r1[s]);
        r1[s].clear();
        h[s] = linear(W1, z[s], 0.25);
    }
    mark("ln1_ffn1");
    boot_sets(h, 4);
    mark("bootstrap_ffn");
    {
        GELUEvaluator gelu(ck_);
        std::vector<const Ciphertext *> hp;
        for (auto &set : h)
            for (auto &c : set) hp.push_back(&c);
        std::vector<Ciphertext> g;
        gelu.gelu_many(hp, g);
        size_t i = 0;
        for (auto &set : h)
            for (auto &c : set) c = std::move(g[i++]);
    }
    mark("gelu");
    std::vector<std::vector<Ciphertext>> r2(S);
    for (size_t s = 0; s < S; s++) {
        r2[s] = linear(W2, h[s], 1.0 / 8);
        h[s].clear();
        add_scaled(r2[s], z[s]);
        z[s].clear();
    }
    mark("ffn2_residual");
    boot_sets(r2, 8 * pre);
    mark("bootstrap_ln2");
    std::vector<std::vector<Ciphertext>> y(S);
    for (size_t s = 0; s < S; s++) y[s] = layer_norm(r2[s]);
    mark("ln2");
    return y;
}

}  // namespace nexus
