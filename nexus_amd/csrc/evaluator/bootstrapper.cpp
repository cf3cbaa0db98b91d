// CKKS bootstrapping: host planning of the CoeffToSlot / SlotToCoeff levels
// and the homomorphic pipeline over the C++ drop-in (include/nexus/bootstrapper.hpp).
#include "../../../include/nexus/bootstrapper.hpp"

#include <algorithm>
#include <cmath>
#include <set>
#include <stdexcept>

namespace nexus {

namespace {

typedef std::complex<double> cd;
typedef std::vector<cd> cvec;
typedef std::map<int, cvec> DiagMap;  // offset mod n -> n-periodic diagonal (length n)

// Twiddle of butterfly stage i (block length len = 2^(i+1)) at in-block index j:
// exp(2 pi i (5^j mod 4 len) / (4 len))  (reference genorigcoeff,
// cuda/src/bootstrapping/Bootstrapper.cu:537-556: theta_0 * 2^(logn-1-i) * 5^j).
std::vector<cd> stage_twiddles(int i)
{
    const uint64_t len = uint64_t(2) << i, m = 4 * len, h = len / 2;
    std::vector<cd> z(h);
    uint64_t e = 1;
    for (uint64_t j = 0; j < h; j++) {
        z[j] = std::polar(1.0, 2.0 * M_PI * (double)e / (double)m);
        e = (e * 5) % m;
    }
    return z;
}

// Diagonal form of one butterfly stage on an n-vector in bit-reversed order:
//   forward: y[s] = x[s] + z x[s+h],        y[s+h] = x[s] - z x[s+h]
//   inverse: y[s] = (x[s] + x[s+h]) / 2,    y[s+h] = (x[s] - x[s+h]) / (2 z)
// (reference genorigcoeff, Bootstrapper.cu:537-590; three diagonals at
// offsets 0, +h, -h).
DiagMap stage_diags(size_t n, int i, bool inverse)
{
    const size_t len = size_t(2) << i, h = len / 2;
    const std::vector<cd> z = stage_twiddles(i);
    DiagMap D;
    cvec &d0 = D[0];
    d0.assign(n, 0.0);
    cvec &dp = D[(int)h];
    dp.assign(n, 0.0);
    cvec &dm = D[(int)((n - h) % n)];  // == dp when h == n/2
    if (dm.empty()) dm.assign(n, 0.0);
    for (size_t b = 0; b < n; b += len)
        for (size_t j = 0; j < h; j++) {
            const size_t s = b + j;
            if (!inverse) {
                d0[s] += 1.0;
                D[(int)h][s] += z[j];
                D[(int)((n - h) % n)][s + h] += 1.0;
                d0[s + h] += -z[j];
            } else {
                const cd iz = std::conj(z[j]);
                d0[s] += 0.5;
                D[(int)h][s] += 0.5;
                D[(int)((n - h) % n)][s + h] += 0.5 * iz;
                d0[s + h] += -0.5 * iz;
            }
        }
    return D;
}

// Diagonals of A*B (B applied first): C_{a+b} += A_a (.) rot(B_b, a).
DiagMap compose(const DiagMap &A, const DiagMap &B, size_t n)
{
    DiagMap C;
    for (const auto &a : A)
        for (const auto &b : B) {
            const int c = (int)(((size_t)a.first + (size_t)b.first) % n);
            cvec &dst = C[c];
            if (dst.empty()) dst.assign(n, 0.0);
            for (size_t s = 0; s < n; s++) dst[s] += a.second[s] * b.second[(s + (size_t)a.first) % n];
        }
    // drop diagonals that cancel
    for (auto it = C.begin(); it != C.end();) {
        double mx = 0;
        for (const cd &v : it->second) mx = std::max(mx, std::abs(v));
        if (mx < 1e-12)
            it = C.erase(it);
        else
            ++it;
    }
    return C;
}

// Turn the composed diagonals of stages [lo, hi) into a BSGS level on N/2 slots.
// in_mask multiplies the element read at slot s + offset, out_mask the output
// slot s (both length N/2 or empty).
LinearLevel make_level(const DiagMap &D, size_t n, size_t Nh, int lo, const cvec &in_mask, const cvec &out_mask,
                       double factor)
{
    LinearLevel lv;
    lv.stride = 1 << lo;
    const long M = (long)(n / (size_t)lv.stride);  // offsets are t * stride, t taken mod M
    std::map<long, const cvec *> byt;
    long tmin = 0, tmax = 0;
    for (const auto &d : D) {
        if (d.first % lv.stride) throw std::logic_error("bootstrap plan: offset is not a multiple of the stride");
        long t = d.first / lv.stride;
        if (t > M / 2) t -= M;
        byt[t] = &d.second;
        tmin = std::min(tmin, t);
        tmax = std::max(tmax, t);
    }
    const long span = tmax - tmin + 1;
    int lg = 0;
    while ((1L << (2 * (lg + 1))) <= span) lg++;  // g = 2^floor(log4 span) .. keep g ~ sqrt(span)
    if ((1L << (2 * lg + 1)) < span) lg++;
    lv.g = 1 << lg;
    for (const auto &e : byt) {
        const long t = e.first;
        const long j = (t >= 0) ? t / lv.g : -((-t + lv.g - 1) / lv.g);
        const long i = t - j * lv.g;
        const cvec &d = *e.second;
        // tiled, masked diagonal over N/2 slots
        cvec full(Nh);
        const long off = t * lv.stride;
        for (size_t s = 0; s < Nh; s++) {
            cd v = d[s % n] * factor;
            if (!in_mask.empty()) v *= in_mask[(size_t)(((long)s + off) % (long)Nh + (long)Nh) % Nh];
            if (!out_mask.empty()) v *= out_mask[s];
            full[s] = v;
        }
        // pre-rotate by -g*j*stride: Dt[s] = D[s - g j stride]
        const long sh = j * lv.g * lv.stride;
        cvec rot(Nh);
        for (size_t s = 0; s < Nh; s++) rot[s] = full[(size_t)((((long)s - sh) % (long)Nh) + (long)Nh) % Nh];
        lv.diag[{(int)j, (int)i}] = std::move(rot);
    }
    return lv;
}

// Split logn stages into three consecutive groups (fewer when logn < 3); the
// group holding the top stage gets the extra stages: its offsets wrap mod n,
// so it has half as many diagonals per stage count.
std::vector<std::pair<int, int>> stage_groups(int logn)
{
    const int parts = std::min(3, logn);
    std::vector<std::pair<int, int>> g;  // [lo, hi) in increasing stage order
    int base = logn / parts, extra = logn % parts, lo = 0;
    for (int p = 0; p < parts; p++) {
        const int sz = base + (p >= parts - extra ? 1 : 0);
        g.push_back({lo, lo + sz});
        lo += sz;
    }
    return g;
}

DiagMap group_diags(size_t n, int lo, int hi, bool inverse)
{
    DiagMap G;
    bool first = true;
    if (!inverse) {
        for (int i = lo; i < hi; i++) {
            DiagMap S = stage_diags(n, i, false);
            G = first ? S : compose(S, G, n);
            first = false;
        }
    } else {
        for (int i = hi - 1; i >= lo; i--) {
            DiagMap S = stage_diags(n, i, true);
            G = first ? S : compose(S, G, n);
            first = false;
        }
    }
    return G;
}

// Least squares min |A x - b| (A m x k, row-major, m >= k) by Householder QR.
std::vector<double> lstsq(std::vector<double> A, std::vector<double> b, size_t m, size_t k)
{
    for (size_t j = 0; j < k; j++) {
        double norm = 0;
        for (size_t i = j; i < m; i++) norm += A[i * k + j] * A[i * k + j];
        norm = std::sqrt(norm);
        if (norm == 0) continue;
        const double alpha = A[j * k + j] > 0 ? -norm : norm;
        std::vector<double> v(m - j);
        for (size_t i = j; i < m; i++) v[i - j] = A[i * k + j];
        v[0] -= alpha;
        double vn = 0;
        for (double x : v) vn += x * x;
        if (vn == 0) continue;
        for (size_t c = j; c < k; c++) {
            double sdot = 0;
            for (size_t i = j; i < m; i++) sdot += v[i - j] * A[i * k + c];
            const double f = 2 * sdot / vn;
            for (size_t i = j; i < m; i++) A[i * k + c] -= f * v[i - j];
        }
        double sdot = 0;
        for (size_t i = j; i < m; i++) sdot += v[i - j] * b[i];
        const double f = 2 * sdot / vn;
        for (size_t i = j; i < m; i++) b[i] -= f * v[i - j];
    }
    std::vector<double> x(k, 0.0);
    for (size_t j = k; j-- > 0;) {
        double s = b[j];
        for (size_t c = j + 1; c < k; c++) s -= A[j * k + c] * x[c];
        x[j] = s / A[j * k + j];
    }
    return x;
}

// Chebyshev coefficients of the EvalMod polynomial: fit of
// cos(2 pi (K u - 1/4) / 2^r) at 40 Chebyshev-distributed points in each
// window [(i - w) / K, (i + w) / K], w = 2^-log_width, |i| < K.
std::vector<double> fit_eval_mod(int K, int degree, int r, int log_width)
{
    if (degree < 1) throw std::invalid_argument("Bootstrapper: degree too small");
    const double w = std::ldexp(1.0, -log_width), R = std::ldexp(1.0, r);
    const int per = 40;
    std::vector<double> xs;
    for (int i = -K + 1; i <= K - 1; i++)
        for (int t = 0; t < per; t++) xs.push_back((i + w * std::cos(M_PI * (t + 0.5) / per)) / K);
    const size_t m = xs.size(), k = (size_t)degree + 1;
    if (m < k) throw std::invalid_argument("Bootstrapper: degree too large for the fit");
    std::vector<double> A(m * k), b(m);
    for (size_t i = 0; i < m; i++) {
        const double x = xs[i];
        double t0 = 1, t1 = x;
        for (size_t j = 0; j < k; j++) {
            A[i * k + j] = j == 0 ? 1.0 : (j == 1 ? x : 0.0);
            if (j >= 2) {
                const double t2 = 2 * x * t1 - t0;
                A[i * k + j] = t2;
                t0 = t1;
                t1 = t2;
            }
        }
        b[i] = std::cos(2.0 * M_PI * (K * x - 0.25) / R);
    }
    return lstsq(std::move(A), std::move(b), m, k);
}

// ceil(log2 k): levels below the input at which T_k is produced
int tdepth(int k)
{
    int d = 0;
    while ((1 << d) < k) d++;
    return d;
}

// split of p = sum c_k T_k by T_P, 2^floor(log2 deg) = P: p = q T_P + r
// (T_{P+j} = 2 T_P T_j - T_{P-j})
void cheb_split(const std::vector<double> &c, int d, int P, std::vector<double> &q, std::vector<double> &r)
{
    q.assign(d - P + 1, 0.0);
    r.assign(P, 0.0);
    for (int k = 0; k < P; k++) r[k] = c[k];
    q[0] = c[P];
    for (int k = P + 1; k <= d; k++) {
        q[k - P] += 2.0 * c[k];
        r[2 * P - k] -= c[k];
    }
}

int trimmed_degree(const std::vector<double> &c)
{
    int d = (int)c.size() - 1;
    while (d > 0 && c[d] == 0.0) d--;
    return d;
}

}  // namespace

std::vector<double> eval_mod_coeffs(int K, int degree, int double_angles, int log_width)
{
    return fit_eval_mod(K, degree, double_angles, log_width);
}

BootPlan make_boot_plan(int logNh, int logn, double K)
{
    if (logn < 1 || logn > logNh) throw std::invalid_argument("bootstrap plan: logn out of range");
    BootPlan P;
    P.logn = logn;
    P.logNh = logNh;
    P.K = K;
    const size_t n = size_t(1) << logn, Nh = size_t(1) << logNh;
    const bool sparse = logn < logNh;
    const double c = (double)(Nh / n);
    const auto groups = stage_groups(logn);

    // CoeffToSlot: inverse stages logn-1 .. 0, i.e. groups from the top down.
    // The last level carries 1/(c K) and the real/imaginary split mask
    // (1/2 on the lower n slots of each 2n period, -i/2 on the upper n;
    // dense packing: 1/2 everywhere, split by conjugation afterwards).
    cvec A(Nh);
    for (size_t s = 0; s < Nh; s++) A[s] = (!sparse || (s % (2 * n)) < n) ? cd(0.5, 0) : cd(0, -0.5);
    for (int gi = (int)groups.size() - 1; gi >= 0; gi--) {
        const auto &g = groups[gi];
        DiagMap D = group_diags(n, g.first, g.second, true);
        const bool last = gi == 0;
        P.cts.push_back(make_level(D, n, Nh, g.first, {}, last ? A : cvec(), last ? 1.0 / (c * K) : 1.0));
    }
    // SlotToCoeff: forward stages 0 .. logn-1.  The first level carries
    // 1/(2 pi) (EvalMod returns sin(2 pi x)) and, for sparse packing, the
    // recombination mask (1 on the lower n slots of each 2n period, i on the
    // upper n) applied to the input it reads.
    cvec B1;
    if (sparse) {
        B1.resize(Nh);
        for (size_t s = 0; s < Nh; s++) B1[s] = ((s % (2 * n)) < n) ? cd(1, 0) : cd(0, 1);
    }
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const auto &g = groups[gi];
        DiagMap D = group_diags(n, g.first, g.second, false);
        const bool first = gi == 0;
        P.stc.push_back(make_level(D, n, Nh, g.first, first ? B1 : cvec(), {}, first ? 1.0 / (2.0 * M_PI) : 1.0));
    }

    std::set<int> steps;
    steps.insert(0);  // conjugation
    auto add_level_steps = [&](const LinearLevel &lv) {
        for (const auto &d : lv.diag) {
            const int j = d.first.first, i = d.first.second;
            if (i) steps.insert(i * lv.stride);
            if (j) steps.insert(j * lv.g * lv.stride);
        }
    };
    for (const auto &lv : P.cts) add_level_steps(lv);
    for (const auto &lv : P.stc) add_level_steps(lv);
    if (sparse)
        for (int i = logn; i < logNh; i++) steps.insert(1 << i);  // SubSum and the final rot(y, n)
    P.rotation_steps.assign(steps.begin(), steps.end());
    return P;
}

// ---------------------------------------------------------------- Bootstrapper

Bootstrapper::Bootstrapper(CKKSEvaluator &ckks, const Params &p) : ck_(ckks), p_(p)
{
    const int logNh = ck_.context->log_n() - 1;
    if (p_.logn == 0) p_.logn = logNh;
    if (p_.cheb_degree < 2) throw std::invalid_argument("Bootstrapper: degree too small");
    if (p_.double_angles < 0) throw std::invalid_argument("Bootstrapper: double_angles < 0");
    plan_ = make_boot_plan(logNh, p_.logn, (double)p_.K);

    // EvalMod polynomial in the Chebyshev basis on [-1, 1]: least-squares fit
    // of cos(2 pi (K u - 1/4) / 2^r) on the windows |K u - i| <= 2^-log_width,
    // |i| < K, where the ModRaise overflow puts the inputs (stands in for the
    // reference's minimax generator on the same windows, RemezCos,
    // cuda/src/bootstrapping/RemezCos.cuh:12-17).
    if (p_.K < 2 || p_.log_width < 1) throw std::invalid_argument("Bootstrapper: K >= 2 and log_width >= 1 required");
    cheb_ = eval_mod_coeffs(p_.K, p_.cheb_degree, p_.double_angles, p_.log_width);
}

// Levels consumed by cheb_eval for p: leaves (degree < 8) take the deepest
// baby step plus one scalar level; a split p = q T_P + r takes one level over
// the deeper of q and T_P, and r runs beside it.  Degree 59 fits in depth 6
// (quotients of degree 27, 11, 3 ride on T_32, T_16, T_8).
int chebyshev_depth(const std::vector<double> &c)
{
    const int d = trimmed_degree(c);
    if (d < 8) {
        int m = 0;
        for (int k = 1; k <= d; k++)
            if (c[k] != 0.0) m = std::max(m, tdepth(k));
        return m + 1;
    }
    int P = 1;
    while (2 * P <= d) P *= 2;
    std::vector<double> q, r;
    cheb_split(c, d, P, q, r);
    return std::max(std::max(chebyshev_depth(q), tdepth(P)) + 1, chebyshev_depth(r));
}

int Bootstrapper::depth() const
{
    return (int)plan_.cts.size() + chebyshev_depth(cheb_) + p_.double_angles + (int)plan_.stc.size();
}

void Bootstrapper::mod_raise(const Ciphertext &in, Ciphertext &out) const
{
    Ciphertext low;
    const Ciphertext *src = &in;
    if (in.coeff_modulus_size() > 1) {
        ck_.evaluator.mod_switch_to(in, ck_.context->data_primes(), low);
        src = &low;
    }
    Ciphertext r(*ck_.context);
    check(nx_mod_raise(src->get(), r.get()));
    // ModRaise leaves m + q_0 I: read it at scale q_0 (Bootstrapper.cu:2997-2998)
    r.set_scale(prime(1));
    out = std::move(r);
}

void Bootstrapper::sub_sum(Ciphertext &ct) const
{
    // Bootstrapper.cu:3004-3008; each rotation in its hoisted form (the
    // digits sigma(lift(c1))), as the batched bootstrap rotates
    for (int i = plan_.logn; i < plan_.logNh; i++) {
        std::vector<Ciphertext> rot;
        ck_.evaluator.rotate_vector_hoisted(ct, {1 << i}, *ck_.galois_keys, rot);
        ck_.evaluator.add_inplace(ct, rot[0]);
    }
}

Ciphertext Bootstrapper::match(const Ciphertext &ct, size_t level, double scale)
{
    // bring ct to `level` limbs at exactly `scale`: multiply by 1 encoded at
    // the scale that the following rescale compensates
    const size_t l = ct.coeff_modulus_size();
    if (l <= level) throw std::logic_error("Bootstrapper: operand below the target level");
    Ciphertext t;
    const double sigma = scale * prime(level + 1) / ct.scale();
    ck_.evaluator.linear_combination({&ct}, {1.0}, {sigma}, chain_of(level + 1), t);  // read in place
    ck_.evaluator.rescale_to_next_inplace(t);
    t.set_scale(scale);
    return t;
}

Ciphertext Bootstrapper::mul_exact(const Ciphertext &a, const Ciphertext &b)
{
    // product of two ciphertexts at the lower of their levels, relinearised and rescaled
    const size_t la = a.coeff_modulus_size(), lb = b.coeff_modulus_size();
    const size_t l = std::min(la, lb);
    Ciphertext xs, ys;
    const Ciphertext *x = &a, *y = &b;
    if (la > l) {
        ck_.evaluator.mod_switch_to(a, chain_of(l), xs);
        x = &xs;
    }
    if (lb > l) {
        ck_.evaluator.mod_switch_to(b, chain_of(l), ys);
        y = &ys;
    }
    Ciphertext r;
    if (&a == &b)
        ck_.evaluator.square(*x, r);
    else
        ck_.evaluator.multiply(*x, *y, r);
    ck_.evaluator.relinearize_inplace(r, *ck_.relin_keys);
    ck_.evaluator.rescale_to_next_inplace(r);
    return r;
}

// Chebyshev basis T_k of the EvalMod input (basis_[1]); T_{a+b} = 2 T_a T_b - T_{a-b}
// with a = 2^(ceil(log2 k) - 1), so T_k sits ceil(log2 k) levels below the input.
// T_{a-b} is brought to the product's level and scale exactly (match), so no
// scale mismatch enters the subtraction.
const Ciphertext &Bootstrapper::basis(int k)
{
    auto it = basis_.find(k);
    if (it != basis_.end()) return it->second;
    int a = 1;
    while (2 * a < k) a *= 2;
    const int b = k - a;
    const Ciphertext &Ta = basis(a);
    const Ciphertext &Tb = basis(b);
    Ciphertext prod = (a == b) ? mul_exact(Ta, Ta) : mul_exact(Ta, Tb);
    const size_t l = prod.coeff_modulus_size();
    if (a == b) {
        ck_.evaluator.add_inplace(prod, prod);  // 2 T_a^2 - 1
        Plaintext one(*ck_.context);
        ck_.encoder.encode(1.0, chain_of(l), prod.scale(), one);
        ck_.evaluator.sub_plain_inplace(prod, one);
    } else {
        // 2 T_a T_b - T_(a-b) in one pass (constants 2 and -1 at scale 1)
        Ciphertext Td = match(basis(a - b), l, prod.scale());
        Ciphertext r;
        ck_.evaluator.linear_combination({&prod, &Td}, {2.0, -1.0}, {1.0, 1.0}, chain_of(l), r);
        prod = std::move(r);
    }
    return basis_.emplace(k, std::move(prod)).first->second;
}

// p(u) = sum_k c_k T_k(u) at exactly (level, scale).  deg < 8: linear
// combination of the baby steps; otherwise p = q T_{2^m} + r (Chebyshev
// division by T_{2^m}, 2^m <= deg < 2^(m+1)).  The quotient is evaluated one
// level higher at the scale that the product's rescale maps onto `scale`.
// Feasible whenever cheb_depth(c) <= (input level - level).
Ciphertext Bootstrapper::cheb_eval(const std::vector<double> &c, size_t level, double scale)
{
    const int d = trimmed_degree(c);
    if (d < 8) {
        const size_t lam = level + 1;
        const double q = prime(lam);
        Ciphertext acc;
        // sum_k c_k T_k at level lam in one pass, the basis read in place
        std::vector<const Ciphertext *> xs;
        std::vector<double> vals, scs;
        for (int k = 1; k <= d; k++) {
            if (c[k] == 0.0) continue;
            const Ciphertext &bk = basis(k);
            if (bk.coeff_modulus_size() < lam) throw std::logic_error("Bootstrapper: basis below the leaf level");
            xs.push_back(&bk);
            vals.push_back(c[k]);
            scs.push_back(scale * q / bk.scale());
        }
        const bool have = !xs.empty();
        if (have) ck_.evaluator.linear_combination(xs, vals, scs, chain_of(lam), acc);
        if (!have) {  // constant polynomial: 0 * T_1 carries the level and scale
            acc = basis(1);
            ck_.evaluator.mod_switch_to_inplace(acc, chain_of(lam));
            Plaintext z(*ck_.context);
            ck_.encoder.encode(0.0, chain_of(lam), scale * q / acc.scale(), z);
            ck_.evaluator.multiply_plain_inplace(acc, z);
        }
        ck_.evaluator.rescale_to_next_inplace(acc);
        acc.set_scale(scale);
        if (c[0] != 0.0) {
            Plaintext c0(*ck_.context);
            ck_.encoder.encode(c[0], chain_of(level), scale, c0);
            ck_.evaluator.add_plain_inplace(acc, c0);
        }
        return acc;
    }
    int P = 1;
    while (2 * P <= d) P *= 2;
    std::vector<double> q, r;
    cheb_split(c, d, P, q, r);
    const Ciphertext &TP = basis(P);
    const size_t lam = level + 1;
    if (TP.coeff_modulus_size() < lam) throw std::logic_error("Bootstrapper: giant step below the product level");
    Ciphertext Q = cheb_eval(q, lam, scale * prime(lam) / TP.scale());
    Ciphertext Tlow;
    const Ciphertext *T = &TP;
    if (TP.coeff_modulus_size() > lam) {
        ck_.evaluator.mod_switch_to(TP, chain_of(lam), Tlow);
        T = &Tlow;
    }
    Ciphertext prod;
    ck_.evaluator.multiply(Q, *T, prod);
    ck_.evaluator.relinearize_inplace(prod, *ck_.relin_keys);
    ck_.evaluator.rescale_to_next_inplace(prod);
    prod.set_scale(scale);
    Ciphertext R = cheb_eval(r, level, scale);
    R.set_scale(scale);
    ck_.evaluator.add_inplace(prod, R);
    return prod;
}

void Bootstrapper::eval_mod(const Ciphertext &u, Ciphertext &out)
{
    // ModularReducer::modular_reduction (cuda/src/bootstrapping/ModularReducer.cu:53-73):
    // y = p(u) ~ cos(2 pi (K u - 1/4) / 2^r); r times y <- 2 y^2 - 1; y ~ sin(2 pi K u)
    basis_.clear();
    basis_.emplace(1, u);
    const size_t l0 = u.coeff_modulus_size();
    const int cd_ = chebyshev_depth(cheb_);
    if (l0 <= (size_t)cd_ + (size_t)p_.double_angles)
        throw std::invalid_argument("Bootstrapper: not enough levels for EvalMod");
    Ciphertext y = cheb_eval(cheb_, l0 - cd_, u.scale());
    basis_.clear();
    for (int i = 0; i < p_.double_angles; i++) {
        ck_.evaluator.square_inplace(y);
        ck_.evaluator.relinearize_inplace(y, *ck_.relin_keys);
        ck_.evaluator.rescale_to_next_inplace(y);
        ck_.evaluator.add_inplace(y, y);
        Plaintext one(*ck_.context);
        ck_.encoder.encode(1.0, y.chain_index(), y.scale(), one);
        ck_.evaluator.sub_plain_inplace(y, one);
    }
    out = std::move(y);
}

// ---- EvalMod over a batch: the single-input steps above, each multiplication
// (with its relinearisation and rescale) and each rescale one batched call.
namespace {

std::vector<const Ciphertext *> ptrs(const std::vector<Ciphertext> &v)
{
    std::vector<const Ciphertext *> p;
    for (const auto &x : v) p.push_back(&x);
    return p;
}

}  // namespace

Bootstrapper::Batch Bootstrapper::match_many(const Batch &ct, size_t level, double scale)
{
    Batch t(ct.size());
    for (size_t i = 0; i < ct.size(); i++) {
        const size_t l = ct[i].coeff_modulus_size();
        if (l <= level) throw std::logic_error("Bootstrapper: operand below the target level");
        // times 1 encoded at the compensating scale, read in place at level + 1
        const double sigma = scale * prime(level + 1) / ct[i].scale();
        ck_.evaluator.linear_combination({&ct[i]}, {1.0}, {sigma}, chain_of(level + 1), t[i]);
    }
    ck_.evaluator.rescale_to_next_many(ptrs(t), t);
    for (auto &x : t) x.set_scale(scale);
    return t;
}

Bootstrapper::Batch Bootstrapper::mul_exact_many(const Batch &a, const Batch &b)
{
    Batch xs(a.size()), ys(a.size());
    std::vector<const Ciphertext *> x, y;
    for (size_t i = 0; i < a.size(); i++) {
        const size_t la = a[i].coeff_modulus_size(), lb = b[i].coeff_modulus_size();
        const size_t l = std::min(la, lb);
        const Ciphertext *pa = &a[i], *pb = &b[i];
        if (la > l) {
            ck_.evaluator.mod_switch_to(a[i], chain_of(l), xs[i]);
            pa = &xs[i];
        }
        if (lb > l) {
            ck_.evaluator.mod_switch_to(b[i], chain_of(l), ys[i]);
            pb = &ys[i];
        }
        x.push_back(pa);
        y.push_back(&a == &b ? pa : pb);
    }
    Batch r;
    ck_.evaluator.multiply_relin_rescale_many(x, y, *ck_.relin_keys, r, true);
    return r;
}

const Bootstrapper::Batch &Bootstrapper::basis_many(int k)
{
    auto it = basis_many_.find(k);
    if (it != basis_many_.end()) return it->second;
    int a = 1;
    while (2 * a < k) a *= 2;
    const int b = k - a;
    const Batch &Ta = basis_many(a);
    const Batch &Tb = basis_many(b);
    Batch prod = (a == b) ? mul_exact_many(Ta, Ta) : mul_exact_many(Ta, Tb);
    const size_t l = prod[0].coeff_modulus_size();
    if (a == b) {
        for (auto &x : prod) ck_.evaluator.add_inplace(x, x);  // 2 T_a^2 - 1
        Plaintext one(*ck_.context);
        ck_.encoder.encode(1.0, chain_of(l), prod[0].scale(), one);
        for (auto &x : prod) ck_.evaluator.sub_plain_inplace(x, one);
    } else {
        // 2 T_a T_b - T_(a-b) in one pass (constants 2 and -1 at scale 1)
        Batch Td = match_many(basis_many(a - b), l, prod[0].scale());
        for (size_t i = 0; i < prod.size(); i++) {
            Ciphertext r;
            ck_.evaluator.linear_combination({&prod[i], &Td[i]}, {2.0, -1.0}, {1.0, 1.0}, chain_of(l), r);
            prod[i] = std::move(r);
        }
    }
    return basis_many_.emplace(k, std::move(prod)).first->second;
}

Bootstrapper::Batch Bootstrapper::cheb_eval_many(const std::vector<double> &c, size_t level, double scale)
{
    const int d = trimmed_degree(c);
    const size_t B = basis_many_.at(1).size();
    if (d < 8) {
        const size_t lam = level + 1;
        const double q = prime(lam);
        Batch acc(B);
        // sum_k c_k T_k at level lam: the basis read in place, the constants
        // encoded at scale * q / T_k's scale, every term in one pass per input
        std::vector<const Batch *> bks;
        std::vector<double> vals, scs;
        for (int k = 1; k <= d; k++) {
            if (c[k] == 0.0) continue;
            const Batch &bk = basis_many(k);
            if (bk[0].coeff_modulus_size() < lam) throw std::logic_error("Bootstrapper: basis below the leaf level");
            bks.push_back(&bk);
            vals.push_back(c[k]);
            scs.push_back(scale * q / bk[0].scale());
        }
        const bool have = !bks.empty();
        for (size_t i = 0; have && i < B; i++) {
            std::vector<const Ciphertext *> xs;
            for (const Batch *bk : bks) xs.push_back(&(*bk)[i]);
            ck_.evaluator.linear_combination(xs, vals, scs, chain_of(lam), acc[i]);
        }
        if (!have) {  // constant polynomial: 0 * T_1 carries the level and scale
            const Batch &b1 = basis_many(1);
            Plaintext z(*ck_.context);
            for (size_t i = 0; i < B; i++) {
                ck_.evaluator.mod_switch_to(b1[i], chain_of(lam), acc[i]);
                if (i == 0) ck_.encoder.encode(0.0, chain_of(lam), scale * q / acc[i].scale(), z);
                ck_.evaluator.multiply_plain_inplace(acc[i], z);
            }
        }
        ck_.evaluator.rescale_to_next_many(ptrs(acc), acc);
        for (auto &x : acc) x.set_scale(scale);
        if (c[0] != 0.0) {
            Plaintext c0(*ck_.context);
            ck_.encoder.encode(c[0], chain_of(level), scale, c0);
            for (auto &x : acc) ck_.evaluator.add_plain_inplace(x, c0);
        }
        return acc;
    }
    int P = 1;
    while (2 * P <= d) P *= 2;
    std::vector<double> q, r;
    cheb_split(c, d, P, q, r);
    const Batch &TP = basis_many(P);
    const size_t lam = level + 1;
    if (TP[0].coeff_modulus_size() < lam) throw std::logic_error("Bootstrapper: giant step below the product level");
    Batch Q = cheb_eval_many(q, lam, scale * prime(lam) / TP[0].scale());
    Batch Tlow(B);
    std::vector<const Ciphertext *> T;
    for (size_t i = 0; i < B; i++) {
        if (TP[i].coeff_modulus_size() > lam) {
            ck_.evaluator.mod_switch_to(TP[i], chain_of(lam), Tlow[i]);
            T.push_back(&Tlow[i]);
        } else {
            T.push_back(&TP[i]);
        }
    }
    Batch prod;
    ck_.evaluator.multiply_relin_rescale_many(ptrs(Q), T, *ck_.relin_keys, prod, true);
    Batch R = cheb_eval_many(r, level, scale);
    for (size_t i = 0; i < B; i++) {
        prod[i].set_scale(scale);
        R[i].set_scale(scale);
        ck_.evaluator.add_inplace(prod[i], R[i]);
    }
    return prod;
}

void Bootstrapper::eval_mod_many(const std::vector<const Ciphertext *> &u, std::vector<Ciphertext> &out)
{
    if (u.empty()) throw std::invalid_argument("Bootstrapper: empty EvalMod batch");
    const size_t l0 = u[0]->coeff_modulus_size();
    const double s0 = u[0]->scale();
    for (const Ciphertext *x : u)
        if (x->coeff_modulus_size() != l0 || x->scale() != s0)
            throw std::invalid_argument("Bootstrapper: EvalMod batch needs one level and scale");
    const int cd_ = chebyshev_depth(cheb_);
    if (l0 <= (size_t)cd_ + (size_t)p_.double_angles)
        throw std::invalid_argument("Bootstrapper: not enough levels for EvalMod");
    basis_many_.clear();
    {
        Batch b1;
        for (const Ciphertext *x : u) b1.push_back(*x);
        basis_many_.emplace(1, std::move(b1));
    }
    Batch y = cheb_eval_many(cheb_, l0 - cd_, s0);
    basis_many_.clear();
    for (int i = 0; i < p_.double_angles; i++) {
        const std::vector<const Ciphertext *> p = ptrs(y);
        ck_.evaluator.multiply_relin_rescale_many(p, p, *ck_.relin_keys, y, true);
        for (auto &x : y) ck_.evaluator.add_inplace(x, x);
        Plaintext one(*ck_.context);
        ck_.encoder.encode(1.0, y[0].chain_index(), y[0].scale(), one);
        for (auto &x : y) ck_.evaluator.sub_plain_inplace(x, one);
    }
    out = std::move(y);
}

Ciphertext Bootstrapper::apply_level(const LinearLevel &lv, int tag, const Ciphertext &x, double out_scale,
                                     double value_factor)
{
    // bsgs_linear_transform (Bootstrapper.cu:1855-1919): baby rotations of x,
    // per giant step a plaintext-weighted sum and one rotation, one rescale.
    const size_t l = x.coeff_modulus_size();
    const double sigma = out_scale * prime(l) / x.scale();
    // baby steps share one ModUp of x (hoisted rotations)
    std::vector<int> bi, bsteps;
    for (const auto &d : lv.diag) {
        const int i = d.first.second;
        if (std::find(bi.begin(), bi.end(), i) != bi.end()) continue;
        bi.push_back(i);
        bsteps.push_back(i * lv.stride);
    }
    std::vector<Ciphertext> rots;
    ck_.evaluator.rotate_vector_hoisted(x, bsteps, *ck_.galois_keys, rots);
    std::map<int, Ciphertext> baby;
    for (size_t k = 0; k < bi.size(); k++) baby.emplace(bi[k], std::move(rots[k]));
    // giant steps: sum_i D~_(j,i) * rot(x, i stride) for every j in one pass
    // over the babies, then their rotations summed with a single ModDown
    std::vector<const Ciphertext *> terms;
    std::map<int, size_t> column;
    for (auto &kv : baby) {
        column[kv.first] = terms.size();
        terms.push_back(&kv.second);
    }
    std::vector<const Plaintext *> weights;
    std::vector<int> gsteps;
    auto it = lv.diag.begin();
    while (it != lv.diag.end()) {
        const int j = it->first.first;
        std::vector<const Plaintext *> row(terms.size(), nullptr);
        for (; it != lv.diag.end() && it->first.first == j; ++it) {
            const int i = it->first.second;
            auto key = std::make_tuple(tag, j, i, l, sigma * value_factor);
            auto pit = ptcache_.find(key);
            if (pit == ptcache_.end()) {
                Plaintext pt(*ck_.context);
                ck_.encoder.encode(it->second, chain_of(l), sigma * value_factor, pt);
                pit = ptcache_.emplace(key, std::move(pt)).first;
            }
            row[column.at(i)] = &pit->second;
        }
        weights.insert(weights.end(), row.begin(), row.end());
        gsteps.push_back(j * lv.g * lv.stride);
    }
    std::vector<Ciphertext> accs;
    ck_.evaluator.multiply_plain_accumulate_multi(terms, weights, accs);
    std::vector<const Ciphertext *> ap;
    for (const auto &a : accs) ap.push_back(&a);
    Ciphertext result;
    ck_.evaluator.rotate_sum(ap, gsteps, *ck_.galois_keys, result);
    ck_.evaluator.rescale_to_next_inplace(result);
    result.set_scale(out_scale);
    return result;
}

void Bootstrapper::coeff_to_slot(const Ciphertext &in, Ciphertext &u_re, Ciphertext &u_im)
{
    // coefftoslot_3 / sflinv_3 (Bootstrapper.cu:2435-2458, 2581-2586)
    Ciphertext x = in;
    for (size_t k = 0; k < plan_.cts.size(); k++) x = apply_level(plan_.cts[k], (int)k, x, x.scale(), 1.0);
    Ciphertext xc;
    ck_.evaluator.complex_conjugate(x, *ck_.galois_keys, xc);
    if (plan_.logn < plan_.logNh) {
        ck_.evaluator.add(x, xc, u_re);  // lower half: Re, upper half: Im (mask folded above)
        u_im = Ciphertext();
    } else {
        // dense packing (coefftoslot_full_3): Re = x + conj x, Im = -i (x - conj x)
        ck_.evaluator.add(x, xc, u_re);
        Ciphertext d;
        ck_.evaluator.sub(x, xc, d);
        const int64_t N = (int64_t)ck_.context->poly_degree();
        ck_.evaluator.multiply_power_of_x(d, 3 * N / 2, u_im);  // X^(3N/2): every slot times -i
    }
}

void Bootstrapper::slot_to_coeff(const Ciphertext &f_re, const Ciphertext *f_im, double input_scale, Ciphertext &out,
                                 double factor)
{
    // slottocoeff_3 / sfl_3 (Bootstrapper.cu:2275-2314, 2589-2594)
    Ciphertext x = f_re;
    if (f_im) {  // dense packing: f_re + i f_im, X^(N/2) multiplies every slot by i
        Ciphertext t;
        const int64_t N = (int64_t)ck_.context->poly_degree();
        ck_.evaluator.multiply_power_of_x(*f_im, N / 2, t);
        t.set_scale(x.scale());
        ck_.evaluator.add_inplace(x, t);
    }
    const double final_scale = p_.final_scale > 0 ? p_.final_scale : input_scale;
    const double q0 = prime(1);
    for (size_t k = 0; k < plan_.stc.size(); k++) {
        const bool last = k + 1 == plan_.stc.size();
        // the last level maps the EvalMod output (m / q_0) back to the message at final_scale
        x = apply_level(plan_.stc[k], 100 + (int)k, x, last ? final_scale : x.scale(),
                        last ? q0 / input_scale * factor : 1.0);
    }
    if (plan_.logn < plan_.logNh) {
        std::vector<Ciphertext> r;
        ck_.evaluator.rotate_vector_hoisted(x, {1 << plan_.logn}, *ck_.galois_keys, r);
        ck_.evaluator.add_inplace(x, r[0]);
    }
    x.set_scale(final_scale);
    out = std::move(x);
}

std::vector<Ciphertext> Bootstrapper::apply_level_many(const LinearLevel &lv, int tag,
                                                       const std::vector<const Ciphertext *> &xs,
                                                       const std::vector<double> &out_scale,
                                                       const std::vector<double> &value_factor)
{
    // apply_level over a batch of inputs of one level: the baby rotations of
    // all inputs hoisted together, each input's giant-step multiply-accumulates
    // in one pass, every input's giant sum with one ModDown (the ModDowns of
    // the batch together), one batched rescale -- word for word apply_level
    // of each input
    const size_t B = xs.size(), l = xs[0]->coeff_modulus_size();
    for (const Ciphertext *x : xs)
        if (x->coeff_modulus_size() != l) throw std::invalid_argument("Bootstrapper: batch needs one level");
    std::vector<int> bi, bsteps;
    for (const auto &d : lv.diag) {
        const int i = d.first.second;
        if (std::find(bi.begin(), bi.end(), i) != bi.end()) continue;
        bi.push_back(i);
        bsteps.push_back(i * lv.stride);
    }
    std::vector<std::vector<Ciphertext>> rots;
    ck_.evaluator.rotate_vector_hoisted_many(xs, bsteps, *ck_.galois_keys, rots);
    std::vector<std::vector<Ciphertext>> accs(B);
    std::vector<std::vector<const Ciphertext *>> groups(B);
    std::vector<std::vector<int>> gsteps(B);
    for (size_t b = 0; b < B; b++) {
        const double sigma = out_scale[b] * prime(l) / xs[b]->scale();
        std::map<int, Ciphertext> baby;
        for (size_t k = 0; k < bi.size(); k++) baby.emplace(bi[k], std::move(rots[b][k]));
        std::vector<const Ciphertext *> terms;
        std::map<int, size_t> column;
        for (auto &kv : baby) {
            column[kv.first] = terms.size();
            terms.push_back(&kv.second);
        }
        std::vector<const Plaintext *> weights;
        auto it = lv.diag.begin();
        while (it != lv.diag.end()) {
            const int j = it->first.first;
            std::vector<const Plaintext *> row(terms.size(), nullptr);
            for (; it != lv.diag.end() && it->first.first == j; ++it) {
                const int i = it->first.second;
                auto key = std::make_tuple(tag, j, i, l, sigma * value_factor[b]);
                auto pit = ptcache_.find(key);
                if (pit == ptcache_.end()) {
                    Plaintext pt(*ck_.context);
                    ck_.encoder.encode(it->second, chain_of(l), sigma * value_factor[b], pt);
                    pit = ptcache_.emplace(key, std::move(pt)).first;
                }
                row[column.at(i)] = &pit->second;
            }
            weights.insert(weights.end(), row.begin(), row.end());
            gsteps[b].push_back(j * lv.g * lv.stride);
        }
        ck_.evaluator.multiply_plain_accumulate_multi(terms, weights, accs[b]);
        for (const auto &a : accs[b]) groups[b].push_back(&a);
    }
    rots.clear();
    std::vector<Ciphertext> sums;
    ck_.evaluator.rotate_sum_many(groups, gsteps, *ck_.galois_keys, sums);
    accs.clear();
    std::vector<Ciphertext> res;
    ck_.evaluator.rescale_to_next_many(ptrs(sums), res);
    for (size_t b = 0; b < B; b++) res[b].set_scale(out_scale[b]);
    return res;
}

void Bootstrapper::bootstrap_many(const std::vector<const Ciphertext *> &in, std::vector<Ciphertext> &out,
                                  const std::vector<double> &factors)
{
    bootstrap_mixed(std::vector<Bootstrapper *>(in.size(), this), in, out, factors);
}

void Bootstrapper::bootstrap_mixed(const std::vector<Bootstrapper *> &bts, const std::vector<const Ciphertext *> &in,
                                   std::vector<Ciphertext> &out, const std::vector<double> &factors)
{
    const size_t B = in.size();
    if (bts.size() != B || factors.size() != B)
        throw std::invalid_argument("Bootstrapper: one bootstrapper and one factor per input");
    if (!B) {
        out.clear();
        return;
    }
    for (size_t i = 0; i < B; i++) {
        if (!bts[i]) throw std::invalid_argument("Bootstrapper: null bootstrapper");
        if (!in[i] || !in[i]->get()) throw std::invalid_argument("Bootstrapper: empty ciphertext");
    }
    Bootstrapper &lead = *bts[0];
    if (lead.ck_.evaluator.per_op()) {
        // per-op mode: one bootstrap per input
        std::vector<Ciphertext> res(B);
        for (size_t i = 0; i < B; i++) bts[i]->bootstrap(*in[i], res[i], factors[i]);
        out = std::move(res);
        return;
    }
    // the inputs of each bootstrapper (slot count) in order of first appearance
    std::vector<Bootstrapper *> plans;
    std::vector<std::vector<size_t>> members;
    for (size_t i = 0; i < B; i++) {
        Bootstrapper *b = bts[i];
        if (b->ck_.context->get() != lead.ck_.context->get())
            throw std::invalid_argument("Bootstrapper: inputs on different contexts");
        if (b->p_.K != lead.p_.K || b->p_.cheb_degree != lead.p_.cheb_degree ||
            b->p_.double_angles != lead.p_.double_angles || b->p_.log_width != lead.p_.log_width)
            throw std::invalid_argument("Bootstrapper: one EvalMod needs the same K, degree, double angles and width");
        auto it = std::find(plans.begin(), plans.end(), b);
        if (it == plans.end()) {
            plans.push_back(b);
            members.emplace_back();
            it = plans.end() - 1;
        }
        members[it - plans.begin()].push_back(i);
    }
    // the linear transforms run over chunks of at most kLevelBatch inputs of one
    // bootstrapper (bounding the baby rotations and giant sums held at once)
    constexpr size_t kLevelBatch = 16;
    {
        std::vector<Bootstrapper *> p2;
        std::vector<std::vector<size_t>> m2;
        for (size_t g = 0; g < plans.size(); g++)
            for (size_t k0 = 0; k0 < members[g].size(); k0 += kLevelBatch) {
                p2.push_back(plans[g]);
                m2.emplace_back(members[g].begin() + k0,
                                members[g].begin() + std::min(members[g].size(), k0 + kLevelBatch));
            }
        plans = std::move(p2);
        members = std::move(m2);
    }
    const int64_t N = (int64_t)lead.ck_.context->poly_degree();
    std::vector<Ciphertext> u_re(B), u_im(B);
    for (size_t g = 0; g < plans.size(); g++) {
        Bootstrapper &bt = *plans[g];
        const std::vector<size_t> &idx = members[g];
        const Evaluator &ev = bt.ck_.evaluator;
        std::vector<Ciphertext> x(idx.size());
        for (size_t k = 0; k < idx.size(); k++) bt.mod_raise(*in[idx[k]], x[k]);
        if (bt.plan_.logn < bt.plan_.logNh)  // SubSum (Bootstrapper.cu:3004-3008), batched per step
            for (int i = bt.plan_.logn; i < bt.plan_.logNh; i++) {
                std::vector<std::vector<Ciphertext>> r;
                ev.rotate_vector_hoisted_many(ptrs(x), {1 << i}, *bt.ck_.galois_keys, r);
                for (size_t k = 0; k < idx.size(); k++) ev.add_inplace(x[k], r[k][0]);
            }
        for (size_t lv = 0; lv < bt.plan_.cts.size(); lv++) {
            std::vector<double> sc, vf(idx.size(), 1.0);
            for (const auto &c : x) sc.push_back(c.scale());
            x = bt.apply_level_many(bt.plan_.cts[lv], (int)lv, ptrs(x), sc, vf);
        }
        for (size_t k = 0; k < idx.size(); k++) {
            Ciphertext xc;
            ev.complex_conjugate(x[k], *bt.ck_.galois_keys, xc);
            const size_t i = idx[k];
            ev.add(x[k], xc, u_re[i]);
            if (bt.plan_.logn == bt.plan_.logNh) {
                // dense packing: Im = -i (x - conj x), X^(3N/2) multiplies every slot by -i
                Ciphertext d;
                ev.sub(x[k], xc, d);
                ev.multiply_power_of_x(d, 3 * N / 2, u_im[i]);
            }
        }
    }
    // one EvalMod over every half of every input
    std::vector<const Ciphertext *> u;
    for (size_t i = 0; i < B; i++) {
        u.push_back(&u_re[i]);
        if (u_im[i].get()) u.push_back(&u_im[i]);
    }
    std::vector<Ciphertext> f;
    lead.eval_mod_many(u, f);
    u_re.clear();
    std::vector<Ciphertext> fre(B), fim(B);
    {
        size_t k = 0;
        for (size_t i = 0; i < B; i++) {
            fre[i] = std::move(f[k++]);
            if (u_im[i].get()) fim[i] = std::move(f[k++]);
        }
    }
    u_im.clear();
    std::vector<Ciphertext> res(B);
    for (size_t g = 0; g < plans.size(); g++) {
        Bootstrapper &bt = *plans[g];
        const std::vector<size_t> &idx = members[g];
        const Evaluator &ev = bt.ck_.evaluator;
        std::vector<Ciphertext> x(idx.size());
        std::vector<double> final_scale(idx.size()), last_vf(idx.size());
        const double q0 = bt.prime(1);
        for (size_t k = 0; k < idx.size(); k++) {
            const size_t i = idx[k];
            x[k] = std::move(fre[i]);
            if (fim[i].get()) {
                Ciphertext t;
                ev.multiply_power_of_x(fim[i], N / 2, t);
                t.set_scale(x[k].scale());
                ev.add_inplace(x[k], t);
            }
            const double input_scale = in[i]->scale();
            final_scale[k] = bt.p_.final_scale > 0 ? bt.p_.final_scale : input_scale;
            last_vf[k] = q0 / input_scale * factors[i];
        }
        for (size_t lv = 0; lv < bt.plan_.stc.size(); lv++) {
            const bool last = lv + 1 == bt.plan_.stc.size();
            std::vector<double> sc(idx.size()), vf(idx.size(), 1.0);
            for (size_t k = 0; k < idx.size(); k++) sc[k] = last ? final_scale[k] : x[k].scale();
            if (last) vf = last_vf;
            x = bt.apply_level_many(bt.plan_.stc[lv], 100 + (int)lv, ptrs(x), sc, vf);
        }
        if (bt.plan_.logn < bt.plan_.logNh) {
            std::vector<std::vector<Ciphertext>> r;
            ev.rotate_vector_hoisted_many(ptrs(x), {1 << bt.plan_.logn}, *bt.ck_.galois_keys, r);
            for (size_t k = 0; k < idx.size(); k++) ev.add_inplace(x[k], r[k][0]);
        }
        for (size_t k = 0; k < idx.size(); k++) {
            x[k].set_scale(final_scale[k]);
            res[idx[k]] = std::move(x[k]);
        }
    }
    out = std::move(res);
}

void Bootstrapper::bootstrap(const Ciphertext &in, Ciphertext &out, double factor)
{
    // bootstrap_3 (Bootstrapper.cu:3358-3370) -> bootstrap_sparse_3 / bootstrap_full_3
    if (!in.get()) throw std::invalid_argument("Bootstrapper: empty ciphertext");
    const double input_scale = in.scale();
    Ciphertext ct;
    mod_raise(in, ct);
    if (plan_.logn < plan_.logNh) sub_sum(ct);
    Ciphertext u_re, u_im;
    coeff_to_slot(ct, u_re, u_im);
    Ciphertext f_re, f_im;
    eval_mod(u_re, f_re);
    if (u_im.get()) eval_mod(u_im, f_im);
    slot_to_coeff(f_re, u_im.get() ? &f_im : nullptr, input_scale, out, factor);
}

}  // namespace nexus

namespace nexus {

// src/argmax.cpp:9-64
void ArgmaxEvaluator::argmax(Ciphertext &x, Ciphertext &res, int len)
{
    if (len < 1 || (len & (len - 1))) throw std::invalid_argument("argmax: len must be a power of two");
    const Evaluator &ev = ckks_->evaluator;
    int log_step = 0;
    while ((1 << log_step) < len) log_step++;

    // [a_0 .. a_len-1, 0 ..] -> [a_0 .. a_len-1, a_0 .. a_len-1, 0 ..]
    Ciphertext tmp;
    ev.rotate_vector(x, -len, *ckks_->galois_keys, tmp);
    ev.add_inplace(x, tmp);

    Ciphertext a = x;
    for (int i = 0; i < log_step; ++i) {
        Ciphertext b, a_plus_b, a_minus_b, a_minus_b_sgn;
        ev.rotate_vector(a, 1 << i, *ckks_->galois_keys, b);
        ev.add(a, b, a_plus_b);
        ev.sub(a, b, a_minus_b);
        Ciphertext sign = ckks_->sgn_eval(a_minus_b, 2, 2);

        // (a - b) sgn(a - b) / 2 (the 1/2 is in sgn_eval's default factor)
        ev.mod_switch_to_inplace(a_minus_b, sign.params_id());
        ev.multiply(a_minus_b, sign, a_minus_b_sgn);
        ev.relinearize_inplace(a_minus_b_sgn, *ckks_->relin_keys);
        ev.rescale_to_next_inplace(a_minus_b_sgn);

        // (a + b) / 2
        Plaintext half(*ckks_->context);
        ckks_->encoder.encode(0.5, a_plus_b.params_id(), a_plus_b.scale(), half);
        ev.multiply_plain_inplace(a_plus_b, half);
        ev.rescale_to_next_inplace(a_plus_b);

        // a = max(a, b)
        a_plus_b.set_scale(a_minus_b_sgn.scale());
        ev.mod_switch_to_inplace(a_plus_b, a_minus_b_sgn.params_id());
        ev.add(a_plus_b, a_minus_b_sgn, a);

        bootstrap(a);
    }

    x.set_scale(a.scale());
    ev.mod_switch_to_inplace(x, a.params_id());
    ev.sub(x, a, res);

    res = ckks_->sgn_eval(res, 2, 2, 1.0);

    Plaintext one(*ckks_->context);
    ckks_->encoder.encode(1.0, res.params_id(), res.scale(), one);
    ev.add_plain_inplace(res, one);
}

// src/argmax.cpp:66-81: to the lowest level, bootstrap at the current scale
void ArgmaxEvaluator::bootstrap(Ciphertext &x)
{
    const size_t L = ckks_->context->data_primes();
    if (x.coeff_modulus_size() > 1) ckks_->evaluator.mod_switch_to_inplace(x, L);
    Ciphertext x0 = x;
    bt_->set_final_scale(x.scale());
    bt_->bootstrap(x0, x);
}

}  // namespace nexus
