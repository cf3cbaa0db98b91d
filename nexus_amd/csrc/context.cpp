#include "context.h"

#include <algorithm>
#include <cmath>

#include "numth.h"

namespace nx {

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        int code = e == hipErrorOutOfMemory ? kNoMemory : kDeviceError;
        fail(code, std::string("HIP error: ") + hipGetErrorString(e) + " in " + what);
    }
}

// ---------------------------------------------------------------- memory
DeviceCache::~DeviceCache() { trim(); }

size_t DeviceCache::size_class(size_t bytes)
{
    // eight classes per power of two above 64 KiB (at most 12.5% slack), so
    // that buffers of nearby sizes are reused instead of cached side by side
    if (bytes <= (size_t)1 << 16) return (bytes + 255) & ~(size_t)255;
    int e = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = (size_t)1 << (e - 3);
    return (bytes + step - 1) & ~(step - 1);
}

void *DeviceCache::try_get(size_t bytes)
{
    if (!bytes) return nullptr;
    bytes = size_class(bytes);
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.find(bytes);
        if (it != free_.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            cached_ -= bytes;
            in_use_ += bytes;
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> g(mu_);
    in_use_ += bytes;
    return p;
}

void *DeviceCache::get(size_t bytes)
{
    if (!bytes) return nullptr;
    void *p = try_get(bytes);
    if (p) return p;
    // release this cache's blocks and retry once
    trim();
    p = try_get(bytes);
    if (!p) NX_HIP(hipErrorOutOfMemory);
    return p;
}

void DeviceCache::put(void *p, size_t bytes)
{
    if (!p) return;
    bytes = size_class(bytes);
    std::lock_guard<std::mutex> g(mu_);
    free_[bytes].push_back(p);
    in_use_ -= bytes;
    cached_ += bytes;
}

void DeviceCache::trim()
{
    std::lock_guard<std::mutex> g(mu_);
    if (!free_.empty()) (void)hipDeviceSynchronize();
    for (auto &kv : free_)
        for (void *p : kv.second) (void)hipFree(p);
    free_.clear();
    cached_ = 0;
}

DevBuf::DevBuf(Context *ctx, size_t words) { reset(ctx, words); }

DevBuf &DevBuf::operator=(DevBuf &&o) noexcept
{
    if (this != &o) {
        release();
        cache_ = std::move(o.cache_);
        p_ = o.p_;
        words_ = o.words_;
        o.p_ = nullptr;
        o.words_ = 0;
    }
    return *this;
}

void DevBuf::reset(Context *ctx, size_t words)
{
    const std::shared_ptr<DeviceCache> &cur = ctx->current_cache();
    if (p_ && cache_ == cur && words == words_) return;
    release();
    cache_ = cur;
    words_ = words;
    if (!words) return;
    p_ = (uint64_t *)cache_->try_get(words * 8);
    if (!p_) {
        // out of device memory: release the cached blocks of every lane of the
        // context (a device-wide synchronisation) and retry
        ctx->trim_caches();
        p_ = (uint64_t *)cache_->get(words * 8);
    }
}

void DevBuf::release()
{
    if (p_ && cache_) cache_->put(p_, words_ * 8);
    p_ = nullptr;
    words_ = 0;
    cache_.reset();
}

// ---------------------------------------------------------------- lanes
namespace {
thread_local const Context *tl_ctx = nullptr;
thread_local Lane *tl_lane = nullptr;
}  // namespace

StreamRef::operator hipStream_t() const
{
    return (tl_ctx == c_ && tl_lane) ? tl_lane->stream : def_;
}

Lane *Context::lane(size_t index)
{
    require(index >= 1 && index <= 64, "lane index out of range", kOutOfRange);
    std::lock_guard<std::mutex> g(lane_mu_);
    if (lanes_.size() < index) lanes_.resize(index);
    std::unique_ptr<Lane> &l = lanes_[index - 1];
    if (!l) {
        l.reset(new Lane());
        NX_HIP(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
        l->cache = std::make_shared<DeviceCache>();
    }
    return l.get();
}

void Context::bind_lane(size_t index)
{
    if (index == 0) {
        if (tl_ctx == this) {
            tl_ctx = nullptr;
            tl_lane = nullptr;
        }
        return;
    }
    Lane *l = lane(index);
    tl_ctx = this;
    tl_lane = l;
}

Lane *Context::bound_lane() const { return tl_ctx == this ? tl_lane : nullptr; }

void Context::trim_caches()
{
    cache->trim();
    std::lock_guard<std::mutex> g(lane_mu_);
    for (auto &l : lanes_)
        if (l) l->cache->trim();
}

const std::shared_ptr<DeviceCache> &Context::current_cache() const
{
    Lane *l = bound_lane();
    return l ? l->cache : cache;
}

// ---------------------------------------------------------------- context
Context::Context(int logn_, const std::vector<int> &bit_sizes, uint32_t hw, size_t sparse, size_t special)
    : logn(logn_), n(size_t(1) << logn_), alpha(special), hamming_weight(hw), sparse_slots(sparse)
{
    require(logn >= 1 && logn <= 17, "poly_modulus_degree is invalid");
    require(alpha >= 1 && bit_sizes.size() >= alpha + 1 && bit_sizes.size() <= (size_t)kMaxPrimes,
            "coeff modulus needs at least one data prime and the special primes");
    require(alpha <= 32, "at most 32 special primes");
    moduli = coeff_modulus_create(n, bit_sizes);
    L = moduli.size() - alpha;
    dnum = (L + alpha - 1) / alpha;
    if (sparse_slots == 0) sparse_slots = n / 2;
    require(sparse_slots <= n / 2 && (sparse_slots & (sparse_slots - 1)) == 0, "sparse_slots is invalid");
    for (uint64_t q : moduli) mods.push_back(make_modconst(q));

    int dev = 0;
    NX_HIP(hipGetDevice(&dev));
    NX_HIP(hipStreamCreateWithFlags(&stream.def_, hipStreamNonBlocking));

    const size_t P = moduli.size();
    std::vector<ulonglong2> fwd(P * n), inv(P * n), ninv(P);
    std::vector<double> fwd_d(P * n), inv_d(P * n), ninv_d(P);  // exact for primes below 2^53
    std::vector<double2> qd(P);
    fwd_host.resize(P);
    inv_host.resize(P);
    for (size_t i = 0; i < P; i++) {
        const uint64_t q = moduli[i];
        ntt_root_tables(logn, q, fwd_host[i], inv_host[i]);
        for (size_t k = 0; k < n; k++) {
            fwd[i * n + k] = make_ulonglong2(fwd_host[i][k], shoup_quotient(fwd_host[i][k], q));
            inv[i * n + k] = make_ulonglong2(inv_host[i][k], shoup_quotient(inv_host[i][k], q));
            fwd_d[i * n + k] = (double)fwd_host[i][k];
            inv_d[i * n + k] = (double)inv_host[i][k];
        }
        uint64_t ni = host_inv_mod(n % q, q);
        ninv[i] = make_ulonglong2(ni, shoup_quotient(ni, q));
        ninv_d[i] = (double)ni;
        qd[i] = make_double2((double)q, 1.0 / (double)q);
    }
    // rescale constants: row `l` (limbs before the rescale), column i < l-1
    std::vector<ulonglong2> resc(P * P, make_ulonglong2(0, 0));
    for (size_t l = 2; l <= L; l++) {
        const uint64_t ql = moduli[l - 1];
        for (size_t i = 0; i + 1 < l; i++) {
            const uint64_t qi = moduli[i];
            uint64_t v = host_inv_mod(ql % qi, qi);
            resc[l * P + i] = make_ulonglong2(v, shoup_quotient(v, qi));
        }
    }
    std::vector<ulonglong2> pinv(P, make_ulonglong2(0, 0));
    for (size_t i = 0; i < L; i++) {
        const uint64_t qi = moduli[i];
        uint64_t pm = 1;
        for (size_t k = 0; k < alpha; k++) pm = host_mul_mod(pm, moduli[L + k] % qi, qi);
        uint64_t v = host_inv_mod(pm, qi);
        pinv[i] = make_ulonglong2(v, shoup_quotient(v, qi));
    }
    if (alpha > 1) build_hybrid_tables();

    // the allocation-free entry points' scratch array: its address on this device
    noalloc_init_current_device();
    NX_HIP(hipMalloc(&d_fwd, P * n * sizeof(ulonglong2)));
    NX_HIP(hipMalloc(&d_inv, P * n * sizeof(ulonglong2)));
    NX_HIP(hipMalloc(&d_ninv, P * sizeof(ulonglong2)));
    NX_HIP(hipMalloc(&d_fwd_d, P * n * sizeof(double)));
    NX_HIP(hipMalloc(&d_inv_d, P * n * sizeof(double)));
    NX_HIP(hipMalloc(&d_ninv_d, P * sizeof(double)));
    NX_HIP(hipMalloc(&d_mod, P * sizeof(ModConst)));
    NX_HIP(hipMalloc(&d_qd, P * sizeof(double2)));
    NX_HIP(hipMalloc(&d_rescale_inv, P * P * sizeof(ulonglong2)));
    NX_HIP(hipMalloc(&d_p_inv, P * sizeof(ulonglong2)));
    NX_HIP(hipMemcpy(d_fwd, fwd.data(), P * n * sizeof(ulonglong2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_inv, inv.data(), P * n * sizeof(ulonglong2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_ninv, ninv.data(), P * sizeof(ulonglong2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_fwd_d, fwd_d.data(), P * n * sizeof(double), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_inv_d, inv_d.data(), P * n * sizeof(double), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_ninv_d, ninv_d.data(), P * sizeof(double), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_mod, mods.data(), P * sizeof(ModConst), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_qd, qd.data(), P * sizeof(double2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_rescale_inv, resc.data(), P * P * sizeof(ulonglong2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_p_inv, pinv.data(), P * sizeof(ulonglong2), hipMemcpyHostToDevice));

    tables.logn = logn;
    tables.num_primes = (int)P;
    tables.max_bits = 0;
    for (uint64_t q : moduli) {
        int b = 0;
        while (b < 64 && (q >> b)) b++;
        tables.max_bits = std::max(tables.max_bits, b);
    }
    tables.fwd = d_fwd;
    tables.inv = d_inv;
    tables.ninv = d_ninv;
    tables.fwd_d = d_fwd_d;
    tables.inv_d = d_inv_d;
    tables.ninv_d = d_ninv_d;
    tables.modup_conv_d = d_modup_conv_d;
    tables.mod_host = moduli.data();
    tables.mod = d_mod;
    tables.qd = d_qd;
    upload_encoder_tables(*this);
}

// Base-conversion constants of hybrid key switching (fast basis extension:
// x -> sum_i [x_i Dhat_i^-1]_{q_i} Dhat_i mod t, Dhat_i = prod_{i' != i} q_i').
// Double-precision copy of a ModDown conversion table over A input primes
// (rows k < A: factors f_ki, row A: the base product mod q_i): rows k < A as
// doubles, row A as q_i - (row A).  Used for output primes below 2^51 when every
// input prime is below 2^51.
static std::vector<double> moddown_conv_doubles(const uint64_t *conv, size_t A, size_t L,
                                                const std::vector<uint64_t> &moduli)
{
    std::vector<double> d((A + 1) * L, 0.0);
    for (size_t i = 0; i < L; i++) {
        for (size_t k = 0; k < A; k++) d[k * L + i] = (double)conv[k * L + i];
        d[A * L + i] = (double)(moduli[i] - conv[A * L + i] % moduli[i]);
    }
    return d;
}

void Context::build_hybrid_tables()
{
    const size_t K = key_limbs();
    std::vector<ulonglong2> uw(dnum * alpha * alpha, make_ulonglong2(0, 0));
    std::vector<uint64_t> uc(dnum * alpha * alpha * K, 0);
    for (size_t j = 0; j < dnum; j++)
        for (size_t s = 1; s <= alpha; s++) {
            if (j * alpha + s > L) break;
            for (size_t i = 0; i < s; i++) {
                const size_t g = j * alpha + i, idx = (j * alpha + s - 1) * alpha + i;
                const uint64_t qg = moduli[g];
                uint64_t dh = 1;
                for (size_t i2 = 0; i2 < s; i2++)
                    if (i2 != i) dh = host_mul_mod(dh, moduli[j * alpha + i2] % qg, qg);
                const uint64_t w = host_inv_mod(dh, qg);
                uw[idx] = make_ulonglong2(w, shoup_quotient(w, qg));
                for (size_t o = 0; o < K; o++) {
                    const uint64_t t = moduli[o];
                    uint64_t v = 1;
                    for (size_t i2 = 0; i2 < s; i2++)
                        if (i2 != i) v = host_mul_mod(v, moduli[j * alpha + i2] % t, t);
                    uc[idx * K + o] = v;
                }
            }
        }
    std::vector<ulonglong2> dw(alpha);
    std::vector<uint64_t> dc((alpha + 1) * L);  // row alpha: P mod q_i
    for (size_t k = 0; k < alpha; k++) {
        const uint64_t pk = moduli[L + k];
        uint64_t ph = 1;
        for (size_t k2 = 0; k2 < alpha; k2++)
            if (k2 != k) ph = host_mul_mod(ph, moduli[L + k2] % pk, pk);
        const uint64_t w = host_inv_mod(ph, pk);
        dw[k] = make_ulonglong2(w, shoup_quotient(w, pk));
        for (size_t i = 0; i < L; i++) {
            const uint64_t qi = moduli[i];
            uint64_t v = 1;
            for (size_t k2 = 0; k2 < alpha; k2++)
                if (k2 != k) v = host_mul_mod(v, moduli[L + k2] % qi, qi);
            dc[k * L + i] = v;
        }
    }
    for (size_t i = 0; i < L; i++) {
        const uint64_t qi = moduli[i];
        uint64_t v = 1;
        for (size_t k = 0; k < alpha; k++) v = host_mul_mod(v, moduli[L + k] % qi, qi);
        dc[alpha * L + i] = v;
    }
    // ModDown by P' = P q_(l-1) (relinearisation and rescale in one step), for
    // every level l >= 2: base b_0 = q_(l-1), b_k = p_(k-1)
    {
        const size_t A = alpha + 1, W = A + 1;
        std::vector<ulonglong2> fw((L + 1) * A, make_ulonglong2(0, 0));
        std::vector<uint64_t> fc((L + 1) * W * L, 0);
        std::vector<ulonglong2> fi((L + 1) * L, make_ulonglong2(0, 0));
        mdr_pmod.assign(L + 1, make_ulonglong2(0, 0));
        for (size_t l = 2; l <= L; l++) {
            const uint64_t ql = moduli[l - 1];
            for (size_t k = 0; k < A; k++) {
                const uint64_t bk = k ? moduli[L + k - 1] : ql;
                // P' / b_k mod b_k
                uint64_t h = k ? ql % bk : 1;
                for (size_t k2 = 0; k2 < alpha; k2++)
                    if (!k || k2 != k - 1) h = host_mul_mod(h, moduli[L + k2] % bk, bk);
                const uint64_t w = host_inv_mod(h, bk);
                fw[l * A + k] = make_ulonglong2(w, shoup_quotient(w, bk));
                for (size_t i = 0; i + 1 < l; i++) {
                    const uint64_t qi = moduli[i];
                    uint64_t v = k ? ql % qi : 1;
                    for (size_t k2 = 0; k2 < alpha; k2++)
                        if (!k || k2 != k - 1) v = host_mul_mod(v, moduli[L + k2] % qi, qi);
                    fc[(l * W + k) * L + i] = v;
                }
            }
            for (size_t i = 0; i + 1 < l; i++) {
                const uint64_t qi = moduli[i];
                uint64_t v = ql % qi;
                for (size_t k2 = 0; k2 < alpha; k2++) v = host_mul_mod(v, moduli[L + k2] % qi, qi);
                fc[(l * W + A) * L + i] = v;  // P' mod q_i
                const uint64_t iv = host_inv_mod(v, qi);
                fi[l * L + i] = make_ulonglong2(iv, shoup_quotient(iv, qi));
            }
            uint64_t pm = 1;
            for (size_t k2 = 0; k2 < alpha; k2++) pm = host_mul_mod(pm, moduli[L + k2] % ql, ql);
            mdr_pmod[l] = make_ulonglong2(pm, shoup_quotient(pm, ql));  // P mod q_(l-1)
        }
        std::vector<double> fcd((L + 1) * W * L, 0.0);
        for (size_t l = 2; l <= L; l++) {
            const std::vector<double> d = moddown_conv_doubles(&fc[l * W * L], A, L, moduli);
            std::copy(d.begin(), d.end(), fcd.begin() + l * W * L);
        }
        NX_HIP(hipMalloc(&d_mdr_conv_d, fcd.size() * sizeof(double)));
        NX_HIP(hipMemcpy(d_mdr_conv_d, fcd.data(), fcd.size() * sizeof(double), hipMemcpyHostToDevice));
        NX_HIP(hipMalloc(&d_mdr_w, fw.size() * sizeof(ulonglong2)));
        NX_HIP(hipMalloc(&d_mdr_conv, fc.size() * sizeof(uint64_t)));
        NX_HIP(hipMalloc(&d_mdr_inv, fi.size() * sizeof(ulonglong2)));
        NX_HIP(hipMemcpy(d_mdr_w, fw.data(), fw.size() * sizeof(ulonglong2), hipMemcpyHostToDevice));
        NX_HIP(hipMemcpy(d_mdr_conv, fc.data(), fc.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        NX_HIP(hipMemcpy(d_mdr_inv, fi.data(), fi.size() * sizeof(ulonglong2), hipMemcpyHostToDevice));
    }
    NX_HIP(hipMalloc(&d_modup_w, uw.size() * sizeof(ulonglong2)));
    NX_HIP(hipMalloc(&d_modup_conv, uc.size() * sizeof(uint64_t)));
    NX_HIP(hipMalloc(&d_moddown_w, dw.size() * sizeof(ulonglong2)));
    NX_HIP(hipMalloc(&d_moddown_conv, dc.size() * sizeof(uint64_t)));
    NX_HIP(hipMemcpy(d_modup_w, uw.data(), uw.size() * sizeof(ulonglong2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_modup_conv, uc.data(), uc.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    {
        // the same factors as doubles (exact: every factor is below its prime, < 2^53 where used)
        std::vector<double> ucd(uc.begin(), uc.end());
        NX_HIP(hipMalloc(&d_modup_conv_d, ucd.size() * sizeof(double)));
        NX_HIP(hipMemcpy(d_modup_conv_d, ucd.data(), ucd.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    NX_HIP(hipMemcpy(d_moddown_w, dw.data(), dw.size() * sizeof(ulonglong2), hipMemcpyHostToDevice));
    NX_HIP(hipMemcpy(d_moddown_conv, dc.data(), dc.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    {
        const std::vector<double> dcd = moddown_conv_doubles(dc.data(), alpha, L, moduli);
        NX_HIP(hipMalloc(&d_moddown_conv_d, dcd.size() * sizeof(double)));
        NX_HIP(hipMemcpy(d_moddown_conv_d, dcd.data(), dcd.size() * sizeof(double), hipMemcpyHostToDevice));
    }
}

Context::~Context()
{
    if (stream.def_) (void)hipStreamSynchronize(stream.def_);
    for (auto &l : lanes_)
        if (l && l->stream) (void)hipStreamSynchronize(l->stream);
    if (tl_ctx == this) {
        tl_ctx = nullptr;
        tl_lane = nullptr;
    }
    (void)hipFree(d_modup_w);
    (void)hipFree(d_modup_conv);
    (void)hipFree(d_modup_conv_d);
    (void)hipFree(d_mdr_w);
    (void)hipFree(d_mdr_conv);
    (void)hipFree(d_mdr_inv);
    (void)hipFree(d_moddown_w);
    (void)hipFree(d_moddown_conv);
    (void)hipFree(d_moddown_conv_d);
    (void)hipFree(d_mdr_conv_d);
    for (auto &kv : galois_) (void)hipFree(kv.second);
    (void)hipFree(d_fwd);
    (void)hipFree(d_inv);
    (void)hipFree(d_ninv);
    (void)hipFree(d_fwd_d);
    (void)hipFree(d_inv_d);
    (void)hipFree(d_ninv_d);
    (void)hipFree(d_mod);
    (void)hipFree(d_qd);
    (void)hipFree(d_rescale_inv);
    (void)hipFree(d_p_inv);
    (void)hipFree(d_enc_index);
    (void)hipFree(d_enc_roots);
    cache->trim();
    for (auto &l : lanes_)
        if (l) {
            l->cache->trim();
            if (l->stream) (void)hipStreamDestroy(l->stream);
        }
    if (stream.def_) (void)hipStreamDestroy(stream.def_);
}

int Context::total_bits(size_t limbs) const
{
    // significant bits of q_0 * ... * q_{limbs-1} (SEAL total_coeff_modulus_bit_count)
    std::vector<uint64_t> prod(limbs + 1, 0);
    prod[0] = 1;
    for (size_t i = 0; i < limbs; i++) {
        u128 carry = 0;
        for (size_t w = 0; w <= limbs; w++) {
            u128 t = (u128)prod[w] * moduli[i] + carry;
            prod[w] = (uint64_t)t;
            carry = t >> 64;
        }
    }
    for (size_t w = limbs + 1; w-- > 0;)
        if (prod[w]) return (int)(w * 64 + 64 - __builtin_clzll(prod[w]));
    return 0;
}

uint32_t Context::elt_from_step(int step) const
{
    // thirdparty/SEAL-4.1-bs/native/src/seal/util/galois.cpp:53-90, generator 5 (galois.h:169)
    const uint32_t m = 2 * (uint32_t)n;
    if (step == 0) return m - 1;
    const bool neg = step < 0;
    const uint32_t pos = (uint32_t)(neg ? -(int64_t)step : step);
    require(pos < n / 2, "step count too large");
    uint32_t s = neg ? (uint32_t)(n / 2) - pos : pos;
    uint64_t e = 1;
    while (s--) e = (e * 5) & (m - 1);
    return (uint32_t)e;
}

const uint32_t *Context::galois_table(uint32_t elt)
{
    std::lock_guard<std::mutex> g(mu_);
    auto it = galois_.find(elt);
    if (it != galois_.end()) return it->second;
    require((elt & 1) && elt < 2 * n, "Galois element is not valid");
    // util/galois.cpp:18-51
    std::vector<uint32_t> t(n);
    const uint32_t mask = (uint32_t)n - 1;
    for (uint32_t i = (uint32_t)n; i < 2 * (uint32_t)n; i++) {
        uint32_t rev = reverse_bits(i, logn + 1);
        uint64_t raw = ((uint64_t)elt * rev) >> 1;
        t[i - n] = reverse_bits((uint32_t)(raw & mask), logn);
    }
    uint32_t *d = nullptr;
    NX_HIP(hipMalloc(&d, n * sizeof(uint32_t)));
    NX_HIP(hipMemcpy(d, t.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    galois_[elt] = d;
    return d;
}

// ---------------------------------------------------------------- NTT helpers
static NttBatch make_batch(uint64_t *data, size_t polys, size_t limbs, size_t poly_stride)
{
    NttBatch b;
    b.data = data;
    b.poly_stride = poly_stride;
    b.polys = (uint32_t)polys;
    b.limbs = (uint32_t)limbs;
    b.prime_by_poly = 0;
    for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)i;
    return b;
}

void ntt_fwd(Context &c, uint64_t *data, size_t polys, size_t limbs, size_t poly_stride)
{
    ntt_forward(c.tables, make_batch(data, polys, limbs, poly_stride), c.stream);
}

void ntt_fwd_epi(Context &c, uint64_t *data, size_t polys, size_t limbs, size_t poly_stride, const NttEpilogue &epi,
                 bool rows_only)
{
    NttBatch b = make_batch(data, polys, limbs, poly_stride);
    b.epi = epi;
    b.rows_only = rows_only ? 1 : 0;
    ntt_forward(c.tables, b, c.stream);
}

void ntt_inv(Context &c, uint64_t *data, size_t polys, size_t limbs, size_t poly_stride)
{
    ntt_inverse(c.tables, make_batch(data, polys, limbs, poly_stride), c.stream);
}

void ntt_fwd_from(Context &c, uint64_t *dst, const uint64_t *src, size_t polys, size_t limbs, size_t dst_stride,
                  size_t src_stride)
{
    NttBatch b = make_batch(dst, polys, limbs, dst_stride);
    b.src = src;
    b.src_poly_stride = src_stride;
    ntt_forward(c.tables, b, c.stream);
}

void ntt_inv_from(Context &c, uint64_t *dst, const uint64_t *src, size_t polys, size_t limbs, size_t dst_stride,
                  size_t src_stride)
{
    NttBatch b = make_batch(dst, polys, limbs, dst_stride);
    b.src = src;
    b.src_poly_stride = src_stride;
    ntt_inverse(c.tables, b, c.stream);
}

}  // namespace nx
