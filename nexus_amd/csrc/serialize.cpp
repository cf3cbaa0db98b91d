// SEAL-compatible serialization of ciphertexts, plaintexts and key-switching
// keys (SEAL 4.1 binary format, compression mode "none"), and SEAL's parms_id.
//
// Container (serialization.h:76-91): a 16-byte header {u16 magic 0xA15E, u8
// header size 16, u8 version major 4, u8 minor 1, u8 compr_mode 0, u16 0,
// u64 total size} followed by the object's members, little endian.
//   Ciphertext (ciphertext.cpp save_members): parms_id[4] u64, is_ntt u8,
//     size u64, N u64, coeff_modulus_size u64, scale f64, correction_factor
//     u64, then the data array as its own container {u64 count, count words}.
//   Plaintext (plaintext.cpp): parms_id[4], coeff_count u64, scale f64, data array.
//   KSwitchKeys (kswitchkeys.cpp): parms_id[4], u64 dim1, per entry u64 dim2
//     and dim2 key-level ciphertexts (PublicKey containers).
// parms_id (encryptionparams.cpp compute_parms_id): BLAKE2b with a 32-byte
// digest of the words {scheme = 2 (CKKS), N, q_0 .. q_{k-1}, plain modulus 0}.
#include <cstring>

#include "context.h"

namespace nx {

namespace {

// ---------------------------------------------------------------- BLAKE2b (RFC 7693)
const uint64_t kIV[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull,
                         0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full,
                         0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};
const uint8_t kSigma[12][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3}};

inline uint64_t rotr(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }

void mix(uint64_t *v, int a, int b, int c, int d, uint64_t x, uint64_t y)
{
    v[a] = v[a] + v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 32);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 24);
    v[a] = v[a] + v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 63);
}

void compress(uint64_t *h, const uint8_t *block, uint64_t t, bool last)
{
    uint64_t m[16], v[16];
    for (int i = 0; i < 16; i++) {
        uint64_t w = 0;
        for (int k = 7; k >= 0; k--) w = (w << 8) | block[i * 8 + k];
        m[i] = w;
    }
    for (int i = 0; i < 8; i++) {
        v[i] = h[i];
        v[i + 8] = kIV[i];
    }
    v[12] ^= t;  // low word of the byte counter (messages here are far below 2^64 bytes)
    if (last) v[14] = ~v[14];
    for (int r = 0; r < 12; r++) {
        const uint8_t *s = kSigma[r];
        mix(v, 0, 4, 8, 12, m[s[0]], m[s[1]]);
        mix(v, 1, 5, 9, 13, m[s[2]], m[s[3]]);
        mix(v, 2, 6, 10, 14, m[s[4]], m[s[5]]);
        mix(v, 3, 7, 11, 15, m[s[6]], m[s[7]]);
        mix(v, 0, 5, 10, 15, m[s[8]], m[s[9]]);
        mix(v, 1, 6, 11, 12, m[s[10]], m[s[11]]);
        mix(v, 2, 7, 8, 13, m[s[12]], m[s[13]]);
        mix(v, 3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}

}  // namespace

void blake2b(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen)
{
    require(outlen >= 1 && outlen <= 64, "digest length must be 1..64");
    uint64_t h[8];
    for (int i = 0; i < 8; i++) h[i] = kIV[i];
    h[0] ^= 0x01010000ull ^ (uint64_t)outlen;  // no key, fanout 1, depth 1
    uint8_t block[128];
    uint64_t t = 0;
    size_t off = 0;
    // all blocks but the last are full; the last (possibly empty) is padded
    while (inlen - off > 128) {
        t += 128;
        compress(h, in + off, t, false);
        off += 128;
    }
    const size_t rest = inlen - off;
    std::memset(block, 0, sizeof(block));
    if (rest) std::memcpy(block, in + off, rest);
    t += rest;
    compress(h, block, t, true);
    for (size_t i = 0; i < outlen; i++) out[i] = (uint8_t)(h[i / 8] >> (8 * (i % 8)));
}

// parms_id of the parameter set whose coefficient modulus is the first `limbs`
// data primes (limbs <= L), or every prime of the key level (limbs = 0).
void parms_id(const Context &c, size_t limbs, uint64_t out[4])
{
    std::vector<uint64_t> w;
    w.push_back(2);  // scheme_type::ckks
    w.push_back(c.n);
    if (limbs == 0)
        for (uint64_t q : c.moduli) w.push_back(q);
    else
        for (size_t i = 0; i < limbs; i++) w.push_back(c.moduli[i]);
    w.push_back(0);  // plain modulus (unset for CKKS)
    std::vector<uint8_t> bytes(w.size() * 8);
    for (size_t i = 0; i < w.size(); i++)
        for (int k = 0; k < 8; k++) bytes[i * 8 + k] = (uint8_t)(w[i] >> (8 * k));
    uint8_t d[32];
    blake2b(d, 32, bytes.data(), bytes.size());
    for (int i = 0; i < 4; i++) {
        uint64_t v = 0;
        for (int k = 7; k >= 0; k--) v = (v << 8) | d[i * 8 + k];
        out[i] = v;
    }
}

namespace {

const uint16_t kMagic = 0xA15E;

struct Writer {
    std::vector<uint8_t> b;
    template <class T>
    void put(T v)
    {
        uint8_t tmp[sizeof(T)];
        std::memcpy(tmp, &v, sizeof(T));
        b.insert(b.end(), tmp, tmp + sizeof(T));
    }
    size_t begin_container()
    {
        const size_t at = b.size();
        put<uint16_t>(kMagic);
        put<uint8_t>(16);
        put<uint8_t>(4);  // SEAL 4.1
        put<uint8_t>(1);
        put<uint8_t>(0);  // compr_mode_type::none
        put<uint16_t>(0);
        put<uint64_t>(0);  // size, patched by end_container
        return at;
    }
    void end_container(size_t at)
    {
        const uint64_t size = b.size() - at;
        std::memcpy(b.data() + at + 8, &size, 8);
    }
    void words(const uint64_t *p, size_t count)
    {
        const size_t at = begin_container();
        put<uint64_t>(count);
        const uint8_t *src = reinterpret_cast<const uint8_t *>(p);
        b.insert(b.end(), src, src + count * 8);
        end_container(at);
    }
};

struct Reader {
    const uint8_t *p;
    size_t n, off = 0;
    template <class T>
    T get()
    {
        require(off + sizeof(T) <= n, "serialized data is truncated");
        T v;
        std::memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
    size_t begin_container()
    {
        const size_t at = off;
        require(get<uint16_t>() == kMagic, "not SEAL-serialized data (magic)");
        require(get<uint8_t>() == 16, "unsupported header size");
        const uint8_t major = get<uint8_t>();
        (void)get<uint8_t>();
        require(major == 4, "unsupported SEAL version");
        require(get<uint8_t>() == 0, "compressed SEAL data is not supported");
        (void)get<uint16_t>();
        const uint64_t size = get<uint64_t>();
        require(size >= 16 && at + size <= n, "serialized size is inconsistent");
        return at + (size_t)size;  // end offset
    }
    void end_container(size_t end) { require(off == end, "serialized size is inconsistent"); }
    std::vector<uint64_t> words(size_t expect)
    {
        const size_t end = begin_container();
        const uint64_t count = get<uint64_t>();
        require(count == expect, "unexpected data size");
        require(off + count * 8 <= n, "serialized data is truncated");
        std::vector<uint64_t> w(count);
        std::memcpy(w.data(), p + off, count * 8);
        off += count * 8;
        end_container(end);
        return w;
    }
};

// data level (limbs) of a parms_id, or 0 for the key level; throws if unknown
size_t limbs_of_parms_id(const Context &c, const uint64_t id[4])
{
    uint64_t t[4];
    parms_id(c, 0, t);
    if (std::memcmp(t, id, 32) == 0) return 0;
    for (size_t l = 1; l <= c.L; l++) {
        parms_id(c, l, t);
        if (std::memcmp(t, id, 32) == 0) return l;
    }
    fail(kInvalidArgument, "parms_id does not belong to this context");
}

void put_ct_members(Writer &w, const Context &c, size_t limbs_for_id, size_t size, size_t limbs, double scale,
                    bool ntt, const uint64_t *host)
{
    uint64_t id[4];
    parms_id(c, limbs_for_id, id);
    for (int i = 0; i < 4; i++) w.put<uint64_t>(id[i]);
    w.put<uint8_t>(ntt ? 1 : 0);
    w.put<uint64_t>(size);
    w.put<uint64_t>(c.n);
    w.put<uint64_t>(limbs);
    w.put<double>(scale);
    w.put<uint64_t>(1);  // correction factor (BGV only)
    w.words(host, size * limbs * c.n);
}

std::vector<uint64_t> download(Context &c, const uint64_t *dev, size_t words)
{
    std::vector<uint64_t> h(words);
    NX_HIP(hipMemcpyAsync(h.data(), dev, words * 8, hipMemcpyDeviceToHost, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));
    return h;
}

void upload(Context &c, uint64_t *dev, const std::vector<uint64_t> &h)
{
    NX_HIP(hipMemcpyAsync(dev, h.data(), h.size() * 8, hipMemcpyHostToDevice, c.stream));
    NX_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

std::vector<uint8_t> save_ciphertext(const Ciphertext &ct)
{
    require(ct.ctx != nullptr && ct.size >= 2, "encrypted is not valid for encryption parameters");
    Context &c = *ct.ctx;
    const std::vector<uint64_t> h = download(c, ct.data.get(), ct.size * ct.limbs * c.n);
    Writer w;
    const size_t at = w.begin_container();
    put_ct_members(w, c, ct.limbs, ct.size, ct.limbs, ct.scale, ct.ntt, h.data());
    w.end_container(at);
    return w.b;
}

void load_ciphertext(Context &c, const uint8_t *data, size_t n, Ciphertext &out)
{
    Reader r{data, n};
    const size_t end = r.begin_container();
    uint64_t id[4];
    for (int i = 0; i < 4; i++) id[i] = r.get<uint64_t>();
    const size_t limbs_id = limbs_of_parms_id(c, id);
    require(limbs_id != 0, "a key-level ciphertext is not a data ciphertext");
    const bool ntt = r.get<uint8_t>() != 0;
    const uint64_t size = r.get<uint64_t>();
    const uint64_t N = r.get<uint64_t>();
    const uint64_t limbs = r.get<uint64_t>();
    const double scale = r.get<double>();
    (void)r.get<uint64_t>();
    require(N == c.n, "poly_modulus_degree mismatch");
    require(limbs == limbs_id, "coeff_modulus_size does not match parms_id");
    require(size >= 2 && size <= 16, "invalid ciphertext size");
    std::vector<uint64_t> h = r.words(size * limbs * c.n);
    r.end_container(end);
    for (size_t i = 0; i < h.size(); i++)
        require(h[i] < c.moduli[(i / c.n) % limbs], "residue out of range");
    Ciphertext ct;
    ct.ctx = &c;
    ct.size = size;
    ct.limbs = limbs;
    ct.scale = scale;
    ct.ntt = ntt;
    ct.data.reset(&c, h.size());
    upload(c, ct.data.get(), h);
    out = std::move(ct);
}

std::vector<uint8_t> save_plaintext(const Plaintext &pt)
{
    require(pt.ctx != nullptr && pt.ntt, "plain is not valid for encryption parameters");
    Context &c = *pt.ctx;
    const std::vector<uint64_t> h = download(c, pt.data.get(), pt.limbs * c.n);
    Writer w;
    const size_t at = w.begin_container();
    uint64_t id[4];
    parms_id(c, pt.limbs, id);
    for (int i = 0; i < 4; i++) w.put<uint64_t>(id[i]);
    w.put<uint64_t>(pt.limbs * c.n);
    w.put<double>(pt.scale);
    w.words(h.data(), h.size());
    w.end_container(at);
    return w.b;
}

void load_plaintext(Context &c, const uint8_t *data, size_t n, Plaintext &out)
{
    Reader r{data, n};
    const size_t end = r.begin_container();
    uint64_t id[4];
    for (int i = 0; i < 4; i++) id[i] = r.get<uint64_t>();
    const size_t limbs = limbs_of_parms_id(c, id);
    require(limbs != 0, "a key-level plaintext is not supported");
    const uint64_t count = r.get<uint64_t>();
    const double scale = r.get<double>();
    require(count == limbs * c.n, "coeff_count does not match parms_id");
    std::vector<uint64_t> h = r.words(count);
    r.end_container(end);
    for (size_t i = 0; i < h.size(); i++) require(h[i] < c.moduli[i / c.n], "residue out of range");
    Plaintext pt;
    pt.ctx = &c;
    pt.limbs = limbs;
    pt.scale = scale;
    pt.ntt = true;
    pt.data.reset(&c, h.size());
    upload(c, pt.data.get(), h);
    out = std::move(pt);
}

// KSwitchKeys: the relinearisation set has one entry (key power 2, index 0);
// Galois sets have N entries indexed by (elt - 1) / 2 (GaloisTool::GetIndexFromElt),
// empty where no key exists.  Each digit is a size-2 key-level ciphertext.
std::vector<uint8_t> save_kswitch_keys(const KSwitchKeys &keys, bool galois)
{
    require(keys.ctx != nullptr, "keys are not initialised");
    Context &c = *keys.ctx;
    const size_t K = c.key_limbs(), digit_words = 2 * K * c.n;
    Writer w;
    const size_t at = w.begin_container();
    uint64_t id[4];
    parms_id(c, 0, id);
    for (int i = 0; i < 4; i++) w.put<uint64_t>(id[i]);
    const size_t dim1 = galois ? c.n : 1;
    w.put<uint64_t>(dim1);
    for (size_t index = 0; index < dim1; index++) {
        const uint32_t elt = galois ? (uint32_t)(2 * index + 1) : 0u;
        auto it = keys.keys.find(elt);
        if (it == keys.keys.end() || (galois && elt == 0)) {
            w.put<uint64_t>(0);
            continue;
        }
        w.put<uint64_t>(c.dnum);
        const std::vector<uint64_t> h = download(c, it->second.get(), c.key_words());
        for (size_t j = 0; j < c.dnum; j++) {
            const size_t pk = w.begin_container();
            put_ct_members(w, c, 0, 2, K, 1.0, true, h.data() + j * digit_words);
            w.end_container(pk);
        }
    }
    w.end_container(at);
    return w.b;
}

void load_kswitch_keys(Context &c, const uint8_t *data, size_t n, bool galois, KSwitchKeys &out)
{
    Reader r{data, n};
    const size_t end = r.begin_container();
    uint64_t id[4];
    for (int i = 0; i < 4; i++) id[i] = r.get<uint64_t>();
    require(limbs_of_parms_id(c, id) == 0, "keys are not at the key level of this context");
    const size_t K = c.key_limbs(), digit_words = 2 * K * c.n;
    const uint64_t dim1 = r.get<uint64_t>();
    require(dim1 <= c.n, "too many key entries");
    out.ctx = &c;
    for (uint64_t index = 0; index < dim1; index++) {
        const uint64_t dim2 = r.get<uint64_t>();
        if (dim2 == 0) continue;
        require(dim2 == c.dnum, "key decomposition count does not match the context");
        std::vector<uint64_t> h(c.key_words());
        for (size_t j = 0; j < dim2; j++) {
            const size_t pk = r.begin_container();
            uint64_t pid[4];
            for (int i = 0; i < 4; i++) pid[i] = r.get<uint64_t>();
            require(limbs_of_parms_id(c, pid) == 0, "key is not at the key level");
            require(r.get<uint8_t>() == 1, "key is not in NTT form");
            require(r.get<uint64_t>() == 2, "key ciphertext size must be 2");
            require(r.get<uint64_t>() == c.n, "poly_modulus_degree mismatch");
            require(r.get<uint64_t>() == K, "key coeff_modulus_size mismatch");
            (void)r.get<double>();
            (void)r.get<uint64_t>();
            std::vector<uint64_t> d = r.words(digit_words);
            r.end_container(pk);
            std::memcpy(h.data() + j * digit_words, d.data(), digit_words * 8);
        }
        const uint32_t elt = galois ? (uint32_t)(2 * index + 1) : 0u;
        DevBuf &b = out.keys[elt];
        out.seeds.erase(elt);  // loaded words: the uniform half is read, not regenerated
        b.reset(&c, h.size());
        upload(c, b.get(), h);
    }
    r.end_container(end);
}

}  // namespace nx
