// Allocation-free rescale for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops): nx_rescale_to_next_noalloc.
//
// The rescale's two scratch buffers (the INTT of the last limb of every
// polynomial, and its centred replicas on the remaining limbs, transformed in
// place) live in one file-scope device array of kScratchWords words.  Every GPU
// the library runs on holds its own copy until the process ends; its address on
// each device is taken once, when a context is created with that device
// current, and only that address is handed to kernels.  All calls on one GPU
// share the array, so calls that overlap on one GPU race.
#include <mutex>

#include "context.h"

namespace nx {

namespace {

constexpr size_t kScratchWords = (size_t)4 << 20;  // 32 MiB
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint64_t scratch[kScratchWords];

std::mutex g_mu;
uint64_t *g_scratch[kMaxDevices] = {};

uint64_t *scratch_of_current_device()
{
    int dev = 0;
    NX_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < kMaxDevices, "device index out of range", kDeviceError);
    std::lock_guard<std::mutex> g(g_mu);
    require(g_scratch[dev] != nullptr, "no context was created on this device", kLogicError);
    return g_scratch[dev];
}

}  // namespace

size_t noalloc_scratch_words() { return kScratchWords; }

void noalloc_init_current_device()
{
    int dev = 0;
    NX_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < kMaxDevices, "device index out of range", kDeviceError);
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[dev]) return;
    void *p = nullptr;
    NX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "scratch array has no device address", kDeviceError);
    g_scratch[dev] = static_cast<uint64_t *>(p);
}

void rescale_noalloc(Context &c, const uint64_t *in, size_t polys, size_t limbs, uint64_t *out)
{
    // the steps of op_rescale_to_next (util/rns.cpp:830-905) on caller memory:
    // INTT of each last limb, + q_last / 2 and centring onto the other limbs,
    // NTT whose last pass stores (x - t) q_last^-1 into out
    require(in != nullptr && out != nullptr, "in and out must be device pointers");
    require(polys >= 1 && polys <= 3, "polys must be 1..3");
    require(limbs >= 2 && limbs <= c.L, "limbs must be 2..L");
    const size_t n = c.n, l = limbs, P = polys;
    require(P * n * l <= kScratchWords, "input too large for the 32 MiB scratch of the allocation-free rescale");
    require(out + P * (l - 1) * n <= in || in + P * l * n <= out, "out must not overlap in");
    uint64_t *last = scratch_of_current_device(), *tmp = last + P * n;
    NttBatch b;
    b.data = last;
    b.poly_stride = n;
    b.polys = (uint32_t)P;
    b.limbs = 1;
    b.prime_by_poly = 0;
    for (int i = 0; i < kMaxPrimes; i++) b.prime_of[i] = (uint8_t)(l - 1);
    b.src = in + (l - 1) * n;
    b.src_poly_stride = l * n;
    ntt_inverse(c.tables, b, c.stream);
    rns_center_last(c.tables, tmp, last, n, (uint32_t)(l - 1), (uint32_t)P, (uint32_t)(l - 1), c.stream);
    NttEpilogue e;
    e.out = out;
    e.out_stride = (l - 1) * n;
    e.x = in;
    e.x_stride = l * n;
    e.inv = c.rescale_inv(l);
    ntt_fwd_epi(c, tmp, P, l - 1, (l - 1) * n, e);
}

}  // namespace nx
