// VALU rate probes: the measured denominators of the integer and
// double-precision rooflines of the modular-arithmetic kernels.  Each thread
// runs eight independent dependency chains so that the issue rate, not the
// latency, bounds the kernel; nothing but one final store touches memory.
//  - kind 0: 64 x 64 -> 128-bit products (low and high halves), the unit of
//    the Shoup and Barrett multiplies of the integer paths;
//  - kind 1: double-precision FMAs, the unit of the double-precision paths.
#include <string>

#include "kernels.h"
#include "prof.h"

#define NX_PROBE_HIP(x)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace nx {

namespace {

constexpr int kProbeBlock = 256;
constexpr int kChains = 8;

__global__ void __launch_bounds__(kProbeBlock) k_probe_mul64(uint64_t *out, uint64_t seed, uint32_t iters)
{
    const uint64_t tid = (uint64_t)blockIdx.x * kProbeBlock + threadIdx.x;
    uint64_t x[kChains];
#pragma unroll
    for (int i = 0; i < kChains; i++) x[i] = seed ^ (tid * kChains + i) * 0x9E3779B97F4A7C15ull;
    const uint64_t a = 0xD1342543DE82EF95ull, b = 0xAF251AF3B0F025B5ull;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < kChains; i++) x[i] = x[i] * a + mulhi64(x[i], b);
    }
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < kChains; i++) r ^= x[i];
    out[tid] = r;
}

__global__ void __launch_bounds__(kProbeBlock) k_probe_fma64(double *out, double seed, uint32_t iters)
{
    const uint64_t tid = (uint64_t)blockIdx.x * kProbeBlock + threadIdx.x;
    double x[kChains];
#pragma unroll
    for (int i = 0; i < kChains; i++) x[i] = seed + (double)(tid % 1024) * 1e-6 + i * 1e-3;
    const double a = 0.999999, b = 1e-7;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < kChains; i++) x[i] = fma(x[i], a, b);
    }
    double r = 0;
#pragma unroll
    for (int i = 0; i < kChains; i++) r += x[i];
    out[tid] = r;
}

}  // namespace

double probe_valu_rate(int kind, hipStream_t s)
{
    if (kind != 0 && kind != 1) throw std::invalid_argument("probe kind must be 0 (int64 products) or 1 (fp64 fma)");
    int dev = 0, cus = 0;
    NX_PROBE_HIP(hipGetDevice(&dev));
    NX_PROBE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 32 workgroups of 4 waves per CU: every SIMD holds 8 waves
    const uint32_t blocks = (uint32_t)(cus > 0 ? cus : 256) * 32u;
    const uint32_t iters = kind == 0 ? 4096u : 32768u;
    const size_t threads = (size_t)blocks * kProbeBlock;
    void *buf = nullptr;
    NX_PROBE_HIP(hipMalloc(&buf, threads * sizeof(uint64_t)));
    hipEvent_t e0, e1;
    NX_PROBE_HIP(hipEventCreate(&e0));
    NX_PROBE_HIP(hipEventCreate(&e1));
    float ms = 0;
    for (int rep = 0; rep < 2; rep++) {  // the first launch loads the code object
        NX_PROBE_HIP(hipEventRecord(e0, s));
        if (kind == 0)
            hipLaunchKernelGGL(k_probe_mul64, dim3(blocks), dim3(kProbeBlock), 0, s, (uint64_t *)buf, 12345ull, iters);
        else
            hipLaunchKernelGGL(k_probe_fma64, dim3(blocks), dim3(kProbeBlock), 0, s, (double *)buf, 0.5, iters);
        NX_PROBE_HIP(hipEventRecord(e1, s));
        NX_PROBE_HIP(hipEventSynchronize(e1));
        NX_PROBE_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    NX_PROBE_HIP(hipEventDestroy(e0));
    NX_PROBE_HIP(hipEventDestroy(e1));
    NX_PROBE_HIP(hipFree(buf));
    const double ops = (double)threads * kChains * iters;
    return ops / (ms * 1e-3);
}

}  // namespace nx
