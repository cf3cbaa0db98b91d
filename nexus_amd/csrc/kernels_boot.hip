// Bootstrapping kernels.
//
// ModRaise (reference src/bootstrapping/Bootstrapper.cpp:2937-2991): the
// coefficient-domain residue x of q0 is lifted to every prime of the chain as
// its centred representative: x mod q_j, plus (q_j - q0 mod q_j) when
// x > floor(q0 / 2).
#include "kernels.h"
#include "prof.h"

namespace nx {

namespace {

__global__ void k_modraise(uint64_t *out, const uint64_t *in0, const ModConst *mod, uint32_t limbs, uint32_t n)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const uint32_t row = blockIdx.y;  // poly * limbs + j
    const uint32_t poly = row / limbs, j = row - poly * limbs;
    const uint64_t q0 = mod[0].q;
    const ModConst m = mod[j];
    const uint64_t x = in0[(uint64_t)poly * n + c];
    uint64_t v = x >= m.q ? barrett64(x, m) : x;
    if (x > (q0 >> 1)) {
        const uint64_t minus_q0 = j == 0 ? 0 : m.q - barrett64(q0, m);
        v += minus_q0;
        v = v >= m.q ? v - m.q : v;
    }
    out[(uint64_t)row * n + c] = v;
}

}  // namespace

void mod_raise(const DeviceTables &t, uint64_t *out, const uint64_t *in0, uint32_t polys, uint32_t limbs,
               hipStream_t s)
{
    const uint32_t n = 1u << t.logn;
    ProfScope prof("modraise", s, 8.0 * n * polys * (1.0 + limbs));
    hipLaunchKernelGGL(k_modraise, dim3((n + 255) / 256, polys * limbs), dim3(256), 0, s, out, in0, t.mod, limbs, n);
}

}  // namespace nx
