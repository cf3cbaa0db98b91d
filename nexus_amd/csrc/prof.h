// Live per-kernel timing with HIP events.  When a kernel class is selected
// (nx_profile_select), every launch of that class is bracketed by two events
// recorded on the launch's own stream, together with the launch's work units
// (RNS limbs processed), so average duration and algorithmic throughput can
// be computed for the real workload.
#pragma once
#include <hip/hip_runtime.h>

namespace nx {

bool prof_active(const char *kernel_class);
// Operation tally (nx_op_tally): `count` operations of `kind` on `limbs` limbs
void tally(const char *kind, size_t limbs, double count = 1);
// units: algorithmic HBM bytes of the launch; flops: its algorithmic FP64
// operations (an FMA counts two), for the compute-bound classes
void prof_record(const char *kernel_class, hipStream_t s, bool begin, double units, double flops = 0);

struct ProfScope {
    const char *cls;
    hipStream_t s;
    double units, flops;
    bool on;
    ProfScope(const char *c, hipStream_t st, double u, double f = 0)
        : cls(c), s(st), units(u), flops(f), on(prof_active(c))
    {
        if (on) prof_record(cls, s, true, units, flops);
    }
    ~ProfScope()
    {
        if (on) prof_record(cls, s, false, units, flops);
    }
};

}  // namespace nx
