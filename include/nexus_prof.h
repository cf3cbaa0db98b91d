/*
 * nexus_prof.h -- live kernel timing for benchmarks.  No counterpart in the
 * reference (which only wall-clocks whole calls, cuda/src/utils.cuh:10-43).
 *
 * Every launch of a selected kernel class is bracketed by two HIP events on
 * the stream it is launched on; the launch also records its algorithmic HBM
 * bytes.  Kernel classes include "ntt_fwd_col", "ntt_fwd_row",
 * "ntt_fwd_sub_mul" (a forward row pass that stores a ModDown or rescale
 * tail), "ntt_fwd_whole", "ntt_inv_row", "ntt_inv_col", "ntt_inv_whole",
 * "ks_inner", "ks_modup_conv", "ks_moddown_conv", "dot_plain", "lincomb",
 * "tensor", "mul_plain", "copy"; nx_profile_classes lists those seen.
 */
#ifndef NEXUS_PROF_H
#define NEXUS_PROF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Comma-separated kernel classes to time, "*" for all, "" to stop.  Resets the statistics. */
int nx_profile_select(const char *classes);
/* Launches, summed event-measured milliseconds and summed algorithmic bytes of one class. */
int nx_profile_read(const char *kernel_class, size_t *launches, double *total_ms, double *bytes);
/* Summed algorithmic double-precision operations (an FMA counts two) of one
 * class; recorded by the compute-bound classes (the conversions fused into
 * the NTT column pass), 0 for the others. */
int nx_profile_read_flops(const char *kernel_class, double *flops);
/* Comma-separated list of the classes with statistics. */
int nx_profile_classes(char *buf, size_t len);
/* Operation tally: start = 1 clears and starts counting the engine's
 * operations by kind and level (key switches "ks", "rescale", ct x ct "tensor",
 * ct x pt "mul_plain", elementwise "add", plaintext "encode"), start = 0 stops
 * and writes {"kind": {"limbs": count, ...}, ...} to buf (*needed: its size
 * with the terminator).  The CPU baseline composes its per-operation timings
 * with this mix. */
int nx_op_tally(int start, char *buf, size_t len, size_t *needed);
/* Measured VALU issue rate of this device, the denominator of the compute
 * roofline of the modular kernels: kind 0 = 64 x 64 -> 128-bit integer
 * products per second (the low and high halves of the Shoup / Barrett
 * multiplies), kind 1 = double-precision FMAs per second (the exact-product
 * path of primes below 2^51).  Eight independent chains per thread, 8 waves
 * per SIMD, no memory traffic. */
int nx_profile_valu_rate(int kind, double *ops_per_s);

#ifdef __cplusplus
}
#endif

#endif /* NEXUS_PROF_H */
