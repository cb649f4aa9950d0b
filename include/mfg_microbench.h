/* mfg_microbench.h -- measurement helper shipped beside the tracer.
 *
 * No counterpart in the reference: it measures the device-scope atomic
 * throughput of the GPU, the peak that the trace kernel's tally atomics
 * (FP64 sums, uint64 counts) are reported against, and the random 128 B
 * line read rate its L2-miss traffic is reported against. Library:
 * molflow_sources_amd/libmfg_microbench.so. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fills f64_per_s and u64_per_s with the rates of FP64 and uint64 atomic
 * adds (no return value, device scope) that a full grid achieves on
 * n_addr distinct 8-byte counters (>= 8192), each thread issuing
 * per_thread adds at addresses 4099 counters apart. Returns 0 on success,
 * 1 on invalid arguments, 2 on a HIP error. */
int mfgb_atomic_rates(int device, uint32_t n_addr, uint32_t per_thread, double* f64_per_s, double* u64_per_s);

/* Fills bytes_per_s with the rate at which a full grid reads whole 128 B
 * lines at pseudo-random line indices of a device array of `bytes` bytes
 * (>= 1 MiB; each thread reads per_thread lines as eight 16 B loads): the
 * access pattern of the tracer's leaf records and nodes past the LDS copy,
 * so the achievable roof for the trace kernel's L2-miss traffic at a
 * working set of that size. Returns 0 on success, 1 on invalid arguments,
 * 2 on a HIP error. */
int mfgb_gather_rate(int device, uint64_t bytes, uint32_t per_thread, double* bytes_per_s);

#ifdef __cplusplus
}
#endif
