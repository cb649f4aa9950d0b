// mfg_microbench.hip -- device-scope atomic throughput of the GPU: the peak
// the trace kernel's tally atomics are compared with. A measurement helper
// beside the tracer, not part of it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mfg_microbench.h"

namespace {

// every thread adds to per_thread counters 4099 apart (n > 4099, so one
// subtraction keeps the index in [0, n))
__global__ void k_atomic_f64(double* buf, uint32_t n, uint32_t per_thread) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a = (uint32_t)((tid * 2654435761ull) % n);
    for (uint32_t k = 0; k < per_thread; ++k) {
        unsafeAtomicAdd(buf + a, 1.0);
        a += 4099u;
        if (a >= n) a -= n;
    }
}

__global__ void k_atomic_u64(unsigned long long* buf, uint32_t n, uint32_t per_thread) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a = (uint32_t)((tid * 2654435761ull) % n);
    for (uint32_t k = 0; k < per_thread; ++k) {
        atomicAdd(buf + a, 1ull);
        a += 4099u;
        if (a >= n) a -= n;
    }
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// each thread reads per_thread whole 128 B lines at hashed line indices
// (< n_lines) and writes one value, so no load is dead
__global__ void k_gather(const uint4* __restrict__ a, uint32_t n_lines, uint32_t per_thread, uint32_t* out) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = mix32(tid * 2654435761u + 1u), s = 0;
    for (uint32_t k = 0; k < per_thread; ++k) {
        const uint4* p = a + (size_t)(h % n_lines) * 8;
        const uint4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3], v4 = p[4], v5 = p[5], v6 = p[6], v7 = p[7];
        s += v0.x ^ v1.y ^ v2.z ^ v3.w ^ v4.x ^ v5.y ^ v6.z ^ v7.w;
        h = mix32(h + k);
    }
    out[tid] = s;
}

}  // namespace

extern "C" int mfgb_atomic_rates(int device, uint32_t n_addr, uint32_t per_thread, double* f64_per_s,
                                 double* u64_per_s) {
    if (n_addr < 8192 || per_thread == 0 || !f64_per_s || !u64_per_s) return 1;
    if (hipSetDevice(device) != hipSuccess) return 2;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 2;
    const int block = 256, grid = prop.multiProcessorCount * 8;
    const double ops = (double)grid * block * per_thread;
    void* buf = nullptr;
    if (hipMalloc(&buf, (size_t)n_addr * 8) != hipSuccess) return 2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 2;
    float ms = 0;
    do {
        if (hipMemset(buf, 0, (size_t)n_addr * 8) != hipSuccess) break;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) break;
        // FP64: one untimed launch, then one timed
        hipLaunchKernelGGL(k_atomic_f64, dim3(grid), dim3(block), 0, 0, (double*)buf, n_addr, per_thread);
        if (hipDeviceSynchronize() != hipSuccess) break;
        if (hipEventRecord(e0, 0) != hipSuccess) break;
        hipLaunchKernelGGL(k_atomic_f64, dim3(grid), dim3(block), 0, 0, (double*)buf, n_addr, per_thread);
        if (hipEventRecord(e1, 0) != hipSuccess) break;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
        *f64_per_s = ops / (ms * 1e-3);
        // uint64 on the same counters
        hipLaunchKernelGGL(k_atomic_u64, dim3(grid), dim3(block), 0, 0, (unsigned long long*)buf, n_addr, per_thread);
        if (hipDeviceSynchronize() != hipSuccess) break;
        if (hipEventRecord(e0, 0) != hipSuccess) break;
        hipLaunchKernelGGL(k_atomic_u64, dim3(grid), dim3(block), 0, 0, (unsigned long long*)buf, n_addr, per_thread);
        if (hipEventRecord(e1, 0) != hipSuccess) break;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
        *u64_per_s = ops / (ms * 1e-3);
        rc = hipGetLastError() == hipSuccess ? 0 : 2;
    } while (false);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}

extern "C" int mfgb_gather_rate(int device, uint64_t bytes, uint32_t per_thread, double* bytes_per_s) {
    if (bytes < (1u << 20) || bytes / 128 > 0xffffffffull || per_thread == 0 || !bytes_per_s) return 1;
    if (hipSetDevice(device) != hipSuccess) return 2;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 2;
    const int block = 256, grid = prop.multiProcessorCount * 16;
    const uint32_t n_lines = (uint32_t)(bytes / 128);
    void* buf = nullptr;
    void* out = nullptr;
    if (hipMalloc(&buf, (size_t)n_lines * 128) != hipSuccess) return 2;
    if (hipMalloc(&out, (size_t)grid * block * sizeof(uint32_t)) != hipSuccess) { (void)hipFree(buf); return 2; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 2;
    float ms = 0;
    do {
        if (hipMemset(buf, 1, (size_t)n_lines * 128) != hipSuccess) break;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) break;
        hipLaunchKernelGGL(k_gather, dim3(grid), dim3(block), 0, 0, (const uint4*)buf, n_lines, per_thread, (uint32_t*)out);
        if (hipDeviceSynchronize() != hipSuccess) break;
        if (hipEventRecord(e0, 0) != hipSuccess) break;
        hipLaunchKernelGGL(k_gather, dim3(grid), dim3(block), 0, 0, (const uint4*)buf, n_lines, per_thread, (uint32_t*)out);
        if (hipEventRecord(e1, 0) != hipSuccess) break;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
        *bytes_per_s = (double)grid * block * per_thread * 128.0 / (ms * 1e-3);
        rc = hipGetLastError() == hipSuccess ? 0 : 2;
    } while (false);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    (void)hipFree(buf);
    return rc;
}
