// probe.hip -- bandwidth probes used only by bench.py to put measured peaks
// next to the published ones: LDS read+write bandwidth of the whole chip and
// an HBM stream copy.  Built into libclifft_probe.so, separate from the
// executor.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double dv2 __attribute__((ext_vector_type(2)));

// Each lane streams 16-byte elements through its own conflict-free LDS
// columns (lane-consecutive addresses), one read and one write per element.
__global__ void __launch_bounds__(256) lds_probe_kernel(double* out, int iters) {
    __shared__ dv2 buf[256 * 16];  // 64 KiB: two workgroups per CU
    const uint32_t t = threadIdx.x;
    for (int j = 0; j < 16; ++j) {
        dv2 v;
        v.x = (double)(t + j);
        v.y = 1.0;
        buf[t + 256 * j] = v;
    }
    __syncthreads();
    dv2 acc;
    acc.x = 0.0;
    acc.y = 0.0;
    for (int it = 0; it < iters; ++it) {
#pragma unroll 8
        for (int j = 0; j < 16; ++j) {
            dv2 v = buf[t + 256 * j];
            acc += v;
            v.y += 1.0;
            buf[t + 256 * j] = v;
        }
    }
    if (acc.x == -1.0) out[blockIdx.x * 256 + t] = acc.y;  // keeps the loop alive
}

// Grid-stride copy with four independent 16-byte loads in flight per lane
// (all four issued before the stores), lane-consecutive addresses.
__global__ void __launch_bounds__(256) copy_probe_kernel(const dv2* __restrict__ src, dv2* __restrict__ dst,
                                                         uint64_t n) {
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * T < n; i += 4 * T) {
        const dv2 a = src[i], b = src[i + T], c = src[i + 2 * T], d = src[i + 3 * T];
        dst[i] = a;
        dst[i + T] = b;
        dst[i + 2 * T] = c;
        dst[i + 3 * T] = d;
    }
    for (; i < n; i += T) dst[i] = src[i];
}

// Grid-stride read of n 16-byte elements (U loads in flight per lane), one
// sum per lane written at the end so the loads stay live.
template <int U>
__global__ void __launch_bounds__(256) read_probe_kernel(const dv2* __restrict__ src, dv2* __restrict__ out, uint64_t n) {
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    dv2 acc;
    acc.x = 0.0;
    acc.y = 0.0;
    for (; i + (U - 1) * T < n; i += U * T) {
        dv2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[i + u * T];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    for (; i < n; i += T) acc += src[i];
    out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Grid-stride write of n 16-byte elements.
__global__ void __launch_bounds__(256) write_probe_kernel(dv2* __restrict__ dst, uint64_t n) {
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    dv2 v;
    v.x = 1.0;
    v.y = 2.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += T) dst[i] = v;
}

// Copy with U independent loads in flight per lane.
template <int U>
__global__ void __launch_bounds__(256) copy_probe_u_kernel(const dv2* __restrict__ src, dv2* __restrict__ dst, uint64_t n) {
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * T < n; i += U * T) {
        dv2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[i + u * T];
#pragma unroll
        for (int u = 0; u < U; ++u) dst[i + u * T] = v[u];
    }
    for (; i < n; i += T) dst[i] = src[i];
}

extern "C" {

// GB/s of one of: mode 0 read-only (8 loads in flight per lane), 1
// write-only, 2 copy with 8 loads in flight; `bytes` per buffer, `blocks`
// workgroups of 256 threads.  Negative on a HIP error.
double clifft_probe_stream(int mode, uint64_t bytes, int blocks, int reps) {
    const uint64_t n = bytes / 16;
    dv2 *a = nullptr, *b = nullptr;
    const uint64_t nb = mode == 0 ? (uint64_t)blocks * 256 : n;
    if (hipMalloc(&a, n * 16) != hipSuccess) return -1.0;
    if (hipMalloc(&b, nb * 16) != hipSuccess) {
        (void)hipFree(a);
        return -1.0;
    }
    (void)hipMemset(a, 0, n * 16);
    const dim3 grid(blocks), block(256);
    auto launch = [&]() {
        if (mode == 0) hipLaunchKernelGGL(read_probe_kernel<8>, grid, block, 0, 0, a, b, n);
        else if (mode == 1) hipLaunchKernelGGL(write_probe_kernel, grid, block, 0, 0, b, n);
        else hipLaunchKernelGGL(copy_probe_u_kernel<8>, grid, block, 0, 0, a, b, n);
    };
    launch();
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, 0);
    const hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return -2.0;
    const double moved = (mode == 2 ? 2.0 : 1.0) * (double)(n * 16) * reps;
    return moved / (ms * 1e-3) / 1e9;
}

// Returns GB/s of LDS traffic (bytes read + written) over `blocks`
// workgroups of 256 threads, or a negative HIP error code.
double clifft_probe_lds(int blocks, int iters) {
    double* out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)) != hipSuccess) return -1.0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(lds_probe_kernel, dim3(blocks), dim3(256), 0, 0, out, iters);  // warm-up
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(lds_probe_kernel, dim3(blocks), dim3(256), 0, 0, out, iters);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess) return -2.0;
    const double bytes = (double)blocks * 256.0 * 16.0 * 32.0 * (double)iters;  // 16 B read + 16 B write
    return bytes / (ms * 1e-3) / 1e9;
}

// Returns GB/s (read + write) of a device-to-device copy of `bytes`.
double clifft_probe_copy(uint64_t bytes, int reps) {
    const uint64_t n = bytes / 16;
    dv2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return -1.0;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return -1.0;
    }
    (void)hipMemset(a, 0, n * 16);
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const dim3 grid(cus * 8), block(256);
    hipLaunchKernelGGL(copy_probe_kernel, grid, block, 0, 0, a, b, n);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_probe_kernel, grid, block, 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return -2.0;
    return 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
}

}  // extern "C"
