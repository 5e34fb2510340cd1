// copy_probe.hip -- plain device copy of a known size, the calibration for
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters: one launch reads and writes
// exactly `bytes` bytes, so counter / bytes is the scale factor applied to the
// counters of the PIC kernels.
//
//   copy_probe [MiB]     (default 1024 MiB, 5 launches)
//   copy_probe --bins N PER
//                        the push kernel's access pattern without its work: N
//                        particles in five SoA arrays (x, y, vx, vy, id; 4 B
//                        each) copied to five others by one 256-thread
//                        workgroup per PER consecutive particles; 40 B per
//                        particle, the practical ceiling of a binned pass
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <string>

__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                         size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void bin_copy_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ vx, const float* __restrict__ vy,
                                                       const unsigned* __restrict__ id, float* __restrict__ ox,
                                                       float* __restrict__ oy, float* __restrict__ ovx,
                                                       float* __restrict__ ovy, unsigned* __restrict__ oid, int n,
                                                       int per) {
    const int b0 = blockIdx.x * per;
    const int end = min(n, b0 + per);
    for (int i = b0 + (int)threadIdx.x; i < end; i += 256) {
        const float a = x[i], b = y[i], c = vx[i], d = vy[i];
        const unsigned e = id[i];
        ox[i] = a;
        oy[i] = b;
        ovx[i] = c;
        ovy[i] = d;
        oid[i] = e;
    }
}

static int bin_copy(int n, int per) {
    void* arr[10];
    for (auto& q : arr) {
        if (hipMalloc(&q, (size_t)n * 4) != hipSuccess || hipMemset(q, 0, (size_t)n * 4) != hipSuccess) {
            fprintf(stderr, "copy_probe: allocation failed\n");
            return 1;
        }
    }
    const int blocks = (n + per - 1) / per;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int it = 0; it < 6; ++it) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(bin_copy_kernel, dim3(blocks), dim3(256), 0, 0, (const float*)arr[0], (const float*)arr[1],
                           (const float*)arr[2], (const float*)arr[3], (const unsigned*)arr[4], (float*)arr[5],
                           (float*)arr[6], (float*)arr[7], (float*)arr[8], (unsigned*)arr[9], n, per);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (it > 0 && ms < best) best = ms;
    }
    printf("copy_probe --bins: %d particles, %d per workgroup, 40 B each, best %.3f ms, %.1f GB/s\n", n, per, best,
           40.0 * n / (best * 1e-3) / 1e9);
    for (auto& q : arr) (void)hipFree(q);
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 3 && std::string(argv[1]) == "--bins") return bin_copy(atoi(argv[2]), atoi(argv[3]));
    const size_t mib = argc > 1 ? (size_t)atol(argv[1]) : 1024;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "copy_probe: allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int it = 0; it < 5; ++it) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    printf("copy_probe: %zu bytes read + %zu bytes written per launch, best %.3f ms, %.1f GB/s\n", bytes, bytes,
           best, 2.0 * bytes / (best * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
