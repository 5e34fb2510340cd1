// field.hip -- grid kernels of the step outside the Poisson solve.
//
//  charge_node (field.hpp), in charge_batch_kernel and in the first residual
//  pass of a single simulation's solve (step_residual_stage, multigrid.hip):
//                  Q20 accumulators -> charge density rho and Poisson rhs
//                  (-rho/eps0); clears the accumulators for the next step.
//                  Replaces cudaMemset(rho) + updateDensity's FP64 result
//                  (main.cu:93-102, kernels.cu:26-69).
//  efield_kernel   E = -grad(phi) into an interleaved double2 grid,
//                  computeElectricField (kernels.cu:134-191).
#include "field.hpp"

namespace pic {

__global__ __launch_bounds__(256) void efield_kernel(const double* __restrict__ phi, double2* __restrict__ E,
                                                     int nx, int ny, double dx, double dy) {
    const int i = blockIdx.x * 16 + threadIdx.x % 16;
    const int j = blockIdx.y * 16 + threadIdx.x / 16;
    if (i >= nx || j >= ny) return;
    const double* r = phi + (size_t)j * nx;
    double ex, ey;
    if (i == 0) ex = -(r[1] - r[0]) / dx;
    else if (i == nx - 1) ex = -(r[i] - r[i - 1]) / dx;
    else ex = -(r[i + 1] - r[i - 1]) / (2.0 * dx);
    if (j == 0) ey = -(phi[(size_t)nx + i] - phi[i]) / dy;
    else if (j == ny - 1) ey = -(phi[(size_t)j * nx + i] - phi[(size_t)(j - 1) * nx + i]) / dy;
    else ey = -(phi[(size_t)(j + 1) * nx + i] - phi[(size_t)(j - 1) * nx + i]) / (2.0 * dy);
    E[(size_t)j * nx + i] = make_double2(ex, ey);
}

__global__ __launch_bounds__(256) void charge_batch_kernel(const ChargeArgs* __restrict__ args,
                                                           const long long* __restrict__ nns,
                                                           const int4* __restrict__ chunks) {
    const int4 c = chunks[blockIdx.x];
    const ChargeArgs& a = args[c.x];
    const size_t i = (size_t)c.y * 256 + threadIdx.x;
    if (i >= (size_t)nns[c.x]) return;
    double r = 0.0, q = 0.0;
    for (int s = 0; s < a.ns; ++s) {
        const double v = (double)a.acc[s][i];
        r = fma(a.cfac[s], v, r);
        q = fma(a.qfac[s], v, q);
        if (a.zero) a.acc[s][i] = 0ull;
    }
    if (a.rho) a.rho[i] = q;
    a.rhs[i] = r;
}

int launch_charge_batch(const ChargeArgs* args, const long long* nns, const int4* chunks, int n_chunks,
                        hipStream_t s) {
    if (n_chunks <= 0) return 0;
    hipLaunchKernelGGL(charge_batch_kernel, dim3(n_chunks), dim3(256), 0, s, args, nns, chunks);
    PIC_HIP(hipGetLastError());
    return 0;
}

int launch_efield(const double* phi, double2* E, int nx, int ny, double dx, double dy, hipStream_t s) {
    hipLaunchKernelGGL(efield_kernel, dim3(div_up(nx, 16), div_up(ny, 16)), dim3(256), 0, s, phi, E, nx, ny, dx,
                       dy);
    PIC_HIP(hipGetLastError());
    return 0;
}

__global__ __launch_bounds__(256) void zero_interior_kernel(double* __restrict__ f, int n) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (size_t)n * n) return;
    const int i = (int)(k % n), j = (int)(k / n);
    if (i > 0 && i < n - 1 && j > 0 && j < n - 1) f[k] = 0.0;
}

int launch_zero_interior(double* f, int n, hipStream_t s) {
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(zero_interior_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, f, n);
    PIC_HIP(hipGetLastError());
    return 0;
}

// sums of squares of 256-value chunks: partial[c] = sum over a[256 c ..] of a^2
__global__ __launch_bounds__(256) void sumsq_kernel(const double* __restrict__ a, size_t n,
                                                    double* __restrict__ partial) {
    __shared__ double sh[256];
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const double v = i < n ? a[i] : 0.0;
    sh[threadIdx.x] = v * v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

int launch_sumsq(const double* a, size_t n, double* partial, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(sumsq_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, partial);
    PIC_HIP(hipGetLastError());
    return 0;
}

}  // namespace pic
