// multigrid.hip -- vertex-centred geometric multigrid for gfx950.
//
// Same V-cycle as the reference (Multigrid.cu:539-569): red-black SOR smoother
// with the odd colour first (:192-227, :501-504), residual (:269-297),
// full-weighting restriction with out-of-range terms dropped (:299-349),
// bilinear prolongation added to the fine level (:351-407), and extra sweeps on
// the coarsest level.  The arithmetic of every node update is the reference's
// expression, so an FP64 solve is bit-identical to a serial one.
//
// gfx950 structure (one launch per level and leg, no grid-wide barrier, which
// replaces the cooperative smoother of :229-267 that fails above the
// co-residency limit):
//   mg_smooth_tile         32x32 output tile + halo 2*nu(+2) staged in LDS;
//                          nu sweeps in LDS (the halo shrinks by one node per
//                          half-sweep), then residual and restriction of the
//                          coarse nodes the tile owns.
//   mg_prolong_smooth_tile prolongation-add over tile + halo 2*nu, nu sweeps.
//   mg_coarse_kernel       one workgroup runs the whole sub-cycle of the levels
//                          with n <= 33 in LDS.
//   ir_residual_kernel     FP64 residual (stored FP32 for the next FP32 cycle)
//                          and per-block sums of squares.
//   ir_correct_kernel      phi += e in FP64.
//   norm_final_kernel      fixed-order reduction of the block sums, relative
//                          residual and the convergence flag (build rule
//                          ||r||/||rhs||; the reference divides by sum rhs^2,
//                          Multigrid.cu:466).
#include "mg.hpp"
#include <algorithm>
#include <cmath>

namespace pic {

#define MG_TILE 32
#define MG_NC 33     // largest level handled by the single-workgroup coarse kernel

template <typename T>
__device__ inline T tfma(T a, T b, T c);
template <> __device__ inline float tfma(float a, float b, float c) { return fmaf(a, b, c); }
template <> __device__ inline double tfma(double a, double b, double c) { return fma(a, b, c); }

// one colour of red-black SOR over a (rows x cols) LDS array whose element
// (0,0) sits at global node (gi0, gj0); only global interior nodes change
template <typename T>
__device__ inline void rb_colour_lds(T* e, const T* r, int W, int Hh, int gi0, int gj0, int n, T h2, T omega,
                                     int colour, int tid, int nthr) {
    const int half = (W - 1) / 2;
    const T om1 = (T)1 - omega;
    for (int k = tid; k < (Hh - 2) * half; k += nthr) {
        const int lj = 1 + k / half, m = k % half;
        const int gj = gj0 + lj;
        const int li = 1 + ((colour ^ (gi0 + 1 + gj)) & 1) + 2 * m;
        if (li > W - 2) continue;
        const int gi = gi0 + li;
        if (gi <= 0 || gi >= n - 1 || gj <= 0 || gj >= n - 1) continue;
        const int idx = lj * W + li;
        const T c = e[idx];
        const T s = e[idx + W] + e[idx - W] + e[idx + 1] + e[idx - 1];
        const T nv = (T)0.25 * (s - h2 * r[idx]);
        e[idx] = tfma(omega, nv, om1 * c);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void mg_smooth_tile(const T* __restrict__ e_in, const T* __restrict__ rhs,
                                                      T* __restrict__ e_out, T* __restrict__ rhs_c, int n, T h2,
                                                      T inv_h2, T omega, int nu, int H) {
    extern __shared__ unsigned char smem_mg[];
    const int W = MG_TILE + 2 * H;
    T* e = reinterpret_cast<T*>(smem_mg);
    T* r = e + W * W;
    T* res = r + W * W;
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * MG_TILE, j0 = blockIdx.y * MG_TILE;
    const int gi0 = i0 - H, gj0 = j0 - H;
    for (int k = tid; k < W * W; k += 256) {
        const int gi = gi0 + k % W, gj = gj0 + k / W;
        T ev = (T)0, rv = (T)0;
        if (gi >= 0 && gi < n && gj >= 0 && gj < n) {
            const size_t g = (size_t)gj * n + gi;
            if (e_in) ev = e_in[g];
            rv = rhs[g];
        }
        e[k] = ev;
        r[k] = rv;
    }
    __syncthreads();
    for (int it = 0; it < nu; ++it)
        for (int colour = 1; colour >= 0; --colour) {
            rb_colour_lds(e, r, W, W, gi0, gj0, n, h2, omega, colour, tid, 256);
            __syncthreads();
        }
    for (int k = tid; k < MG_TILE * MG_TILE; k += 256) {
        const int gi = i0 + k % MG_TILE, gj = j0 + k / MG_TILE;
        if (gi < n && gj < n) e_out[(size_t)gj * n + gi] = e[(gj - gj0) * W + (gi - gi0)];
    }
    if (!rhs_c) return;
    const int RW = MG_TILE + 2;
    for (int k = tid; k < RW * RW; k += 256) {
        const int gi = i0 - 1 + k % RW, gj = j0 - 1 + k / RW;
        T v = (T)0;
        if (gi > 0 && gi < n - 1 && gj > 0 && gj < n - 1) {
            const int idx = (gj - gj0) * W + (gi - gi0);
            const T c = e[idx];
            const T lap = (e[idx - W] + e[idx + W] + e[idx - 1] + e[idx + 1] - (T)4 * c) * inv_h2;
            v = r[idx] - lap;
        }
        res[k] = v;
    }
    __syncthreads();
    const int nc = (n + 1) / 2, TC = MG_TILE / 2;
    for (int k = tid; k < TC * TC; k += 256) {
        const int ic = i0 / 2 + k % TC, jc = j0 / 2 + k / TC;
        if (ic >= nc || jc >= nc) continue;
        const int i = 2 * ic, j = 2 * jc;
        if (i >= n || j >= n) {
            rhs_c[(size_t)jc * nc + ic] = (T)0;
            continue;
        }
        const int f = (j - j0 + 1) * RW + (i - i0 + 1);
        T s = (T)4 * res[f];
        if (i > 0) s += (T)2 * res[f - 1];
        if (i < n - 1) s += (T)2 * res[f + 1];
        if (j > 0) s += (T)2 * res[f - RW];
        if (j < n - 1) s += (T)2 * res[f + RW];
        if (i > 0 && j > 0) s += res[f - RW - 1];
        if (i < n - 1 && j > 0) s += res[f - RW + 1];
        if (i > 0 && j < n - 1) s += res[f + RW - 1];
        if (i < n - 1 && j < n - 1) s += res[f + RW + 1];
        rhs_c[(size_t)jc * nc + ic] = s * (T)0.0625;
    }
}

template <typename T>
__device__ inline T prolong_val(const T* __restrict__ coarse, int nc, int gi, int gj) {
    const int ic = gi / 2, jc = gj / 2;
    const size_t c = (size_t)jc * nc + ic;
    const bool io = gi & 1, jo = gj & 1;
    if (!io && !jo) return coarse[c];
    if (io && !jo) return (ic + 1 < nc) ? (T)0.5 * (coarse[c] + coarse[c + 1]) : coarse[c];
    if (!io && jo) return (jc + 1 < nc) ? (T)0.5 * (coarse[c] + coarse[c + nc]) : coarse[c];
    return (ic + 1 < nc && jc + 1 < nc)
               ? (T)0.25 * (coarse[c] + coarse[c + 1] + coarse[c + nc] + coarse[c + nc + 1])
               : coarse[c];
}

template <typename T>
__global__ __launch_bounds__(256) void mg_prolong_smooth_tile(const T* __restrict__ e_in,
                                                              const T* __restrict__ coarse,
                                                              const T* __restrict__ rhs, T* __restrict__ e_out,
                                                              int n, T h2, T omega, int nu, int H) {
    extern __shared__ unsigned char smem_mg[];
    const int W = MG_TILE + 2 * H;
    T* e = reinterpret_cast<T*>(smem_mg);
    T* r = e + W * W;
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * MG_TILE, j0 = blockIdx.y * MG_TILE;
    const int gi0 = i0 - H, gj0 = j0 - H;
    const int nc = (n + 1) / 2;
    for (int k = tid; k < W * W; k += 256) {
        const int gi = gi0 + k % W, gj = gj0 + k / W;
        T ev = (T)0, rv = (T)0;
        if (gi >= 0 && gi < n && gj >= 0 && gj < n) {
            const size_t g = (size_t)gj * n + gi;
            ev = e_in[g];
            rv = rhs[g];
            if (gi / 2 < nc && gj / 2 < nc) ev = ev + prolong_val(coarse, nc, gi, gj);
        }
        e[k] = ev;
        r[k] = rv;
    }
    __syncthreads();
    for (int it = 0; it < nu; ++it)
        for (int colour = 1; colour >= 0; --colour) {
            rb_colour_lds(e, r, W, W, gi0, gj0, n, h2, omega, colour, tid, 256);
            __syncthreads();
        }
    for (int k = tid; k < MG_TILE * MG_TILE; k += 256) {
        const int gi = i0 + k % MG_TILE, gj = j0 + k / MG_TILE;
        if (gi < n && gj < n) e_out[(size_t)gj * n + gi] = e[(gj - gj0) * W + (gi - gi0)];
    }
}

// ---- the single-workgroup sub-cycle for the small levels ---------------------
template <typename T>
__device__ inline void lds_smooth(T* e, const T* r, int n, T h, T omega, int iters) {
    const T h2 = h * h;
    for (int it = 0; it < iters; ++it)
        for (int colour = 1; colour >= 0; --colour) {
            rb_colour_lds(e, r, n, n, 0, 0, n, h2, omega, colour, threadIdx.x, blockDim.x);
            __syncthreads();
        }
}

template <typename T>
__global__ __launch_bounds__(1024) void mg_coarse_kernel(const T* __restrict__ e_in, const T* __restrict__ rhs_top,
                                                         T* __restrict__ e_out, int n_top, int Lc, T h_top, int pre,
                                                         int post, int csweeps, T omega) {
    extern __shared__ unsigned char smem_mg[];
    T* base = reinterpret_cast<T*>(smem_mg);
    T* E[16];
    T* R[16];
    int N[16];
    int off = 0, n = n_top;
    for (int l = 0; l < Lc; ++l) {
        N[l] = n;
        E[l] = base + off;
        off += n * n;
        R[l] = base + off;
        off += n * n;
        n = (n + 1) / 2;
    }
    T* res = base + off;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k < n_top * n_top; k += nt) {
        E[0][k] = e_in ? e_in[k] : (T)0;
        R[0][k] = rhs_top[k];
    }
    __syncthreads();
    T h = h_top;
    for (int l = 0; l < Lc - 1; ++l) {
        const int nf = N[l], ncs = N[l + 1];
        lds_smooth(E[l], R[l], nf, h, omega, pre);
        const T inv_h2 = (T)1 / (h * h);
        for (int k = tid; k < nf * nf; k += nt) {
            const int i = k % nf, j = k / nf;
            T v = (T)0;
            if (!(i == 0 || i == nf - 1 || j == 0 || j == nf - 1)) {
                const T* e = E[l];
                const T lap = (e[k - nf] + e[k + nf] + e[k - 1] + e[k + 1] - (T)4 * e[k]) * inv_h2;
                v = R[l][k] - lap;
            }
            res[k] = v;
        }
        __syncthreads();
        for (int k = tid; k < ncs * ncs; k += nt) {
            const int ic = k % ncs, jc = k / ncs;
            const int i = 2 * ic, j = 2 * jc;
            T s = (T)0;
            if (i < nf && j < nf) {
                const int f = j * nf + i;
                s = (T)4 * res[f];
                if (i > 0) s += (T)2 * res[f - 1];
                if (i < nf - 1) s += (T)2 * res[f + 1];
                if (j > 0) s += (T)2 * res[f - nf];
                if (j < nf - 1) s += (T)2 * res[f + nf];
                if (i > 0 && j > 0) s += res[f - nf - 1];
                if (i < nf - 1 && j > 0) s += res[f - nf + 1];
                if (i > 0 && j < nf - 1) s += res[f + nf - 1];
                if (i < nf - 1 && j < nf - 1) s += res[f + nf + 1];
                s = s * (T)0.0625;
            }
            R[l + 1][k] = s;
            E[l + 1][k] = (T)0;
        }
        __syncthreads();
        h = h * (T)2;
    }
    lds_smooth(E[Lc - 1], R[Lc - 1], N[Lc - 1], h, omega, csweeps);
    for (int l = Lc - 2; l >= 0; --l) {
        h = h / (T)2;
        const int nf = N[l], ncs = N[l + 1];
        for (int k = tid; k < nf * nf; k += nt) {
            const int i = k % nf, j = k / nf;
            if (i / 2 < ncs && j / 2 < ncs) E[l][k] = E[l][k] + prolong_val(E[l + 1], ncs, i, j);
        }
        __syncthreads();
        lds_smooth(E[l], R[l], nf, h, omega, post);
    }
    for (int k = tid; k < n_top * n_top; k += nt) e_out[k] = E[0][k];
}

// ---- FP64 refinement and norms --------------------------------------------------
__device__ inline double block_sum256(double v, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = sh[tid] + sh[tid + s];
        __syncthreads();
    }
    return sh[0];
}

template <bool WRITE_R32>
__global__ __launch_bounds__(256) void ir_residual_kernel(const double* __restrict__ phi,
                                                          const double* __restrict__ f, float* __restrict__ r32,
                                                          double* __restrict__ partial, int n, double inv_h2) {
    __shared__ double sh[256];
    const int i = blockIdx.x * 16 + threadIdx.x % 16, j = blockIdx.y * 16 + threadIdx.x / 16;
    double r = 0.0;
    if (i < n && j < n) {
        const size_t idx = (size_t)j * n + i;
        if (!(i == 0 || i == n - 1 || j == 0 || j == n - 1)) {
            const double c = phi[idx];
            const double lap = (phi[idx - n] + phi[idx + n] + phi[idx - 1] + phi[idx + 1] - 4.0 * c) * inv_h2;
            r = f[idx] - lap;
        }
        if (WRITE_R32) r32[idx] = (float)r;
    }
    const double s = block_sum256(r * r, sh);
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void sumsq_kernel(const double* __restrict__ v, size_t nn,
                                                    double* __restrict__ partial) {
    __shared__ double sh[256];
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const double x = (i < nn) ? v[i] : 0.0;
    const double s = block_sum256(x * x, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void ir_correct_kernel(double* __restrict__ phi, const float* __restrict__ e,
                                                         size_t nn) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nn) phi[i] = phi[i] + (double)e[i];
}

// which = 0: store the rhs norm; which = 1: residual norm, rel and flag
__global__ __launch_bounds__(256) void norm_final_kernel(const double* __restrict__ partial, int np, MGNorm* out,
                                                         int which, double tol) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < np; k += 256) acc += partial[k];
    const double s = block_sum256(acc, sh);
    if (threadIdx.x == 0) {
        if (which == 0) {
            out->rhs_sq = s;
        } else {
            out->res_sq = s;
            const double rn = sqrt(s), fn = sqrt(out->rhs_sq);
            const double rel = fn > 0.0 ? rn / fn : rn;
            out->rel = rel;
            out->converged = rel < tol ? 1 : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
MGSolver::MGSolver(int n, const MGParams& p, hipStream_t s) : n_(n), L_(0), lc_(0), p_(p), stream_(s) {}

MGSolver::~MGSolver() {
    for (int b = 0; b < 2; ++b) {
        for (auto* q : ef_[b]) if (q) (void)hipFree(q);
        for (auto* q : ed_[b]) if (q) (void)hipFree(q);
    }
    for (auto* q : rf_) if (q) (void)hipFree(q);
    for (auto* q : rd_) if (q) (void)hipFree(q);
    if (r32_) (void)hipFree(r32_);
    if (phi_tmp_) (void)hipFree(phi_tmp_);
    if (partials_) (void)hipFree(partials_);
    if (d_norm_) (void)hipFree(d_norm_);
    if (h_norm_) (void)hipHostFree(h_norm_);
}

int MGSolver::init() {
    PIC_REQUIRE(n_ >= 3, "multigrid: grid must have at least 3 nodes per side");
    int L = 1, m = n_;
    if (p_.levels > 0) {
        L = p_.levels;
    } else {
        while (m > 3) { m = (m + 1) / 2; ++L; }
    }
    ns_.assign(L, 0);
    ns_[0] = n_;
    for (int l = 1; l < L; ++l) {
        PIC_REQUIRE(ns_[l - 1] % 2 == 1, "multigrid: every level above the coarsest needs an odd node count");
        ns_[l] = (ns_[l - 1] + 1) / 2;
        PIC_REQUIRE(ns_[l] >= 2, "multigrid: too many levels for this grid");
    }
    PIC_REQUIRE(L <= 16, "multigrid: at most 16 levels");
    PIC_REQUIRE(p_.pre >= 0 && p_.pre <= 3 && p_.post >= 0 && p_.post <= 3,
                "multigrid: pre/post sweeps must be in [0,3]");
    L_ = L;
    lc_ = L - 1;
    for (int l = 0; l < L; ++l)
        if (ns_[l] <= MG_NC) { lc_ = l; break; }
    for (int b = 0; b < 2; ++b) {
        ef_[b].assign(L, nullptr);
        ed_[b].assign(L, nullptr);
    }
    rf_.assign(L, nullptr);
    rd_.assign(L, nullptr);
    const bool mixed = p_.mode == PIC_MG_MIXED;
    for (int l = 0; l < L; ++l) {
        const size_t sz = (size_t)ns_[l] * ns_[l];
        for (int b = 0; b < 2; ++b) {
            if (mixed) PIC_HIP(hipMalloc(&ef_[b][l], sz * sizeof(float)));
            else if (l > 0) PIC_HIP(hipMalloc(&ed_[b][l], sz * sizeof(double)));
        }
        if (l > 0) {
            if (mixed) PIC_HIP(hipMalloc(&rf_[l], sz * sizeof(float)));
            else PIC_HIP(hipMalloc(&rd_[l], sz * sizeof(double)));
        }
    }
    const size_t nn = (size_t)n_ * n_;
    if (mixed) PIC_HIP(hipMalloc(&r32_, nn * sizeof(float)));
    else PIC_HIP(hipMalloc(&phi_tmp_, nn * sizeof(double)));
    const int gb = div_up(n_, 16);
    n_partials_ = std::max(gb * gb, (int)((nn + 255) / 256));
    PIC_HIP(hipMalloc(&partials_, (size_t)n_partials_ * sizeof(double)));
    PIC_HIP(hipMalloc(&d_norm_, sizeof(MGNorm)));
    PIC_HIP(hipMemset(d_norm_, 0, sizeof(MGNorm)));
    PIC_HIP(hipHostMalloc(&h_norm_, sizeof(MGNorm), hipHostMallocDefault));
    ok_ = true;
    return 0;
}

template <typename T>
static size_t coarse_lds_bytes(int n_top, int Lc) {
    size_t tot = 0;
    int n = n_top;
    for (int l = 0; l < Lc; ++l) { tot += 2 * (size_t)n * n; n = (n + 1) / 2; }
    tot += (size_t)n_top * n_top;
    return tot * sizeof(T);
}

template <typename T>
int MGSolver::vcycle(const T* e0_in, T* e0_out, T* e0_tmp, const T* rhs0, double h) {
    const bool mixed = p_.mode == PIC_MG_MIXED;
    auto E = [&](int b, int l) -> T* { return (T*)(mixed ? (void*)ef_[b][l] : (void*)ed_[b][l]); };
    auto R = [&](int l) -> T* { return (T*)(mixed ? (void*)rf_[l] : (void*)rd_[l]); };
    const T omega = (T)p_.omega;
    std::vector<double> hs(L_);
    hs[0] = h;
    for (int l = 1; l < L_; ++l) hs[l] = hs[l - 1] * 2.0;
    const int lc = lc_;
    // down leg
    for (int l = 0; l < lc; ++l) {
        const int n = ns_[l];
        const int H = 2 * p_.pre + 2;
        const int W = MG_TILE + 2 * H;
        const size_t lds = (2 * (size_t)W * W + (MG_TILE + 2) * (MG_TILE + 2)) * sizeof(T);
        const T* ein = (l == 0) ? e0_in : nullptr;
        const T* rhs = (l == 0) ? rhs0 : R(l);
        T* eout = (l == 0) ? e0_tmp : E(1, l);
        const double hh = hs[l];
        hipLaunchKernelGGL(mg_smooth_tile<T>, dim3(div_up(n, MG_TILE), div_up(n, MG_TILE)), dim3(256), lds, stream_,
                           ein, rhs, eout, R(l + 1), n, (T)(hh * hh), (T)(1.0 / (hh * hh)), omega, p_.pre, H);
        ++launches_;
    }
    // coarse part
    {
        const int n = ns_[lc];
        const T* ein = (lc == 0) ? e0_in : nullptr;
        const T* rhs = (lc == 0) ? rhs0 : R(lc);
        T* eout = (lc == 0) ? e0_out : E(0, lc);
        if (n <= MG_NC) {
            const size_t lds = coarse_lds_bytes<T>(n, L_ - lc);
            hipLaunchKernelGGL(mg_coarse_kernel<T>, dim3(1), dim3(1024), lds, stream_, ein, rhs, eout, n, L_ - lc,
                               (T)hs[lc], p_.pre, p_.post, p_.coarse_sweeps, omega);
            ++launches_;
        } else {
            // a large coarsest level: tiled sweeps in chunks of up to 3, ping-pong
            T* bufs[2] = {eout, (lc == 0) ? e0_tmp : E(1, lc)};
            const T* src = ein;
            // never write the buffer a launch reads (tiles read their neighbours' halos)
            int left = p_.coarse_sweeps, which = (src == bufs[0]) ? 1 : 0;
            bool wrote = false;
            const double hh = hs[lc];
            while (left > 0 || !wrote) {
                const int nu = std::min(left, 3);
                const int H = 2 * nu;
                const int W = MG_TILE + 2 * H;
                const size_t lds = (2 * (size_t)W * W + (MG_TILE + 2) * (MG_TILE + 2)) * sizeof(T);
                hipLaunchKernelGGL(mg_smooth_tile<T>, dim3(div_up(n, MG_TILE), div_up(n, MG_TILE)), dim3(256), lds,
                                   stream_, src, rhs, bufs[which], (T*)nullptr, n, (T)(hh * hh),
                                   (T)(1.0 / (hh * hh)), omega, nu, H);
                ++launches_;
                src = bufs[which];
                which ^= 1;
                left -= nu;
                wrote = true;
            }
            if (src != eout) {
                PIC_HIP(hipMemcpyAsync(eout, src, (size_t)n * n * sizeof(T), hipMemcpyDeviceToDevice, stream_));
            }
        }
    }
    // up leg
    for (int l = lc - 1; l >= 0; --l) {
        const int n = ns_[l];
        const int H = 2 * p_.post;
        const int W = MG_TILE + 2 * H;
        const size_t lds = 2 * (size_t)W * W * sizeof(T);
        const T* ein = (l == 0) ? e0_tmp : E(1, l);
        const T* rhs = (l == 0) ? rhs0 : R(l);
        T* eout = (l == 0) ? e0_out : E(0, l);
        const double hh = hs[l];
        hipLaunchKernelGGL(mg_prolong_smooth_tile<T>, dim3(div_up(n, MG_TILE), div_up(n, MG_TILE)), dim3(256), lds,
                           stream_, ein, E(0, l + 1), rhs, eout, n, (T)(hh * hh), omega, p_.post, H);
        ++launches_;
    }
    PIC_HIP(hipGetLastError());
    return 0;
}

int MGSolver::reduce(const double* partials, int np, int which) {
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(256), 0, stream_, partials, np, d_norm_, which, p_.tol);
    ++launches_;
    return 0;
}

int MGSolver::solve(double* phi, const double* rhs, double h, MGStats* st) {
    PIC_REQUIRE(ok_, "multigrid: solver not initialised");
    launches_ = 0;
    const size_t nn = (size_t)n_ * n_;
    const int gb = div_up(n_, 16);
    const double inv_h2 = 1.0 / (h * h);
    const bool mixed = p_.mode == PIC_MG_MIXED;
    // ||rhs||
    hipLaunchKernelGGL(sumsq_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, stream_, rhs, nn, partials_);
    ++launches_;
    reduce(partials_, (int)((nn + 255) / 256), 0);
    auto residual = [&]() {
        if (mixed)
            hipLaunchKernelGGL(ir_residual_kernel<true>, dim3(gb, gb), dim3(256), 0, stream_, phi, rhs, r32_,
                               partials_, n_, inv_h2);
        else
            hipLaunchKernelGGL(ir_residual_kernel<false>, dim3(gb, gb), dim3(256), 0, stream_, phi, rhs,
                               (float*)nullptr, partials_, n_, inv_h2);
        ++launches_;
        reduce(partials_, gb * gb, 1);
    };
    auto fetch = [&]() -> int {
        PIC_HIP(hipMemcpyAsync(h_norm_, d_norm_, sizeof(MGNorm), hipMemcpyDeviceToHost, stream_));
        PIC_HIP(hipStreamSynchronize(stream_));
        return 0;
    };
    residual();
    if (int rc = fetch()) return rc;
    int cycles = 0;
    while (!h_norm_->converged && cycles < p_.max_cycles) {
        if (mixed) {
            if (int rc = vcycle<float>(nullptr, ef_[0][0], ef_[1][0], r32_, h)) return rc;
            hipLaunchKernelGGL(ir_correct_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, stream_, phi,
                               ef_[0][0], nn);
            ++launches_;
        } else {
            if (int rc = vcycle<double>(phi, phi, phi_tmp_, rhs, h)) return rc;
        }
        residual();
        if (int rc = fetch()) return rc;
        ++cycles;
    }
    PIC_HIP(hipGetLastError());
    if (st) {
        st->cycles = cycles;
        st->rel_residual = h_norm_->rel;
        st->rhs_norm = std::sqrt(h_norm_->rhs_sq);
    }
    return 0;
}

}  // namespace pic
