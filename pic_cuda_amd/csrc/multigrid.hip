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
#include <cstring>

namespace pic {

#define MG_TILE 32
// largest level handled by the single-workgroup coarse kernel (its LDS stays
// below 64 KB: FP32 65^2, FP64 33^2)
#define MG_NC_F32 65
#define MG_NC_F64 33

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
__global__ __launch_bounds__(256) void mg_smooth_stage(const MGLevelTask<T>* __restrict__ tasks,
                                                       const MGTile* __restrict__ tiles) {
    const MGTile tile = tiles[blockIdx.x];
    const MGLevelTask<T>& tk = tasks[tile.task];
    if (tk.gate && tk.gate->converged) return;   // cycle enqueued after convergence: no-op
    const T* __restrict__ e_in = tk.e_in;
    const T* __restrict__ rhs = tk.rhs;
    T* __restrict__ e_out = tk.e_out;
    T* __restrict__ rhs_c = tk.rhs_c;
    const int n = tk.n, nu = tk.nu, H = tk.H;
    const T h2 = tk.h2, inv_h2 = tk.inv_h2, omega = tk.omega;
    extern __shared__ unsigned char smem_mg[];
    const int W = MG_TILE + 2 * H;
    T* e = reinterpret_cast<T*>(smem_mg);
    T* r = e + W * W;
    T* res = r + W * W;
    const int tid = threadIdx.x;
    const int i0 = tile.tx * MG_TILE, j0 = tile.ty * MG_TILE;
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
__global__ __launch_bounds__(256) void mg_prolong_stage(const MGLevelTask<T>* __restrict__ tasks,
                                                        const MGTile* __restrict__ tiles) {
    const MGTile tile = tiles[blockIdx.x];
    const MGLevelTask<T>& tk = tasks[tile.task];
    if (tk.gate && tk.gate->converged) return;
    const T* __restrict__ e_in = tk.e_in;
    const T* __restrict__ coarse = tk.coarse;
    const T* __restrict__ rhs = tk.rhs;
    T* __restrict__ e_out = tk.e_out;
    const int n = tk.n, nu = tk.nu, H = tk.H;
    const T h2 = tk.h2, omega = tk.omega;
    extern __shared__ unsigned char smem_mg[];
    const int W = MG_TILE + 2 * H;
    T* e = reinterpret_cast<T*>(smem_mg);
    T* r = e + W * W;
    const int tid = threadIdx.x;
    const int i0 = tile.tx * MG_TILE, j0 = tile.ty * MG_TILE;
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
__global__ __launch_bounds__(1024) void mg_coarse_stage(const MGCoarseTask<T>* __restrict__ tasks) {
    const MGCoarseTask<T>& tk = tasks[blockIdx.x];
    if (tk.gate && tk.gate->converged) return;
    const T* __restrict__ e_in = tk.e_in;
    const T* __restrict__ rhs_top = tk.rhs;
    T* __restrict__ e_out = tk.e_out;
    const int n_top = tk.n_top, Lc = tk.Lc, pre = tk.pre, post = tk.post, csweeps = tk.csweeps;
    const T h_top = tk.h_top, omega = tk.omega;
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

// FP64 residual r = rhs - A phi (zero on the boundary), optionally stored in
// FP32 as the next cycle's rhs, and per-block sums of r^2.  which == 2: skipped
// once the case has stopped.
template <bool WRITE_R32>
__global__ __launch_bounds__(256) void ir_residual_stage(const MGFineTask* __restrict__ tasks,
                                                         const MGTile* __restrict__ tiles, int which) {
    const MGTile tile = tiles[blockIdx.x];
    const MGFineTask& tk = tasks[tile.task];
    if (which == 2 && tk.norm->converged) return;
    __shared__ double sh[256];
    const int n = tk.n;
    const double* __restrict__ phi = tk.phi;
    const int i = tile.tx * 16 + threadIdx.x % 16, j = tile.ty * 16 + threadIdx.x / 16;
    double r = 0.0;
    if (i < n && j < n) {
        const size_t idx = (size_t)j * n + i;
        if (!(i == 0 || i == n - 1 || j == 0 || j == n - 1)) {
            const double c = phi[idx];
            const double lap = (phi[idx - n] + phi[idx + n] + phi[idx - 1] + phi[idx + 1] - 4.0 * c) * tk.inv_h2;
            r = tk.rhs[idx] - lap;
        }
        if (WRITE_R32) tk.r32[idx] = (float)r;
    }
    const double s = block_sum256(r * r, sh);
    if (threadIdx.x == 0) tk.partial[tile.ty * tk.gx + tile.tx] = s;
}

__global__ __launch_bounds__(256) void sumsq_stage(const MGFineTask* __restrict__ tasks,
                                                   const MGTile* __restrict__ tiles) {
    const MGTile tile = tiles[blockIdx.x];
    const MGFineTask& tk = tasks[tile.task];
    __shared__ double sh[256];
    const size_t nn = (size_t)tk.n * tk.n;
    const size_t i = (size_t)tile.tx * 256 + threadIdx.x;
    const double x = (i < nn) ? tk.rhs[i] : 0.0;
    const double s = block_sum256(x * x, sh);
    if (threadIdx.x == 0) tk.partial[tile.tx] = s;
}

__global__ __launch_bounds__(256) void ir_correct_stage(const MGFineTask* __restrict__ tasks,
                                                        const MGTile* __restrict__ tiles) {
    const MGTile tile = tiles[blockIdx.x];
    const MGFineTask& tk = tasks[tile.task];
    if (tk.norm->converged) return;
    const size_t nn = (size_t)tk.n * tk.n;
    const size_t i = (size_t)tile.tx * 256 + threadIdx.x;
    if (i < nn) tk.phi[i] = tk.phi[i] + (double)tk.e32[i];
}

// One workgroup per case: fixed-order sum of the block partials.  which = 0:
// start of a solve (rhs norm, clear the flag and the cycle count); 1: initial
// residual test; 2: residual test at the end of a cycle, counted, and a no-op
// once the flag is set, so cycles enqueued after convergence change nothing.
// Relative residual ||r||/||rhs|| (the reference divides by sum rhs^2,
// Multigrid.cu:466).
__global__ __launch_bounds__(256) void norm_final_stage(const MGFineTask* __restrict__ tasks, int which, double tol,
                                                        int max_cycles) {
    const MGFineTask& tk = tasks[blockIdx.x];
    MGNorm* out = tk.norm;
    if (which == 2 && out->converged) return;
    __shared__ double sh[256];
    const int np = which == 0 ? tk.nchunks : tk.nblocks;
    double acc = 0.0;
    for (int k = threadIdx.x; k < np; k += 256) acc += tk.partial[k];
    const double tot = block_sum256(acc, sh);
    if (threadIdx.x == 0) {
        if (which == 0) {
            out->rhs_sq = tot;
            out->converged = 0;
            out->cycles = 0;
            out->max_cycles = max_cycles;
        } else {
            out->res_sq = tot;
            const double rn = sqrt(tot), fn = sqrt(out->rhs_sq);
            const double rel = fn > 0.0 ? rn / fn : rn;
            out->rel = rel;
            if (which == 2) out->cycles += 1;
            out->converged = (rel < tol || out->cycles >= out->max_cycles) ? 1 : 0;
            out->hit_tol = rel < tol ? 1 : 0;
        }
        *tk.mirror = *out;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void mg_copy_stage(const MGLevelTask<T>* __restrict__ tasks,
                                                     const MGTile* __restrict__ tiles) {
    const MGTile tile = tiles[blockIdx.x];
    const MGLevelTask<T>& tk = tasks[tile.task];
    if (tk.gate && tk.gate->converged) return;
    const size_t nn = (size_t)tk.n * tk.n;
    const size_t i = (size_t)tile.tx * 256 + threadIdx.x;
    if (i < nn) tk.e_out[i] = tk.e_in[i];
}

// ---------------------------------------------------------------------------
// MGSolver: the pyramid of one grid
// ---------------------------------------------------------------------------
MGSolver::MGSolver(int n, const MGParams& p, hipStream_t s) : n_(n), L_(0), lc_(0), p_(p), stream_(s) {}

MGSolver::~MGSolver() {
    delete solo_;
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
}

int MGSolver::init() {
    PIC_REQUIRE(n_ >= 3, "multigrid: grid must have at least 3 nodes per side");
    int L = 1, m = n_;
    if (p_.levels > 0) {
        L = p_.levels;
    } else {
        // coarsen while the level has an odd node count above 3; the coarsest
        // level must then be 3x3 (one unknown) or 4x4 (2x2 unknowns)
        while (m > 3 && (m & 1)) { m = (m + 1) / 2; ++L; }
        PIC_REQUIRE(m == 3 || m == 4, "multigrid: with automatic levels the grid needs 2^k*{1,3}+1 nodes per side");
    }
    PIC_REQUIRE(L <= 16, "multigrid: at most 16 levels");
    ns_.assign(L, 0);
    ns_[0] = n_;
    for (int l = 1; l < L; ++l) {
        PIC_REQUIRE(ns_[l - 1] % 2 == 1, "multigrid: every level above the coarsest needs an odd node count");
        ns_[l] = (ns_[l - 1] + 1) / 2;
        PIC_REQUIRE(ns_[l] >= 2, "multigrid: too many levels for this grid");
    }
    PIC_REQUIRE(p_.pre >= 0 && p_.pre <= 3 && p_.post >= 0 && p_.post <= 3,
                "multigrid: pre/post sweeps must be in [0,3]");
    PIC_REQUIRE(p_.coarse_sweeps >= 0, "multigrid: coarse sweeps must be non-negative");
    L_ = L;
    lc_ = L - 1;
    nc_ = p_.mode == PIC_MG_MIXED ? MG_NC_F32 : MG_NC_F64;
    for (int l = 0; l < L; ++l)
        if (ns_[l] <= nc_) { lc_ = l; break; }
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
    ok_ = true;
    return 0;
}

int MGSolver::solve(double* phi, const double* rhs, double h, MGStats* st) {
    PIC_REQUIRE(ok_, "multigrid: solver not initialised");
    if (!solo_ || phi != solo_phi_ || rhs != solo_rhs_ || h != solo_h_) {
        delete solo_;
        solo_ = new MGBatch(stream_);
        if (int rc = solo_->build({this}, {phi}, {rhs}, {h})) {
            delete solo_;
            solo_ = nullptr;
            return rc;
        }
        solo_phi_ = phi;
        solo_rhs_ = rhs;
        solo_h_ = h;
    }
    if (int rc = solo_->solve()) return rc;
    launches_ = solo_->last_launches();
    if (st) {
        const MGNorm& nm = solo_->norm(0);
        st->cycles = nm.cycles;
        st->rel_residual = nm.rel;
        st->rhs_norm = std::sqrt(nm.rhs_sq);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// MGBatch: one launch per stage for all member grids
// ---------------------------------------------------------------------------
template <typename T>
static size_t coarse_lds_bytes(int n_top, int Lc) {
    size_t tot = 0;
    int n = n_top;
    for (int l = 0; l < Lc; ++l) { tot += 2 * (size_t)n * n; n = (n + 1) / 2; }
    tot += (size_t)n_top * n_top;
    return tot * sizeof(T);
}

static size_t smooth_lds(int H, size_t es) {
    const size_t W = MG_TILE + 2 * H;
    return (2 * W * W + (MG_TILE + 2) * (MG_TILE + 2)) * es;
}

MGBatch::~MGBatch() { free_all(); }

void MGBatch::free_all() {
    (void)hipStreamSynchronize(stream_);
    for (auto& st : cycle_) {
        if (st.tasks) (void)hipFree(st.tasks);
        if (st.tiles) (void)hipFree(st.tiles);
    }
    cycle_.clear();
    if (d_fine_) (void)hipFree(d_fine_);
    if (d_chunk_tiles_) (void)hipFree(d_chunk_tiles_);
    if (d_res_tiles_) (void)hipFree(d_res_tiles_);
    if (d_mirror_) (void)hipFree(d_mirror_);
    if (h_mirror_) (void)hipHostFree(h_mirror_);
    d_fine_ = nullptr;
    d_chunk_tiles_ = d_res_tiles_ = nullptr;
    d_mirror_ = nullptr;
    h_mirror_ = nullptr;
}

template <typename T>
static int upload_vec(void** dst, const std::vector<T>& v) {
    PIC_HIP(hipMalloc(dst, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) PIC_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
int MGBatch::add_level_stage(int kind, const std::vector<MGLevelTask<T>>& tasks, size_t lds) {
    if (tasks.empty()) return 0;
    std::vector<MGTile> tiles;
    for (int k = 0; k < (int)tasks.size(); ++k) {
        const int n = tasks[k].n;
        if (kind == 3) {
            const int chunks = (int)(((size_t)n * n + 255) / 256);
            for (int c = 0; c < chunks; ++c) tiles.push_back(MGTile{k, c, 0, 0});
        } else {
            const int g = div_up(n, MG_TILE);
            for (int ty = 0; ty < g; ++ty)
                for (int tx = 0; tx < g; ++tx) tiles.push_back(MGTile{k, tx, ty, 0});
        }
    }
    Stage st;
    st.kind = kind;
    st.fp64 = sizeof(T) == 8;
    st.lds = lds;
    st.n_items = (int)tiles.size();
    if (int rc = upload_vec(&st.tasks, tasks)) return rc;
    if (int rc = upload_vec((void**)&st.tiles, tiles)) return rc;
    cycle_.push_back(st);
    return 0;
}

template <typename T>
int MGBatch::build_cycle(const std::vector<double>& h) {
    const int M = (int)m_.size();
    const T omega = (T)p_.omega;
    // per member: level-0 input, output, scratch and rhs
    auto E = [&](MGSolver* s, int b, int l) -> T* { return (T*)(mixed_ ? (void*)s->ef_[b][l] : (void*)s->ed_[b][l]); };
    auto R = [&](MGSolver* s, int l) -> T* { return (T*)(mixed_ ? (void*)s->rf_[l] : (void*)s->rd_[l]); };
    auto in0 = [&](int k) -> const T* { return mixed_ ? nullptr : (const T*)(void*)phi_[k]; };
    auto out0 = [&](int k) -> T* { return mixed_ ? (T*)(void*)m_[k]->ef_[0][0] : (T*)(void*)phi_[k]; };
    auto tmp0 = [&](int k) -> T* { return mixed_ ? (T*)(void*)m_[k]->ef_[1][0] : (T*)(void*)m_[k]->phi_tmp_; };
    auto rhs0 = [&](int k) -> const T* { return mixed_ ? (const T*)(void*)m_[k]->r32_ : (const T*)(void*)rhs_[k]; };
    auto hl = [&](int k, int l) { return h[k] * std::ldexp(1.0, l); };
    int D = 0;
    for (auto* s : m_) D = std::max(D, s->lc_);
    // down leg
    for (int d = 0; d < D; ++d) {
        std::vector<MGLevelTask<T>> tasks;
        for (int k = 0; k < M; ++k) {
            MGSolver* s = m_[k];
            if (s->lc_ <= d) continue;
            const double hh = hl(k, d);
            MGLevelTask<T> t{};
            t.e_in = d == 0 ? in0(k) : nullptr;
            t.rhs = d == 0 ? rhs0(k) : R(s, d);
            t.e_out = d == 0 ? tmp0(k) : E(s, 1, d);
            t.rhs_c = R(s, d + 1);
            t.n = s->ns_[d];
            t.nu = p_.pre;
            t.H = 2 * p_.pre + 2;
            t.h2 = (T)(hh * hh);
            t.inv_h2 = (T)(1.0 / (hh * hh));
            t.omega = omega;
            t.gate = s->d_norm_;
            tasks.push_back(t);
        }
        if (int rc = add_level_stage<T>(0, tasks, smooth_lds(2 * p_.pre + 2, sizeof(T)))) return rc;
    }
    // coarse part: the single-workgroup cycle, or tiled sweeps on a large coarsest level
    {
        std::vector<MGCoarseTask<T>> ct;
        size_t lds = 0;
        for (int k = 0; k < M; ++k) {
            MGSolver* s = m_[k];
            const int lc = s->lc_;
            if (s->ns_[lc] > s->nc_) continue;
            MGCoarseTask<T> t{};
            t.e_in = lc == 0 ? in0(k) : nullptr;
            t.rhs = lc == 0 ? rhs0(k) : R(s, lc);
            t.e_out = lc == 0 ? out0(k) : E(s, 0, lc);
            t.n_top = s->ns_[lc];
            t.Lc = s->L_ - lc;
            t.pre = p_.pre;
            t.post = p_.post;
            t.csweeps = s->coarse_sweeps_eff();
            t.h_top = (T)hl(k, lc);
            t.omega = omega;
            t.gate = s->d_norm_;
            lds = std::max(lds, coarse_lds_bytes<T>(t.n_top, t.Lc));
            ct.push_back(t);
        }
        if (!ct.empty()) {
            Stage st;
            st.kind = 2;
            st.fp64 = sizeof(T) == 8;
            st.lds = lds;
            st.n_items = (int)ct.size();
            if (int rc = upload_vec(&st.tasks, ct)) return rc;
            cycle_.push_back(st);
        }
        // large coarsest levels: chunks of up to 3 sweeps, ping-pong, then a copy
        std::vector<int> big;
        for (int k = 0; k < M; ++k)
            if (m_[k]->ns_[m_[k]->lc_] > m_[k]->nc_) big.push_back(k);
        if (!big.empty()) {
            std::vector<const T*> src(M, nullptr);
            std::vector<int> which(M, 0);
            std::vector<T*> b0(M), b1(M);
            for (int k : big) {
                MGSolver* s = m_[k];
                const int lc = s->lc_;
                b0[k] = lc == 0 ? out0(k) : E(s, 0, lc);
                b1[k] = lc == 0 ? tmp0(k) : E(s, 1, lc);
                src[k] = lc == 0 ? in0(k) : nullptr;
                which[k] = (src[k] == b0[k]) ? 1 : 0;   // never write the buffer a launch reads
            }
            int left = p_.coarse_sweeps;
            bool wrote = false;
            while (left > 0 || !wrote) {
                const int nu = std::min(left, 3);
                std::vector<MGLevelTask<T>> tasks;
                for (int k : big) {
                    MGSolver* s = m_[k];
                    const int lc = s->lc_;
                    const double hh = hl(k, lc);
                    MGLevelTask<T> t{};
                    t.e_in = src[k];
                    t.rhs = lc == 0 ? rhs0(k) : R(s, lc);
                    t.e_out = which[k] ? b1[k] : b0[k];
                    t.n = s->ns_[lc];
                    t.nu = nu;
                    t.H = 2 * nu;
                    t.h2 = (T)(hh * hh);
                    t.inv_h2 = (T)(1.0 / (hh * hh));
                    t.omega = omega;
                    t.gate = s->d_norm_;
                    tasks.push_back(t);
                    src[k] = t.e_out;
                    which[k] ^= 1;
                }
                if (int rc = add_level_stage<T>(0, tasks, smooth_lds(2 * nu, sizeof(T)))) return rc;
                left -= nu;
                wrote = true;
            }
            std::vector<MGLevelTask<T>> copies;
            for (int k : big) {
                if (src[k] == b0[k]) continue;
                MGLevelTask<T> t{};
                t.e_in = src[k];
                t.e_out = b0[k];
                t.n = m_[k]->ns_[m_[k]->lc_];
                t.gate = m_[k]->d_norm_;
                copies.push_back(t);
            }
            if (int rc = add_level_stage<T>(3, copies, 0)) return rc;
        }
    }
    // up leg
    for (int d = D - 1; d >= 0; --d) {
        std::vector<MGLevelTask<T>> tasks;
        for (int k = 0; k < M; ++k) {
            MGSolver* s = m_[k];
            if (s->lc_ <= d) continue;
            const double hh = hl(k, d);
            MGLevelTask<T> t{};
            t.e_in = d == 0 ? tmp0(k) : E(s, 1, d);
            t.coarse = E(s, 0, d + 1);
            t.rhs = d == 0 ? rhs0(k) : R(s, d);
            t.e_out = d == 0 ? out0(k) : E(s, 0, d);
            t.n = s->ns_[d];
            t.nu = p_.post;
            t.H = 2 * p_.post;
            t.h2 = (T)(hh * hh);
            t.inv_h2 = (T)(1.0 / (hh * hh));
            t.omega = omega;
            t.gate = s->d_norm_;
            tasks.push_back(t);
        }
        const size_t W = MG_TILE + 2 * (2 * p_.post);
        if (int rc = add_level_stage<T>(1, tasks, 2 * W * W * sizeof(T))) return rc;
    }
    return 0;
}

int MGBatch::build(const std::vector<MGSolver*>& members, const std::vector<double*>& phi,
                   const std::vector<const double*>& rhs, const std::vector<double>& h) {
    free_all();
    PIC_REQUIRE(!members.empty() && members.size() == phi.size() && members.size() == rhs.size() &&
                    members.size() == h.size(), "multigrid batch: inconsistent members");
    m_ = members;
    phi_ = phi;
    rhs_ = rhs;
    p_ = members[0]->p_;
    for (auto* s : members) {
        PIC_REQUIRE(s->ok_, "multigrid batch: member not initialised");
        const MGParams& q = s->p_;
        PIC_REQUIRE(q.mode == p_.mode && q.pre == p_.pre && q.post == p_.post && q.coarse_sweeps == p_.coarse_sweeps &&
                        q.omega == p_.omega && q.tol == p_.tol && q.max_cycles == p_.max_cycles,
                    "multigrid batch: members need identical cycle parameters");
    }
    mixed_ = p_.mode == PIC_MG_MIXED;
    const int M = (int)m_.size();
    PIC_HIP(hipMalloc(&d_mirror_, M * sizeof(MGNorm)));
    PIC_HIP(hipMemset(d_mirror_, 0, M * sizeof(MGNorm)));
    PIC_HIP(hipHostMalloc(&h_mirror_, M * sizeof(MGNorm), hipHostMallocDefault));
    std::memset(h_mirror_, 0, M * sizeof(MGNorm));
    std::vector<MGFineTask> fine(M);
    std::vector<MGTile> chunks, res;
    for (int k = 0; k < M; ++k) {
        MGSolver* s = m_[k];
        const int n = s->n_;
        const size_t nn = (size_t)n * n;
        MGFineTask& f = fine[k];
        f.phi = phi[k];
        f.rhs = rhs[k];
        f.r32 = s->r32_;
        f.e32 = mixed_ ? s->ef_[0][0] : nullptr;
        f.partial = s->partials_;
        f.norm = s->d_norm_;
        f.mirror = d_mirror_ + k;
        f.n = n;
        f.gx = div_up(n, 16);
        f.nblocks = f.gx * f.gx;
        f.nchunks = (int)((nn + 255) / 256);
        f.inv_h2 = 1.0 / (h[k] * h[k]);
        for (int c = 0; c < f.nchunks; ++c) chunks.push_back(MGTile{k, c, 0, 0});
        for (int ty = 0; ty < f.gx; ++ty)
            for (int tx = 0; tx < f.gx; ++tx) res.push_back(MGTile{k, tx, ty, 0});
    }
    if (int rc = upload_vec((void**)&d_fine_, fine)) return rc;
    if (int rc = upload_vec((void**)&d_chunk_tiles_, chunks)) return rc;
    if (int rc = upload_vec((void**)&d_res_tiles_, res)) return rc;
    n_chunk_ = (int)chunks.size();
    n_res_ = (int)res.size();
    return mixed_ ? build_cycle<float>(h) : build_cycle<double>(h);
}

int MGBatch::launch_stage(const Stage& st) {
    const dim3 g(st.n_items), b(st.kind == 2 ? 1024 : 256);
    if (st.fp64) {
        auto* t = (const MGLevelTask<double>*)st.tasks;
        if (st.kind == 0) hipLaunchKernelGGL(mg_smooth_stage<double>, g, b, st.lds, stream_, t, st.tiles);
        else if (st.kind == 1) hipLaunchKernelGGL(mg_prolong_stage<double>, g, b, st.lds, stream_, t, st.tiles);
        else if (st.kind == 2)
            hipLaunchKernelGGL(mg_coarse_stage<double>, g, b, st.lds, stream_, (const MGCoarseTask<double>*)st.tasks);
        else hipLaunchKernelGGL(mg_copy_stage<double>, g, b, 0, stream_, t, st.tiles);
    } else {
        auto* t = (const MGLevelTask<float>*)st.tasks;
        if (st.kind == 0) hipLaunchKernelGGL(mg_smooth_stage<float>, g, b, st.lds, stream_, t, st.tiles);
        else if (st.kind == 1) hipLaunchKernelGGL(mg_prolong_stage<float>, g, b, st.lds, stream_, t, st.tiles);
        else if (st.kind == 2)
            hipLaunchKernelGGL(mg_coarse_stage<float>, g, b, st.lds, stream_, (const MGCoarseTask<float>*)st.tasks);
        else hipLaunchKernelGGL(mg_copy_stage<float>, g, b, 0, stream_, t, st.tiles);
    }
    ++launches_;
    return 0;
}

// what: 0 rhs norm, 1 residual (which 1 or 2), 2 correction
int MGBatch::enqueue_fine(int what, int which) {
    const int M = (int)m_.size();
    if (what == 0) {
        hipLaunchKernelGGL(sumsq_stage, dim3(n_chunk_), dim3(256), 0, stream_, d_fine_, d_chunk_tiles_);
        hipLaunchKernelGGL(norm_final_stage, dim3(M), dim3(256), 0, stream_, d_fine_, 0, p_.tol, p_.max_cycles);
        launches_ += 2;
    } else if (what == 1) {
        if (mixed_)
            hipLaunchKernelGGL(ir_residual_stage<true>, dim3(n_res_), dim3(256), 0, stream_, d_fine_, d_res_tiles_,
                               which);
        else
            hipLaunchKernelGGL(ir_residual_stage<false>, dim3(n_res_), dim3(256), 0, stream_, d_fine_, d_res_tiles_,
                               which);
        hipLaunchKernelGGL(norm_final_stage, dim3(M), dim3(256), 0, stream_, d_fine_, which, p_.tol, p_.max_cycles);
        launches_ += 2;
    } else {
        hipLaunchKernelGGL(ir_correct_stage, dim3(n_chunk_), dim3(256), 0, stream_, d_fine_, d_chunk_tiles_);
        ++launches_;
    }
    return 0;
}

// The host enqueues cycles in batches (first batch: the previous solve's
// largest cycle count) and reads the norms of all members once per batch.  The
// device flags turn cycles beyond a member's convergence into no-ops, so each
// result equals testing after every cycle, alone or in a batch.
int MGBatch::solve() {
    PIC_REQUIRE(!m_.empty(), "multigrid batch: not built");
    launches_ = 0;
    const int M = (int)m_.size();
    enqueue_fine(0, 0);
    enqueue_fine(1, 1);
    int enqueued = 0;
    int batch = std::max(1, last_cycles_);
    for (;;) {
        batch = std::min(batch, p_.max_cycles - enqueued);
        for (int b = 0; b < batch; ++b) {
            for (const auto& st : cycle_)
                if (int rc = launch_stage(st)) return rc;
            if (mixed_) enqueue_fine(2, 0);
            enqueue_fine(1, 2);
        }
        enqueued += std::max(batch, 0);
        PIC_HIP(hipMemcpyAsync(h_mirror_, d_mirror_, M * sizeof(MGNorm), hipMemcpyDeviceToHost, stream_));
        PIC_HIP(hipStreamSynchronize(stream_));
        bool all = true;
        for (int k = 0; k < M; ++k) all = all && h_mirror_[k].converged;
        if (all || enqueued >= p_.max_cycles) break;
        batch = 1;
    }
    PIC_HIP(hipGetLastError());
    int mx = 0;
    for (int k = 0; k < M; ++k) mx = std::max(mx, h_mirror_[k].cycles);
    last_cycles_ = std::max(1, mx);
    return 0;
}

}  // namespace pic
