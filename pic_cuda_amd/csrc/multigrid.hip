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
// co-residency limit; a device-wide barrier costs more than a launch here):
//   mg_smooth_stage   32x32 output tile + halo 2*nu(+2) staged in LDS; nu
//                     sweeps in LDS, each half-sweep skipping the outer ring
//                     that can no longer be exact; then residual and
//                     restriction of the coarse nodes the tile owns.  The
//                     sweep count is a template parameter, so the tile width
//                     is a constant and every global load of the tile is in
//                     flight before the first wait.
//   mg_prolong_stage  coarse patch + fine tile + halo 2*nu in LDS,
//                     prolongation-add, nu sweeps.
//   mg_coarse_stage   one workgroup runs the whole sub-cycle of the levels
//                     with n <= 65 (FP32) / 33 (FP64) in LDS; in mixed mode
//                     it can instead solve its top level exactly in FP64 with
//                     a 2D sine transform on the matrix cores.
//   ir_residual_stage FP64 residual on 64x16 tiles, per-tile sums of r^2; in
//                     mixed mode fused with the correction phi += e of the
//                     previous FP32 cycle (phi alternates between two buffers);
//                     the first one of a solve also sums rhs^2 and can copy
//                     phi (the warm start) for the caller.
//   norm_final_stage  fixed-order reduction of the tile sums, relative
//                     residual and the convergence flag (build rule
//                     ||r||/||rhs||; the reference divides by sum rhs^2,
//                     Multigrid.cu:466).
#include "mg.hpp"
#include "step_device.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>
#include <mutex>
#include <map>
#include <type_traits>

namespace pic {

#define MG_TILE 32
// largest level handled by the single-workgroup coarse kernel: 65^2 in both
// precisions (FP64: the levels' e and rhs and the residual scratch take
// 125 KB of LDS, within gfx950's 160 KB per workgroup)
#define MG_NC_F32 65
#define MG_NC_F64 65
// direct coarse solve: levels up to 65 nodes (63 unknowns a side); the two
// FP64 LDS buffers have row stride MG_DS (63 x 63 each: 63.5 KB in all, and
// 16 consecutive rows fall in distinct bank pairs)
#define MG_ND 65
#define MG_DS 63
#define MG_DLDS (2 * MG_DS * MG_DS * 8)
// big direct solve (coarse_direct 2): levels up to 257 nodes, unknowns a
// side padded to a multiple of 64.  An exact solve at 513^2 (P = 512, chains
// of 16 steps) measured the same as a tiled 513^2 level over the exact 257^2
// solve on one grid (1025^2 steps, the 2049^2 Poisson solve) and slower in a
// batch, where the tiled stages of many cases share launches and the 8x
// larger transform does not.
#define MG_NDB 257
static inline int dst_pad(int m) { return (m + 63) / 64 * 64; }
// tiled launches of at most this many tiles (one per CU of an MI355X) run
// 1024 threads per tile, up to four times as many 512
#define MG_SMALL_LAUNCH 256
// tiles a side of a tiled stage on a level of n nodes (the kernels' edge_nodes
// write the boundary row and column this leaves out when n = 32 g + 1)
static inline int mg_tiles(int n) { return n <= 1 ? 1 : div_up(n - 1, MG_TILE); }
// largest finest level (nodes a side) of a batch that fuses the finest
// up/down stages between V-cycles
#define MG_FUSE_MAX 1025
// residual tiles: 64 columns x 16 rows, 256 threads with 4 rows each
#define MG_RX 64
#define MG_RY 16

template <typename T>
__device__ inline T tfma(T a, T b, T c);
template <> __device__ inline float tfma(float a, float b, float c) { return fmaf(a, b, c); }
template <> __device__ inline double tfma(double a, double b, double c) { return fma(a, b, c); }

// Work-item addressing of a stage.  A single-grid stage carries its task by
// value and derives its tile from blockIdx, so no dependent load precedes the
// data; a batched stage reads its (case, tile) entry and then the case's task.
template <typename Task>
struct StageRef {
    Task t0;
    const Task* tasks;     // null: the single task t0
    const MGTile* tiles;   // batched tiled stages
    int gx;                // single task: tiles per row (0: one tile per block, tile.tx = blockIdx.x)
};

// The dispatcher deals workgroups to the 8 XCDs in turn (workgroup b to XCD
// b % 8), and each XCD has its own L2.  Renumbered so that XCD x takes one
// contiguous run of the grid's items (counts q or q + 1, in order), a tile's
// halo rows were mostly loaded by tiles of the same XCD.  A bijection on
// [0, gridDim.x) whatever the partition, so only the schedule depends on it.
__device__ inline int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int x = b & 7, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + (b >> 3);
}

// the item of block number b (after the XCD renumbering)
template <typename Task>
__device__ inline void stage_item_at(const StageRef<Task>& a, Task& tk, MGTile& tile, int b) {
    if (a.tasks) {
        tile = a.tiles[b];
        tk = a.tasks[tile.task];
    } else {
        tk = a.t0;
        tile.task = 0;
        tile.tx = a.gx ? b % a.gx : b;
        tile.ty = a.gx ? b / a.gx : 0;
    }
}

template <typename Task>
__device__ inline void stage_item(const StageRef<Task>& a, Task& tk, MGTile& tile) {
    stage_item_at(a, tk, tile, xcd_block());
}

// Stop flag of a task without one: always zero.  Reading the flag through an
// unconditional load lets its latency overlap the tile's data loads.
__device__ const int g_never_stop = 0;

template <typename T>
__device__ inline const int* stop_flag(const MGLevelTask<T>& tk) {
    return tk.gate ? &tk.gate->converged : &g_never_stop;
}

// one task per block (coarse cycle, norms)
template <typename Task>
__device__ inline void stage_case(const StageRef<Task>& a, Task& tk) {
    if (a.tasks) tk = a.tasks[blockIdx.x];
    else tk = a.t0;
}

// The 5-point Laplacian times h^2 at node idx.  FP64 (the reference's V-cycle,
// bitwise with it): the sum of the neighbours minus 4 c.  FP32 (the inner
// cycles of the mixed solve): the sum of the four differences to the centre.
// Neighbours of a smooth correction lie within a factor of two of the centre,
// so each difference is exact in FP32 and only the small sum rounds; in the
// reference form the sum of the neighbours rounds at the size of the
// correction itself, and h^-2 amplifies that rounding into the residual,
// which limited what two FP32 V-cycles gain per FP64 refinement step.
template <typename T>
__device__ inline T lap_h2(const T* e, int idx, int stride) {
    const T c = e[idx];
    if constexpr (std::is_same<T, float>::value) {
        return ((e[idx + stride] - c) + (e[idx - stride] - c)) + ((e[idx + 1] - c) + (e[idx - 1] - c));
    } else {
        return e[idx - stride] + e[idx + stride] + e[idx - 1] + e[idx + 1] - (T)4 * c;
    }
}

// the SOR update of one node: FP64 the reference's expression; FP32 the same
// update as c + (omega/4) (lap_h2 - h^2 r)
template <typename T>
__device__ inline T sor_node(const T* e, const T* r, int idx, int stride, T h2, T omega) {
    const T c = e[idx];
    if constexpr (std::is_same<T, float>::value) {
        return tfma(omega * (T)0.25, tfma(-h2, r[idx], lap_h2(e, idx, stride)), c);
    } else {
        const T s = e[idx + stride] + e[idx - stride] + e[idx + 1] + e[idx - 1];
        const T nv = (T)0.25 * (s - h2 * r[idx]);
        return tfma(omega, nv, ((T)1 - omega) * c);
    }
}

// One colour of red-black SOR on a W x W LDS tile whose element (0,0) is the
// global node (gi0, gj0).  Only rows and columns [lo, W-1-lo] are updated: after
// s half-sweeps the outer s rings are stale, so half-sweep s uses lo = s + 1.
// Colour c updates the nodes with (gi + gj) odd for c = 1, even for c = 0.
// W is even, so every row of the span holds span/2 nodes of each colour:
// thread t takes colour slot t % (span/2) of rows t / (span/2), then every
// rows-per-pass row after it (index arithmetic once per half-sweep, not per
// node).  EDGE: the tile reaches the domain boundary, whose nodes are skipped;
// an interior tile tests no bounds.
template <typename T, int W, int BS, bool EDGE>
__device__ inline void tile_colour(T* e, const T* r, const uint8_t* mk, int gi0, int gj0, int n, T h2, T omega,
                                   int colour, int lo) {
    static_assert(W % 2 == 0, "tile width must be even");
    const int span = W - 2 * lo;
    const int halfc = span / 2;
    const int rp = BS / halfc;             // rows per pass
    const int t = threadIdx.x;
    if (t >= rp * halfc) return;
    const int m = t % halfc, q0 = t / halfc;
    const int c0 = (colour + gi0 + gj0) & 1;
    for (int q = q0; q < span; q += rp) {
        const int lj = lo + q;
        const int li = lo + ((c0 + q) & 1) + 2 * m;
        if (EDGE) {
            const int gi = gi0 + li, gj = gj0 + lj;
            if (gi <= 0 || gi >= n - 1 || gj <= 0 || gj >= n - 1) continue;
        }
        const int idx = lj * W + li;
        if (mk && mk[idx]) continue;   // electrode node: held fixed
        e[idx] = sor_node(e, r, idx, W, h2, omega);
    }
}

// the tile's half-sweeps, with the bound tests only where the tile's LDS
// window reaches the domain boundary
template <typename T, int W, int BS, int NSWEEP>
__device__ inline void tile_sweeps(T* e, const T* r, const uint8_t* mk, int gi0, int gj0, int n, T h2, T omega) {
    const bool edge = gi0 <= 0 || gj0 <= 0 || gi0 + W - 1 >= n - 1 || gj0 + W - 1 >= n - 1;
#pragma unroll
    for (int s = 0; s < 2 * NSWEEP; ++s) {
        if (edge) tile_colour<T, W, BS, true>(e, r, mk, gi0, gj0, n, h2, omega, 1 - (s & 1), s + 1);
        else tile_colour<T, W, BS, false>(e, r, mk, gi0, gj0, n, h2, omega, 1 - (s & 1), s + 1);
        __syncthreads();
    }
}

// Residual on the tile + 1 ring and its full-weighting restriction to the
// coarse nodes (2ic, 2jc) inside the tile; e and r are the W-wide LDS tile
// whose element (0,0) is node (gi0, gj0).  The residual replaces r in place
// (each node's update reads only its own r and the unchanged e), so the stage
// needs no LDS of its own for it.  Fixed (electrode) nodes have zero residual.
template <typename T, int W, int BS>
__device__ inline void restrict_tile(const T* e, T* r, const uint8_t* mk, int i0, int j0, int gi0, int gj0, int n,
                                     T inv_h2, const gptr<T> rhs_c) {
    const int tid = threadIdx.x;
    constexpr int RW = MG_TILE + 2;
    constexpr int RP = BS / RW;            // rows per pass; thread t takes column t % RW
    if (tid < RP * RW) {
        const int c = tid % RW;
        const int gi = i0 - 1 + c;
        const bool col_in = gi > 0 && gi < n - 1;
        for (int q = tid / RW; q < RW; q += RP) {
            const int gj = j0 - 1 + q;
            T v = (T)0;
            const int idx = (gj - gj0) * W + (gi - gi0);
            if (col_in && gj > 0 && gj < n - 1 && !(mk && mk[idx])) {
                const T lap = lap_h2(e, idx, W) * inv_h2;
                v = r[idx] - lap;
            }
            r[idx] = v;
        }
    }
    __syncthreads();
    const int nc = (n + 1) / 2;
    constexpr int TC = MG_TILE / 2;
    if (tid >= TC * TC) return;   // one thread per coarse node of the tile (256)
    const int ic = i0 / 2 + tid % TC, jc = j0 / 2 + tid / TC;
    if (ic >= nc || jc >= nc) return;
    const int i = 2 * ic, j = 2 * jc;
    if (i >= n || j >= n) {
        rhs_c[(size_t)jc * nc + ic] = (T)0;
        return;
    }
    const int f = (j - gj0) * W + (i - gi0);
    T s = (T)4 * r[f];
    if (i > 0) s += (T)2 * r[f - 1];
    if (i < n - 1) s += (T)2 * r[f + 1];
    if (j > 0) s += (T)2 * r[f - W];
    if (j < n - 1) s += (T)2 * r[f + W];
    if (i > 0 && j > 0) s += r[f - W - 1];
    if (i < n - 1 && j > 0) s += r[f - W + 1];
    if (i > 0 && j < n - 1) s += r[f + W - 1];
    if (i < n - 1 && j < n - 1) s += r[f + W + 1];
    rhs_c[(size_t)jc * nc + ic] = s * (T)0.0625;
}

// bilinear prolongation of the coarse value at fine node (gi, gj), with the
// coarse patch in LDS: element (0,0) of the CW-wide patch is coarse (ic0, jc0)
template <typename T>
__device__ inline T prolong_lds(const T* C, int CW, int ic0, int jc0, int nc, int gi, int gj) {
    const int ic = gi >> 1, jc = gj >> 1;
    const int c = (jc - jc0) * CW + (ic - ic0);
    const bool io = gi & 1, jo = gj & 1;
    if (!io && !jo) return C[c];
    if (io && !jo) return (ic + 1 < nc) ? (T)0.5 * (C[c] + C[c + 1]) : C[c];
    if (!io && jo) return (jc + 1 < nc) ? (T)0.5 * (C[c] + C[c + CW]) : C[c];
    return (ic + 1 < nc && jc + 1 < nc) ? (T)0.25 * (C[c] + C[c + 1] + C[c + CW] + C[c + CW + 1]) : C[c];
}

// The tiles of a level cover nodes [0, 32 g) a side with g = ceil((n - 1) / 32)
// (mg_tiles), which leaves out the last boundary row and column when
// n = 32 g + 1 -- every level of a 2^k + 1 ladder from 33 nodes up, where a
// tile of them would hold nothing else (2049^2: 4225 tiles -> 4096).  A
// boundary node is never smoothed, so its output is its input (plus the
// coarse correction in an up stage): the tiles next to that row and column
// write it, from global memory (a stage with no halo has no copy of it in
// LDS), and set the coarse rhs of the coarse boundary nodes they would have
// owned to zero (a boundary rhs is never read).
template <typename T>
__device__ inline void edge_nodes(int i0, int j0, int n, const gptr<const T> e_in, const T* C, int CW, int ic0,
                                  int jc0, const gptr<const uint8_t> gmask, const gptr<T> e_out, const gptr<T> rhs_c) {
    const bool right = i0 + MG_TILE == n - 1, bottom = j0 + MG_TILE == n - 1;
    if (!right && !bottom) return;
    const int tid = threadIdx.x;
    int gi = -1, gj = -1;
    if (tid < MG_TILE) {
        if (right) { gi = n - 1; gj = j0 + tid; }
    } else if (tid >= 64 && tid < 64 + MG_TILE) {
        if (bottom) { gi = i0 + tid - 64; gj = n - 1; }
    } else if (tid == 128) {
        if (right && bottom) { gi = n - 1; gj = n - 1; }
    }
    const int nc = (n + 1) / 2;
    if (gi >= 0 && gi < n && gj < n) {
        const size_t g = (size_t)gj * n + gi;
        T v = e_in ? e_in[g] : (T)0;
        if (C && !(gmask && gmask[g])) v = v + prolong_lds(C, CW, ic0, jc0, nc, gi, gj);
        e_out[g] = v;
        if (rhs_c && !(gi & 1) && !(gj & 1)) rhs_c[(size_t)(gj >> 1) * nc + (gi >> 1)] = (T)0;
    }
}

// Down-leg stage: load tile + halo, NU sweeps, store the tile; with RESTRICT
// also the residual on tile + 1 ring and the restriction to the coarse nodes
// (2ic, 2jc) inside the tile.  Without RESTRICT (tiled sweeps of a large
// coarsest level) the halo is 2*NU.  With electrodes the mask tile is staged
// in LDS too: fixed nodes are not smoothed and have zero residual.
template <typename T, int NU, bool RESTRICT, bool MASKED, int BS>
__global__ __launch_bounds__(BS) void mg_smooth_stage(const StageRef<MGLevelTask<T>> a) {
    constexpr int H = RESTRICT ? 2 * NU + 2 : 2 * NU;
    constexpr int W = MG_TILE + 2 * H;
    constexpr int WW = W * W;
    constexpr int NL = (WW + BS - 1) / BS;
    MGTile tile;
    MGLevelTask<T> tk;
    stage_item(a, tk, tile);
    // stop flag (a cycle enqueued after convergence is a no-op): read here,
    // tested once the tile's loads are issued
    const int stop = *G(stop_flag(tk));
    const gptr<const T> e_in = G(tk.e_in);
    const gptr<const T> rhs = G(tk.rhs);
    const gptr<T> e_out = G(tk.e_out);
    const gptr<T> rhs_c = G(tk.rhs_c);
    const int n = tk.n;
    const T h2 = tk.h2, inv_h2 = tk.inv_h2, omega = tk.omega;
    extern __shared__ __align__(16) unsigned char smem_mg[];
    T* e = reinterpret_cast<T*>(smem_mg);
    T* r = e + WW;
    // electrodes: only the MASKED instantiation carries the mask code
    const gptr<const uint8_t> gmask = MASKED ? G(tk.mask) : nullptr;
    uint8_t* mk = (MASKED && gmask) ? reinterpret_cast<uint8_t*>(r + WW) : nullptr;
    const int tid = threadIdx.x;
    const int i0 = tile.tx * MG_TILE, j0 = tile.ty * MG_TILE;
    const int gi0 = i0 - H, gj0 = j0 - H;
    T ev[NL], rv[NL];
    uint8_t mv[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int k = tid + BS * q;
        const int lj = k / W, li = k - lj * W;
        const int gi = gi0 + li, gj = gj0 + lj;
        ev[q] = (T)0;
        rv[q] = (T)0;
        mv[q] = 0;
        if (k < WW && gi >= 0 && gi < n && gj >= 0 && gj < n) {
            const size_t g = (size_t)gj * n + gi;
            if (e_in) ev[q] = e_in[g];
            rv[q] = rhs[g];
            if (gmask) mv[q] = gmask[g];
        }
    }
    if (stop) return;
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int k = tid + BS * q;
        if (k < WW) {
            e[k] = ev[q];
            r[k] = rv[q];
            if (mk) mk[k] = mv[q];
        }
    }
    __syncthreads();
    tile_sweeps<T, W, BS, NU>(e, r, mk, gi0, gj0, n, h2, omega);
#pragma unroll
    for (int q = 0; q < (MG_TILE * MG_TILE) / BS; ++q) {
        const int k = tid + BS * q;
        const int gi = i0 + k % MG_TILE, gj = j0 + k / MG_TILE;
        if (gi < n && gj < n) e_out[(size_t)gj * n + gi] = e[(gj - gj0) * W + (gi - gi0)];
    }
    edge_nodes<T>(i0, j0, n, e_in, nullptr, 0, 0, 0, gmask, e_out, RESTRICT ? rhs_c : G((T*)nullptr));
    if constexpr (RESTRICT) {
        if (rhs_c) restrict_tile<T, W, BS>(e, r, mk, i0, j0, gi0, gj0, n, inv_h2, rhs_c);
    }
}

// Up-leg stage: e = e_in + P(coarse) on tile + halo 2*NU, NU sweeps, store the
// tile.  With RESTRICT (the finest level between two V-cycles of one
// refinement step) NU is the post-sweeps of this cycle plus the pre-sweeps of
// the next, the halo is 2*NU+2, and the residual is restricted into rhs_c as
// the next cycle's down leg would: the same arithmetic in one launch fewer.
template <typename T, int NU, bool RESTRICT, bool MASKED, int BS>
__global__ __launch_bounds__(BS) void mg_prolong_stage(const StageRef<MGLevelTask<T>> a) {
    constexpr int H = RESTRICT ? 2 * NU + 2 : 2 * NU;
    constexpr int W = MG_TILE + 2 * H;
    constexpr int WW = W * W;
    constexpr int NL = (WW + BS - 1) / BS;
    constexpr int CW = W / 2 + 2;
    constexpr int CC = CW * CW;
    constexpr int NCL = (CC + BS - 1) / BS;
    MGTile tile;
    MGLevelTask<T> tk;
    stage_item(a, tk, tile);
    const int stop = *G(stop_flag(tk));   // tested once the loads are issued
    const gptr<const T> e_in = G(tk.e_in);
    const gptr<const T> coarse = G(tk.coarse);
    const gptr<const T> rhs = G(tk.rhs);
    const gptr<T> e_out = G(tk.e_out);
    const gptr<T> rhs_c = G(tk.rhs_c);
    const int n = tk.n;
    const T h2 = tk.h2, inv_h2 = tk.inv_h2, omega = tk.omega;
    extern __shared__ __align__(16) unsigned char smem_mg[];
    T* e = reinterpret_cast<T*>(smem_mg);
    T* r = e + WW;
    T* C = r + WW;

    const gptr<const uint8_t> gmask = MASKED ? G(tk.mask) : nullptr;
    uint8_t* mk = (MASKED && gmask) ? reinterpret_cast<uint8_t*>(C + CC) : nullptr;
    const int tid = threadIdx.x;
    const int i0 = tile.tx * MG_TILE, j0 = tile.ty * MG_TILE;
    const int gi0 = i0 - H, gj0 = j0 - H;   // even: H and i0 are
    const int nc = (n + 1) / 2;
    const int ic0 = gi0 >> 1, jc0 = gj0 >> 1;
    T ev[NL], rv[NL], cv[NCL];
    uint8_t mv[NL];
#pragma unroll
    for (int q = 0; q < NCL; ++q) {
        const int k = tid + BS * q;
        const int cj = k / CW, ci = k - cj * CW;
        const int ic = ic0 + ci, jc = jc0 + cj;
        cv[q] = (T)0;
        if (k < CC && ic >= 0 && ic < nc && jc >= 0 && jc < nc) cv[q] = coarse[(size_t)jc * nc + ic];
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int k = tid + BS * q;
        const int lj = k / W, li = k - lj * W;
        const int gi = gi0 + li, gj = gj0 + lj;
        ev[q] = (T)0;
        rv[q] = (T)0;
        mv[q] = 0;
        if (k < WW && gi >= 0 && gi < n && gj >= 0 && gj < n) {
            const size_t g = (size_t)gj * n + gi;
            ev[q] = e_in[g];
            rv[q] = rhs[g];
            if (gmask) mv[q] = gmask[g];
        }
    }
    if (stop) return;
#pragma unroll
    for (int q = 0; q < NCL; ++q) {
        const int k = tid + BS * q;
        if (k < CC) C[k] = cv[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int k = tid + BS * q;
        if (k < WW) {
            const int lj = k / W, li = k - lj * W;
            const int gi = gi0 + li, gj = gj0 + lj;
            T v = ev[q];
            // electrode nodes take no correction
            if (gi >= 0 && gi < n && gj >= 0 && gj < n && (gi >> 1) < nc && (gj >> 1) < nc && !mv[q])
                v = v + prolong_lds(C, CW, ic0, jc0, nc, gi, gj);
            e[k] = v;
            r[k] = rv[q];
            if (mk) mk[k] = mv[q];
        }
    }
    __syncthreads();
    tile_sweeps<T, W, BS, NU>(e, r, mk, gi0, gj0, n, h2, omega);
#pragma unroll
    for (int q = 0; q < (MG_TILE * MG_TILE) / BS; ++q) {
        const int k = tid + BS * q;
        const int gi = i0 + k % MG_TILE, gj = j0 + k / MG_TILE;
        if (gi < n && gj < n) e_out[(size_t)gj * n + gi] = e[(gj - gj0) * W + (gi - gi0)];
    }
    edge_nodes<T>(i0, j0, n, e_in, C, CW, ic0, jc0, gmask, e_out, RESTRICT ? rhs_c : G((T*)nullptr));
    if constexpr (RESTRICT) {
        if (rhs_c) restrict_tile<T, W, BS>(e, r, mk, i0, j0, gi0, gj0, n, inv_h2, rhs_c);
    }
}

// ---- the single-workgroup sub-cycle for the small levels ---------------------
// 1024 threads as 32 x 32: thread (tx, ty) takes rows ty, ty+32, ... and
// columns tx, tx+32, ... (or every other column for one colour).
//
// Electrodes: a bitmap of the top level's mask in LDS; node (i, j) of level l
// is fixed when the top-level node (i 2^l, j 2^l) is (the mask of a coarse
// level is that of the fine node it sits on).
struct CoarseMask {
    const unsigned* bits;   // null: no electrodes
    int n_top;
    __device__ bool fixed(int l, int i, int j) const {
        if (!bits) return false;
        const int k = (j << l) * n_top + (i << l);
        return (bits[k >> 5] >> (k & 31)) & 1u;
    }
};

template <typename T>
__device__ inline void lds_sweeps(T* e, const T* r, int n, T h, T omega, int iters, const CoarseMask& cm, int l) {
    const T h2 = h * h;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int it = 0; it < iters; ++it)
        for (int colour = 1; colour >= 0; --colour) {
            for (int j = 1 + ty; j < n - 1; j += 32) {
                const int ib = 1 + ((colour + j + 1) & 1);   // (i + j) & 1 == colour
                for (int i = ib + 2 * tx; i < n - 1; i += 64) {
                    if (cm.fixed(l, i, j)) continue;
                    const int idx = j * n + i;
                    e[idx] = sor_node(e, r, idx, n, h2, omega);
                }
            }
            __syncthreads();
        }
}

// ---- mixed mode: exact FP64 solve of the coarse kernel's level ---------------
// lap_h e = r with e = 0 on the boundary, on n <= 65 nodes (m = n - 2 <= 63
// unknowns a side), by the 2D discrete sine transform in FP64:
//   e = S (h^2 W o (S R S)) S,   S[a][b] = sin(pi (a+1)(b+1) / (m+1)),
//   W[p][q] = (2/(m+1))^2 / (-4 (sin^2(pi (p+1) / (2(m+1))) + sin^2(pi (q+1) / (2(m+1))))),
// since S^2 = (m+1)/2 I and the columns of S are the eigenvectors of the 1D
// second difference.  tab holds S then W, each 64 x 64 and zero outside
// m x m.  The four products run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16 (i % 4), element i / 4), halved by the mirror
// symmetry S[a][m-1-b] = (-1)^a S[a][b] (and S symmetric): with S on the
// right, P = X S has P[i][b] = E + O and P[i][m-1-b] = E - O, where E and O
// sum over the even and the odd k only, for the (m+1)/2 columns b < (m+1)/2;
// with S on the left the same holds for the rows.  Wavefront w < 8 computes
// E and O of one 16 x 16 tile: rb = w / 2 the block of the LDS operand's free
// index, cb = w % 2 the block of the halved index, S's entries at the even and
// the odd k of column block cb in registers (S symmetric: the same registers
// are the A operand when S is on the left).  X and Y hold m x m FP64 values
// with row stride MG_DS in LDS.
typedef double mg_v4d __attribute__((ext_vector_type(4)));

// E and O of one 16 x 16 tile: the LDS operand is the A side (LDS_LEFT:
// A[base + i][k]) or the B side (B[k][base + j]), Fe / Fo the other at
// k = 8 kb + 2 lk (+ 1).  All 16 LDS values are loaded first (entries beyond m
// read a clamped in-range address and are replaced by zero), then the two
// MFMA chains run side by side.
template <bool LDS_LEFT>
__device__ inline void pair_product(const double* A, int base, const double (&Fe)[8], const double (&Fo)[8], int m,
                                    int li, int lk, mg_v4d& E, mg_v4d& O) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = 8 * (q >> 1) + 2 * lk + (q & 1);
        const int row = LDS_LEFT ? base + li : k;
        const int col = LDS_LEFT ? k : base + li;
        const double x = A[min(row, m - 1) * MG_DS + min(col, m - 1)];
        v[q] = (row < m && col < m) ? x : 0.0;
    }
    E = mg_v4d{0.0, 0.0, 0.0, 0.0};
    O = mg_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        if (kb * 8 >= m) break;   // k >= 8 kb >= m: zero padding
        if (LDS_LEFT) {
            E = __builtin_amdgcn_mfma_f64_16x16x4f64(v[2 * kb], Fe[kb], E, 0, 0, 0);
            O = __builtin_amdgcn_mfma_f64_16x16x4f64(v[2 * kb + 1], Fo[kb], O, 0, 0, 0);
        } else {
            E = __builtin_amdgcn_mfma_f64_16x16x4f64(Fe[kb], v[2 * kb], E, 0, 0, 0);
            O = __builtin_amdgcn_mfma_f64_16x16x4f64(Fo[kb], v[2 * kb + 1], O, 0, 0, 0);
        }
    }
}

template <typename T>
__device__ void coarse_direct_solve(const MGCoarseTask<T>& tk, double* X, double* Y) {
    const int n = tk.n_top, m = n - 2, hm = (m + 1) / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rb = (wave >> 1) & 3, cb = wave & 1;
    const int li = lane & 15, lk = lane >> 4;
    const bool live = wave < 8 && rb * 16 < m && cb * 16 < hm;   // tiles beyond stay zero
    const gptr<const double> tab = G(tk.dtab);
    const gptr<const T> r = G(tk.rhs);
    const gptr<T> e = G(tk.e_out);
    // the gate (the case converged: a no-op) is read with the data loads and
    // tested once they are issued
    const int stop = tk.gate ? *G(&tk.gate->converged) : 0;
    // R into X (m*m <= 3969 < 4 * 1024): all loads first, from clamped indices
    float rv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = min((int)threadIdx.x + 1024 * u, m * m - 1);
        const int j = t / m, k = t - j * m;
        rv[u] = (float)r[(j + 1) * n + (k + 1)];
    }
    // S at the even / odd k of column block cb; h^2 W at the output rows
    // a = cb*16 + 4q + lk and m-1-a (product 2), column rb*16 + li
    double Fe[8], Fo[8], wsc[4], wsm[4];
    const double h2 = (double)tk.h_top * (double)tk.h_top;
    if (wave < 8) {
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            Fe[kb] = tab[(8 * kb + 2 * lk) * 64 + cb * 16 + li];
            Fo[kb] = tab[(8 * kb + 2 * lk + 1) * 64 + cb * 16 + li];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int aa = cb * 16 + 4 * q + lk, am = max(m - 1 - aa, 0);
            wsc[q] = tab[4096 + aa * 64 + rb * 16 + li] * h2;
            wsm[q] = tab[4096 + am * 64 + rb * 16 + li] * h2;
        }
    }
    if (stop) return;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = threadIdx.x + 1024 * u;
        const int j = t / m, k = t - j * m;
        if (t < m * m) X[j * MG_DS + k] = (double)rv[u];
    }
    __syncthreads();
    mg_v4d Ev, Ov;
    // T1 = R S -> Y: rows i = rb*16 + 4q + lk, columns b = cb*16 + li and m-1-b
    if (live) {
        pair_product<true>(X, rb * 16, Fe, Fo, m, li, lk, Ev, Ov);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = rb * 16 + 4 * q + lk, b = cb * 16 + li;
            if (i < m && b < hm) {
                Y[i * MG_DS + b] = Ev[q] + Ov[q];
                if (m - 1 - b != b) Y[i * MG_DS + (m - 1 - b)] = Ev[q] - Ov[q];
            }
        }
    }
    __syncthreads();
    // T2 = S T1, times h^2 W -> X: rows a = cb*16 + 4q + lk and m-1-a, columns j = rb*16 + li
    if (live) {
        pair_product<false>(Y, rb * 16, Fe, Fo, m, li, lk, Ev, Ov);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int a = cb * 16 + 4 * q + lk, j = rb * 16 + li;
            if (a < hm && j < m) {
                X[a * MG_DS + j] = (Ev[q] + Ov[q]) * wsc[q];
                if (m - 1 - a != a) X[(m - 1 - a) * MG_DS + j] = (Ev[q] - Ov[q]) * wsm[q];
            }
        }
    }
    __syncthreads();
    // T3 = T2 S -> Y, as T1
    if (live) {
        pair_product<true>(X, rb * 16, Fe, Fo, m, li, lk, Ev, Ov);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = rb * 16 + 4 * q + lk, b = cb * 16 + li;
            if (i < m && b < hm) {
                Y[i * MG_DS + b] = Ev[q] + Ov[q];
                if (m - 1 - b != b) Y[i * MG_DS + (m - 1 - b)] = Ev[q] - Ov[q];
            }
        }
    }
    __syncthreads();
    // E = S T3, as T2 unscaled: E[a][j] is node (j + 1, a + 1)
    if (live) {
        pair_product<false>(Y, rb * 16, Fe, Fo, m, li, lk, Ev, Ov);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int a = cb * 16 + 4 * q + lk, j = rb * 16 + li;
            if (a < hm && j < m) {
                e[(a + 1) * n + (j + 1)] = (T)(Ev[q] + Ov[q]);
                if (m - 1 - a != a) e[(m - a) * n + (j + 1)] = (T)(Ev[q] - Ov[q]);
            }
        }
    }
    // the boundary of the correction is zero
    for (int t = threadIdx.x; t < 4 * (n - 1); t += 1024) {
        const int side = t / (n - 1), p = t - side * (n - 1);
        const int idx = side == 0 ? p : side == 1 ? p * n + (n - 1) : side == 2 ? (n - 1) * n + (n - 1 - p)
                                                                          : (n - 1 - p) * n;
        e[idx] = (T)0;
    }
}

template <typename T, bool MASKED>
__global__ __launch_bounds__(1024) void mg_coarse_stage(const StageRef<MGCoarseTask<T>> a) {
    MGCoarseTask<T> tk;
    stage_case(a, tk);
    extern __shared__ __align__(16) unsigned char smem_mg[];
    if constexpr (sizeof(T) == 4) {
        if (tk.dtab) {   // tests the gate itself
            double* X = reinterpret_cast<double*>(smem_mg);
            coarse_direct_solve(tk, X, X + MG_DS * MG_DS);
            return;
        }
    }
    if (tk.gate && *G(&tk.gate->converged)) return;
    const gptr<const T> e_in = G(tk.e_in);
    const gptr<const T> rhs_top = G(tk.rhs);
    const gptr<T> e_out = G(tk.e_out);
    const int n_top = tk.n_top, Lc = tk.Lc, csweeps = tk.csweeps;
    const T h_top = tk.h_top, omega = tk.omega;
    T* base = reinterpret_cast<T*>(smem_mg);
    // level l: n = s_n[l], e at base + s_off[l], rhs right after it; the
    // residual scratch follows the last level (s_off[Lc])
    __shared__ int s_n[17], s_off[17];
    if ((int)threadIdx.x <= Lc) {
        int n = n_top, off = 0;
        for (int l = 0; l < (int)threadIdx.x; ++l) {
            off += 2 * n * n;
            n = (n + 1) / 2;
        }
        s_n[threadIdx.x] = n;
        s_off[threadIdx.x] = off;
    }
    auto E = [&](int l) { return base + s_off[l]; };
    auto R = [&](int l) { return base + s_off[l] + s_n[l] * s_n[l]; };
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < n_top; j += 32)
        for (int i = tx; i < n_top; i += 32) {
            const int k = j * n_top + i;
            base[k] = e_in ? e_in[k] : (T)0;
            base[n_top * n_top + k] = rhs_top[k];
        }
    // electrode bitmap of the top level (n_top <= 65)
    __shared__ unsigned s_mb[MASKED ? (65 * 65 + 31) / 32 : 1];
    const gptr<const uint8_t> gmask = MASKED ? G(tk.mask) : nullptr;
    if (MASKED && gmask) {
        const int nw = (n_top * n_top + 31) / 32;
        for (int w = threadIdx.x; w < nw; w += blockDim.x) {
            unsigned v = 0u;
            for (int b = 0; b < 32; ++b) {
                const int k = w * 32 + b;
                if (k < n_top * n_top && gmask[k]) v |= 1u << b;
            }
            s_mb[w] = v;
        }
    }
    const CoarseMask cm{(MASKED && gmask) ? s_mb : nullptr, n_top};
    __syncthreads();
    T h = h_top;
    for (int l = 0; l < Lc - 1; ++l) {
        const int nf = s_n[l], ncs = s_n[l + 1];
        lds_sweeps(E(l), R(l), nf, h, omega, (int)((tk.pre_bits >> (2 * l)) & 3u), cm, l);
        const T inv_h2 = (T)1 / (h * h);
        const T* e = E(l);
        const T* rl = R(l);
        T* res = base + s_off[Lc];
        for (int j = ty; j < nf; j += 32)
            for (int i = tx; i < nf; i += 32) {
                const int k = j * nf + i;
                T v = (T)0;
                if (!(i == 0 || i == nf - 1 || j == 0 || j == nf - 1) && !cm.fixed(l, i, j)) {
                    const T lap = lap_h2(e, k, nf) * inv_h2;
                    v = rl[k] - lap;
                }
                res[k] = v;
            }
        __syncthreads();
        for (int jc = ty; jc < ncs; jc += 32)
            for (int ic = tx; ic < ncs; ic += 32) {
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
                const int kc = jc * ncs + ic;
                R(l + 1)[kc] = s;
                E(l + 1)[kc] = (T)0;
            }
        __syncthreads();
        h = h * (T)2;
    }
    lds_sweeps(E(Lc - 1), R(Lc - 1), s_n[Lc - 1], h, omega, csweeps, cm, Lc - 1);
    for (int l = Lc - 2; l >= 0; --l) {
        h = h / (T)2;
        const int nf = s_n[l], ncs = s_n[l + 1];
        const T* C = E(l + 1);
        T* el = E(l);
        for (int j = ty; j < nf; j += 32)
            for (int i = tx; i < nf; i += 32) {
                if ((i >> 1) >= ncs || (j >> 1) >= ncs || cm.fixed(l, i, j)) continue;
                const int k = j * nf + i;
                el[k] = el[k] + prolong_lds(C, ncs, 0, 0, ncs, i, j);
            }
        __syncthreads();
        lds_sweeps(el, R(l), nf, h, omega, (int)((tk.post_bits >> (2 * l)) & 3u), cm, l);
    }
    for (int j = ty; j < n_top; j += 32)
        for (int i = tx; i < n_top; i += 32) e_out[j * n_top + i] = base[j * n_top + i];
}

// ---- mixed mode: exact FP64 solve of a level of 66..257 nodes -----------------
// The same transform as coarse_direct_solve, e = S (h^2 W o (S R S)) S, for
// m <= 255 unknowns a side, as four products of P x P matrices (P = m rounded
// up to a multiple of 64, zero padded), one launch each, S alternately on the
// left and on the right:
//   phase 0: Y = S R                (R: the level's FP32 rhs, interior nodes)
//   phase 1: T1 = (Y S) o h^2 W
//   phase 2: V = S T1
//   phase 3: e  = V S               (FP32, interior nodes; the boundary set to zero)
// The operand layout of v_mfma_f64_16x16x4f64 (A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k) reads a row-major B and a column-major A with
// consecutive lanes at consecutive addresses, so phases 0 and 2 store their
// product transposed (T0 = Y^T, then V^T), which phases 1 and 3 read as their
// A operand; S is symmetric, so its A-side entries S[i][k] are read as S[k][i].
// Each product is halved by the mirror symmetry of S (S[k][m-1-b] =
// (-1)^k S[k][b]): with S on the left, rows a and m-1-a of S X are E + O and
// E - O, where E and O sum over the even and the odd k only, so only the
// nb = (m+1)/2 rows a < nb are computed; with S on the right the same holds for
// the columns.  A workgroup owns a 16 x 16 tile of E and O: free block tile.tx
// (columns with S on the left, rows with S on the right) and halved block
// tile.ty.  Its four wavefronts take a quarter of the k range each (E and O
// chains of P/32 steps; every operand of a wavefront is loaded before its first
// product), and the four partial tiles are added in LDS in a fixed order, so
// the result does not depend on scheduling.  Only entries inside m x m are
// written; T0 and T1 are zeroed once when allocated, so their padding stays
// zero and the padded products add exact zeros.
template <int PHASE, int NS>
__global__ __launch_bounds__(256) void mg_dst_stage(const StageRef<MGDstTask> a) {
    MGTile tile;
    MGDstTask tk;
    stage_item(a, tk, tile);
    const int stop = tk.gate ? *G(&tk.gate->converged) : 0;
    constexpr bool LEFT = PHASE == 0 || PHASE == 2;   // S is the left operand
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int P = tk.P, m = tk.m, n = tk.n;
    const int nb = (m + 1) >> 1;           // computed rows (left) / columns (right)
    const int f0 = tile.tx * 16, h0 = tile.ty * 16;
    // steps of each chain: P/4 values of k, half of them even.  The launch's
    // NS is the largest of its cases; a smaller case loads its last step's
    // operands again (in-range addresses, no branch anywhere in the chain) and
    // zeroes them, so its extra products add exact zeros.
    const int ns = P >> 5;
    const int k0 = wave * (P >> 2);
    const gptr<const double> S = G(tk.S);
    const gptr<const double> T0 = G((const double*)tk.T0);
    const gptr<const double> T1 = G((const double*)tk.T1);
    const gptr<const float> R = G(tk.rhs);
    double ae[NS], be[NS], ao[NS], bo[NS];
    float re[NS], ro[NS];   // phase 0: R as loaded, converted when used (no wait behind each load)
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        const int ke = k0 + 8 * min(t, ns - 1) + 2 * lk, ko = ke + 1;
        if (LEFT) {
            // A[i][k] = S[h0 + li][k] = S[k][h0 + li]; B[k][j] = R or T1 at column f0 + li
            const int i = h0 + li, j = f0 + li;
            ae[t] = S[(size_t)ke * P + i];
            ao[t] = S[(size_t)ko * P + i];
            if (PHASE == 0) {
                // R: interior node (k + 1, j + 1), clamped into the grid (zeroed beyond m when used)
                const int jc = min(j, m - 1), kec = min(ke, m - 1), koc = min(ko, m - 1);
                re[t] = R[(size_t)(kec + 1) * n + (jc + 1)];
                ro[t] = R[(size_t)(koc + 1) * n + (jc + 1)];
            } else {
                be[t] = T1[(size_t)ke * P + j];
                bo[t] = T1[(size_t)ko * P + j];
            }
        } else {
            // A[i][k] = X[f0 + li][k] = T0[k][f0 + li] (X^T stored); B[k][j] = S[k][h0 + li]
            const int i = f0 + li, j = h0 + li;
            ae[t] = T0[(size_t)ke * P + i];
            ao[t] = T0[(size_t)ko * P + i];
            be[t] = S[(size_t)ke * P + j];
            bo[t] = S[(size_t)ko * P + j];
        }
    }
    if (stop) return;
    mg_v4d accE = mg_v4d{0.0, 0.0, 0.0, 0.0}, accO = mg_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        if (PHASE == 0) {
            const int ke = k0 + 8 * min(t, ns - 1) + 2 * lk, j = f0 + li;
            be[t] = (j < m && ke < m) ? (double)re[t] : 0.0;
            bo[t] = (j < m && ke + 1 < m) ? (double)ro[t] : 0.0;
        }
        if (t >= ns) ae[t] = ao[t] = 0.0;
        accE = __builtin_amdgcn_mfma_f64_16x16x4f64(ae[t], be[t], accE, 0, 0, 0);
        accO = __builtin_amdgcn_mfma_f64_16x16x4f64(ao[t], bo[t], accO, 0, 0, 0);
    }
    // partial tiles: D[r][c] sits in lane c + 16 (r % 4), element r / 4
    __shared__ double red[2][4][4][64];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        red[0][wave][q][lane] = accE[q];
        red[1][wave][q][lane] = accO[q];
    }
    __syncthreads();
    // thread -> tile entry (r, c), ordered so that consecutive lanes store to
    // consecutive addresses: r fastest for the transposed stores of phases 0
    // and 2, c fastest for the row-major stores of phases 1 and 3
    const int u = threadIdx.x & 15, v = threadIdx.x >> 4;
    const int r = LEFT ? u : v, c = LEFT ? v : u;
    const int sl = c + 16 * (r & 3), sq = r >> 2;
    const double vE = ((red[0][0][sq][sl] + red[0][1][sq][sl]) + red[0][2][sq][sl]) + red[0][3][sq][sl];
    const double vO = ((red[1][0][sq][sl] + red[1][1][sq][sl]) + red[1][2][sq][sl]) + red[1][3][sq][sl];
    // the computed entry (row, col) of the product and its mirror (row2, col2)
    int row, col, row2, col2;
    if (LEFT) {
        row = h0 + r;
        row2 = m - 1 - row;
        col = col2 = f0 + c;
    } else {
        row = row2 = f0 + r;
        col = h0 + c;
        col2 = m - 1 - col;
    }
    const int hidx = LEFT ? row : col;     // index along the halved side
    const int fidx = LEFT ? col : row;     // index along the free side
    if (hidx < nb && fidx < m) {
        const double v1 = vE + vO, v2 = vE - vO;
        const bool mirror = m - 1 - hidx != hidx;
        if (LEFT) {
            // stored transposed: T0[col][row]
            G(tk.T0)[(size_t)col * P + row] = v1;
            if (mirror) G(tk.T0)[(size_t)col2 * P + row2] = v2;
        } else if (PHASE == 1) {
            const gptr<const double> W = G(tk.W);
            G(tk.T1)[(size_t)row * P + col] = v1 * (tk.h2 * W[(size_t)row * P + col]);
            if (mirror) G(tk.T1)[(size_t)row2 * P + col2] = v2 * (tk.h2 * W[(size_t)row2 * P + col2]);
        } else {
            G(tk.e_out)[(size_t)(row + 1) * n + (col + 1)] = (float)v1;
            if (mirror) G(tk.e_out)[(size_t)(row2 + 1) * n + (col2 + 1)] = (float)v2;
        }
    }
    if (PHASE == 3 && tile.tx == 0 && tile.ty == 0) {
        // the boundary of the correction is zero
        for (int t = threadIdx.x; t < 4 * (n - 1); t += 256) {
            const int side = t / (n - 1), p = t - side * (n - 1);
            const int idx = side == 0 ? p : side == 1 ? p * n + (n - 1) : side == 2 ? (n - 1) * n + (n - 1 - p)
                                                                              : (n - 1 - p) * n;
            G(tk.e_out)[idx] = 0.0f;
        }
    }
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

// FP64 residual r = rhs - A phi (zero on the boundary) on a 64x16 tile, and
// the tile's sum of r^2.  Thread (tx, ty) owns column tx, rows 4ty..4ty+3.
//   MODE 0: phi as is (FP64 V-cycles).
//   MODE 1: phi as is, r stored FP32 as the next FP32 cycle's rhs.
//   MODE 2: the correction of the FP32 cycle just done is applied first:
//           phi_new = phi_src + e (every node value, neighbours included, is
//           formed the same way), written to the other phi buffer (parity 0:
//           phi -> phi_tmp, 1: phi_tmp -> phi); r stored FP32.
// which == 1: the residual before the first cycle; it also writes the tile's
// sum of rhs^2 (every node, the boundary included) after the r^2 sums and,
// with phi_save, copies phi.
// CHARGE (with which == 1): the rhs of the tile's nodes comes from the step's
// charge conversion, done here node by node (rhs and rho written, the
// accumulators cleared) instead of in a launch of its own, with the values
// that launch would write.
template <int MODE, bool CHARGE>
__device__ inline void residual_tile(const MGFineTask& tk, const MGTile& tile, int which, int parity,
                                     const ChargeArgs& chg) {
    __shared__ double sh[256];
    const int n = tk.n;
    const gptr<const double> src = G((MODE == 2 && parity) ? tk.phi_tmp : tk.phi);
    const gptr<double> dst = G((MODE == 2) ? (parity ? tk.phi : tk.phi_tmp) : nullptr);
    const gptr<const float> e32 = G(tk.e32);
    const gptr<const double> rhs = G(tk.rhs);
    const gptr<float> r32 = G(tk.r32);
    const gptr<const uint8_t> mask = G(tk.mask);   // electrodes: zero residual
    const double inv_h2 = tk.inv_h2;
    const int i = tile.tx * MG_RX + (threadIdx.x & (MG_RX - 1));
    const int jb = tile.ty * MG_RY + (threadIdx.x / MG_RX) * 4;
    auto node = [&](int ii, int jj) -> double {
        const size_t g = (size_t)jj * n + ii;
        double v = src[g];
        if (MODE == 2) v = v + (double)e32[g];
        return v;
    };
    double acc = 0.0, acc_rhs = 0.0;
    const gptr<double> save = G(which == 1 ? tk.phi_save : nullptr);
    if (i < n) {
        double col[6], lf[4], rt[4], rh[4], fv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = jb + q;
            fv[q] = (CHARGE && j < n) ? charge_node(chg, (size_t)j * n + i) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int j = jb - 1 + q;
            col[q] = (j >= 0 && j < n) ? node(i, j) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = jb + q;
            const bool in = j > 0 && j < n - 1 && i > 0 && i < n - 1;
            lf[q] = in ? node(i - 1, j) : 0.0;
            rt[q] = in ? node(i + 1, j) : 0.0;
            rh[q] = in ? (CHARGE ? fv[q] : rhs[(size_t)j * n + i]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = jb + q;
            if (j >= n) break;
            const size_t g = (size_t)j * n + i;
            const double c = col[q + 1];
            double r = 0.0;
            if (!(i == 0 || i == n - 1 || j == 0 || j == n - 1) && !(mask && mask[g])) {
                const double lap = (col[q] + col[q + 2] + lf[q] + rt[q] - 4.0 * c) * inv_h2;
                r = rh[q] - lap;
            }
            if (MODE == 2) dst[g] = c;
            if (MODE >= 1) r32[g] = (float)r;
            acc += r * r;
            if (which == 1) {
                const double f = CHARGE ? fv[q] : rhs[g];
                acc_rhs += f * f;
                if (save) save[g] = c;
            }
        }
    }
    const double s = block_sum256(acc, sh);
    const int tix = tile.ty * tk.gx + tile.tx;
    if (threadIdx.x == 0) G(tk.partial)[tix] = s;
    if (which == 1) {
        __syncthreads();
        const double s2 = block_sum256(acc_rhs, sh);
        if (threadIdx.x == 0) G(tk.partial)[tk.nblocks + tix] = s2;
    }
}

// One residual tile per workgroup (residual_tile); with which == 2, the
// residual at the end of a cycle, a no-op once the case stopped.
template <int MODE>
__global__ __launch_bounds__(256) void ir_residual_stage(const StageRef<MGFineTask> a, int which, int parity) {
    MGTile tile;
    MGFineTask tk;
    stage_item(a, tk, tile);
    if (which == 2 && *G(&tk.norm->converged)) return;
    residual_tile<MODE, false>(tk, tile, which, parity, ChargeArgs{});
}

// The first residual of a step's solve (which == 1) with the step's charge
// conversion and end of step in the same launch, for a solve of one case: one
// block more than the residual tiles; the block renumbered first ends the step
// (clears, flips, spill status and re-binning flag, as end_step_kernel), the
// others take the residual tiles (tile numbers and partial sums as in
// ir_residual_stage, so the norms are the same bits).
template <int MODE, typename T>
__global__ __launch_bounds__(256) void step_residual_stage(const StageRef<MGFineTask> a, const ChargeArgs chg,
                                                           const CaseDev<T>* __restrict__ cases, CaseState* states,
                                                           int* __restrict__ status_out,
                                                           int* __restrict__ rebin_out) {
    const int b = xcd_block();
    if (b == 0) {
        end_step_block(cases, states, 0, 1, status_out, rebin_out);
        return;
    }
    MGTile tile;
    MGFineTask tk;
    stage_item_at(a, tk, tile, b - 1);
    residual_tile<MODE, true>(tk, tile, 1, 0, chg);
}

// End of a mixed-mode solve: after an odd number of executed cycles the
// solution sits in phi_tmp; copy it back.
__global__ __launch_bounds__(256) void ir_finish_stage(const StageRef<MGFineTask> a) {
    MGTile tile;
    MGFineTask tk;
    stage_item(a, tk, tile);
    if (!(*G(&tk.norm->steps) & 1)) return;
    const size_t nn = (size_t)tk.n * tk.n;
    const size_t i = (size_t)tile.tx * 256 + threadIdx.x;
    if (i < nn) G(tk.phi)[i] = G(tk.phi_tmp)[i];
}

// Electrodes: the Dirichlet nodes of the mask take their values (solid nodes
// keep theirs), at the start of every solve.
__global__ __launch_bounds__(256) void dirichlet_stage(const StageRef<MGFineTask> a) {
    MGTile tile;
    MGFineTask tk;
    stage_item(a, tk, tile);
    if (!tk.mask) return;
    const size_t nn = (size_t)tk.n * tk.n;
    const size_t i = (size_t)tile.tx * 256 + threadIdx.x;
    if (i < nn && G(tk.mask)[i] == 1) G(tk.phi)[i] = G(tk.phi_bc)[i];
}

// One workgroup per case: fixed-order sum of the block partials.  which = 0:
// start of a solve (rhs norm, clear the flag and the cycle count); 1: initial
// residual test; 2: residual test at the end of a cycle, counted, and a no-op
// once the flag is set, so cycles enqueued after convergence change nothing.
// Relative residual ||r||/||rhs|| (the reference divides by sum rhs^2,
// Multigrid.cu:466).
// Thread t's share of a fixed-order sum: p[t], p[t + 256], p[t + 512], ... added
// in that order.  The loads of eight terms are issued before their adds (in
// range: clamped), so their latencies overlap instead of adding up.
__device__ inline double strided_sum(const gptr<const double> p, int np) {
    double acc = 0.0;
    for (int base = threadIdx.x; base < np; base += 8 * 256) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = p[min(base + 256 * q, np - 1)];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (base + 256 * q < np) acc += v[q];
    }
    return acc;
}

__global__ __launch_bounds__(256) void norm_final_stage(const StageRef<MGFineTask> a, int which, double tol,
                                                        int max_cycles, int inner) {
    MGFineTask tk;
    stage_case(a, tk);
    MGNorm* out = tk.norm;
    if (which == 2 && out->converged) return;
    __shared__ double sh[256];
    const int np = tk.nblocks;
    const double tot = block_sum256(strided_sum(G((const double*)tk.partial), np), sh);
    double tot_rhs = 0.0;
    if (which == 1) {   // a new solve: the rhs norm from the same pass, counters reset
        __syncthreads();
        tot_rhs = block_sum256(strided_sum(G((const double*)tk.partial) + np, np), sh);
    }
    if (threadIdx.x == 0) {
        if (which == 1) {
            out->rhs_sq = tot_rhs;
            out->converged = 0;
            out->cycles = 0;
            out->steps = 0;
            out->max_cycles = max_cycles;
        }
        out->res_sq = tot;
        const double rn = sqrt(tot), fn = sqrt(out->rhs_sq);
        const double rel = fn > 0.0 ? rn / fn : rn;
        out->rel = rel;
        if (which == 2) {
            out->cycles += inner;   // V-cycles of the refinement step just done
            out->steps += 1;
        }
        out->converged = (rel < tol || out->cycles >= out->max_cycles) ? 1 : 0;
        out->hit_tol = rel < tol ? 1 : 0;
        *tk.mirror = *out;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void mg_copy_stage(const StageRef<MGLevelTask<T>> a) {
    MGTile tile;
    MGLevelTask<T> tk;
    stage_item(a, tk, tile);
    if (tk.gate && tk.gate->converged) return;
    const size_t nn = (size_t)tk.n * tk.n;
    const size_t i = (size_t)tile.tx * 256 + threadIdx.x;
    if (i < nn) G(tk.e_out)[i] = G(tk.e_in)[i];
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
    if (dst_t_) (void)hipFree(dst_t_);
    for (auto* q : mask_) (void)hipFree(q);
    if (phi_bc_) (void)hipFree(phi_bc_);
}

int MGSolver::set_mask(const uint8_t* mask, const double* phi_bc) {
    PIC_REQUIRE(ok_, "multigrid: solver not initialised");
    PIC_HIP(hipStreamSynchronize(stream_));
    for (auto* q : mask_) (void)hipFree(q);
    mask_.clear();
    if (phi_bc_) (void)hipFree(phi_bc_);
    phi_bc_ = nullptr;
    delete solo_;   // its launch tables hold the old mask pointers
    solo_ = nullptr;
    if (!mask) return 0;
    const size_t nn = (size_t)n_ * n_;
    std::vector<uint8_t> cur(mask, mask + nn);
    for (size_t k = 0; k < nn; ++k) PIC_REQUIRE(cur[k] <= 2, "electrode mask values are 0, 1 or 2");
    bool dirichlet = false;
    for (size_t k = 0; k < nn; ++k) dirichlet = dirichlet || cur[k] == 1;
    PIC_REQUIRE(!dirichlet || phi_bc, "Dirichlet nodes need phi_bc");
    mask_.assign(L_, nullptr);
    for (int l = 0; l < L_; ++l) {
        const int n = ns_[l];
        if (l > 0) {
            // coarse node (ic, jc) takes the mask of the fine node (2ic, 2jc)
            const int nf = ns_[l - 1];
            std::vector<uint8_t> c((size_t)n * n);
            for (int jc = 0; jc < n; ++jc)
                for (int ic = 0; ic < n; ++ic) c[(size_t)jc * n + ic] = cur[(size_t)(2 * jc) * nf + 2 * ic];
            cur.swap(c);
        }
        PIC_HIP(hipMalloc(&mask_[l], (size_t)n * n));
        PIC_HIP(hipMemcpy(mask_[l], cur.data(), (size_t)n * n, hipMemcpyHostToDevice));
    }
    PIC_HIP(hipMalloc(&phi_bc_, nn * sizeof(double)));
    if (phi_bc) PIC_HIP(hipMemcpy(phi_bc_, phi_bc, nn * sizeof(double), hipMemcpyHostToDevice));
    else PIC_HIP(hipMemset(phi_bc_, 0, nn * sizeof(double)));
    return 0;
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
    for (int l = 0; l < 16 && p_.per_level; ++l)
        PIC_REQUIRE(p_.pre_l[l] >= 0 && p_.pre_l[l] <= 3 && p_.post_l[l] >= 0 && p_.post_l[l] <= 3,
                    "multigrid: per-level sweeps must be in [0,3]");
    PIC_REQUIRE(p_.coarse_sweeps >= 0, "multigrid: coarse sweeps must be non-negative");
    L_ = L;
    lc_ = L - 1;
    nc_ = p_.mode == PIC_MG_MIXED ? MG_NC_F32 : MG_NC_F64;
    for (int l = 0; l < L; ++l)
        if (ns_[l] <= nc_) { lc_ = l; break; }
    lcd_ = -1;
    if (p_.mode == PIC_MG_MIXED)
        for (int l = 0; l < L; ++l)
            if (ns_[l] <= MG_NDB) { lcd_ = l; break; }
    if (p_.mode == PIC_MG_MIXED && p_.coarse_direct == 2 && lcd_ >= 0 && lcd_ < lc_) {
        const size_t P = dst_pad(ns_[lcd_] - 2);
        PIC_HIP(hipMalloc(&dst_t_, 2 * P * P * sizeof(double)));
        // the products write only inside m x m: the padding stays zero
        PIC_HIP(hipMemsetAsync(dst_t_, 0, 2 * P * P * sizeof(double), stream_));
    }
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
    PIC_HIP(hipMalloc(&phi_tmp_, nn * sizeof(double)));
    n_partials_ = 2 * div_up(n_, MG_RX) * div_up(n_, MG_RY);   // r^2 and rhs^2 per residual tile
    PIC_HIP(hipMalloc(&partials_, (size_t)n_partials_ * sizeof(double)));
    PIC_HIP(hipMalloc(&d_norm_, sizeof(MGNorm)));
    PIC_HIP(hipMemset(d_norm_, 0, sizeof(MGNorm)));
    ok_ = true;
    return 0;
}

int MGSolver::prepare(double* phi, const double* rhs, double h) {
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
    return 0;
}

static void fill_stats(const MGNorm& nm, MGStats* st) {
    if (!st) return;
    st->cycles = nm.cycles;
    st->rel_residual = nm.rel;
    st->rhs_norm = std::sqrt(nm.rhs_sq);
}

int MGSolver::solve(double* phi, const double* rhs, double h, MGStats* st) {
    if (int rc = prepare(phi, rhs, h)) return rc;
    if (int rc = solo_->solve()) return rc;
    launches_ = solo_->last_launches();
    fill_stats(solo_->norm(0), st);
    return 0;
}

int MGSolver::enqueue_fixed(int cycles, const MGStepFuse* fuse) {
    PIC_REQUIRE(solo_, "multigrid: prepare() first");
    return solo_->enqueue_fixed(cycles, fuse);
}

int MGSolver::complete_fixed(MGStats* st) {
    PIC_REQUIRE(solo_, "multigrid: prepare() first");
    if (int rc = solo_->complete_fixed()) return rc;
    launches_ = solo_->last_launches();
    fill_stats(solo_->norm(0), st);
    return 0;
}

int MGSolver::last_cycles() const { return solo_ ? solo_->last_cycles() : 1; }

const int* MGSolver::stop_flag() const { return d_norm_ ? &d_norm_->converged : nullptr; }

bool MGSolver::fixed_stopped() const { return solo_ && solo_->norm(0).converged != 0; }

void MGSolver::set_phi_save(double* dst) {
    if (dst != phi_save_) {   // the launch tables of the batch of one hold the pointer
        delete solo_;
        solo_ = nullptr;
    }
    phi_save_ = dst;
}

// ---------------------------------------------------------------------------
// MGBatch: one launch per stage for all member grids
// ---------------------------------------------------------------------------
// the levels' e and rhs and the residual scratch, or the two FP64 buffers of
// the direct solve
template <typename T>
static size_t coarse_lds_bytes(int n_top, int Lc, bool direct) {
    if (direct) return MG_DLDS;
    size_t tot = 0;
    int n = n_top;
    for (int l = 0; l < Lc; ++l) { tot += 2 * (size_t)n * n; n = (n + 1) / 2; }
    tot += (size_t)n_top * n_top;
    return tot * sizeof(T);
}

// LDS of a tiled stage: e and rhs over tile + halo (the residual before a
// restriction replaces rhs in place), plus the coarse patch of a prolongation
// (kinds 1 and 4)
static size_t tiled_lds(int kind, int nu, bool restr, size_t es, bool masked) {
    const bool res = (kind == 0 && restr) || kind == 4;
    const size_t H = res ? 2 * nu + 2 : 2 * nu;
    const size_t W = MG_TILE + 2 * H;
    size_t tot = 2 * W * W;
    if (kind == 1 || kind == 4) tot += (W / 2 + 2) * (W / 2 + 2);
    return tot * es + (masked ? W * W : 0);   // + the electrode mask tile (bytes)
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
    if (h_mirror_) (void)hipHostFree(h_mirror_);
    if (d_dtab_) (void)hipFree(d_dtab_);
    d_dtab_ = nullptr;
    if (d_btab_) (void)hipFree(d_btab_);
    d_btab_ = nullptr;
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
int MGBatch::add_level_stage(int kind, const std::vector<MGLevelTask<T>>& tasks, int nu, bool restr,
                             std::vector<Stage>* dst) {
    if (tasks.empty()) return 0;
    std::vector<MGTile> tiles;
    for (int k = 0; k < (int)tasks.size(); ++k) {
        const int n = tasks[k].n;
        if (kind == 3) {
            const int chunks = (int)(((size_t)n * n + 255) / 256);
            for (int c = 0; c < chunks; ++c) tiles.push_back(MGTile{k, c, 0, 0});
        } else {
            const int g = mg_tiles(n);
            for (int ty = 0; ty < g; ++ty)
                for (int tx = 0; tx < g; ++tx) tiles.push_back(MGTile{k, tx, ty, 0});
        }
    }
    Stage st;
    st.kind = kind;
    st.nu = nu;
    st.restr = restr;
    st.fp64 = sizeof(T) == 8;
    st.n_items = (int)tiles.size();
    st.n_tasks = (int)tasks.size();
    for (const auto& t : tasks) st.masked = st.masked || t.mask != nullptr;
    st.lds = kind == 3 ? 0 : tiled_lds(kind, nu, restr, sizeof(T), st.masked);
    // a launch of at most one tile per CU (the small levels) is bound by the
    // latency of one workgroup's sweeps: 1024 threads per tile; up to four
    // tiles per CU, 512 (the CU's 32 wavefronts either way); 256 beyond
    st.bs = 256;
    if (kind == 0 || kind == 1) st.bs = st.n_items <= MG_SMALL_LAUNCH ? 1024 : st.n_items <= 4 * MG_SMALL_LAUNCH ? 512 : 256;
    if (kind == 4 && st.n_items <= 4 * MG_SMALL_LAUNCH) st.bs = 512;
    st.gx = kind == 3 ? 0 : mg_tiles(tasks[0].n);
    static_assert(sizeof(MGLevelTask<T>) <= sizeof(st.t0), "task does not fit the by-value slot");
    std::memcpy(st.t0, &tasks[0], sizeof(MGLevelTask<T>));
    if (int rc = upload_vec(&st.tasks, tasks)) return rc;
    if (int rc = upload_vec((void**)&st.tiles, tiles)) return rc;
    (dst ? dst : &cycle_)->push_back(st);
    return 0;
}

// the four product launches of the big direct solve, over the (case, tile) items
int MGBatch::add_dst_stages(const std::vector<MGDstTask>& tasks, std::vector<Stage>* dst) {
    if (tasks.empty()) return 0;
    std::vector<MGTile> tiles;
    for (int k = 0; k < (int)tasks.size(); ++k)
        for (int ty = 0; ty < tasks[k].nh; ++ty)
            for (int tx = 0; tx < tasks[k].gx; ++tx) tiles.push_back(MGTile{k, tx, ty, 0});
    int ns = 2;   // chain steps of the largest case: P / 32
    for (const auto& t : tasks) ns = std::max(ns, t.P / 32);
    for (int phase = 0; phase < 4; ++phase) {
        Stage st;
        st.kind = 5;
        st.nu = phase;
        st.ns = ns;
        st.n_items = (int)tiles.size();
        st.n_tasks = (int)tasks.size();
        st.gx = tasks[0].gx;
        static_assert(sizeof(MGDstTask) <= sizeof(st.t0), "task does not fit the by-value slot");
        std::memcpy(st.t0, &tasks[0], sizeof(MGDstTask));
        if (int rc = upload_vec(&st.tasks, tasks)) return rc;
        if (int rc = upload_vec((void**)&st.tiles, tiles)) return rc;
        (dst ? dst : &cycle_)->push_back(st);
    }
    return 0;
}

// The stages of one refinement step (FP64 mode: one V-cycle): inner_ V-cycles,
// the later ones starting from the correction so far.  Between two V-cycles
// the finest up-leg stage of a member with tiled levels also runs the next
// cycle's pre-sweeps and restriction (stage kind 4), so that cycle's down leg
// starts one level lower; its level-0 correction then alternates between the
// two level-0 buffers (final_e0 tells where it ends).  Only batches whose
// finest levels have at most MG_FUSE_MAX nodes a side fuse: on larger grids the
// fused stage's bigger tile (fewer workgroups resident per CU) costs more than
// the launch it saves (2049^2: 67.9 us against 27.9 + 27.9 us).
template <typename T>
int MGBatch::build_step(const std::vector<double>& h) {
    std::vector<Stage>* dst = &cycle_;
    const int M = (int)m_.size();
    const T omega = (T)p_.omega;
    auto E = [&](MGSolver* s, int b, int l) -> T* { return (T*)(mixed_ ? (void*)s->ef_[b][l] : (void*)s->ed_[b][l]); };
    auto R = [&](MGSolver* s, int l) -> T* { return (T*)(mixed_ ? (void*)s->rf_[l] : (void*)s->rd_[l]); };
    auto out0 = [&](int k) -> T* { return mixed_ ? (T*)(void*)m_[k]->ef_[0][0] : (T*)(void*)phi_[k]; };
    auto tmp0 = [&](int k) -> T* { return mixed_ ? (T*)(void*)m_[k]->ef_[1][0] : (T*)(void*)m_[k]->phi_tmp_; };
    auto rhs0 = [&](int k) -> const T* { return mixed_ ? (const T*)(void*)m_[k]->r32_ : (const T*)(void*)rhs_[k]; };
    auto hl = [&](int k, int l) { return h[k] * std::ldexp(1.0, l); };
    int D = 0;
    for (auto* s : m_) D = std::max(D, s->coarse_level());
    auto level_task = [&](int k, int d) {
        MGSolver* s = m_[k];
        const double hh = hl(k, d);
        MGLevelTask<T> t{};
        t.rhs = d == 0 ? rhs0(k) : R(s, d);
        t.n = s->ns_[d];
        t.h2 = (T)(hh * hh);
        t.inv_h2 = (T)(1.0 / (hh * hh));
        t.omega = omega;
        t.gate = s->d_norm_;
        t.mask = s->masked() ? s->mask_[d] : nullptr;
        return t;
    };
    // level-0 buffers of members with tiled levels: the pre-smoothed
    // correction (input of the finest up stage) and the up stage's output
    std::vector<T*> pre0(M), post0(M);
    for (int k = 0; k < M; ++k) {
        pre0[k] = tmp0(k);
        post0[k] = out0(k);
    }
    final_e0_.assign(M, 0);
    bool fuse_ok = mixed_;
    for (auto* s : m_) fuse_ok = fuse_ok && s->ns_[0] <= MG_FUSE_MAX;
    for (int c = 0; c < inner_; ++c) {
        const bool first = c == 0, last = c == inner_ - 1;
        // the level-0 input of a cycle: FP64 mode phi; mixed mode zero, or the
        // correction so far in the later cycles
        auto in0 = [&](int k) -> const T* {
            return mixed_ ? (first ? nullptr : (const T*)out0(k)) : (const T*)(void*)phi_[k];
        };
        // down leg (after the first cycle, level 0 was done by the fused stage)
        for (int d = (first || !fuse_ok) ? 0 : 1; d < D; ++d) {
            std::vector<MGLevelTask<T>> tasks;
            for (int k = 0; k < M; ++k) {
                MGSolver* s = m_[k];
                if (s->coarse_level() <= d) continue;
                MGLevelTask<T> t = level_task(k, d);
                t.e_in = d == 0 ? in0(k) : nullptr;
                t.e_out = d == 0 ? pre0[k] : E(s, 1, d);
                t.rhs_c = R(s, d + 1);
                t.nu = p_.pre_at(d);
                t.H = 2 * t.nu + 2;
                tasks.push_back(t);
            }
            if (int rc = add_level_stage<T>(0, tasks, p_.pre_at(d), true, dst)) return rc;
        }
        // coarse part: the exact solve of a level of 66..257 nodes over several
        // workgroups, the single-workgroup cycle, or tiled sweeps on a large coarsest level
        if constexpr (sizeof(T) == 4) {
            std::vector<MGDstTask> dt;
            for (int k = 0; k < M; ++k) {
                MGSolver* s = m_[k];
                if (!s->big_direct() || !btab_[k]) continue;
                const int lc = s->lcd_;
                MGDstTask t{};
                t.rhs = lc == 0 ? (const float*)rhs0(k) : (const float*)R(s, lc);
                t.e_out = lc == 0 ? (float*)out0(k) : (float*)E(s, 0, lc);
                t.n = s->ns_[lc];
                t.m = t.n - 2;
                t.P = dst_pad(t.m);
                t.gx = t.P / 16;
                t.nh = div_up((t.m + 1) / 2, 16);
                t.S = btab_[k];
                t.W = btab_[k] + (size_t)t.P * t.P;
                t.T0 = s->dst_t_;
                t.T1 = s->dst_t_ + (size_t)t.P * t.P;
                const double hh = hl(k, lc);
                t.h2 = hh * hh;
                t.gate = s->d_norm_;
                dt.push_back(t);
            }
            if (int rc = add_dst_stages(dt, dst)) return rc;
        }
        {
            std::vector<MGCoarseTask<T>> ct;
            size_t lds = 0;
            for (int k = 0; k < M; ++k) {
                MGSolver* s = m_[k];
                if (s->big_direct()) continue;
                const int lc = s->lc_;
                if (s->ns_[lc] > s->nc_) continue;
                MGCoarseTask<T> t{};
                t.e_in = lc == 0 ? in0(k) : nullptr;
                t.rhs = lc == 0 ? rhs0(k) : R(s, lc);
                t.e_out = lc == 0 ? out0(k) : E(s, 0, lc);
                t.n_top = s->ns_[lc];
                t.Lc = s->L_ - lc;
                t.pre_bits = t.post_bits = 0u;
                for (int l = 0; l < 16; ++l) {
                    const unsigned a = (unsigned)(lc + l < 16 ? p_.pre_at(lc + l) : p_.pre);
                    const unsigned b = (unsigned)(lc + l < 16 ? p_.post_at(lc + l) : p_.post);
                    t.pre_bits |= (a & 3u) << (2 * l);
                    t.post_bits |= (b & 3u) << (2 * l);
                }
                t.csweeps = s->coarse_sweeps_eff();
                t.h_top = (T)hl(k, lc);
                t.omega = omega;
                t.gate = s->d_norm_;
                t.mask = s->masked() ? s->mask_[lc] : nullptr;
                t.dtab = sizeof(T) == 4 ? dtab_[k] : nullptr;
                lds = std::max(lds, coarse_lds_bytes<T>(t.n_top, t.Lc, t.dtab != nullptr));
                ct.push_back(t);
            }
            if (!ct.empty()) {
                Stage st;
                st.kind = 2;
                st.fp64 = sizeof(T) == 8;
                st.lds = lds;
                st.n_items = (int)ct.size();
                st.n_tasks = (int)ct.size();
                for (const auto& t : ct) st.masked = st.masked || t.mask != nullptr;
                static_assert(sizeof(MGCoarseTask<T>) <= sizeof(st.t0), "task does not fit the by-value slot");
                std::memcpy(st.t0, &ct[0], sizeof(MGCoarseTask<T>));
                if (int rc = upload_vec(&st.tasks, ct)) return rc;
                dst->push_back(st);
            }
            // large coarsest levels: chunks of up to 3 sweeps, ping-pong, then a copy
            std::vector<int> big;
            for (int k = 0; k < M; ++k)
                if (!m_[k]->big_direct() && m_[k]->ns_[m_[k]->lc_] > m_[k]->nc_) big.push_back(k);
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
                        MGLevelTask<T> t = level_task(k, lc);
                        t.e_in = src[k];
                        t.e_out = which[k] ? b1[k] : b0[k];
                        t.nu = nu;
                        t.H = 2 * nu;
                        tasks.push_back(t);
                        src[k] = t.e_out;
                        which[k] ^= 1;
                    }
                    if (int rc = add_level_stage<T>(0, tasks, nu, false, dst)) return rc;
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
                if (int rc = add_level_stage<T>(3, copies, 0, false, dst)) return rc;
            }
        }
        // up leg; level 0 fused with the next cycle's down-leg stage unless last
        for (int d = D - 1; d >= 0; --d) {
            const bool fuse = d == 0 && !last && fuse_ok;
            std::vector<MGLevelTask<T>> tasks;
            for (int k = 0; k < M; ++k) {
                MGSolver* s = m_[k];
                if (s->coarse_level() <= d) continue;
                MGLevelTask<T> t = level_task(k, d);
                t.e_in = d == 0 ? pre0[k] : E(s, 1, d);
                t.coarse = E(s, 0, d + 1);
                t.e_out = d == 0 ? post0[k] : E(s, 0, d);
                t.nu = fuse ? p_.post_at(0) + p_.pre_at(0) : p_.post_at(d);
                t.H = fuse ? 2 * t.nu + 2 : 2 * t.nu;
                if (fuse) t.rhs_c = R(s, 1);
                tasks.push_back(t);
            }
            if (tasks.empty()) continue;
            if (fuse) {
                if (int rc = add_level_stage<T>(4, tasks, tasks[0].nu, true, dst)) return rc;
            } else if (int rc = add_level_stage<T>(1, tasks, p_.post_at(d), false, dst)) {
                return rc;
            }
        }
        // after a fused stage the next cycle's finest up stage reads what it wrote
        for (int k = 0; k < M; ++k)
            if (m_[k]->coarse_level() > 0 && !last && fuse_ok) std::swap(pre0[k], post0[k]);
    }
    for (int k = 0; k < M; ++k) final_e0_[k] = (m_[k]->coarse_level() > 0 && post0[k] != out0(k)) ? 1 : 0;
    return 0;
}

// The tables of the sine-transform solve for m unknowns a side laid out P x P
// (S then W, zero outside m x m), computed in long double once per (m, P,
// form) and kept for the process: a batch rebuild, which every admission to a
// serving ensemble causes, then only copies them (computing a 255^2 set took
// 5 ms of host time per size and rebuild).  reduced: the big solve's form,
// the sine argument (a+1)(b+1) taken mod 2(m+1).
static const std::vector<double>& dst_tables(int m, int P, bool reduced) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, bool>, std::vector<double>> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(m, P, reduced);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<double> tab(2 * (size_t)P * P, 0.0);
    double* S = tab.data();
    double* W = S + (size_t)P * P;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> sig(m);
    for (int p = 0; p < m; ++p) {
        const long double sp = std::sin(pi * (p + 1) / (2.0L * (m + 1)));
        sig[p] = sp * sp;
    }
    const long double norm = (2.0L / (m + 1)) * (2.0L / (m + 1));
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) {
            if (reduced) {
                // (a+1)(b+1) reduced mod 2(m+1) keeps the sine argument in [0, 2 pi)
                // (sin of a multiple of pi is exactly zero, so the middle
                // column's odd-k sum in the halved products is zero as well)
                const long long t = (long long)(a + 1) * (b + 1) % (2LL * (m + 1));
                S[(size_t)a * P + b] = t % (m + 1) == 0 ? 0.0
                                                         : (double)std::sin(pi * (long double)t / (long double)(m + 1));
            } else {
                S[(size_t)a * P + b] = (double)std::sin(pi * (long double)((a + 1) * (b + 1)) / (long double)(m + 1));
            }
            W[(size_t)a * P + b] = (double)(norm / (-4.0L * (sig[a] + sig[b])));
        }
    return cache.emplace(key, std::move(tab)).first->second;
}

// Direct coarse solve (mixed mode, members without electrodes whose coarse
// kernel level has 3..65 nodes): the sine-transform tables of that size, one
// set per size (S then W, 64 x 64 each, zero outside m x m).
// Tables of the big direct solve, one set per size: S then W, each P x P,
// zero outside m x m (m = n - 2 unknowns a side, P = m rounded up to 64).
int MGBatch::build_big_tables() {
    const int M = (int)m_.size();
    std::vector<int> sizes;
    std::vector<size_t> offs;
    size_t total = 0;
    for (int k = 0; k < M; ++k) {
        const MGSolver* s = m_[k];
        if (!s->big_direct()) continue;
        const int n = s->ns_[s->lcd_];
        if (std::find(sizes.begin(), sizes.end(), n) != sizes.end()) continue;
        sizes.push_back(n);
        offs.push_back(total);
        const size_t P = dst_pad(n - 2);
        total += 2 * P * P;
    }
    if (sizes.empty()) return 0;
    std::vector<double> tab(total, 0.0);
    for (size_t q = 0; q < sizes.size(); ++q) {
        const int m = sizes[q] - 2, P = dst_pad(m);
        const std::vector<double>& t = dst_tables(m, P, true);
        std::copy(t.begin(), t.end(), tab.begin() + offs[q]);
    }
    if (int rc = upload_vec((void**)&d_btab_, tab)) return rc;
    for (int k = 0; k < M; ++k) {
        const MGSolver* s = m_[k];
        if (!s->big_direct()) continue;
        const size_t q = std::find(sizes.begin(), sizes.end(), s->ns_[s->lcd_]) - sizes.begin();
        btab_[k] = d_btab_ + offs[q];
    }
    return 0;
}

int MGBatch::build_direct_tables() {
    const int M = (int)m_.size();
    dtab_.assign(M, nullptr);
    btab_.assign(M, nullptr);
    if (!mixed_ || !p_.coarse_direct) return 0;
    if (int rc = build_big_tables()) return rc;
    std::vector<int> sizes, want(M, 0);
    for (int k = 0; k < M; ++k) {
        const MGSolver* s = m_[k];
        const int n = s->ns_[s->lc_];
        if (s->masked() || n > MG_ND || n < 3) continue;
        want[k] = n;
        if (std::find(sizes.begin(), sizes.end(), n) == sizes.end()) sizes.push_back(n);
    }
    if (sizes.empty()) return 0;
    const size_t per = 2 * 64 * 64;
    std::vector<double> tab(sizes.size() * per, 0.0);
    for (size_t q = 0; q < sizes.size(); ++q) {
        const std::vector<double>& t = dst_tables(sizes[q] - 2, 64, false);
        std::copy(t.begin(), t.end(), tab.begin() + q * per);
    }
    if (int rc = upload_vec((void**)&d_dtab_, tab)) return rc;
    for (int k = 0; k < M; ++k)
        if (want[k]) {
            const size_t q = std::find(sizes.begin(), sizes.end(), want[k]) - sizes.begin();
            dtab_[k] = d_dtab_ + q * per;
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
        bool same_levels = q.per_level == p_.per_level;
        for (int l = 0; l < 16 && same_levels && q.per_level; ++l)
            same_levels = q.pre_l[l] == p_.pre_l[l] && q.post_l[l] == p_.post_l[l];
        PIC_REQUIRE(q.mode == p_.mode && q.pre == p_.pre && q.post == p_.post && q.coarse_sweeps == p_.coarse_sweeps &&
                        q.omega == p_.omega && q.tol == p_.tol && q.max_cycles == p_.max_cycles && same_levels &&
                        q.inner_eff() == p_.inner_eff() && q.coarse_direct == p_.coarse_direct,
                    "multigrid batch: members need identical cycle parameters");
    }
    mixed_ = p_.mode == PIC_MG_MIXED;
    inner_ = p_.inner_eff();
    max_outer_ = std::max(1, (p_.max_cycles + inner_ - 1) / inner_);
    const int M = (int)m_.size();
    // the norm kernel writes the norms straight into pinned, coherent host
    // memory, read after each synchronization
    PIC_HIP(hipHostMalloc(&h_mirror_, M * sizeof(MGNorm), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(h_mirror_, 0, M * sizeof(MGNorm));
    PIC_HIP(hipHostGetDevicePointer((void**)&d_mirror_, h_mirror_, 0));
    // the stages of a refinement step first: they decide which level-0 buffer
    // holds each member's correction at its end
    if (int rc = build_direct_tables()) return rc;
    if (int rc = mixed_ ? build_step<float>(h) : build_step<double>(h)) return rc;
    std::vector<MGFineTask> fine(M);
    any_mask_ = false;
    std::vector<MGTile> chunks, res;
    for (int k = 0; k < M; ++k) {
        MGSolver* s = m_[k];
        const int n = s->n_;
        const size_t nn = (size_t)n * n;
        MGFineTask& f = fine[k];
        f.phi = phi[k];
        f.phi_tmp = s->phi_tmp_;
        f.rhs = rhs[k];
        f.r32 = s->r32_;
        f.e32 = mixed_ ? s->ef_[final_e0_[k]][0] : nullptr;
        f.partial = s->partials_;
        f.phi_save = s->phi_save_;
        f.norm = s->d_norm_;
        f.mirror = d_mirror_ + k;
        f.n = n;
        f.gx = div_up(n, MG_RX);
        const int gy = div_up(n, MG_RY);
        f.nblocks = f.gx * gy;
        f.nchunks = (int)((nn + 255) / 256);
        f.inv_h2 = 1.0 / (h[k] * h[k]);
        f.mask = s->masked() ? s->mask_[0] : nullptr;
        f.phi_bc = s->phi_bc_;
        any_mask_ = any_mask_ || s->masked();
        for (int c = 0; c < f.nchunks; ++c) chunks.push_back(MGTile{k, c, 0, 0});
        for (int ty = 0; ty < gy; ++ty)
            for (int tx = 0; tx < f.gx; ++tx) res.push_back(MGTile{k, tx, ty, 0});
    }
    if (int rc = upload_vec((void**)&d_fine_, fine)) return rc;
    h_fine_ = fine;
    if (int rc = upload_vec((void**)&d_chunk_tiles_, chunks)) return rc;
    if (int rc = upload_vec((void**)&d_res_tiles_, res)) return rc;
    n_chunk_ = (int)chunks.size();
    n_res_ = (int)res.size();
    return 0;
}

struct MGBatchStageView {
    int kind, nu, ns, bs;
    bool restr, masked;
    const void* tasks;
    const MGTile* tiles;
    int n_items;
    size_t lds;
    int n_tasks, gx;
    const unsigned char* t0;
};

template <typename Task>
static StageRef<Task> make_ref(const void* t0, const void* tasks, const MGTile* tiles, int n_tasks, int gx) {
    StageRef<Task> r;
    std::memcpy(&r.t0, t0, sizeof(Task));
    const bool single = n_tasks == 1;
    r.tasks = single ? nullptr : (const Task*)tasks;
    r.tiles = tiles;
    r.gx = gx;
    return r;
}

template <typename T, int NU, bool M, int BS>
static void launch_tiled_bs(int kind, bool restr, dim3 g, size_t lds, hipStream_t s, const StageRef<MGLevelTask<T>>& r) {
    const dim3 b(BS);
    if (kind == 1) hipLaunchKernelGGL((mg_prolong_stage<T, NU, false, M, BS>), g, b, lds, s, r);
    else if (restr) hipLaunchKernelGGL((mg_smooth_stage<T, NU, true, M, BS>), g, b, lds, s, r);
    else hipLaunchKernelGGL((mg_smooth_stage<T, NU, false, M, BS>), g, b, lds, s, r);
}

template <typename T, int NU, bool M>
static void launch_tiled_nu(int kind, bool restr, int bs, dim3 g, size_t lds, hipStream_t s,
                            const StageRef<MGLevelTask<T>>& r) {
    if (bs == 1024) launch_tiled_bs<T, NU, M, 1024>(kind, restr, g, lds, s, r);
    else if (bs == 512) launch_tiled_bs<T, NU, M, 512>(kind, restr, g, lds, s, r);
    else launch_tiled_bs<T, NU, M, 256>(kind, restr, g, lds, s, r);
}

template <typename T, int NU>
static void launch_tiled_m(int kind, bool restr, bool masked, int bs, dim3 g, size_t lds, hipStream_t s,
                           const StageRef<MGLevelTask<T>>& r) {
    if (masked) launch_tiled_nu<T, NU, true>(kind, restr, bs, g, lds, s, r);
    else launch_tiled_nu<T, NU, false>(kind, restr, bs, g, lds, s, r);
}

// the finest up-leg stage fused with the next V-cycle's down-leg stage (mixed
// mode only, so FP32): NU = post + pre sweeps, 0..6
template <int NU, int BS>
static void launch_fused_bs(bool masked, dim3 g, size_t lds, hipStream_t s, const StageRef<MGLevelTask<float>>& r) {
    if (masked) hipLaunchKernelGGL((mg_prolong_stage<float, NU, true, true, BS>), g, dim3(BS), lds, s, r);
    else hipLaunchKernelGGL((mg_prolong_stage<float, NU, true, false, BS>), g, dim3(BS), lds, s, r);
}

template <int NU>
static void launch_fused_m(bool masked, int bs, dim3 g, size_t lds, hipStream_t s,
                           const StageRef<MGLevelTask<float>>& r) {
    if (bs == 512) launch_fused_bs<NU, 512>(masked, g, lds, s, r);
    else launch_fused_bs<NU, 256>(masked, g, lds, s, r);
}

static void launch_fused(int nu, bool masked, int bs, dim3 g, size_t lds, hipStream_t s,
                         const StageRef<MGLevelTask<float>>& r) {
    switch (nu) {
    case 0: launch_fused_m<0>(masked, bs, g, lds, s, r); break;
    case 1: launch_fused_m<1>(masked, bs, g, lds, s, r); break;
    case 2: launch_fused_m<2>(masked, bs, g, lds, s, r); break;
    case 3: launch_fused_m<3>(masked, bs, g, lds, s, r); break;
    case 4: launch_fused_m<4>(masked, bs, g, lds, s, r); break;
    case 5: launch_fused_m<5>(masked, bs, g, lds, s, r); break;
    default: launch_fused_m<6>(masked, bs, g, lds, s, r); break;
    }
}

template <typename T>
static void launch_tiled(int kind, int nu, bool restr, bool masked, int bs, dim3 g, size_t lds, hipStream_t s,
                         const StageRef<MGLevelTask<T>>& r) {
    switch (nu) {
    case 0: launch_tiled_m<T, 0>(kind, restr, masked, bs, g, lds, s, r); break;
    case 1: launch_tiled_m<T, 1>(kind, restr, masked, bs, g, lds, s, r); break;
    case 2: launch_tiled_m<T, 2>(kind, restr, masked, bs, g, lds, s, r); break;
    default: launch_tiled_m<T, 3>(kind, restr, masked, bs, g, lds, s, r); break;
    }
}

// the big direct solve's product of one phase, chains of NS steps
template <int NS>
static void launch_dst(int phase, dim3 g, hipStream_t s, const StageRef<MGDstTask>& r) {
    const dim3 b(256);
    switch (phase) {
    case 0: hipLaunchKernelGGL((mg_dst_stage<0, NS>), g, b, 0, s, r); break;
    case 1: hipLaunchKernelGGL((mg_dst_stage<1, NS>), g, b, 0, s, r); break;
    case 2: hipLaunchKernelGGL((mg_dst_stage<2, NS>), g, b, 0, s, r); break;
    default: hipLaunchKernelGGL((mg_dst_stage<3, NS>), g, b, 0, s, r); break;
    }
}

template <typename T>
static void launch_stage_t(const MGBatchStageView& st, hipStream_t s) {
    const dim3 g(st.n_items);
    if (st.kind == 0 || st.kind == 1) {
        const auto r = make_ref<MGLevelTask<T>>(st.t0, st.tasks, st.tiles, st.n_tasks, st.gx);
        launch_tiled<T>(st.kind, st.nu, st.restr, st.masked, st.bs, g, st.lds, s, r);
    } else if (st.kind == 4) {
        if constexpr (sizeof(T) == 4) {
            const auto r = make_ref<MGLevelTask<float>>(st.t0, st.tasks, st.tiles, st.n_tasks, st.gx);
            launch_fused(st.nu, st.masked, st.bs, g, st.lds, s, r);
        }
    } else if (st.kind == 5) {
        const auto r = make_ref<MGDstTask>(st.t0, st.tasks, st.tiles, st.n_tasks, st.gx);
        static_assert((MG_NDB - 1) / 32 == 8, "chain lengths instantiated up to P = 256");
        switch (st.ns) {
        case 2: launch_dst<2>(st.nu, g, s, r); break;
        case 4: launch_dst<4>(st.nu, g, s, r); break;
        case 6: launch_dst<6>(st.nu, g, s, r); break;
        default: launch_dst<8>(st.nu, g, s, r); break;
        }
    } else if (st.kind == 2) {
        const auto r = make_ref<MGCoarseTask<T>>(st.t0, st.tasks, nullptr, st.n_tasks, 0);
        if (st.masked) hipLaunchKernelGGL((mg_coarse_stage<T, true>), g, dim3(1024), st.lds, s, r);
        else hipLaunchKernelGGL((mg_coarse_stage<T, false>), g, dim3(1024), st.lds, s, r);
    } else {
        const auto r = make_ref<MGLevelTask<T>>(st.t0, st.tasks, st.tiles, st.n_tasks, 0);
        hipLaunchKernelGGL(mg_copy_stage<T>, g, dim3(256), 0, s, r);
    }
}

int MGBatch::launch_stage(const Stage& st) {
    const MGBatchStageView v{st.kind, st.nu, st.ns, st.bs, st.restr, st.masked, st.tasks, st.tiles, st.n_items, st.lds,
                             st.n_tasks, st.gx, st.t0};
    if (st.fp64) launch_stage_t<double>(v, stream_);
    else launch_stage_t<float>(v, stream_);
    ++launches_;
    return 0;
}

// what: 0 the Dirichlet values of electrodes into phi (a no-op without
// electrodes); 1 residual and its norm (which 1: before the first cycle, with
// the rhs norm and the counters of a new solve; 2: after a cycle, whose index
// has the given parity); 3 mixed mode: the solution back into phi after an odd
// number of cycles
int MGBatch::enqueue_fine(int what, int which, int parity) {
    const int M = (int)m_.size();
    const auto chunks = make_ref<MGFineTask>(&h_fine_[0], d_fine_, d_chunk_tiles_, M, 0);
    const auto cases = make_ref<MGFineTask>(&h_fine_[0], d_fine_, nullptr, M, 0);
    if (what == 0) {
        if (any_mask_) {
            hipLaunchKernelGGL(dirichlet_stage, dim3(n_chunk_), dim3(256), 0, stream_, chunks);
            ++launches_;
        }
    } else if (what == 1) {
        const auto res = make_ref<MGFineTask>(&h_fine_[0], d_fine_, d_res_tiles_, M, h_fine_[0].gx);
        const dim3 g(n_res_), b(256);
        if (!mixed_) hipLaunchKernelGGL(ir_residual_stage<0>, g, b, 0, stream_, res, which, 0);
        else if (which == 1) hipLaunchKernelGGL(ir_residual_stage<1>, g, b, 0, stream_, res, which, 0);
        else hipLaunchKernelGGL(ir_residual_stage<2>, g, b, 0, stream_, res, which, parity);
        hipLaunchKernelGGL(norm_final_stage, dim3(M), dim3(256), 0, stream_, cases, which, p_.tol,
                           max_outer_ * inner_, inner_);
        launches_ += 2;
    } else {
        hipLaunchKernelGGL(ir_finish_stage, dim3(n_chunk_), dim3(256), 0, stream_, chunks);
        ++launches_;
    }
    return 0;
}

// The first residual (which 1) of a one-case solve as step_residual_stage,
// which also converts the step's charge and ends the step, then its norm.
int MGBatch::enqueue_step_residual(const MGStepFuse& f) {
    const auto res = make_ref<MGFineTask>(&h_fine_[0], d_fine_, d_res_tiles_, 1, h_fine_[0].gx);
    const auto cases = make_ref<MGFineTask>(&h_fine_[0], d_fine_, nullptr, 1, 0);
    const dim3 g(n_res_ + 1), b(256);
    if (f.fp32) {
        const auto* c = (const CaseDev<float>*)f.cases;
        if (mixed_) hipLaunchKernelGGL((step_residual_stage<1, float>), g, b, 0, stream_, res, f.charge, c, f.states,
                                       f.status, f.rebin);
        else hipLaunchKernelGGL((step_residual_stage<0, float>), g, b, 0, stream_, res, f.charge, c, f.states,
                                f.status, f.rebin);
    } else {
        const auto* c = (const CaseDev<double>*)f.cases;
        if (mixed_) hipLaunchKernelGGL((step_residual_stage<1, double>), g, b, 0, stream_, res, f.charge, c, f.states,
                                       f.status, f.rebin);
        else hipLaunchKernelGGL((step_residual_stage<0, double>), g, b, 0, stream_, res, f.charge, c, f.states,
                                f.status, f.rebin);
    }
    hipLaunchKernelGGL(norm_final_stage, dim3(1), dim3(256), 0, stream_, cases, 1, p_.tol, max_outer_ * inner_,
                       inner_);
    launches_ += 2;
    PIC_HIP(hipGetLastError());
    return 0;
}

// Some member ended after an odd number of refinement steps, so its solution
// is in phi_tmp and the gated copy back is needed (read after a synchronization).
bool MGBatch::any_odd_steps() const {
    for (int k = 0; k < (int)m_.size(); ++k)
        if (h_mirror_[k].steps & 1) return true;
    return false;
}

// One refinement step (FP64 mode: one V-cycle): its inner_ V-cycles.
int MGBatch::launch_refinement() {
    for (const auto& st : cycle_)
        if (int rc = launch_stage(st)) return rc;
    return 0;
}

// The host enqueues cycles in batches (first batch: the previous solve's
// largest cycle count) and reads the norms of all members once per batch.  The
// device flags turn cycles beyond a member's convergence into no-ops, so each
// result equals testing after every cycle, alone or in a batch.  The executed
// cycles of a member are a prefix of the enqueued ones, so in mixed mode the
// buffer holding its solution follows from its own cycle count.
int MGBatch::solve() {
    PIC_REQUIRE(!m_.empty(), "multigrid batch: not built");
    launches_ = 0;
    const int M = (int)m_.size();
    enqueue_fine(0, 0, 0);
    enqueue_fine(1, 1, 0);
    int enqueued = 0;   // refinement steps
    int batch = std::max(1, (last_cycles_ + inner_ - 1) / inner_);
    for (;;) {
        batch = std::min(batch, max_outer_ - enqueued);
        for (int b = 0; b < batch; ++b) {
            if (int rc = launch_refinement()) return rc;
            enqueue_fine(1, 2, (enqueued + b) & 1);
        }
        enqueued += std::max(batch, 0);
        PIC_HIP(hipStreamSynchronize(stream_));
        bool all = true;
        for (int k = 0; k < M; ++k) all = all && h_mirror_[k].converged;
        if (all || enqueued >= max_outer_) break;
        batch = 1;
    }
    if (mixed_ && any_odd_steps()) enqueue_fine(3, 0, 0);
    PIC_HIP(hipGetLastError());
    int mx = 0;
    for (int k = 0; k < M; ++k) mx = std::max(mx, h_mirror_[k].cycles);
    last_cycles_ = std::max(1, mx);
    return 0;
}

int MGBatch::enqueue_fixed(int cycles, const MGStepFuse* fuse) {
    PIC_REQUIRE(!m_.empty(), "multigrid batch: not built");
    PIC_REQUIRE(!fuse || m_.size() == 1, "multigrid batch: a fused step residual needs a batch of one");
    launches_ = 0;
    // cycles: V-cycles, run as whole refinement steps
    const int K = std::max(1, std::min((cycles + inner_ - 1) / inner_, max_outer_));
    enqueue_fine(0, 0, 0);
    if (fuse) {
        if (int rc = enqueue_step_residual(*fuse)) return rc;
    } else {
        enqueue_fine(1, 1, 0);
    }
    for (int b = 0; b < K; ++b) {
        if (int rc = launch_refinement()) return rc;
        enqueue_fine(1, 2, b & 1);
    }
    if (mixed_) enqueue_fine(3, 0, 0);
    fixed_cycles_ = K;
    return 0;
}

// The mirror holds the norms after the fixed cycles.  A member that has not
// stopped continues with cycle index `enqueued`: its solution is in phi (the
// copy-back ran after an odd count) or it is the cycle's source by parity.
int MGBatch::complete_fixed() {
    const int M = (int)m_.size();
    int enqueued = fixed_cycles_;
    for (;;) {
        bool all = true;
        for (int k = 0; k < M; ++k) all = all && h_mirror_[k].converged;
        if (all || enqueued >= max_outer_) break;
        if (int rc = launch_refinement()) return rc;
        enqueue_fine(1, 2, enqueued & 1);
        ++enqueued;
        PIC_HIP(hipStreamSynchronize(stream_));
    }
    if (enqueued > fixed_cycles_ && mixed_ && any_odd_steps()) enqueue_fine(3, 0, 0);
    PIC_HIP(hipGetLastError());
    int mx = 0;
    for (int k = 0; k < M; ++k) mx = std::max(mx, h_mirror_[k].cycles);
    last_cycles_ = std::max(1, mx);
    return 0;
}

}  // namespace pic
