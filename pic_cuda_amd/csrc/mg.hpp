// mg.hpp -- geometric multigrid Poisson solver for gfx950 (host side).
//
// Solves lap_h(phi) = rhs on square vertex-centred grids with Dirichlet
// boundary values (the boundary nodes of phi are never changed), like
// solve_multigrid (Multigrid.hpp:3-11, Multigrid.cu:573-642).
//
// Modes:
//   PIC_MG_FP64   V-cycles in FP64 directly on phi (the reference's scheme).
//   PIC_MG_MIXED  FP64 iterative refinement: r = rhs - A phi in FP64 (stored
//                 FP32), one FP32 V-cycle for A e = r, phi += e in FP64.
//
// MGSolver owns the level pyramid of one grid.  MGBatch runs the solves of
// several MGSolvers together: every stage of the cycle (a level of the down
// leg, the coarse cycle, a level of the up leg, the refinement kernels) is one
// launch over a table of (case, tile) items, and every case stops on its own
// device flag.  A single solve is a batch of one.
#pragma once
#include "pic_common.hpp"
#include "pic_device.hpp"
#include "field.hpp"
#include <algorithm>
#include <vector>

namespace pic {

enum { PIC_MG_FP64 = 0, PIC_MG_MIXED = 1 };

struct MGParams {
    int levels;         // 0: coarsen to 3x3
    int pre, post;      // sweeps per level
    int coarse_sweeps;  // sweeps on the coarsest level
    double omega;       // SOR factor
    double tol;         // ||r|| / ||rhs||
    int max_cycles;
    int mode;
    int inner = 1;      // mixed mode: FP32 V-cycles per FP64 refinement step
    int coarse_direct = 0;   // mixed mode: exact FP64 solve of the coarse kernel's level (1: n <= 65;
                             // 2: the first level of at most 257 nodes, several workgroups)
    int per_level = 0;  // 1: pre_l/post_l give each level's sweeps (index 0 = finest)
    int pre_l[16] = {}, post_l[16] = {};
    int pre_at(int l) const { return per_level ? pre_l[l] : pre; }
    int post_at(int l) const { return per_level ? post_l[l] : post; }
    int inner_eff() const { return mode == PIC_MG_MIXED ? std::max(1, std::min(inner, 4)) : 1; }
};

struct MGStats {
    int cycles;
    double rel_residual;
    double rhs_norm;
};

// norms and the stop flag of one case, written by the reduction kernel
struct MGNorm {
    double res_sq;
    double rhs_sq;
    double rel;
    int converged;     // stop flag: tolerance met or max_cycles done
    int cycles;        // V-cycles executed in this solve
    int max_cycles;
    int hit_tol;       // tolerance met
    int steps;         // refinement steps executed (mixed mode: the phi buffer parity)
};

// one (case, level) entry of a tiled stage
template <typename T>
struct MGLevelTask {
    const T* e_in;     // null: zero initial guess
    const T* rhs;
    T* e_out;
    T* rhs_c;          // smooth stage: coarse rhs to restrict into (may be null)
    const T* coarse;   // prolongation stage: coarse correction
    int n;
    int nu;            // sweeps
    int H;             // halo
    int pad;
    T h2, inv_h2, omega;
    const MGNorm* gate;
    const uint8_t* mask;   // embedded electrodes of this level (null: none); nonzero = fixed node
};

// one case of the single-workgroup coarse cycle
template <typename T>
struct MGCoarseTask {
    const T* e_in;
    const T* rhs;
    T* e_out;
    int n_top, Lc, csweeps, pad;
    const uint8_t* mask;   // electrodes of the top level (coarser levels: injection)
    // FP64 sine-transform tables of the top level (S then W, 64 x 64 each,
    // zero padded): solve it exactly; null: the reference's V-cycle below it
    const double* dtab;
    // sweeps of the kernel's level l (0 = its top) in bits 2l..2l+1 (packed so
    // that no array is indexed at run time)
    unsigned pre_bits, post_bits;
    T h_top, omega;
    const MGNorm* gate;
};

// one case of the FP64 refinement / norm kernels
struct MGFineTask {
    double* phi;
    double* phi_tmp;   // mixed mode: phi alternates between phi and phi_tmp
    const double* rhs;
    float* r32;        // mixed mode: FP32 residual (the cycle's rhs)
    const float* e32;  // mixed mode: the cycle's correction
    double* partial;   // per-tile sums of r^2, then (first residual) of rhs^2
    double* phi_save;  // first residual pass: copy of phi (the warm start), or null
    MGNorm* norm;
    MGNorm* mirror;    // this case's entry of the batch's norms in pinned host memory
    int n;
    int gx;            // residual tiles per row
    int nblocks;       // residual partials of this case
    int nchunks;       // 256-node chunks of this case
    double inv_h2;
    const uint8_t* mask;   // electrodes of the fine level (null: none)
    const double* phi_bc;  // Dirichlet values of the mask's 1-nodes
};

struct MGTile {
    int task, tx, ty, pad;
};

// one case of the multi-workgroup exact FP64 solve of a level of 66..257
// nodes (mixed mode, coarse_direct 2): four P x P products, one launch each
struct MGDstTask {
    const float* rhs;     // the level's rhs, n x n
    float* e_out;         // the level's correction, n x n (boundary written zero)
    const double* S;      // P x P: S[a][b] = sin(pi (a+1)(b+1) / (m+1)), zero beyond m
    const double* W;      // P x P: (2/(m+1))^2 / (-4 (sig_a + sig_b)), zero beyond m
    double* T0;           // P x P FP64 intermediates
    double* T1;
    int n, m, P, gx;      // nodes, unknowns a side, padded size (multiple of 64), P / 16
    int nh, pad;          // 16-blocks of the halved side: ceil(((m + 1) / 2) / 16)
    double h2;
    const MGNorm* gate;
};

class MGBatch;

// A step's charge conversion and end of step, run by the first residual pass
// of the step's solve (a solve of one case): the conversion's arguments, the
// case's descriptor (CaseDev<float> when fp32, else CaseDev<double>) and
// state, and where the end of step writes the spill fill and the re-binning
// flag (either may be null).
struct MGStepFuse {
    ChargeArgs charge;
    const void* cases;
    CaseState* states;
    int* status;
    int* rebin;
    bool fp32;
};

class MGSolver {
public:
    MGSolver(int n, const MGParams& p, hipStream_t s);
    ~MGSolver();
    int init();
    int solve(double* phi, const double* rhs, double h, MGStats* st);
    // the same solve split for graph capture: prepare (allocations) before
    // capture, enqueue_fixed (no host synchronization) inside it, and
    // complete_fixed once the stream has run it
    int prepare(double* phi, const double* rhs, double h);
    // fuse (may be null): the first residual pass also converts the step's
    // charge into rhs and ends the step
    int enqueue_fixed(int cycles, const MGStepFuse* fuse = nullptr);
    int complete_fixed(MGStats* st);
    int last_cycles() const;
    // the device stop flag of the solve (set once the tolerance is met or the
    // cycle limit reached; a kernel enqueued after the solve can gate on it),
    // and its value after the cycles of enqueue_fixed, read once the stream has
    // run them and before complete_fixed
    const int* stop_flag() const;
    bool fixed_stopped() const;
    // embedded electrodes (host arrays of n*n): mask 0 solved, 1 Dirichlet at
    // phi_bc, 2 solid (held at its value); null mask removes them
    int set_mask(const uint8_t* mask, const double* phi_bc);
    bool masked() const { return !mask_.empty(); }
    const MGParams& params() const { return p_; }
    int n() const { return n_; }
    int levels() const { return L_; }
    int last_launches() const { return launches_; }
    void set_stream(hipStream_t s) { stream_ = s; }
    // n*n device buffer that receives phi as each solve starts (its warm
    // start, written by the first residual pass), or null
    void set_phi_save(double* dst);
    // sweeps on the coarsest level: a 4x4 coarsest level of an automatic ladder
    // (2x2 unknowns, Gauss-Seidel factor 1/4 per sweep) gets at least 32, which
    // solves it to rounding; a 3x3 level is solved exactly by one sweep
    // the level the cycle ends on: the first level of at most 257 nodes when
    // it is solved exactly by the multi-workgroup sine transform (mixed mode,
    // coarse_direct 2, no electrodes, more than 65 nodes), else the coarse
    // kernel's level
    bool big_direct() const {
        return p_.mode == PIC_MG_MIXED && p_.coarse_direct == 2 && mask_.empty() && lcd_ >= 0 && lcd_ < lc_;
    }
    int coarse_level() const { return big_direct() ? lcd_ : lc_; }
    int coarse_sweeps_eff() const {
        return (p_.levels <= 0 && !ns_.empty() && ns_.back() == 4) ? (p_.coarse_sweeps > 32 ? p_.coarse_sweeps : 32)
                                                                 : p_.coarse_sweeps;
    }

private:
    friend class MGBatch;
    int n_, L_, lc_;
    int lcd_ = -1;                 // first level of at most 257 nodes (mixed mode)
    int nc_ = 33;                  // largest level of the coarse kernel
    double* dst_t_ = nullptr;      // big direct solve: two P x P FP64 intermediates
    MGParams p_;
    hipStream_t stream_;
    std::vector<int> ns_;
    // FP32 pyramid (mixed) and FP64 pyramid (fp64 mode): e[l][0..1], rhs[l]
    std::vector<float*> ef_[2], rf_;
    std::vector<double*> ed_[2], rd_;
    float* r32_ = nullptr;
    double* phi_tmp_ = nullptr;
    std::vector<uint8_t*> mask_;     // per level (injection), empty: no electrodes
    double* phi_bc_ = nullptr;
    double* phi_save_ = nullptr;
    double* partials_ = nullptr;
    int n_partials_ = 0;
    MGNorm* d_norm_ = nullptr;
    int launches_ = 0;
    bool ok_ = false;
    MGBatch* solo_ = nullptr;      // the batch of one used by solve()
    double* solo_phi_ = nullptr;
    const double* solo_rhs_ = nullptr;
    double solo_h_ = 0.0;
};

class MGBatch {
public:
    explicit MGBatch(hipStream_t s) : stream_(s) {}
    ~MGBatch();
    // members: solvers with identical parameters; phi/rhs/h per member
    int build(const std::vector<MGSolver*>& members, const std::vector<double*>& phi,
              const std::vector<const double*>& rhs, const std::vector<double>& h);
    // solve every member (at least one V-cycle batch is enqueued per sync)
    int solve();
    // a solve of a fixed number of cycles with no host synchronization (it can
    // be captured in a graph), and the completion after the stream has run it:
    // members that have not stopped get further cycles until they do
    int enqueue_fixed(int cycles, const MGStepFuse* fuse = nullptr);
    int complete_fixed();
    int last_cycles() const { return last_cycles_; }
    int last_launches() const { return launches_; }
    const MGNorm& norm(int k) const { return h_mirror_[k]; }
    int size() const { return (int)m_.size(); }

private:
    struct Stage {
        int kind = 0;      // 0 smooth, 1 prolong, 2 coarse, 3 copy, 4 prolong + next cycle's smooth,
                           // 5 a product of the big direct solve (phase in nu)
        int nu = 0;        // sweeps of a tiled stage
        int ns = 0;        // kind 5: steps of each product chain (P / 32 of the largest case)
        int bs = 256;      // kinds 0 and 1: threads per tile
        bool restr = false;   // smooth stage also restricts the residual
        bool masked = false;  // some task has electrodes
        bool fp64 = false;
        void* tasks = nullptr;
        MGTile* tiles = nullptr;
        int n_items = 0;   // launch blocks
        size_t lds = 0;
        int n_tasks = 0;
        int gx = 0;        // single task: tiles per row
        alignas(16) unsigned char t0[128];   // single task: passed by value
    };
    template <typename T>
    int build_step(const std::vector<double>& h);
    int build_direct_tables();
    int build_big_tables();
    template <typename T>
    int add_level_stage(int kind, const std::vector<MGLevelTask<T>>& tasks, int nu, bool restr,
                        std::vector<Stage>* dst = nullptr);
    int add_dst_stages(const std::vector<MGDstTask>& tasks, std::vector<Stage>* dst);
    int launch_stage(const Stage& st);
    int launch_refinement();   // the inner V-cycles of one refinement step
    bool any_odd_steps() const;
    int enqueue_fine(int what, int which, int parity);
    int enqueue_step_residual(const MGStepFuse& f);
    void free_all();

    hipStream_t stream_;
    std::vector<MGSolver*> m_;
    std::vector<double*> phi_;
    std::vector<const double*> rhs_;
    MGParams p_{};
    bool mixed_ = false;
    std::vector<Stage> cycle_;          // stages of one refinement step (inner_ V-cycles)
    // per member: the level-0 buffer (ef_[x][0]) holding the correction at the
    // end of a refinement step
    std::vector<int> final_e0_;
    int inner_ = 1;                     // V-cycles per refinement step
    int max_outer_ = 1;                 // refinement steps allowed (max_cycles / inner, rounded up)
    MGFineTask* d_fine_ = nullptr;      // per-case refinement tasks
    std::vector<MGFineTask> h_fine_;
    MGTile* d_chunk_tiles_ = nullptr;   // (case, chunk of 256 nodes)
    MGTile* d_res_tiles_ = nullptr;     // (case, 64x16 residual tile)
    int n_chunk_ = 0, n_res_ = 0;
    MGNorm* h_mirror_ = nullptr;        // the norms of all members: pinned, coherent host memory
    MGNorm* d_mirror_ = nullptr;        // its device address
    int last_cycles_ = 1;
    bool any_mask_ = false;             // some member has electrodes
    // direct coarse solve: per member its tables (null: none)
    std::vector<const double*> dtab_;
    double* d_dtab_ = nullptr;
    // big direct solve: per member its S and W tables (null: none)
    std::vector<const double*> btab_;
    double* d_btab_ = nullptr;
    int fixed_cycles_ = 0;
    int launches_ = 0;
};

}  // namespace pic
