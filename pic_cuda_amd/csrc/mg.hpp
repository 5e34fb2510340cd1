// mg.hpp -- geometric multigrid Poisson solver for gfx950 (host side).
//
// Solves lap_h(phi) = rhs on a square vertex-centred grid with Dirichlet
// boundary values (the boundary nodes of phi are never changed), like
// solve_multigrid (Multigrid.hpp:3-11, Multigrid.cu:573-642).
//
// Modes:
//   PIC_MG_FP64   V-cycles in FP64 directly on phi (the reference's scheme).
//   PIC_MG_MIXED  FP64 iterative refinement: r = rhs - A phi in FP64 (stored
//                 FP32), one FP32 V-cycle for A e = r, phi += e in FP64.
#pragma once
#include "pic_common.hpp"
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
};

struct MGStats {
    int cycles;
    double rel_residual;
    double rhs_norm;
};

// norms and the convergence flag, written by the reduction kernel
struct MGNorm {
    double res_sq;
    double rhs_sq;
    double rel;
    int converged;
    int pad;
};

class MGSolver {
public:
    MGSolver(int n, const MGParams& p, hipStream_t s);
    ~MGSolver();
    int init();                                              // allocate (returns status)
    int solve(double* phi, const double* rhs, double h, MGStats* st);
    const MGParams& params() const { return p_; }
    int n() const { return n_; }
    int levels() const { return L_; }
    void set_stream(hipStream_t s) { stream_ = s; }
    // kernel launch counts of the last solve (for the measurement report)
    int last_launches() const { return launches_; }

private:
    template <typename T>
    int vcycle(const T* e0_in, T* e0_out, T* e0_tmp, const T* rhs0, double h);
    int reduce(const double* v_partials, int np, int which);

    int n_, L_, lc_;
    MGParams p_;
    hipStream_t stream_;
    std::vector<int> ns_;
    // FP32 pyramid (mixed) and FP64 pyramid (fp64 mode): e[l][0..1], rhs[l]
    std::vector<float*> ef_[2], rf_;
    std::vector<double*> ed_[2], rd_;
    float* r32_ = nullptr;
    double* phi_tmp_ = nullptr;
    double* partials_ = nullptr;
    int n_partials_ = 0;
    MGNorm* d_norm_ = nullptr;
    MGNorm* h_norm_ = nullptr;   // pinned
    int launches_ = 0;
    bool ok_ = false;
};

}  // namespace pic
