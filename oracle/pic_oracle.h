/*
 * pic_oracle.h -- serial CPU restatement of the 2D electrostatic PIC time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * pic_cuda_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every routine follows a function of the reference (dimad285/PIC_CUDA) and
 * cites it.  Where the reference is nondeterministic or defective the oracle
 * applies the documented replacement:
 *   - cuRAND XORWOW seeded with time(NULL) (mcc.cu:129-145)  -> Philox4x32-10
 *     keyed on (particle id, step);
 *   - FP64 atomicAdd charge deposit (kernels.cu:65-68)        -> Q20 fixed-point
 *     CIC weights summed in uint64 per species;
 *   - head/tail stamp compaction (particles.cu:217-261)      -> stable stream
 *     compaction (set semantics; survivor order = input order);
 *   - elastic MCC with P = 1-exp(-n sigma v dt) (mcc.cu:56-85) -> null-collision
 *     MCC over a process list;
 *   - convergence test ||r||/sum(rhs^2) (Multigrid.cu:466, defect) -> ||r||/||rhs||.
 *
 * Units inside the particle phase: positions in cell units (x/dx, y/dy, so the
 * domain is [0, nx-1] x [0, ny-1]); velocities in m/s; E in V/m; phi in V.
 *
 * Parity pins (checked in tests/): Philox against the Random123 known-answer
 * vectors and rocRAND's philox4x32_10 header; det_log against libm; E-field
 * against closed forms; deposit against charge conservation and hand-worked
 * weights; multigrid against the discrete sine eigenfunction and a dense direct
 * solve; wall test against brute-force segment geometry; MCC against binomial
 * statistics and energy bookkeeping.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_QBITS 20                 /* fixed-point fraction bits of CIC weights */
#define ORC_MAX_PROC 4

enum { ORC_ELASTIC = 0, ORC_EXCITATION = 1, ORC_IONIZATION = 2, ORC_CHARGE_EXCHANGE = 3 };

/* ---- random numbers -------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]);
double orc_det_log(double x);

/* ---- fields ------------------------------------------------------------ */
/* E = -grad(phi), reference kernels.cu:134-191 (one-sided at edges/corners). */
void orc_efield(const double* phi, int nx, int ny, double dx, double dy,
                double* ex, double* ey);

/* Q20 CIC deposit of n particles (cell units) into acc[ny*nx], kernels.cu:26-69. */
void orc_deposit(int prec, const void* x, const void* y, int64_t n, int nx, int ny,
                 uint64_t* acc);

/* rhs[i] = sum_s cfac[s] * (double)acc[s][i]   (cfac already holds -q w / (dx dy eps0 2^2Q)) */
void orc_charge_to_rhs(const uint64_t* const* acc, const double* cfac, int nspec,
                       int64_t nn, double* rhs);

/* ---- multigrid, reference Multigrid.cu:192-407, 539-642 ------------------ */
typedef struct {
    int levels;          /* 0: coarsen until n <= 3 */
    int pre, post;       /* sweeps per level (reference: 3, 3) */
    int coarse_sweeps;   /* sweeps on the coarsest level (reference: 20) */
    double omega;        /* SOR factor (reference: 1.25) */
    double tol;          /* relative residual ||r||/||rhs|| */
    int max_cycles;      /* reference: 50 */
    int per_level;       /* 1: pre_level/post_level give the sweeps of each level
                            above the coarsest (index 0 = finest) */
    int pre_level[16], post_level[16];
} orc_mg_params;

typedef struct {
    int cycles;
    double rel_residual;
    double rhs_norm;
} orc_mg_stats;

int orc_mg_solve(double* phi, const double* rhs, int nx, int ny, double h,
                 const orc_mg_params* p, orc_mg_stats* st);
/* with embedded electrodes: mask[nx*ny] 0 solved, 1 Dirichlet at phi_bc, 2 solid */
int orc_mg_solve_masked(double* phi, const double* rhs, int nx, int ny, double h, const orc_mg_params* p,
                        const uint8_t* mask, const double* phi_bc, orc_mg_stats* st);
/* one red-black SOR smoothing call on a level (Multigrid.cu:192-227, :476-506) */
void orc_mg_smooth(double* phi, const double* rhs, int n, double h, double omega, int iters);
void orc_mg_residual(double* res, const double* phi, const double* rhs, int n, double h);
void orc_mg_restrict(double* coarse, const double* fine, int nf);
void orc_mg_prolong(double* fine, const double* coarse, int nc);

/* ---- particle push ----------------------------------------------------- */
typedef struct {
    int prec;                    /* 4: float particles, 8: double particles */
    int nx, ny;                  /* grid nodes */
    double dx, dy;               /* physical spacing (E stencil) */
    double kq;                   /* q dt / m */
    double dtdx, dtdy;           /* dt/dx, dt/dy (m/s -> cells per step) */
    int do_kick, do_mcc, do_drift;
    int box_walls;               /* absorb when leaving [0,nx-1]x[0,ny-1] */
    int n_walls;                 /* extra wall segments, cell units x1,y1,x2,y2 */
    const double* walls;
    double wall_eps;             /* parallel cut of the segment test */
    uint32_t seed, step, case_id, species;
    /* null-collision MCC */
    int n_proc;
    int proc_type[ORC_MAX_PROC];
    double proc_thr[ORC_MAX_PROC];   /* eV */
    const double* tab;           /* n_proc * K entries: n_gas * sigma_k(E_j), E_j = j*dE */
    int K;
    double inv_de;
    double nu_max;
    double p_null;
    double efac;                 /* E[eV] = efac * v^2 */
    double gas_vth;              /* sqrt(kT_gas / M) per component, for CX and born ions */
    /* products of ionization */
    uint32_t ion_species;
    double ion_efac_unused;
    /* energy-banded majorant: band 0 below emax*2^-14, b = 15 + floor(log2(E/emax))
       above, 15 from emax/2 on; a candidate has u0 < band_p_null[band] */
    double inv_emax;
    double band_nu_max[16];
    double band_p_null[16];
} orc_push_params;

enum {
    ORC_C_ABSORBED = 0, ORC_C_NULL_CAND = 1, ORC_C_PROC0 = 2, /* 2..5 per process */
    ORC_C_BORN_SELF = 6, ORC_C_BORN_OTHER = 7, ORC_C_BORN_ABSORBED = 8,
    ORC_C_NUMAX_EXCEEDED = 9, ORC_C_REJECT_FALLBACK = 10, ORC_NCOUNT = 16
};

/*
 * One fused particle step for a flat species array (kick, MCC, drift, walls,
 * deposit, stable compaction).  Arrays x,y,vx,vy are float or double per prec,
 * id is uint32.  Capacity cap >= n + born electrons.  Survivors are compacted in
 * place (input order); electrons born by ionization are appended after them.
 * Ions born by ionization are written to the *_o arrays (capacity cap_o).
 * Returns the new count; *n_other receives the born-ion count.
 */
int64_t orc_push(const orc_push_params* p, const double* phi,
                 void* x, void* y, void* vx, void* vy, uint32_t* id, int64_t n, int64_t cap,
                 uint64_t* acc_self,
                 void* xo, void* yo, void* vxo, void* vyo, uint32_t* ido, int64_t cap_o,
                 int64_t* n_other, uint64_t* acc_other, uint64_t* counters);

/* Bourke segment-segment test, collisions.cu:7-54 (inclusive endpoints). */
int orc_segment_hits(double x1, double y1, double x2, double y2,
                     const double* walls, int n_walls, double eps);

/* Uniform positions in the open box and Maxwellian velocities (polar method),
   Philox counter {i, 0xFFFFFFFF, c, case_id}, key {seed, species}. */
void orc_init_particles(int prec, int64_t n, int nx, int ny, double vth,
                        uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0,
                        void* x, void* y, void* vx, void* vy, uint32_t* id);

/* Uniform positions in [x0,x1)x[y0,y1) (cell units) and uniform velocities in
   [v0,v1) per component (initializeParticles, main.cu:17-32, made
   reproducible): Philox counter {i, 0xFFFFFFFE, 0, case_id}, key {seed, species}. */
void orc_init_uniform(int prec, int64_t n, double x0, double x1, double y0, double y1, double v0, double v1,
                      uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0,
                      void* x, void* y, void* vx, void* vy, uint32_t* id);

/* Wall-clock helper for the CPU baseline (OpenMP build only uses >1 thread). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
