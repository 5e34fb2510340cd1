/*
 * pic_amd.h -- C ABI of the MI355X (gfx950) 2D electrostatic PIC time step.
 *
 * Drop-in boundary for the step of dimad285/PIC_CUDA.  Every entry point names
 * the reference interface it replaces.  Conventions:
 *   - every function returns an int status: 0 on success, negative on error
 *     (PIC_ERR_*); pic_last_error() gives the message of the last failure on
 *     the calling thread.  Nothing exits the process (the reference calls exit()
 *     on allocation failure, particles.cu:135-136).
 *   - particle positions cross the boundary in CELL units (x/dx, y/dy: the
 *     domain is [0, m-1] x [0, n-1]), velocities in m/s.  The C++ and Python
 *     mirrors of Particles2D convert from domain units.
 *   - grid fields are FP64, row-major, m (x, fast) by n (y) nodes.
 *   - the library's device work runs in order on one HIP stream of its own
 *     (default priority); device pointers handed in are used on that stream.
 *     Calls are not thread-safe: call the library from one host thread.
 */
#ifndef PIC_AMD_H
#define PIC_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_OK 0
#define PIC_ERR_ARG -1
#define PIC_ERR_HIP -2
#define PIC_ERR_STATE -3
#define PIC_ERR_CAPACITY -4
#define PIC_ERR_IO -5

#define PIC_PREC_FP32 4
#define PIC_PREC_FP64 8

/* collision process types (mcc.hpp:6-16 lists the channels; mcc.cu implements elastic) */
#define PIC_PROC_ELASTIC 0
#define PIC_PROC_EXCITATION 1
#define PIC_PROC_IONIZATION 2
#define PIC_PROC_CHARGE_EXCHANGE 3

/* per-species event counters, index into the 16-entry array of pic_sim_counters */
#define PIC_CNT_ABSORBED 0
#define PIC_CNT_NULL_CANDIDATES 1
#define PIC_CNT_PROCESS0 2          /* 2..5: real collisions of process 0..3 */
#define PIC_CNT_BORN_SELF 6
#define PIC_CNT_BORN_OTHER 7
#define PIC_CNT_BORN_ABSORBED 8
#define PIC_CNT_NUMAX_EXCEEDED 9
#define PIC_CNT_REJECT_FALLBACK 10
#define PIC_CNT_SPILLED 11
#define PIC_CNT_LOST 12
#define PIC_CNT_GLOBAL_DEPOSIT 13

#define PIC_FIELD_RHO 0
#define PIC_FIELD_PHI 1
#define PIC_FIELD_EX 2
#define PIC_FIELD_EY 3
#define PIC_FIELD_RHS 4

#define PIC_MG_FP64 0
#define PIC_MG_MIXED 1
#define PIC_MG_PRESET_REFERENCE 0   /* Multigrid.cu defaults: SOR 1.25, 3/3/20, 1e-5, 50 cycles, FP64 */
#define PIC_MG_PRESET_STEP 1        /* 2+2 Gauss-Seidel to 1e-8, FP64 refinement, exact coarse solve <= 257^2 */
#define PIC_MG_PRESET_POISSON 2     /* 2+2 red-black SOR 1.2 to 1e-10, FP64 refinement, exact coarse solve <= 257^2 */
#define PIC_MG_PRESET_STEP_SOR 3    /* 2+2 red-black SOR 1.2 to 1e-8, FP64 refinement: the simulation
                                       default; over-relaxation cuts the V-cycle contraction from
                                       about 0.06 (Gauss-Seidel) to about 0.01 per cycle */

typedef struct pic_particles pic_particles;
typedef struct pic_grid pic_grid;
typedef struct pic_walls pic_walls;
typedef struct pic_mg pic_mg;
typedef struct pic_xs pic_xs;
typedef struct pic_sim pic_sim;

/* ---- library ---------------------------------------------------------- */
const char* pic_last_error(void);
int pic_version(int* major, int* minor);
int pic_device_synchronize(void);
/* device memory for callers without an allocator of their own (the float
   pair stream of pic_particles_export_positions, say); copies to the host
   wait for the library's work first */
int pic_device_alloc(int64_t bytes, void** out);
int pic_device_free(void* ptr);
int pic_device_to_host(void* host, const void* dev, int64_t bytes);
/* host -> device copy (after the library's work), for filling the staging
   arrays of pic_particles_staging without a HIP runtime of one's own */
int pic_host_to_device(void* dev, const void* host, int64_t bytes);

/* ---- Particles2D: particles.cuh:3-39, particles.cu:112-261 --------------- */
/* Particles2D(int num_particles) particles.cu:112-150.  Species 0 exists with
   m = 1, q = 0.1 like the reference constructor (particles.cu:146-149).
   max_particles is the capacity PER SPECIES.  precision: PIC_PREC_FP32/64.
   A species holds about 1.6 max_particles + 2^20 slots in each of its two
   buffers, at most 2^30 (FP32) or 2^29 (FP64): the particle kernels address
   each array with 32-bit byte offsets. */
int pic_particles_create(int64_t max_particles, int precision, pic_particles** out);
/* ~Particles2D() particles.cu:153-171 */
int pic_particles_destroy(pic_particles* p);
/* add_species(double m, double q) particles.cu:174-190 (argument order m, q) */
int pic_particles_add_species(pic_particles* p, double m, double q, int* index);
/* change_species(int s, double m, double q) particles.cu:193-197 */
int pic_particles_change_species(pic_particles* p, int species, double m, double q);
/* macro-particle weight (real particles per macro-particle per metre of depth) */
int pic_particles_set_weight(pic_particles* p, int species, double weight);
int pic_particles_num_species(const pic_particles* p, int* n);
/* public member last_alive (particles.cuh:20); species -1 sums all species */
int pic_particles_last_alive(pic_particles* p, int species, int64_t* n);
/* host -> device append (replaces the cudaMemcpy of main.cu:232-238); id may be NULL */
int pic_particles_upload(pic_particles* p, int species, int64_t n, const double* x, const double* y,
                         const double* vx, const double* vy, const uint32_t* id);
/* device -> host copy of a species (order unspecified; sort by id to compare) */
int pic_particles_download(pic_particles* p, int species, int64_t cap, double* x, double* y, double* vx,
                           double* vy, uint32_t* id, int64_t* n);
/* The public device arrays of Particles2D (particles.cuh:7-20: r, r_old, v,
   type) for callers that fill them with cudaMemcpy, as main.cu:232-238 does:
   a staging area of n_max entries in the reference's layout -- r, r_old, v as
   (x, y) double pairs in domain units, type as the species index of each
   entry (zeroed: species 0).  r_old is allocated only when asked for and is
   never read (the step does not use old positions).  A larger n_max
   reallocates (the pointers change); any output pointer may be NULL. */
int pic_particles_staging(pic_particles* p, int64_t n_max, double** r, double** r_old, double** v,
                          int32_t** type);
/* particles.last_alive = n (main.cu:238): the first n staged entries become
   particles of their species, with ids continuing from the last staged batch.
   Without a simulation the hand-over waits for pic_sim_create, whose grid gives
   the cell size; in a running simulation it happens at once and the next step
   re-bins.  The type array is zeroed again.  PIC_ERR_ARG for a type that is
   not a species, PIC_ERR_CAPACITY when a species has no room. */
int pic_particles_set_staged(pic_particles* p, int64_t n);
/* the same hand-over now, with the cell size given (x / dx, y / dy) */
int pic_particles_commit_staging(pic_particles* p, int64_t n, double dx, double dy);
/* convertPositionsToFloat (kernels.cu:268-273) as used by updateVBO
   (render.h:384-405): the positions of every live particle of `species` as
   float (x, y) pairs packed into the device buffer dst_device (room for cap
   pairs; NULL to count), in cell units (units 0) or domain units (1, needs the
   particles attached to a simulation); the order is unspecified. *n = live
   particles. Reads the store only: a simulation being shown is not synced or
   changed. Ordered on the library stream and finished before return. */
int pic_particles_export_positions(pic_particles* p, int species, float* dst_device, int64_t cap, int units,
                                   int64_t* n);
/* pic_particles_export_positions without any allocation: no counterpart in
   the reference; for callers that cannot allocate device memory (code under
   graph capture, a render or serving loop that forbids hipMalloc).  Same
   arguments, results and ordering guarantees as pic_particles_export_positions;
   its work-item table and counter live in a static device array of
   256 B + 65536 work items (1 MiB per GPU, held until the process ends), so
   a species with more than 65536 bins + spill chunks of 4096 slots is
   rejected with PIC_ERR_CAPACITY.  Every call on a GPU shares that array:
   calls that overlap on one GPU (from several host threads) race. */
int pic_particles_export_positions_static(pic_particles* p, int species, float* dst_device, int64_t cap, int units,
                                          int64_t* n);
/* seeded initial state on the device (replaces initializeParticles, main.cu:17-32):
   uniform in the open box [0,nx-1]x[0,ny-1], Maxwellian sigma = vth per component */
int pic_particles_init_maxwellian(pic_particles* p, int species, int64_t n, int nx, int ny, double vth,
                                  uint32_t seed, uint32_t case_id, uint32_t id0);
/* initializeParticles (main.cu:17-32) with its distributions, made
   reproducible: positions uniform in [x_lo,x_hi)x[y_lo,y_hi) (cell units),
   each velocity component uniform in [v_lo,v_hi) (m/s); keyed on (seed,
   species, id, case) like init_maxwellian, on a stream of its own */
int pic_particles_init_uniform(pic_particles* p, int species, int64_t n, double x_lo, double x_hi, double y_lo,
                               double y_hi, double v_lo, double v_hi, uint32_t seed, uint32_t case_id, uint32_t id0);
/* removeCollidedParticles() particles.cu:217-261: compaction is fused into
   pic_sim_step; kept for the interface and returns immediately */
int pic_particles_remove_collided(pic_particles* p);

/* ---- Grid2D: grid.cuh:7-24, grid.cu:10-53 ------------------------------- */
/* Grid2D(int M, int N, double X, double Y): dx = X/(M-1).  Fields are zeroed
   (the reference leaves them uninitialised). */
int pic_grid_create(int m, int n, double X, double Y, pic_grid** out);
int pic_grid_destroy(pic_grid* g);
int pic_grid_info(const pic_grid* g, int* m, int* n, double* dx, double* dy);
/* public members rho, phi, E (device pointers; E is interleaved Ex,Ey) */
int pic_grid_device_ptrs(pic_grid* g, double** rho, double** phi, double** E);
int pic_grid_download(pic_grid* g, int field, double* out);
int pic_grid_upload(pic_grid* g, int field, const double* in);
/* computeElectricField kernels.cu:134-191 into the E member */
int pic_grid_compute_E(pic_grid* g);

/* ---- Boundaries2D: boundaries.h:7-23, boundaries.cu:3-25 ----------------- */
int pic_walls_create(double width, double height, int nx, int ny, pic_walls** out);
/* add_wall(x1,y1,x2,y2) in domain units */
int pic_walls_add(pic_walls* w, double x1, double y1, double x2, double y2);
/* the four domain edges as absorbing walls (evaluated by a box test) */
int pic_walls_add_box(pic_walls* w);
/* compile_walls() */
int pic_walls_compile(pic_walls* w);
int pic_walls_count(const pic_walls* w, int* n);
int pic_walls_destroy(pic_walls* w);

/* ---- multigrid: Multigrid.hpp:3-54, Multigrid.cu:30-642 ------------------ */
typedef struct {
    int levels;          /* 0: coarsen to a 3x3 grid */
    int pre, post;       /* sweeps per level, 0..3 */
    int coarse_sweeps;   /* sweeps on the coarsest level */
    double omega;        /* SOR factor */
    double tol;          /* relative residual ||r||/||rhs|| */
    int max_cycles;      /* V-cycles (mixed mode: rounded up to whole refinement steps) */
    int mode;            /* PIC_MG_FP64 or PIC_MG_MIXED */
    int inner;           /* mixed mode: FP32 V-cycles per FP64 refinement step, 1..4 (0: 1);
                            the convergence test runs after each refinement step */
    int coarse_direct;   /* mixed mode, no electrodes: 1 solves the first level with at most
                            65 x 65 nodes exactly (FP64 sine transform on the matrix cores of
                            one workgroup) instead of recursing and sweeping below it; 2 does
                            the same on the first level with at most 257 x 257 nodes, as four
                            FP64 matrix-core products over the whole GPU (grids whose levels
                            skip 66..257 nodes fall back to 1); 0 keeps the sweeps.
                            FP64 mode ignores it (its cycle is the reference's,
                            Multigrid.cu:539-569) */
} pic_mg_params;

typedef struct {
    int cycles;
    double rel_residual;
    double rhs_norm;
    double ms;           /* wall time of the solve */
    int launches;        /* kernel launches of the solve */
} pic_mg_stats;

int pic_mg_default_params(int preset, pic_mg_params* out);
/* explicit solver context (replaces the process-global pool, Multigrid.cu:30-187) */
int pic_mg_create(int n, const pic_mg_params* params, pic_mg** out);
/* solve lap(phi) = rhs; phi_d in/out (warm start, boundary kept), device pointers */
int pic_mg_solve(pic_mg* mg, double* phi_d, const double* rhs_d, double h, pic_mg_stats* st);
int pic_mg_destroy(pic_mg* mg);
/* solve_multigrid(phi, rho, nx, ny, levels, h, omega) Multigrid.hpp:3-11 with the
   reference's other settings (3/3/20 sweeps, tol 1e-5, 50 cycles, FP64) */
int pic_solve_multigrid(double* phi_d, double* rho_d, int nx, int ny, int levels, double h, double omega);

/* ---- cross sections and MCC: mcc.hpp:6-21, mcc.cu:17-150 ----------------- */
/* read_cross_section(filename) mcc.cu:17-41: "E sigma" pairs, '#' comments.
   energies/sigmas may be NULL to query the count. */
int pic_read_cross_section(const char* path, int cap, double* energies, double* sigmas, int* n);
/* background gas: density (m^-3) and temperature (eV) */
int pic_xs_create(double gas_density, double gas_temperature_eV, pic_xs** out);
/* tabulated process (energies eV ascending, sigma m^2); product_species receives
   the ion of an ionization (ignored otherwise) */
int pic_xs_add_process(pic_xs* xs, int species, int type, double threshold_eV, int n, const double* energy_eV,
                       const double* sigma_m2, int product_species);
/* analytic process: constant sigma (elastic, charge exchange) or the Lotz-like
   shape sigma = peak * e * ln(E/thr)/(E/thr) above thr (excitation, ionization) */
int pic_xs_add_analytic(pic_xs* xs, int species, int type, double threshold_eV, double peak_m2,
                        int product_species);
/* uniform energy table the processes are sampled on (default 2048 points, 200 eV) */
int pic_xs_set_table(pic_xs* xs, int points, double emax_eV);
int pic_xs_destroy(pic_xs* xs);
/* ---- cross-section files beyond mcc.cu:17-41 (no counterpart in the
   reference, which reads one two-column table per process) ---------------- */
/* as pic_read_cross_section, but numbers may also be written m*10^{e} or
   m*10^e (csf/Ar_conv.txt, whose rows read_cross_section drops); both
   spellings of a value give the same double */
int pic_read_table(const char* path, int cap, double* energies, double* sigmas, int* n);
/* every block of an LXCat file as a tabulated process of `species`:
   ELASTIC / EFFECTIVE / ISOTROPIC -> elastic, BACKSCAT -> charge exchange,
   EXCITATION / IONIZATION -> inelastic with the block's threshold (ionization
   products go to product_species); ATTACHMENT blocks are skipped and counted.
   All or nothing: a malformed block, or more than 4 processes for the
   species, adds none. */
int pic_xs_load_lxcat(pic_xs* xs, const char* path, int species, int product_species, int* n_added,
                      int* n_skipped);
typedef struct {
    int species, type, product_species;
    int n_points;          /* table rows; 0 for an analytic process */
    double threshold_eV;
    double peak_m2;        /* analytic processes */
} pic_xs_process;
int pic_xs_process_count(const pic_xs* xs, int* n);
/* process k; energies/sigmas (capacity cap) may be NULL */
int pic_xs_get_process(const pic_xs* xs, int k, pic_xs_process* out, int cap, double* energies, double* sigmas);
/* elasticCollisions(particles, density, dt, xs) mcc.cu:114-150: one null-collision
   elastic pass over a species with a tabulated sigma, keyed on (id, step) */
int pic_elastic_collisions(pic_particles* p, int species, double density, double dt, int n,
                           const double* energy_eV, const double* sigma_m2, uint32_t seed, uint32_t step);

/* ---- step driver: updateSimulationCUDA main.cu:37-140 -------------------- */
typedef struct {
    double dt;           /* s */
    double eps0;         /* F/m */
    uint32_t seed;
    uint32_t case_id;
    int bin_w, bin_h;    /* particle bin tile in cells */
    int halo;            /* deposit window halo in nodes */
    int mcc;             /* run MCC when a pic_xs is attached */
    int rebin_every;     /* full re-binning period in steps (0: only on spill) */
    double wall_eps;     /* parallel cut of the segment wall test (cell units^2) */
    pic_mg_params mg;
} pic_sim_params;

typedef struct {
    int64_t steps;
    int64_t alive;
    int last_mg_cycles;
    double last_mg_rel;
    int64_t total_mg_cycles;
    int rebins;
    double ms_push, ms_field, ms_total;
    int64_t graph_steps;     /* steps run as a graph launch */
    int64_t graph_captures;  /* graph (re)captures */
    double ms_graph;         /* device time of the graph steps */
    int64_t ahead_steps;     /* eager steps whose push ran ahead, behind the previous step's solve */
    int64_t ahead_held;      /* pushes enqueued ahead and held back by their gate (solve not stopped
                                after the predicted cycles, or re-binning due) and enqueued again */
} pic_sim_report;

/* null-collision constants of a species (inputs of the oracle's MCC).  The
   majorant is taken per energy band: band 0 holds E < emax*2^-14, band b in
   1..14 holds emax*2^(b-15) <= E < emax*2^(b-14), band 15 holds E >= emax/2;
   a particle is a collision candidate when its first uniform is below the
   p_null of its band. */
typedef struct {
    int n_proc;
    int K;
    int type[4];
    double threshold[4];
    double inv_de;
    double nu_max;       /* largest band majorant */
    double p_null;       /* 1 - exp(-nu_max dt) */
    double efac;
    double gas_vth;
    int product_species;
    double emax;                 /* table end, eV */
    double band_nu_max[16];      /* majorant of each band (x1.001) */
    double band_p_null[16];      /* 1 - exp(-band_nu_max dt) */
} pic_mcc_info;

int pic_sim_default_params(pic_sim_params* out);
int pic_sim_create(pic_particles* p, pic_grid* g, pic_walls* w, pic_xs* xs, const pic_sim_params* params,
                   pic_sim** out);
/* updateSimulationCUDA(particles, grid, boundaries, step, dt, steps, timing) */
int pic_sim_step(pic_sim* s, int n_steps);
/* complete the owed half of the last step (kick and collisions) */
int pic_sim_sync(pic_sim* s);
int pic_sim_counters(pic_sim* s, int species, uint64_t out[16]);
int pic_sim_stats(pic_sim* s, pic_sim_report* out);

/* Energy and momentum diagnostics of the synchronized state (the owed half
   step is completed first).  No counterpart in the reference, whose
   diagnostics.h holds stage timings only; energies are per unit length in z. */
typedef struct {
    int n_species;
    double count[4];      /* live particles of each species */
    double kinetic[4];    /* sum of w m (vx^2 + vy^2) / 2, J/m */
    double px[4], py[4];  /* sum of w m v, kg/s */
    double field_energy;  /* eps0/2 sum over the nodes of |E|^2 dx dy, J/m (E as computeElectricField) */
    double phi_l2;        /* sqrt of the sum over the nodes of phi^2 */
} pic_diagnostics;
int pic_sim_diagnostics(pic_sim* s, pic_diagnostics* out);
int pic_sim_mcc_info(pic_sim* s, int species, pic_mcc_info* out, double* table, int cap);
/* the same constants and table (n_proc x K rows of n_gas sigma on the uniform
   energy grid) that pic_sim_create builds for `species` of a cross-section
   set, from the species masses (kg) and dt; host only, no device needed */
int pic_xs_build(const pic_xs* xs, int species, int n_species, const double* masses, double dt, pic_mcc_info* out,
                 int cap, double* table);
/* re-bin now with a new tile (0 keeps the current one) */
int pic_sim_rebin(pic_sim* s, int bin_w, int bin_h);
/* graph mode (no counterpart in the reference, which launches every kernel
   from the host): each step after the first is one hipGraph launch (push,
   charge, a multigrid solve of the last step's cycle count) followed by one
   host synchronization; a solve that needs more cycles gets them eagerly.
   Results equal the eager path's.  Off by default. */
int pic_sim_use_graph(pic_sim* s, int on);
/* per-bin candidate walls (detect_intersections, collisions.cu:7-54, tests
   every wall for every particle): each bin lists the walls within margin + 1
   cells of it; a path that stays within `margin` cells (0..64, default 2) of
   its bin tests that list, a longer one the lists of the bins its bounding
   box meets (up to 16), a path beyond that every wall; same results in every
   case.  On by default; rebuilt whenever the bin tile changes.
   wall_list_stats reports the bins, the list entries and the longest list
   (all 0 without internal walls or with the lists off). */
int pic_sim_set_wall_lists(pic_sim* s, int on, int margin);
/* one frame of every species' positions (domain units) into a binary file,
   replaced atomically (a path under /dev/shm serves a shared-memory reader):
   "PICF", u32 version 1, u64 step, u32 species, u32 units 1, f64 width,
   f64 height, then per species u64 count and count float pairs; little-endian.
   No counterpart in the reference, which draws the pairs with OpenGL. */
int pic_sim_write_frame(pic_sim* s, const char* path);
int pic_sim_wall_list_stats(pic_sim* s, int* n_bins, long long* n_entries, int* max_per_bin);
int pic_sim_destroy(pic_sim* s);

/* ---- tuning (no counterpart in the reference, whose tiles and sweeps are
   fixed in the source) --------------------------------------------------- */
#define PIC_TUNE_BINS 1    /* particle bin tile (the push kernel's LDS tiles) */
#define PIC_TUNE_MG 2      /* SOR factor, V-cycles per refinement step, sweeps per level */
typedef struct {
    int bin_w, bin_h;        /* bin tile chosen */
    double push_ms;          /* its trial push time */
    double omega;            /* SOR factor chosen */
    int levels;              /* multigrid levels of the grid */
    int pre[16], post[16];   /* sweeps chosen per level, finest first (the
                                coarsest level uses coarse_sweeps) */
    double solve_ms;         /* a solve of the current charge from a zero
                                interior with them */
    int solve_cycles;
    int trials;              /* configurations timed */
    int inner;               /* FP32 V-cycles per refinement step chosen (mixed mode) */
} pic_tune_report;
/* Times candidate bin tiles with trial pushes that leave the simulation's
   state unchanged, and multigrid settings with solves of the current charge
   into scratch memory, and keeps the fastest.  Not for ensemble members. */
int pic_sim_autotune(pic_sim* s, int what, pic_tune_report* out);
/* sweeps of each multigrid level (finest first, n entries, each in [0,3]);
   levels from n on keep the uniform pre/post */
int pic_sim_set_level_sweeps(pic_sim* s, int n, const int* pre, const int* post);
int pic_mg_set_level_sweeps(pic_mg* m, int n, const int* pre, const int* post);

/* ---- embedded electrodes (Multigrid_embedded.cu:7-190) -------------------
   mask: n*n host bytes, 0 FLUID (solved), 1 DIRICHLET (held at phi_bc),
   2 SOLID (held at its value); phi_bc: n*n host doubles (may be NULL without
   Dirichlet nodes).  A NULL mask removes the electrodes.  The mask of a coarse
   level is that of the fine node it sits on; fixed nodes are not smoothed,
   have zero residual and take no correction.  The Dirichlet values are
   written into phi at the start of every solve.  Particles are not absorbed by
   electrodes: give them wall segments (pic_walls_add) for that. */
int pic_mg_set_electrodes(pic_mg* m, const uint8_t* mask, const double* phi_bc);
int pic_sim_set_electrodes(pic_sim* s, const uint8_t* mask, const double* phi_bc);

/* ---- ensembles of independent plasmas (no counterpart in the reference: it
   steps one plasma).  Members step together, one launch per stage for all of
   them; each member computes exactly what it computes alone. ------------- */
typedef struct pic_ensemble pic_ensemble;
int pic_ensemble_create(pic_ensemble** out);
/* members need the same particle precision and multigrid cycle parameters */
int pic_ensemble_add(pic_ensemble* e, pic_sim* s);
int pic_ensemble_remove(pic_ensemble* e, pic_sim* s);
int pic_ensemble_size(const pic_ensemble* e, int* n);
/* n batched steps of updateSimulationCUDA (main.cu:37-140) for every member */
int pic_ensemble_step(pic_ensemble* e, int n_steps);
/* graph mode for the batched step, as pic_sim_use_graph */
int pic_ensemble_use_graph(pic_ensemble* e, int on);
int pic_ensemble_destroy(pic_ensemble* e);

/* ---- job server (no counterpart in the reference): continuous batching of
   PIC jobs over an ensemble, admitted by HBM footprint. --------------------- */
typedef struct pic_server pic_server;
typedef struct {
    int cells;           /* cells per side (grid of cells+1 nodes) */
    double ppc;          /* macro-particles per cell and species */
    int steps;
    uint32_t seed;       /* Philox seed of the job */
    double te_eV;        /* electron temperature (0: 2 eV) */
    double density;      /* plasma density m^-3 (0: 1e15) */
    double gas_density;  /* Ar background for MCC (0: 1e21) */
    double arrival_s;    /* earliest admission time on the caller's clock */
    int mcc;             /* null-collision MCC on */
} pic_job_desc;
typedef struct {
    int state;           /* 0 waiting, 1 running, 2 finished, 3 released, -1 failed */
    double arrival_s, admit_s, done_s;
    int steps_done;
    double step_ms_mean, step_ms_max;
    int64_t particles, alive_end;
    double footprint_bytes;
} pic_job_result;
/* base: step-driver parameters shared by all jobs (dt and seed are per job);
   keep_finished: finished jobs keep their objects until pic_server_release */
int pic_server_create(double hbm_budget_bytes, int max_active, int precision, const pic_sim_params* base,
                      int keep_finished, pic_server** out);
int pic_server_submit(pic_server* sv, const pic_job_desc* job, int* job_id);
/* admit arrived jobs that fit, one batched step of every running job, retire
   the finished ones; now_s is the caller's clock */
int pic_server_tick(pic_server* sv, double now_s, int* n_running, int* n_waiting);
int pic_server_job_result(pic_server* sv, int job_id, pic_job_result* out);
int pic_server_job_step_ms(pic_server* sv, int job_id, int cap, float* out, int* n);
int pic_server_job_handles(pic_server* sv, int job_id, pic_particles** p, pic_grid** g, pic_sim** s);
int pic_server_release(pic_server* sv, int job_id);
int pic_server_used_bytes(const pic_server* sv, double* used);
int pic_server_destroy(pic_server* sv);

#ifdef __cplusplus
}
#endif
#endif
