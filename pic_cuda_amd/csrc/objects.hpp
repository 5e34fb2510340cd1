// objects.hpp -- host objects behind the C ABI (pic_amd.h).
#pragma once
#include "pic_device.hpp"
#include "mg.hpp"
#include "field.hpp"
#include "push.hpp"
#include "../../include/pic_amd.h"
#include <vector>
#include <string>
#include <memory>

struct pic_sim;

// graph mode: consecutive steps that need fewer multigrid cycles than the
// captured graph runs before the graph is captured again with the lower count
#define PIC_GRAPH_SHRINK 8

namespace pic {

hipStream_t lib_stream();          // the library's stream (default priority)
constexpr double kElementaryCharge = 1.602176634e-19;

// one species store: two SoA buffers + per-parity bin layouts
struct SpeciesStore {
    double m = 1.0, q = 0.1, w = 1.0;
    int64_t capacity = 0;          // slots per buffer
    void* x[2] = {nullptr, nullptr};
    void* y[2] = {nullptr, nullptr};
    void* vx[2] = {nullptr, nullptr};
    void* vy[2] = {nullptr, nullptr};
    uint32_t* id[2] = {nullptr, nullptr};
    int32_t* d_start[2] = {nullptr, nullptr};
    int32_t* d_cap[2] = {nullptr, nullptr};
    int32_t* d_count[2] = {nullptr, nullptr};
    int32_t* d_mcount[2] = {nullptr, nullptr};
    int layout_alloc = 0;          // entries allocated in the layout arrays
    int nb = 0;                    // bins of the current layout (spill is index nb)
    std::vector<int32_t> h_start, h_cap;
    unsigned long long* d_counters = nullptr;
};

}  // namespace pic

struct pic_particles {
    int prec = 4;
    int64_t max_per_species = 0;
    std::vector<pic::SpeciesStore> sp;
    pic::CaseDev<float> hf{};
    pic::CaseDev<double> hd{};
    void* d_case = nullptr;        // CaseDev<T> on the device (own slot or an ensemble slot)
    pic::CaseState* d_state = nullptr;
    void* own_case = nullptr;      // the slots allocated with the particles
    pic::CaseState* own_state = nullptr;
    int h_cur = 0;                 // host mirror of the buffer parity
    pic::WorkItem* d_items = nullptr;   // scratch item table for store-level kernels
    int items_alloc = 0;
    int32_t* d_scratch = nullptr;  // counters / histogram scratch
    int scratch_alloc = 0;
    long long* d_ll = nullptr;
    pic_sim* attached = nullptr;
    // the reference's public device arrays (Particles2D::r, r_old, v, type):
    // a staging area of stage_cap entries; stage_pending entries are handed to
    // their species when a simulation takes the particles
    double* stage_r = nullptr;
    double* stage_r_old = nullptr;
    double* stage_v = nullptr;
    int32_t* stage_type = nullptr;
    int64_t stage_cap = 0;
    int64_t stage_pending = 0;
    uint32_t stage_next_id = 0;    // ids of staged particles continue from here
};

struct pic_grid {
    int m = 0, n = 0;
    double X = 0, Y = 0, dx = 0, dy = 0;
    double* rho = nullptr;
    double* phi = nullptr;
    double* E = nullptr;           // interleaved Ex, Ey
    double* rhs = nullptr;
};

struct pic_walls {
    double width = 0, height = 0;
    int nx = 0, ny = 0;
    std::vector<double> walls;     // domain units x1,y1,x2,y2
    bool box = false;
    bool compiled = false;
    double* d_walls = nullptr;     // compile_walls(): device copy (domain units)
};

struct pic_xs {
    struct Proc {
        int species, type, product;
        double thr, peak;
        std::vector<double> e, s;  // tabulated; empty for analytic
    };
    double density = 0, t_gas = 0;
    int points = 2048;
    double emax = 200.0;
    std::vector<Proc> procs;
};

struct pic_mg {
    std::unique_ptr<pic::MGSolver> solver;
};

struct pic_sim {
    pic_particles* p = nullptr;
    pic_grid* g = nullptr;
    pic_walls* w = nullptr;
    pic_xs* xs = nullptr;
    double dt = 0, eps0 = 0;
    int bin_w = 16, bin_h = 16, halo = 2, mcc = 1, rebin_every = 0;
    double wall_eps = 1e-10;
    pic::MGParams mgp{};
    std::unique_ptr<pic::MGSolver> mg;
    std::vector<unsigned long long*> acc;
    std::vector<void*> tabs;
    std::vector<std::vector<double>> tab_host;
    std::vector<pic_mcc_info> info;
    double* d_walls_cell = nullptr;
    // per-bin candidate walls of the current bin tile: offsets (bins + 1)
    // then wall indices, one allocation; rebuilt on every re-binning
    int* d_wall_lists = nullptr;
    bool wall_lists = true;
    int wall_margin = 2;
    long long wall_entries = 0;
    int wall_max_bin = 0;
    // frame export staging (float pairs), grown on demand
    float* d_frame = nullptr;
    int64_t frame_cap = 0;
    pic::WorkItem* d_items = nullptr;
    int n_items = 0;
    int spill_chunk = 4096;
    bool pending = false;
    bool need_rebin = false;
    int64_t steps = 0;
    int64_t total_cycles = 0;
    int last_cycles = 0;
    double last_rel = 0;
    int rebins = 0;
    double ms_push = 0, ms_field = 0, ms_total = 0, ms_graph = 0;
    int* h_status = nullptr;   // pinned, coherent host memory
    int* d_status = nullptr;   // its device address
    // device flag of the end of a step: some species' spill region is more
    // than half full (the re-binning rule), read by a push enqueued ahead
    int* d_rebin = nullptr;
    pic::ChargeArgs charge{};
    // push timing of alternate steps (0-1, 3-4) and the end of a step's solve (2)
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    struct pic_ensemble* ens = nullptr;   // batch this simulation belongs to
    // potential before the last step's solve: with the charge, the last
    // warm-started solve, which the tuner times
    double* phi_prev = nullptr;
    bool have_prev = false;
    // graph mode: the instantiated step and the launch parameters it was
    // captured with (a change of any of them captures it again)
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    const void* g_items = nullptr;
    int g_n_items = 0, g_flags = 0, g_cycles = 0;
    int g_shrink = 0;   // consecutive graph steps that needed fewer cycles than captured
    size_t g_lds = 0;
    int64_t graph_steps = 0, graph_captures = 0;
    int64_t ahead_steps = 0, ahead_held = 0;
};

namespace pic {
// shared helpers of capi.cpp used by ensemble.cpp
void items_for_species(const pic_particles* p, int s, int chunk, std::vector<WorkItem>& out, int case_idx = 0);
int upload_case(pic_particles* p);
int upload_items(WorkItem** dptr, int* alloc, const std::vector<WorkItem>& items);
// PIC_ERR_CAPACITY (with the message set) when a step dropped particles of a
// species for want of spill slots; fill is the spill slot counter after the step
int spill_status(const SpeciesStore& ss, int fill, int species);
}  // namespace pic
size_t cur_lds(pic_sim* S);
int mcc_active(pic_sim* S);
int sim_rebin(pic_sim* S, int bw, int bh);
void sim_drop_graph(pic_sim* S);
// the simulation's ensemble (if any) rebuilds its batch before the next step
void ensemble_invalidate(pic_sim* S);
// replace the simulation's multigrid solver (new parameters)
int sim_set_mg(pic_sim* S, const pic::MGParams& prm);
