// ensemble.cpp -- batched stepping of independent plasmas and the job server.
//
// pic_ensemble: simulations (pic_sim) join a batch whose case descriptors and
// states live in contiguous device arrays.  One batched step is one launch of
// each stage for all members: the particle pass over the concatenated
// (case, species, bin) work items, the end-of-step flip (one block per case),
// the charge conversion over (case, chunk) items and the multigrid cycles of
// MGBatch (every stage one launch, every case gated on its own convergence
// flag).  Each case computes exactly what it computes alone.
//
// pic_server: a request loop over an ensemble.  Jobs arrive with a time stamp;
// on every tick the server admits arrived jobs whose HBM footprint fits the
// budget (and a slot is free), advances all admitted jobs by one batched step,
// and retires the finished ones, recording step and completion latencies.
#include "objects.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>

using namespace pic;

struct pic_ensemble {
    std::vector<pic_sim*> m;
    int prec = 0;
    void* d_cases = nullptr;
    CaseState* d_states = nullptr;
    WorkItem* d_items = nullptr;
    int items_alloc = 0;
    int n_items = 0;
    ChargeArgs* d_charge = nullptr;
    long long* d_nns = nullptr;
    int4* d_chunks = nullptr;
    int n_chunks = 0;
    int* h_status = nullptr;   // pinned, coherent host memory
    int* d_status = nullptr;   // its device address
    std::unique_ptr<MGBatch> mg;
    std::vector<int> flags;        // flags currently stored in each case state
    bool dirty = true;
    size_t lds = 0;
    int64_t steps = 0;
    double ms = 0.0;
    // graph mode: the instantiated batched step and what it was captured for
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    int64_t gen = 0, g_gen = -1;   // layout generation (bumped by every rebuild)
    int g_cycles = 0;
    int g_shrink = 0;   // consecutive graph steps that needed fewer cycles than captured
    int64_t graph_steps = 0, graph_captures = 0;
};

static void drop_graph(pic_ensemble* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
}

static void free_arrays(pic_ensemble* e) {
    (void)hipStreamSynchronize(lib_stream());
    (void)hipFree(e->d_cases);
    (void)hipFree(e->d_states);
    (void)hipFree(e->d_items);
    (void)hipFree(e->d_charge);
    (void)hipFree(e->d_nns);
    (void)hipFree(e->d_chunks);
    if (e->h_status) (void)hipHostFree(e->h_status);
    e->d_cases = nullptr;
    e->d_states = nullptr;
    e->d_items = nullptr;
    e->items_alloc = 0;
    e->d_charge = nullptr;
    e->d_nns = nullptr;
    e->d_chunks = nullptr;
    e->d_status = nullptr;
    e->h_status = nullptr;
}

// move every member's descriptor and state into fresh contiguous arrays and
// rebuild the item tables and the multigrid batch
static int rebuild(pic_ensemble* e) {
    hipStream_t sm = lib_stream();
    PIC_HIP(hipStreamSynchronize(sm));
    drop_graph(e);
    e->gen++;
    const int M = (int)e->m.size();
    void* old_cases = e->d_cases;
    CaseState* old_states = e->d_states;
    e->d_cases = nullptr;
    e->d_states = nullptr;
    free_arrays(e);
    e->mg.reset();
    if (M == 0) {
        (void)hipFree(old_cases);
        (void)hipFree(old_states);
        e->dirty = false;
        return 0;
    }
    const size_t csz = e->prec == 4 ? sizeof(CaseDev<float>) : sizeof(CaseDev<double>);
    PIC_HIP(hipMalloc(&e->d_cases, M * csz));
    PIC_HIP(hipMalloc(&e->d_states, M * sizeof(CaseState)));
    for (int k = 0; k < M; ++k) {
        pic_particles* p = e->m[k]->p;
        PIC_HIP(hipMemcpyAsync(e->d_states + k, p->d_state, sizeof(CaseState), hipMemcpyDeviceToDevice, sm));
        p->d_case = (char*)e->d_cases + k * csz;
        p->d_state = e->d_states + k;
        if (int rc = upload_case(p)) return rc;
    }
    PIC_HIP(hipStreamSynchronize(sm));
    (void)hipFree(old_cases);
    (void)hipFree(old_states);
    std::vector<WorkItem> items;
    std::vector<ChargeArgs> charge(M);
    std::vector<long long> nns(M);
    std::vector<int4> chunks;
    std::vector<MGSolver*> solvers;
    std::vector<double*> phis;
    std::vector<const double*> rhss;
    std::vector<double> hs;
    e->lds = 0;
    e->flags.assign(M, -1);
    for (int k = 0; k < M; ++k) {
        pic_sim* S = e->m[k];
        for (int s = 0; s < (int)S->p->sp.size(); ++s) items_for_species(S->p, s, S->spill_chunk, items, k);
        charge[k] = S->charge;
        nns[k] = (long long)S->g->m * S->g->n;
        for (long long c = 0; c < (nns[k] + 255) / 256; ++c) chunks.push_back(make_int4(k, (int)c, 0, 0));
        solvers.push_back(S->mg.get());
        phis.push_back(S->g->phi);
        rhss.push_back(S->g->rhs);
        hs.push_back(S->g->dx);
        e->lds = std::max(e->lds, cur_lds(S));
    }
    if (int rc = upload_items(&e->d_items, &e->items_alloc, items)) return rc;
    e->n_items = (int)items.size();
    PIC_HIP(hipMalloc(&e->d_charge, M * sizeof(ChargeArgs)));
    PIC_HIP(hipMemcpy(e->d_charge, charge.data(), M * sizeof(ChargeArgs), hipMemcpyHostToDevice));
    PIC_HIP(hipMalloc(&e->d_nns, M * sizeof(long long)));
    PIC_HIP(hipMemcpy(e->d_nns, nns.data(), M * sizeof(long long), hipMemcpyHostToDevice));
    PIC_HIP(hipMalloc(&e->d_chunks, chunks.size() * sizeof(int4)));
    PIC_HIP(hipMemcpy(e->d_chunks, chunks.data(), chunks.size() * sizeof(int4), hipMemcpyHostToDevice));
    e->n_chunks = (int)chunks.size();
    PIC_HIP(hipHostMalloc(&e->h_status, M * PIC_MAX_SPECIES * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    PIC_HIP(hipHostGetDevicePointer((void**)&e->d_status, e->h_status, 0));
    std::memset(e->h_status, 0, M * PIC_MAX_SPECIES * sizeof(int));
    e->mg.reset(new MGBatch(sm));
    if (int rc = e->mg->build(solvers, phis, rhss, hs)) return rc;
    e->dirty = false;
    return 0;
}

extern "C" int pic_ensemble_create(pic_ensemble** out) {
    PIC_REQUIRE(out, "out is NULL");
    PIC_REQUIRE(lib_stream() != nullptr, "could not create a HIP stream");
    *out = new pic_ensemble();
    return 0;
}

extern "C" int pic_ensemble_add(pic_ensemble* e, pic_sim* S) {
    PIC_REQUIRE(e && S, "NULL argument");
    PIC_REQUIRE(!S->ens, "the simulation already belongs to an ensemble");
    PIC_REQUIRE(e->m.empty() || e->prec == S->p->prec, "ensemble members need the same particle precision");
    if (!e->m.empty()) {
        const MGParams& a = e->m[0]->mgp;
        const MGParams& b = S->mgp;
        PIC_REQUIRE(a.mode == b.mode && a.pre == b.pre && a.post == b.post && a.coarse_sweeps == b.coarse_sweeps &&
                        a.omega == b.omega && a.tol == b.tol && a.max_cycles == b.max_cycles &&
                        a.inner_eff() == b.inner_eff() && a.coarse_direct == b.coarse_direct,
                    "ensemble members need identical multigrid cycle parameters");
        PIC_REQUIRE(S->spill_chunk == e->m[0]->spill_chunk, "ensemble members need the same spill chunk");
    }
    e->prec = S->p->prec;
    S->have_prev = false;   // batched steps do not record the member's pre-solve potential
    e->m.push_back(S);
    S->ens = e;
    e->dirty = true;
    return 0;
}

extern "C" int pic_ensemble_remove(pic_ensemble* e, pic_sim* S) {
    PIC_REQUIRE(e && S, "NULL argument");
    auto it = std::find(e->m.begin(), e->m.end(), S);
    PIC_REQUIRE(it != e->m.end(), "the simulation is not a member of this ensemble");
    pic_particles* p = S->p;
    hipStream_t sm = lib_stream();
    PIC_HIP(hipStreamSynchronize(sm));
    if (p->d_state != p->own_state) {
        PIC_HIP(hipMemcpy(p->own_state, p->d_state, sizeof(CaseState), hipMemcpyDeviceToDevice));
        p->d_case = p->own_case;
        p->d_state = p->own_state;
        if (int rc = upload_case(p)) return rc;
    }
    S->ens = nullptr;
    e->m.erase(it);
    e->dirty = true;
    return 0;
}

extern "C" int pic_ensemble_size(const pic_ensemble* e, int* n) {
    PIC_REQUIRE(e && n, "NULL argument");
    *n = (int)e->m.size();
    return 0;
}

// push, end of step, spill status to the host, charge; then the multigrid
// solve: with cycles > 0 a fixed-cycle solve with no host synchronization (the
// body graph mode captures), with 0 the adaptive solve
template <typename T>
static int enqueue_body(pic_ensemble* e, int cycles) {
    hipStream_t sm = lib_stream();
    const int M = (int)e->m.size();
    bool walls = false;
    for (const pic_sim* S : e->m) walls = walls || S->d_walls_cell != nullptr;
    if (int rc = launch_push<T>((CaseDev<T>*)e->d_cases, e->d_states, e->d_items, e->n_items, e->m[0]->spill_chunk,
                                -1, -1, e->lds, sm, walls))
        return rc;
    if (int rc = launch_end_step<T>((CaseDev<T>*)e->d_cases, e->d_states, M, 1, sm, e->d_status)) return rc;
    if (int rc = launch_charge_batch(e->d_charge, e->d_nns, e->d_chunks, e->n_chunks, sm)) return rc;
    return cycles > 0 ? e->mg->enqueue_fixed(cycles) : e->mg->solve();
}

template <typename T>
static int ens_step_t(pic_ensemble* e) {
    hipStream_t sm = lib_stream();
    const int M = (int)e->m.size();
    for (int k = 0; k < M; ++k) {
        pic_sim* S = e->m[k];
        if (S->need_rebin || (S->rebin_every > 0 && S->steps > 0 && S->steps % S->rebin_every == 0)) {
            if (int rc = sim_rebin(S, S->bin_w, S->bin_h)) return rc;
            e->dirty = true;
        }
    }
    if (e->dirty)
        if (int rc = rebuild(e)) return rc;
    for (int k = 0; k < M; ++k) {
        pic_sim* S = e->m[k];
        const int f = PIC_F_DRIFT | PIC_F_DEPOSIT | (S->pending ? (PIC_F_KICK | (mcc_active(S) ? PIC_F_MCC : 0)) : 0);
        if (f != e->flags[k]) {
            PIC_HIP(hipMemcpyAsync(&e->d_states[k].flags, &f, sizeof(int), hipMemcpyHostToDevice, sm));
            PIC_HIP(hipStreamSynchronize(sm));
            e->flags[k] = f;
        }
    }
    bool all_pending = true;
    for (auto* S : e->m) all_pending = all_pending && S->pending;
    if (e->use_graph && all_pending) {
        // one graph launch: the captured body below with a solve of the last
        // step's cycle count, then completion of members that need more
        const int cycles = e->mg->last_cycles();
        // kept with up to one refinement step of gated no-op cycles to spare
        // until a lower count holds for PIC_GRAPH_SHRINK steps (the rule of
        // the single-simulation graph)
        const int step = e->m.empty() ? 1 : e->m[0]->mgp.inner_eff();
        e->g_shrink = (e->gexec && cycles < e->g_cycles) ? e->g_shrink + 1 : 0;
        const bool cycles_fit = e->g_cycles >= cycles && e->g_cycles <= cycles + step && e->g_shrink < PIC_GRAPH_SHRINK;
        if (e->gexec && (e->g_gen != e->gen || !cycles_fit)) drop_graph(e);
        if (!e->gexec) {
            PIC_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_body<T>(e, cycles);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(sm, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            PIC_HIP(ec);
            const hipError_t ei = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PIC_HIP(ei);
            e->g_gen = e->gen;
            e->g_cycles = cycles;
            e->graph_captures++;
        }
        PIC_HIP(hipGraphLaunch(e->gexec, sm));
        PIC_HIP(hipStreamSynchronize(sm));
        e->graph_steps++;
        if (int rc = e->mg->complete_fixed()) return rc;
    } else {
        if (int rc = enqueue_body<T>(e, 0)) return rc;
    }
    for (auto* S : e->m) S->p->h_cur ^= 1;
    int err = 0;   // the first member that lost particles; every member's state is still advanced
    for (int k = 0; k < M; ++k) {
        pic_sim* S = e->m[k];
        const MGNorm& nm = e->mg->norm(k);
        S->pending = true;
        S->steps++;
        S->last_cycles = nm.cycles;
        S->last_rel = nm.rel;
        S->total_cycles += nm.cycles;
        for (int s = 0; s < (int)S->p->sp.size(); ++s) {
            const SpeciesStore& ss = S->p->sp[s];
            if (e->h_status[k * PIC_MAX_SPECIES + s] > ss.h_cap[ss.nb] / 2) S->need_rebin = true;
            if (!err) err = spill_status(ss, e->h_status[k * PIC_MAX_SPECIES + s], s);
        }
    }
    e->steps++;
    return err;
}

extern "C" int pic_ensemble_step(pic_ensemble* e, int n_steps) {
    PIC_REQUIRE(e, "ensemble is NULL");
    PIC_REQUIRE(n_steps >= 0, "n_steps must be non-negative");
    if (e->m.empty()) return 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n_steps; ++i) {
        int rc = e->prec == 4 ? ens_step_t<float>(e) : ens_step_t<double>(e);
        if (rc) return rc;
    }
    e->ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

extern "C" int pic_ensemble_use_graph(pic_ensemble* e, int on) {
    PIC_REQUIRE(e, "ensemble is NULL");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    e->use_graph = on != 0;
    if (!e->use_graph) drop_graph(e);
    return 0;
}

extern "C" int pic_ensemble_destroy(pic_ensemble* e) {
    if (!e) return 0;
    drop_graph(e);
    while (!e->m.empty())
        if (int rc = pic_ensemble_remove(e, e->m.back())) return rc;
    free_arrays(e);
    e->mg.reset();
    delete e;
    return 0;
}

void ensemble_invalidate(pic_sim* S) {
    if (S->ens) S->ens->dirty = true;
}

// a simulation leaving its ensemble when destroyed
int ensemble_forget(pic_sim* S) {
    if (!S->ens) return 0;
    return pic_ensemble_remove(S->ens, S);
}

// ============================================================================
// job server
// ============================================================================
namespace {

constexpr double kEps0 = 8.8541878128e-12, kMe = 9.1093837015e-31, kAmu = 1.66053906660e-27;

struct Job {
    pic_job_desc d{};
    int state = 0;                 // 0 waiting, 1 active, 2 finished, 3 released, -1 failed
    pic_particles* p = nullptr;
    pic_grid* g = nullptr;
    pic_walls* w = nullptr;
    pic_xs* xs = nullptr;
    pic_sim* s = nullptr;
    double t_admit = 0, t_done = 0;
    int steps_done = 0;
    double step_ms_sum = 0, step_ms_max = 0;
    std::vector<float> step_ms;
    size_t footprint = 0;
    int64_t alive_end = 0;
};

void destroy_job(Job& j) {
    if (j.s) pic_sim_destroy(j.s);
    if (j.p) pic_particles_destroy(j.p);
    if (j.g) pic_grid_destroy(j.g);
    if (j.w) pic_walls_destroy(j.w);
    if (j.xs) pic_xs_destroy(j.xs);
    j.s = nullptr;
    j.p = nullptr;
    j.g = nullptr;
    j.w = nullptr;
    j.xs = nullptr;
}

int64_t job_particles(const pic_job_desc& d) {
    return std::max<int64_t>(1000, (int64_t)std::llround(d.ppc * (double)d.cells * (double)d.cells));
}

// bytes of device memory a job holds: two species stores of two buffers, the
// grid fields and accumulators, and the multigrid pyramid
size_t job_footprint(const pic_job_desc& d, int prec) {
    const int64_t n = job_particles(d);
    const int64_t cap = (int64_t)(n * 1.02) + 4096;
    const int64_t slots = cap + cap / 4 + 65536;
    const size_t per_slot = 4 * (size_t)prec + 4;
    const size_t nn = (size_t)(d.cells + 1) * (d.cells + 1);
    return 2 * 2 * (size_t)slots * per_slot + nn * (8 * 3 + 16 + 2 * 8) + nn * 24;
}

}  // namespace

struct pic_server {
    double budget = 0;
    size_t used = 0;
    int max_active = 0;
    int prec = 4;
    int keep = 0;
    pic_sim_params base{};
    std::vector<Job> jobs;
    std::deque<int> waiting;
    std::vector<int> active;
    pic_ensemble* ens = nullptr;
    int64_t ticks = 0;
};

static int build_job(pic_server* sv, Job& j) {
    const pic_job_desc& d = j.d;
    const double dens = d.density > 0 ? d.density : 1e15;
    const double te = d.te_eV > 0 ? d.te_eV : 2.0;
    const double debye = std::sqrt(kEps0 * te / (dens * kElementaryCharge));
    const double dx = 0.5 * debye;
    const double L = d.cells * dx;
    const double wpe = std::sqrt(dens * kElementaryCharge * kElementaryCharge / (kEps0 * kMe));
    const double dt = 0.2 / wpe;
    const double m_ar = 39.948 * kAmu;
    const int nn = d.cells + 1;
    const int64_t n = job_particles(d);
    const int64_t cap = (int64_t)(n * 1.02) + 4096;
    int rc = pic_grid_create(nn, nn, L, L, &j.g);
    if (!rc) rc = pic_particles_create(cap, sv->prec, &j.p);
    if (!rc) rc = pic_particles_change_species(j.p, 0, kMe, -kElementaryCharge);
    int ion = 1;
    if (!rc) rc = pic_particles_add_species(j.p, m_ar, kElementaryCharge, &ion);
    const double w = dens * L * L / (double)n;
    if (!rc) rc = pic_particles_set_weight(j.p, 0, w);
    if (!rc) rc = pic_particles_set_weight(j.p, ion, w);
    if (!rc) rc = pic_particles_init_maxwellian(j.p, 0, n, nn, nn, std::sqrt(te * kElementaryCharge / kMe), d.seed, 0, 0);
    if (!rc)
        rc = pic_particles_init_maxwellian(j.p, ion, n, nn, nn, std::sqrt(0.03 * kElementaryCharge / m_ar), d.seed, 0, 0);
    if (!rc) rc = pic_walls_create(L, L, nn, nn, &j.w);
    if (!rc) rc = pic_walls_add_box(j.w);
    if (!rc && d.mcc) {
        rc = pic_xs_create(d.gas_density > 0 ? d.gas_density : 1e21, 0.03, &j.xs);
        if (!rc) rc = pic_xs_add_analytic(j.xs, 0, PIC_PROC_ELASTIC, 0.0, 1e-19, -1);
        if (!rc) rc = pic_xs_add_analytic(j.xs, 0, PIC_PROC_EXCITATION, 11.5, 2e-21, -1);
        if (!rc) rc = pic_xs_add_analytic(j.xs, 0, PIC_PROC_IONIZATION, 15.8, 3e-20, ion);
        if (!rc) rc = pic_xs_add_analytic(j.xs, ion, PIC_PROC_CHARGE_EXCHANGE, 0.0, 5e-19, -1);
    }
    pic_sim_params prm = sv->base;
    prm.dt = dt;
    prm.seed = d.seed;
    prm.case_id = 0;
    prm.mcc = d.mcc;
    if (!rc) rc = pic_sim_create(j.p, j.g, j.w, j.xs, &prm, &j.s);
    if (rc) {
        std::string msg = pic::last_error();
        destroy_job(j);
        pic::set_error("building job " + std::to_string(d.seed) + ": " + msg);
    }
    return rc;
}

extern "C" int pic_server_create(double hbm_budget_bytes, int max_active, int precision, const pic_sim_params* base,
                                 int keep_finished, pic_server** out) {
    PIC_REQUIRE(out && base, "NULL argument");
    PIC_REQUIRE(hbm_budget_bytes > 0 && max_active > 0, "bad budget or slot count");
    PIC_REQUIRE(precision == 4 || precision == 8, "bad precision");
    auto* sv = new pic_server();
    sv->budget = hbm_budget_bytes;
    sv->max_active = max_active;
    sv->prec = precision;
    sv->keep = keep_finished;
    sv->base = *base;
    if (int rc = pic_ensemble_create(&sv->ens)) {
        delete sv;
        return rc;
    }
    *out = sv;
    return 0;
}

extern "C" int pic_server_submit(pic_server* sv, const pic_job_desc* d, int* job_id) {
    PIC_REQUIRE(sv && d, "NULL argument");
    PIC_REQUIRE(d->cells >= 2 && d->steps >= 0 && d->ppc > 0, "bad job description");
    Job j;
    j.d = *d;
    j.footprint = job_footprint(*d, sv->prec);
    PIC_REQUIRE((double)j.footprint <= sv->budget, "job footprint exceeds the HBM budget");
    sv->jobs.push_back(j);
    const int id = (int)sv->jobs.size() - 1;
    sv->waiting.push_back(id);
    if (job_id) *job_id = id;
    return 0;
}

static int retire(pic_server* sv, int id, double now) {
    Job& j = sv->jobs[id];
    if (int rc = pic_ensemble_remove(sv->ens, j.s)) return rc;
    if (int rc = pic_sim_sync(j.s)) return rc;
    int64_t n = 0;
    if (int rc = pic_particles_last_alive(j.p, -1, &n)) return rc;
    j.alive_end = n;
    j.t_done = now;
    j.state = 2;
    if (!sv->keep) {
        destroy_job(j);
        j.state = 3;
        sv->used -= j.footprint;
    }
    return 0;
}

extern "C" int pic_server_tick(pic_server* sv, double now_s, int* n_active, int* n_waiting) {
    PIC_REQUIRE(sv, "server is NULL");
    // admission: arrived jobs that fit the remaining budget and a free slot, in arrival order
    for (auto it = sv->waiting.begin(); it != sv->waiting.end();) {
        Job& j = sv->jobs[*it];
        if (j.d.arrival_s > now_s) { ++it; continue; }
        if ((int)sv->active.size() >= sv->max_active) break;
        if ((double)(sv->used + j.footprint) > sv->budget) { ++it; continue; }
        if (int rc = build_job(sv, j)) {
            j.state = -1;
            it = sv->waiting.erase(it);
            return rc;
        }
        if (int rc = pic_ensemble_add(sv->ens, j.s)) {
            // the job's objects were built but are not counted in `used`: free them
            destroy_job(j);
            j.state = -1;
            it = sv->waiting.erase(it);
            return rc;
        }
        j.state = 1;
        j.t_admit = now_s;
        sv->used += j.footprint;
        sv->active.push_back(*it);
        it = sv->waiting.erase(it);
    }
    // jobs of zero steps finish on admission
    for (size_t k = 0; k < sv->active.size();) {
        Job& j = sv->jobs[sv->active[k]];
        if (j.steps_done >= j.d.steps) {
            if (int rc = retire(sv, sv->active[k], now_s)) return rc;
            sv->active.erase(sv->active.begin() + k);
        } else {
            ++k;
        }
    }
    if (!sv->active.empty()) {
        auto t0 = std::chrono::steady_clock::now();
        if (int rc = pic_ensemble_step(sv->ens, 1)) return rc;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const double t_end = now_s + ms * 1e-3;
        for (size_t k = 0; k < sv->active.size();) {
            Job& j = sv->jobs[sv->active[k]];
            j.steps_done++;
            j.step_ms_sum += ms;
            j.step_ms_max = std::max(j.step_ms_max, ms);
            j.step_ms.push_back((float)ms);
            if (j.steps_done >= j.d.steps) {
                if (int rc = retire(sv, sv->active[k], t_end)) return rc;
                sv->active.erase(sv->active.begin() + k);
            } else {
                ++k;
            }
        }
        sv->ticks++;
    }
    if (n_active) *n_active = (int)sv->active.size();
    if (n_waiting) *n_waiting = (int)sv->waiting.size();
    return 0;
}

extern "C" int pic_server_job_result(pic_server* sv, int id, pic_job_result* out) {
    PIC_REQUIRE(sv && out && id >= 0 && id < (int)sv->jobs.size(), "bad job id");
    const Job& j = sv->jobs[id];
    *out = pic_job_result{};
    out->state = j.state;
    out->arrival_s = j.d.arrival_s;
    out->admit_s = j.t_admit;
    out->done_s = j.t_done;
    out->steps_done = j.steps_done;
    out->step_ms_mean = j.steps_done ? j.step_ms_sum / j.steps_done : 0.0;
    out->step_ms_max = j.step_ms_max;
    out->particles = 2 * job_particles(j.d);
    out->alive_end = j.alive_end;
    out->footprint_bytes = (double)j.footprint;
    return 0;
}

extern "C" int pic_server_job_step_ms(pic_server* sv, int id, int cap, float* out, int* n) {
    PIC_REQUIRE(sv && n && id >= 0 && id < (int)sv->jobs.size(), "bad job id");
    const Job& j = sv->jobs[id];
    *n = (int)j.step_ms.size();
    if (out) {
        PIC_REQUIRE(cap >= *n, "capacity too small");
        std::copy(j.step_ms.begin(), j.step_ms.end(), out);
    }
    return 0;
}

extern "C" int pic_server_job_handles(pic_server* sv, int id, pic_particles** p, pic_grid** g, pic_sim** s) {
    PIC_REQUIRE(sv && id >= 0 && id < (int)sv->jobs.size(), "bad job id");
    const Job& j = sv->jobs[id];
    PIC_REQUIRE(j.state == 1 || j.state == 2, "job holds no objects (waiting or released)");
    if (p) *p = j.p;
    if (g) *g = j.g;
    if (s) *s = j.s;
    return 0;
}

extern "C" int pic_server_release(pic_server* sv, int id) {
    PIC_REQUIRE(sv && id >= 0 && id < (int)sv->jobs.size(), "bad job id");
    Job& j = sv->jobs[id];
    PIC_REQUIRE(j.state == 2, "only finished jobs can be released");
    destroy_job(j);
    j.state = 3;
    sv->used -= j.footprint;
    return 0;
}

extern "C" int pic_server_used_bytes(const pic_server* sv, double* used) {
    PIC_REQUIRE(sv && used, "NULL argument");
    *used = (double)sv->used;
    return 0;
}

extern "C" int pic_server_destroy(pic_server* sv) {
    if (!sv) return 0;
    for (auto& j : sv->jobs) {
        if (j.s && j.s->ens) pic_ensemble_remove(sv->ens, j.s);
        destroy_job(j);
    }
    pic_ensemble_destroy(sv->ens);
    delete sv;
    return 0;
}
