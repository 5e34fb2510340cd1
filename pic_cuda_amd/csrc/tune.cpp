// tune.cpp -- the autotuner: bin tile of the particle kernel and SOR factor and
// sweeps per level of the multigrid, chosen by timing on the device.
//
// Bin tiles are timed with trial pushes: the push reads the current buffer
// parity and writes the other one, so a trial that is followed by clearing the
// output parity's bin counts, zeroing the charge accumulators and restoring the
// counters leaves the simulation exactly as it was (the random stream is keyed
// on the particle id and step, so the real step draws the same numbers).
// Multigrid settings are timed on the last step's solve: the current charge
// from the potential the step started from (from a zero interior before the
// first step), into a scratch potential.
#include "objects.hpp"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

using namespace pic;

namespace {

struct Counters {
    std::vector<std::vector<unsigned long long>> v;
};

int save_counters(pic_sim* S, Counters& c) {
    c.v.assign(S->p->sp.size(), std::vector<unsigned long long>(PIC_NCOUNT));
    for (size_t s = 0; s < S->p->sp.size(); ++s)
        PIC_HIP(hipMemcpy(c.v[s].data(), S->p->sp[s].d_counters, PIC_NCOUNT * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost));
    return 0;
}

int restore(pic_sim* S, const Counters& c) {
    hipStream_t sm = lib_stream();
    const size_t nn = (size_t)S->g->m * S->g->n;
    PIC_HIP(hipStreamSynchronize(sm));
    for (size_t s = 0; s < S->p->sp.size(); ++s)
        PIC_HIP(hipMemcpy(S->p->sp[s].d_counters, c.v[s].data(), PIC_NCOUNT * sizeof(unsigned long long),
                          hipMemcpyHostToDevice));
    for (auto* a : S->acc)
        if (a) PIC_HIP(hipMemset(a, 0, nn * sizeof(unsigned long long)));
    return 0;
}

// fastest of `reps` trial pushes with the current bin layout
template <typename T>
int trial_push(pic_sim* S, int reps, double* best) {
    pic_particles* p = S->p;
    hipStream_t sm = lib_stream();
    const int flags = PIC_F_DRIFT | PIC_F_DEPOSIT | (S->pending ? (PIC_F_KICK | (mcc_active(S) ? PIC_F_MCC : 0)) : 0);
    *best = 1e30;
    for (int r = 0; r < reps; ++r) {
        PIC_HIP(hipEventRecord(S->ev[0], sm));
        if (int rc = launch_push<T>((CaseDev<T>*)p->d_case, p->d_state, S->d_items, S->n_items, S->spill_chunk, flags,
                                    -1, cur_lds(S), sm, S->d_walls_cell != nullptr))
            return rc;
        PIC_HIP(hipEventRecord(S->ev[1], sm));
        if (int rc = launch_clear_out<T>((CaseDev<T>*)p->d_case, p->d_state, 1, sm)) return rc;
        PIC_HIP(hipStreamSynchronize(sm));
        float ms = 0.f;
        PIC_HIP(hipEventElapsedTime(&ms, S->ev[0], S->ev[1]));
        *best = std::min(*best, (double)ms);
    }
    return 0;
}

template <typename T>
int tune_bins(pic_sim* S, pic_tune_report* out) {
    static const int cand[][2] = {{8, 8}, {16, 8}, {16, 16}, {32, 16}, {32, 32}, {64, 32}};
    const int cells_x = S->g->m - 1, cells_y = S->g->n - 1;
    Counters saved;
    if (int rc = save_counters(S, saved)) return rc;
    const int bw0 = S->bin_w, bh0 = S->bin_h;
    int best_w = bw0, best_h = bh0;
    double best = 1e30;
    for (const auto& c : cand) {
        if (c[0] > cells_x || c[1] > cells_y) continue;
        if (int rc = sim_rebin(S, c[0], c[1])) return rc;
        double ms = 0;
        if (int rc = trial_push<T>(S, 3, &ms)) return rc;
        out->trials++;
        if (ms < best) {
            best = ms;
            best_w = c[0];
            best_h = c[1];
        }
    }
    if (int rc = restore(S, saved)) return rc;
    if (best_w != S->bin_w || best_h != S->bin_h)
        if (int rc = sim_rebin(S, best_w, best_h)) return rc;
    out->bin_w = best_w;
    out->bin_h = best_h;
    out->push_ms = best;
    return 0;
}

// fastest of three solves (a first one builds the launch tables, untimed)
int time_solve(pic_sim* S, const MGParams& prm, double* scratch, double* ms, int* cycles) {
    hipStream_t sm = lib_stream();
    const int n = S->g->m;
    const size_t nn = (size_t)n * n;
    MGSolver mg(n, prm, sm);
    if (int rc = mg.init()) return rc;
    *ms = 1e30;
    for (int r = 0; r < 4; ++r) {
        if (S->have_prev) {
            PIC_HIP(hipMemcpyAsync(scratch, S->phi_prev, nn * sizeof(double), hipMemcpyDeviceToDevice, sm));
        } else {
            PIC_HIP(hipMemcpyAsync(scratch, S->g->phi, nn * sizeof(double), hipMemcpyDeviceToDevice, sm));
            if (int rc = launch_zero_interior(scratch, n, sm)) return rc;
        }
        PIC_HIP(hipStreamSynchronize(sm));
        MGStats st{};
        const auto t0 = std::chrono::steady_clock::now();
        if (int rc = mg.solve(scratch, S->g->rhs, S->g->dx, &st)) return rc;   // ends with a synchronization
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (r > 0 && el < *ms) {
            *ms = el;
            *cycles = st.cycles;
        }
    }
    return 0;
}

int tune_mg(pic_sim* S, pic_tune_report* out) {
    const int n = S->g->m;
    const int L = S->mg->levels();
    const int nl = std::max(0, std::min(L - 1, 16));   // levels above the coarsest
    double* scratch = nullptr;
    PIC_HIP(hipMalloc(&scratch, (size_t)n * n * sizeof(double)));
    struct Free {
        double* p;
        ~Free() { (void)hipFree(p); }
    } guard{scratch};
    MGParams best = S->mgp;
    double best_ms = 1e30;
    int best_cyc = 0;
    auto trial = [&](const MGParams& p, bool* better) -> int {
        double ms = 0;
        int cyc = 0;
        if (int rc = time_solve(S, p, scratch, &ms, &cyc)) return rc;
        out->trials++;
        // a change has to win by 2% (timing noise)
        *better = ms < best_ms * 0.98;
        if (*better) {
            best = p;
            best_ms = ms;
            best_cyc = cyc;
        }
        return 0;
    };
    bool b = false;
    {
        MGParams p = S->mgp;
        p.per_level = 0;
        best = p;
        if (int rc = trial(p, &b)) return rc;
    }
    // SOR factor with the uniform sweeps
    for (double om : {1.0, 1.1, 1.15, 1.2, 1.25, 1.3}) {
        MGParams p = best;
        if (p.omega == om) continue;
        p.omega = om;
        if (int rc = trial(p, &b)) return rc;
    }
    // FP32 V-cycles per refinement step (mixed mode)
    if (best.mode == PIC_MG_MIXED)
        for (int in : {1, 2, 3}) {
            MGParams p = best;
            if (p.inner_eff() == in) continue;
            p.inner = in;
            if (int rc = trial(p, &b)) return rc;
        }
    // uniform sweep counts
    static const int uni[][2] = {{1, 1}, {2, 1}, {1, 2}, {2, 2}, {3, 2}, {2, 3}, {3, 3}};
    for (const auto& u : uni) {
        MGParams p = best;
        if (p.pre == u[0] && p.post == u[1]) continue;
        p.pre = u[0];
        p.post = u[1];
        if (int rc = trial(p, &b)) return rc;
    }
    // per level, finest first: pre then post sweeps
    {
        MGParams p = best;
        p.per_level = 1;
        for (int l = 0; l < 16; ++l) {
            p.pre_l[l] = p.pre;
            p.post_l[l] = p.post;
        }
        best = p;
    }
    for (int l = 0; l < nl; ++l)
        for (int leg = 0; leg < 2; ++leg)
            for (int v = 0; v <= 3; ++v) {
                MGParams p = best;
                int& slot = leg == 0 ? p.pre_l[l] : p.post_l[l];
                if (slot == v) continue;
                slot = v;
                if (int rc = trial(p, &b)) return rc;
            }
    if (int rc = sim_set_mg(S, best)) return rc;
    out->omega = best.omega;
    out->inner = best.inner_eff();
    out->levels = L;
    for (int l = 0; l < 16; ++l) {
        out->pre[l] = best.pre_at(l);
        out->post[l] = best.post_at(l);
    }
    out->solve_ms = best_ms;
    out->solve_cycles = best_cyc;
    return 0;
}

}  // namespace

extern "C" int pic_sim_autotune(pic_sim* S, int what, pic_tune_report* out) {
    PIC_REQUIRE(S && out, "NULL argument");
    PIC_REQUIRE(!S->ens, "the simulation is in an ensemble");
    PIC_REQUIRE((what & ~(PIC_TUNE_BINS | PIC_TUNE_MG)) == 0, "unknown tuning flags");
    std::memset(out, 0, sizeof(*out));
    out->bin_w = S->bin_w;
    out->bin_h = S->bin_h;
    out->omega = S->mgp.omega;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    sim_drop_graph(S);
    if (what & PIC_TUNE_BINS) {
        int rc = S->p->prec == 4 ? tune_bins<float>(S, out) : tune_bins<double>(S, out);
        if (rc) return rc;
    }
    if (what & PIC_TUNE_MG)
        if (int rc = tune_mg(S, out)) return rc;
    return 0;
}
