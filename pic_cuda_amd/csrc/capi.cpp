// capi.cpp -- host side of the C ABI declared in include/pic_amd.h.
//
// Objects mirror the reference's host interfaces (Particles2D, Grid2D,
// Boundaries2D, solve_multigrid, cross_sections/elasticCollisions and the step
// driver updateSimulationCUDA) and own their device memory.  All work runs in
// order on one HIP stream of the library.  Errors are returned as status codes.
#include "objects.hpp"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

namespace pic {

static thread_local std::string t_err;
void set_error(const std::string& msg) { t_err = msg; }
const char* last_error() { return t_err.c_str(); }

hipStream_t lib_stream() {
    static hipStream_t s = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (hipStreamCreate(&s) != hipSuccess) s = nullptr;
    });
    return s;
}

template <typename T> CaseDev<T>& hcase(pic_particles* p);
template <> CaseDev<float>& hcase<float>(pic_particles* p) { return p->hf; }
template <> CaseDev<double>& hcase<double>(pic_particles* p) { return p->hd; }

int upload_case(pic_particles* p) {
    hipStream_t s = lib_stream();
    if (p->prec == 4) PIC_HIP(hipMemcpyAsync(p->d_case, &p->hf, sizeof(p->hf), hipMemcpyHostToDevice, s));
    else PIC_HIP(hipMemcpyAsync(p->d_case, &p->hd, sizeof(p->hd), hipMemcpyHostToDevice, s));
    PIC_HIP(hipStreamSynchronize(s));
    return 0;
}

template <typename T>
static void fill_species_ptrs(pic_particles* p, int s) {
    SpeciesDev<T>& S = hcase<T>(p).sp[s];
    SpeciesStore& st = p->sp[s];
    for (int par = 0; par < 2; ++par) {
        S.x[par] = (T*)st.x[par];
        S.y[par] = (T*)st.y[par];
        S.vx[par] = (T*)st.vx[par];
        S.vy[par] = (T*)st.vy[par];
        S.id[par] = st.id[par];
        S.start[par] = st.d_start[par];
        S.cap[par] = st.d_cap[par];
        S.count[par] = st.d_count[par];
        S.mcount[par] = st.d_mcount[par];
        S.blk[par] = (char*)st.x[par];
        S.lay[par] = st.d_count[par];
        S.nb[par] = st.nb;
    }
    S.pcap = (uint32_t)st.capacity;
    S.counters = st.d_counters;
    S.key = (uint32_t)s;
    S.other = -1;
}

static int ensure_scratch(pic_particles* p, int n) {
    if (p->scratch_alloc >= n) return 0;
    if (p->d_scratch) PIC_HIP(hipFree(p->d_scratch));
    p->d_scratch = nullptr;
    PIC_HIP(hipMalloc(&p->d_scratch, (size_t)n * sizeof(int32_t)));
    p->scratch_alloc = n;
    return 0;
}

int upload_items(WorkItem** dptr, int* alloc, const std::vector<WorkItem>& items) {
    if ((int)items.size() > *alloc) {
        if (*dptr) PIC_HIP(hipFree(*dptr));
        *dptr = nullptr;
        const int n = std::max<int>((int)items.size(), 64);
        PIC_HIP(hipMalloc(dptr, (size_t)n * sizeof(WorkItem)));
        *alloc = n;
    }
    if (!items.empty()) {
        PIC_HIP(hipMemcpyAsync(*dptr, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                               lib_stream()));
        PIC_HIP(hipStreamSynchronize(lib_stream()));
    }
    return 0;
}

void items_for_species(const pic_particles* p, int s, int chunk, std::vector<WorkItem>& out, int case_idx) {
    const SpeciesStore& st = p->sp[s];
    for (int b = 0; b < st.nb; ++b) out.push_back(WorkItem{case_idx, s, b, 0});
    const int spill_cap = st.h_cap[st.nb];
    const int chunks = std::max(1, (spill_cap + chunk - 1) / chunk);
    for (int k = 0; k < chunks; ++k) out.push_back(WorkItem{case_idx, s, st.nb, k});
}

static int alloc_layout_parity(SpeciesStore& st, int par, int entries) {
    // one block: count, mcount, start, cap (entries each)
    if (st.d_count[par]) PIC_HIP(hipFree(st.d_count[par]));
    st.d_start[par] = st.d_cap[par] = st.d_count[par] = st.d_mcount[par] = nullptr;
    int32_t* b = nullptr;
    PIC_HIP(hipMalloc(&b, 4 * (size_t)entries * sizeof(int32_t)));
    st.d_count[par] = b;
    st.d_mcount[par] = b + entries;
    st.d_start[par] = b + 2 * (size_t)entries;
    st.d_cap[par] = b + 3 * (size_t)entries;
    PIC_HIP(hipMemset(b, 0, 2 * (size_t)entries * sizeof(int32_t)));
    return 0;
}

static int create_species(pic_particles* p, double m, double q) {
    PIC_REQUIRE((int)p->sp.size() < PIC_MAX_SPECIES, "at most 4 species");
    p->sp.emplace_back();
    SpeciesStore& st = p->sp.back();
    st.m = m;
    st.q = q;
    st.w = 1.0;
    // bins get 1.5x their population plus 64 slots at every re-binning
    const int64_t cap = p->max_per_species + p->max_per_species / 2 + p->max_per_species / 8 + (1 << 20);
    PIC_REQUIRE(cap < (int64_t)INT32_MAX, "capacity per species must stay below 2^31 slots");
    PIC_REQUIRE(cap * (int64_t)p->prec < ((int64_t)1 << 32), "capacity per species must stay below 2^32 bytes per array");
    st.capacity = cap;
    const size_t eb = (size_t)p->prec;
    for (int par = 0; par < 2; ++par) {
        // one block: x, y, vx, vy, id (cap slots each)
        char* b = nullptr;
        PIC_HIP(hipMalloc(&b, cap * (4 * eb + sizeof(uint32_t))));
        st.x[par] = b;
        st.y[par] = b + cap * eb;
        st.vx[par] = b + 2 * cap * eb;
        st.vy[par] = b + 3 * cap * eb;
        st.id[par] = (uint32_t*)(b + 4 * cap * eb);
        if (int rc = alloc_layout_parity(st, par, 1)) return rc;
    }
    st.layout_alloc = 1;
    st.nb = 0;
    st.h_start.assign(1, 0);
    st.h_cap.assign(1, (int32_t)cap);
    for (int par = 0; par < 2; ++par) {
        PIC_HIP(hipMemcpy(st.d_start[par], st.h_start.data(), sizeof(int32_t), hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(st.d_cap[par], st.h_cap.data(), sizeof(int32_t), hipMemcpyHostToDevice));
    }
    PIC_HIP(hipMalloc(&st.d_counters, PIC_NCOUNT * sizeof(unsigned long long)));
    PIC_HIP(hipMemset(st.d_counters, 0, PIC_NCOUNT * sizeof(unsigned long long)));
    const int s = (int)p->sp.size() - 1;
    if (p->prec == 4) {
        fill_species_ptrs<float>(p, s);
        p->hf.ns = (int)p->sp.size();
    } else {
        fill_species_ptrs<double>(p, s);
        p->hd.ns = (int)p->sp.size();
    }
    return upload_case(p);
}

static int read_count(pic_particles* p, int s, int par, int b, int32_t* out) {
    PIC_HIP(hipMemcpyAsync(out, p->sp[s].d_count[par] + b, sizeof(int32_t), hipMemcpyDeviceToHost, lib_stream()));
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    return 0;
}

}  // namespace pic

using namespace pic;

// forward declarations of the sim helpers used by the particle API
static int sim_sync_impl(pic_sim* s);

// ============================================================================
// library
// ============================================================================
extern "C" const char* pic_last_error(void) { return pic::last_error(); }

extern "C" int pic_version(int* major, int* minor) {
    if (major) *major = 1;
    if (minor) *minor = 0;
    return 0;
}

extern "C" int pic_device_synchronize(void) {
    PIC_HIP(hipDeviceSynchronize());
    return 0;
}

extern "C" int pic_device_alloc(int64_t bytes, void** out) {
    PIC_REQUIRE(out && bytes > 0, "bad size or NULL argument");
    *out = nullptr;
    PIC_HIP(hipMalloc(out, (size_t)bytes));
    return 0;
}

extern "C" int pic_device_free(void* ptr) {
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    PIC_HIP(hipFree(ptr));
    return 0;
}

extern "C" int pic_device_to_host(void* host, const void* dev, int64_t bytes) {
    PIC_REQUIRE(host && dev && bytes >= 0, "NULL argument or negative size");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    if (bytes) PIC_HIP(hipMemcpy(host, dev, (size_t)bytes, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int pic_host_to_device(void* dev, const void* host, int64_t bytes) {
    PIC_REQUIRE(host && dev && bytes >= 0, "NULL argument or negative size");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    if (bytes) PIC_HIP(hipMemcpy(dev, host, (size_t)bytes, hipMemcpyHostToDevice));
    return 0;
}

// ============================================================================
// Particles2D
// ============================================================================
extern "C" int pic_particles_create(int64_t max_particles, int precision, pic_particles** out) {
    PIC_REQUIRE(out, "out is NULL");
    PIC_REQUIRE(max_particles > 0, "max_particles must be positive");
    PIC_REQUIRE(precision == 4 || precision == 8, "precision must be PIC_PREC_FP32 or PIC_PREC_FP64");
    PIC_REQUIRE(lib_stream() != nullptr, "could not create a HIP stream");
    auto* p = new pic_particles();
    p->prec = precision;
    p->max_per_species = max_particles;
    if (hipMalloc(&p->d_case, precision == 4 ? sizeof(CaseDev<float>) : sizeof(CaseDev<double>)) != hipSuccess ||
        hipMalloc(&p->d_state, sizeof(CaseState)) != hipSuccess || hipMalloc(&p->d_ll, sizeof(long long)) != hipSuccess) {
        set_error("pic_particles_create: device allocation failed");
        delete p;
        return PIC_ERR_HIP;
    }
    p->own_case = p->d_case;
    p->own_state = p->d_state;
    CaseState cs{0u, 0, 0, 0};
    if (hipMemcpy(p->d_state, &cs, sizeof(cs), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("pic_particles_create: state upload failed");
        return PIC_ERR_HIP;
    }
    // reference constructor: species 0 with q = 0.1, m = 1 (particles.cu:146-149)
    if (int rc = create_species(p, 1.0, 0.1)) {
        pic_particles_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

extern "C" int pic_particles_destroy(pic_particles* p) {
    if (!p) return 0;
    PIC_REQUIRE(!p->attached, "destroy the pic_sim before its particles");
    (void)hipStreamSynchronize(lib_stream());
    for (auto& st : p->sp) {
        for (int par = 0; par < 2; ++par) {
            (void)hipFree(st.x[par]);         // the particle block
            (void)hipFree(st.d_count[par]);   // the layout block
        }
        (void)hipFree(st.d_counters);
    }
    (void)hipFree(p->own_case);
    (void)hipFree(p->own_state);
    (void)hipFree(p->d_items);
    (void)hipFree(p->d_scratch);
    (void)hipFree(p->d_ll);
    (void)hipFree(p->stage_r);
    (void)hipFree(p->stage_r_old);
    (void)hipFree(p->stage_v);
    (void)hipFree(p->stage_type);
    delete p;
    return 0;
}

extern "C" int pic_particles_add_species(pic_particles* p, double m, double q, int* index) {
    PIC_REQUIRE(p, "particles is NULL");
    PIC_REQUIRE(!p->attached, "species cannot be added while a pic_sim uses the particles");
    PIC_REQUIRE(m > 0, "mass must be positive");
    if (int rc = create_species(p, m, q)) return rc;
    if (index) *index = (int)p->sp.size() - 1;
    return 0;
}

extern "C" int pic_particles_change_species(pic_particles* p, int species, double m, double q) {
    PIC_REQUIRE(p && species >= 0 && species < (int)p->sp.size(), "bad species");
    PIC_REQUIRE(!p->attached, "species cannot change while a pic_sim uses the particles");
    PIC_REQUIRE(m > 0, "mass must be positive");
    p->sp[species].m = m;
    p->sp[species].q = q;
    return 0;
}

extern "C" int pic_particles_set_weight(pic_particles* p, int species, double weight) {
    PIC_REQUIRE(p && species >= 0 && species < (int)p->sp.size(), "bad species");
    PIC_REQUIRE(!p->attached, "weights cannot change while a pic_sim uses the particles");
    p->sp[species].w = weight;
    return 0;
}

extern "C" int pic_particles_num_species(const pic_particles* p, int* n) {
    PIC_REQUIRE(p && n, "NULL argument");
    *n = (int)p->sp.size();
    return 0;
}

extern "C" int pic_particles_last_alive(pic_particles* p, int species, int64_t* n) {
    PIC_REQUIRE(p && n, "NULL argument");
    PIC_REQUIRE(species >= -1 && species < (int)p->sp.size(), "bad species");
    int64_t tot = 0;
    for (int s = 0; s < (int)p->sp.size(); ++s) {
        if (species >= 0 && s != species) continue;
        int rc = p->prec == 4 ? launch_counts<float>((CaseDev<float>*)p->d_case, p->d_state, 0, s, p->d_ll, lib_stream())
                              : launch_counts<double>((CaseDev<double>*)p->d_case, p->d_state, 0, s, p->d_ll, lib_stream());
        if (rc) return rc;
        long long v = 0;
        PIC_HIP(hipMemcpyAsync(&v, p->d_ll, sizeof(v), hipMemcpyDeviceToHost, lib_stream()));
        PIC_HIP(hipStreamSynchronize(lib_stream()));
        tot += v;
    }
    *n = tot;
    return 0;
}

template <typename T>
static int upload_t(pic_particles* p, int s, int64_t n, const double* x, const double* y, const double* vx,
                    const double* vy, const uint32_t* id) {
    SpeciesStore& st = p->sp[s];
    const int par = p->h_cur, b = st.nb;
    int32_t cnt = 0;
    if (int rc = read_count(p, s, par, b, &cnt)) return rc;
    cnt = std::min(cnt, st.h_cap[b]);
    if ((int64_t)cnt + n > st.h_cap[b]) {
        set_error("pic_particles_upload: not enough free slots (capacity per species exceeded)");
        return PIC_ERR_CAPACITY;
    }
    std::vector<T> bx(n), by(n), bvx(n), bvy(n);
    std::vector<uint32_t> bid(n);
    // ids continue after the particles already held when none are given
    for (int64_t i = 0; i < n; ++i) {
        bx[i] = (T)x[i];
        by[i] = (T)y[i];
        bvx[i] = (T)vx[i];
        bvy[i] = (T)vy[i];
        bid[i] = id ? id[i] : (uint32_t)(cnt + i);
    }
    const size_t off = (size_t)st.h_start[b] + cnt;
    hipStream_t sm = lib_stream();
    PIC_HIP(hipMemcpyAsync((T*)st.x[par] + off, bx.data(), n * sizeof(T), hipMemcpyHostToDevice, sm));
    PIC_HIP(hipMemcpyAsync((T*)st.y[par] + off, by.data(), n * sizeof(T), hipMemcpyHostToDevice, sm));
    PIC_HIP(hipMemcpyAsync((T*)st.vx[par] + off, bvx.data(), n * sizeof(T), hipMemcpyHostToDevice, sm));
    PIC_HIP(hipMemcpyAsync((T*)st.vy[par] + off, bvy.data(), n * sizeof(T), hipMemcpyHostToDevice, sm));
    PIC_HIP(hipMemcpyAsync(st.id[par] + off, bid.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, sm));
    const int32_t nc = cnt + (int32_t)n;
    PIC_HIP(hipMemcpyAsync(st.d_count[par] + b, &nc, sizeof(nc), hipMemcpyHostToDevice, sm));
    PIC_HIP(hipStreamSynchronize(sm));
    return 0;
}

extern "C" int pic_particles_upload(pic_particles* p, int species, int64_t n, const double* x, const double* y,
                                    const double* vx, const double* vy, const uint32_t* id) {
    PIC_REQUIRE(p && species >= 0 && species < (int)p->sp.size(), "bad species");
    PIC_REQUIRE(n >= 0, "n must be non-negative");
    if (n == 0) return 0;
    PIC_REQUIRE(x && y && vx && vy, "NULL array");
    if (p->attached) {
        if (int rc = sim_sync_impl(p->attached)) return rc;
    }
    return p->prec == 4 ? upload_t<float>(p, species, n, x, y, vx, vy, id)
                        : upload_t<double>(p, species, n, x, y, vx, vy, id);
}

// ---- the reference's public device arrays (particles.cuh:7-20) ------------
extern "C" int pic_particles_staging(pic_particles* p, int64_t n_max, double** r, double** r_old, double** v,
                                     int32_t** type) {
    PIC_REQUIRE(p, "particles is NULL");
    PIC_REQUIRE(n_max >= 0 && n_max < INT32_MAX, "bad staging size");
    PIC_REQUIRE(p->stage_pending == 0, "staged particles are waiting for a simulation");
    if (n_max > p->stage_cap) {
        PIC_HIP(hipStreamSynchronize(lib_stream()));
        (void)hipFree(p->stage_r);
        (void)hipFree(p->stage_r_old);
        (void)hipFree(p->stage_v);
        (void)hipFree(p->stage_type);
        p->stage_r = p->stage_r_old = p->stage_v = nullptr;
        p->stage_type = nullptr;
        p->stage_cap = 0;
        PIC_HIP(hipMalloc(&p->stage_r, (size_t)n_max * 2 * sizeof(double)));
        PIC_HIP(hipMalloc(&p->stage_v, (size_t)n_max * 2 * sizeof(double)));
        PIC_HIP(hipMalloc(&p->stage_type, (size_t)n_max * sizeof(int32_t)));
        // species 0 until a caller writes the types
        PIC_HIP(hipMemset(p->stage_type, 0, (size_t)n_max * sizeof(int32_t)));
        p->stage_cap = n_max;
    }
    if (r_old && !p->stage_r_old && p->stage_cap > 0)
        PIC_HIP(hipMalloc(&p->stage_r_old, (size_t)p->stage_cap * 2 * sizeof(double)));
    if (r) *r = p->stage_r;
    if (r_old) *r_old = p->stage_r_old;
    if (v) *v = p->stage_v;
    if (type) *type = p->stage_type;
    return 0;
}

template <typename T>
static int commit_staging_t(pic_particles* p, int64_t n, double dx, double dy) {
    const int ns = (int)p->sp.size(), par = p->h_cur;
    constexpr int NC = PIC_MAX_SPECIES + 1;
    if (int rc = ensure_scratch(p, 2 * NC)) return rc;
    hipStream_t sm = lib_stream();
    int32_t* counts = p->d_scratch;
    int32_t* fill = p->d_scratch + NC;
    PIC_HIP(hipMemsetAsync(p->d_scratch, 0, 2 * NC * sizeof(int32_t), sm));
    if (int rc = launch_staged_count(p->stage_type, n, ns, counts, sm)) return rc;
    int32_t hc[NC] = {};
    PIC_HIP(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, sm));
    PIC_HIP(hipStreamSynchronize(sm));
    if (hc[PIC_MAX_SPECIES]) {
        set_error("staged particles: " + std::to_string(hc[PIC_MAX_SPECIES]) + " entries of type[] are not a species");
        return PIC_ERR_ARG;
    }
    StagedDst<T> d{};
    int32_t nc[PIC_MAX_SPECIES] = {};
    for (int s = 0; s < ns; ++s) {
        SpeciesStore& st = p->sp[s];
        const int b = st.nb;
        int32_t cnt = 0;
        if (int rc = read_count(p, s, par, b, &cnt)) return rc;
        if (cnt < 0 || (int64_t)cnt + hc[s] > st.h_cap[b]) {
            set_error("staged particles: not enough free slots for species " + std::to_string(s) +
                      " (capacity per species exceeded)");
            return PIC_ERR_CAPACITY;
        }
        const size_t off = (size_t)st.h_start[b] + cnt;
        d.x[s] = (T*)st.x[par] + off;
        d.y[s] = (T*)st.y[par] + off;
        d.vx[s] = (T*)st.vx[par] + off;
        d.vy[s] = (T*)st.vy[par] + off;
        d.id[s] = st.id[par] + off;
        nc[s] = cnt + hc[s];
    }
    d.fill = fill;
    if (int rc = launch_staged_append<T>(p->stage_r, p->stage_v, p->stage_type, n, dx, dy, ns, d, p->stage_next_id, sm))
        return rc;
    for (int s = 0; s < ns; ++s)
        PIC_HIP(hipMemcpyAsync(p->sp[s].d_count[par] + p->sp[s].nb, &nc[s], sizeof(int32_t), hipMemcpyHostToDevice, sm));
    // the next fill starts from species 0 again
    PIC_HIP(hipMemsetAsync(p->stage_type, 0, (size_t)n * sizeof(int32_t), sm));
    PIC_HIP(hipStreamSynchronize(sm));
    p->stage_next_id += (uint32_t)n;
    p->stage_pending = 0;
    return 0;
}

static int commit_staging(pic_particles* p, int64_t n, double dx, double dy) {
    PIC_REQUIRE(n >= 0 && n <= p->stage_cap, "more staged particles than the staging area holds");
    PIC_REQUIRE(dx > 0 && dy > 0, "cell size must be positive");
    if (n == 0) {
        p->stage_pending = 0;
        return 0;
    }
    return p->prec == 4 ? commit_staging_t<float>(p, n, dx, dy) : commit_staging_t<double>(p, n, dx, dy);
}

extern "C" int pic_particles_commit_staging(pic_particles* p, int64_t n, double dx, double dy) {
    PIC_REQUIRE(p, "particles is NULL");
    if (p->attached) {
        if (int rc = sim_sync_impl(p->attached)) return rc;
    }
    if (int rc = commit_staging(p, n, dx, dy)) return rc;
    if (p->attached) p->attached->need_rebin = true;   // bin them before the next step
    return 0;
}

extern "C" int pic_particles_set_staged(pic_particles* p, int64_t n) {
    PIC_REQUIRE(p, "particles is NULL");
    PIC_REQUIRE(n >= 0 && n <= p->stage_cap, "more staged particles than the staging area holds");
    if (!p->attached) {
        p->stage_pending = n;   // handed over when a simulation takes the particles
        return 0;
    }
    return pic_particles_commit_staging(p, n, p->attached->g->dx, p->attached->g->dy);
}

template <typename T>
static int download_t(pic_particles* p, int s, int64_t cap, double* x, double* y, double* vx, double* vy,
                      uint32_t* id, int64_t* nout) {
    SpeciesStore& st = p->sp[s];
    std::vector<WorkItem> items;
    items_for_species(p, s, 4096, items);
    if (int rc = upload_items(&p->d_items, &p->items_alloc, items)) return rc;
    if (int rc = ensure_scratch(p, 1)) return rc;
    hipStream_t sm = lib_stream();
    PIC_HIP(hipMemsetAsync(p->d_scratch, 0, sizeof(int32_t), sm));
    if (int rc = launch_gather<T>((CaseDev<T>*)p->d_case, p->d_state, p->d_items, (int)items.size(), 4096,
                                  p->d_scratch, sm))
        return rc;
    int32_t n = 0;
    PIC_HIP(hipMemcpyAsync(&n, p->d_scratch, sizeof(n), hipMemcpyDeviceToHost, sm));
    PIC_HIP(hipStreamSynchronize(sm));
    if (nout) *nout = n;
    if (!x) return 0;   // count query
    if (n > cap) {
        set_error("pic_particles_download: output capacity too small");
        return PIC_ERR_CAPACITY;
    }
    const int par = p->h_cur ^ 1;
    std::vector<T> b(n);
    void* src[4] = {st.x[par], st.y[par], st.vx[par], st.vy[par]};
    double* dst[4] = {x, y, vx, vy};
    for (int k = 0; k < 4; ++k) {
        if (!dst[k]) continue;
        PIC_HIP(hipMemcpy(b.data(), src[k], (size_t)n * sizeof(T), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) dst[k][i] = (double)b[i];
    }
    if (id) PIC_HIP(hipMemcpy(id, st.id[par], (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int pic_particles_download(pic_particles* p, int species, int64_t cap, double* x, double* y, double* vx,
                                      double* vy, uint32_t* id, int64_t* n) {
    PIC_REQUIRE(p && species >= 0 && species < (int)p->sp.size(), "bad species");
    if (p->attached) {
        if (int rc = sim_sync_impl(p->attached)) return rc;
    }
    return p->prec == 4 ? download_t<float>(p, species, cap, x, y, vx, vy, id, n)
                        : download_t<double>(p, species, cap, x, y, vx, vy, id, n);
}

template <typename T>
static int export_t(pic_particles* p, int s, float* dst, int64_t cap, double sx, double sy, int64_t* nout) {
    std::vector<WorkItem> items;
    items_for_species(p, s, 4096, items);
    if (int rc = upload_items(&p->d_items, &p->items_alloc, items)) return rc;
    if (int rc = ensure_scratch(p, 1)) return rc;
    hipStream_t sm = lib_stream();
    PIC_HIP(hipMemsetAsync(p->d_scratch, 0, sizeof(int32_t), sm));
    if (int rc = launch_export<T>((CaseDev<T>*)p->d_case, p->d_state, p->d_items, (int)items.size(), 4096,
                                  p->d_scratch, (float2*)dst, dst ? cap : 0, sx, sy, sm))
        return rc;
    int32_t n = 0;
    PIC_HIP(hipMemcpyAsync(&n, p->d_scratch, sizeof(n), hipMemcpyDeviceToHost, sm));
    PIC_HIP(hipStreamSynchronize(sm));
    *nout = n;
    if (dst && n > cap) {
        set_error("pic_particles_export_positions: output capacity too small");
        return PIC_ERR_CAPACITY;
    }
    return 0;
}

extern "C" int pic_particles_export_positions(pic_particles* p, int species, float* dst_device, int64_t cap,
                                              int units, int64_t* n) {
    PIC_REQUIRE(p && n && species >= 0 && species < (int)p->sp.size(), "bad species or NULL argument");
    PIC_REQUIRE(units == 0 || units == 1, "units: 0 cells, 1 domain");
    PIC_REQUIRE(cap >= 0, "negative capacity");
    double sx = 1.0, sy = 1.0;
    if (units == 1) {
        PIC_REQUIRE(p->attached, "domain units need the particles attached to a simulation (its grid spacing)");
        sx = p->attached->g->dx;
        sy = p->attached->g->dy;
    }
    return p->prec == 4 ? export_t<float>(p, species, dst_device, cap, sx, sy, n)
                        : export_t<double>(p, species, dst_device, cap, sx, sy, n);
}

// ---------------------------------------------------------------------------
// Position export that allocates nothing: its work-item table and counter live
// in `scratch`, a file-scope device array (one copy per GPU, held until the
// process ends), instead of buffers the library grows with hipMalloc.
// ---------------------------------------------------------------------------
#define PIC_STATIC_EXPORT_ITEMS 65536
#define PIC_STATIC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[256 + PIC_STATIC_EXPORT_ITEMS * sizeof(WorkItem)];

// the device address of this GPU's copy of scratch, looked up once per device
// with that device current (the array's name is not a device address in host code)
static int scratch_ptr(unsigned char** out) {
    static std::mutex mu;
    static unsigned char* ptrs[PIC_STATIC_MAX_DEVICES] = {};
    int dev = 0;
    PIC_HIP(hipGetDevice(&dev));
    PIC_REQUIRE(dev >= 0 && dev < PIC_STATIC_MAX_DEVICES, "device index beyond the static export's table");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        PIC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptrs[dev] = (unsigned char*)p;
    }
    *out = ptrs[dev];
    return 0;
}

template <typename T>
static int export_static_t(pic_particles* p, int s, float* dst, int64_t cap, double sx, double sy, int64_t* nout) {
    const SpeciesStore& st = p->sp[s];
    const int chunk = 4096;
    const int chunks = std::max(1, (st.h_cap[st.nb] + chunk - 1) / chunk);
    const int n_items = st.nb + chunks;
    if (n_items > PIC_STATIC_EXPORT_ITEMS) {
        set_error("pic_particles_export_positions_static: the species needs " + std::to_string(n_items) +
                  " work items (bins + spill chunks), more than the static maximum of " +
                  std::to_string(PIC_STATIC_EXPORT_ITEMS) + "; use pic_particles_export_positions");
        return PIC_ERR_CAPACITY;
    }
    unsigned char* base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    int32_t* counter = (int32_t*)base;
    WorkItem* items = (WorkItem*)(base + 256);
    hipStream_t sm = lib_stream();
    // the items kernel clears the counter the previous call left behind
    if (int rc = launch_species_items(items, counter, s, st.nb, chunks, sm)) return rc;
    if (int rc = launch_export<T>((CaseDev<T>*)p->d_case, p->d_state, items, n_items, chunk, counter, (float2*)dst,
                                  dst ? cap : 0, sx, sy, sm))
        return rc;
    int32_t n = 0;
    PIC_HIP(hipMemcpyAsync(&n, counter, sizeof(n), hipMemcpyDeviceToHost, sm));
    PIC_HIP(hipStreamSynchronize(sm));
    *nout = n;
    if (dst && n > cap) {
        set_error("pic_particles_export_positions_static: output capacity too small");
        return PIC_ERR_CAPACITY;
    }
    return 0;
}

extern "C" int pic_particles_export_positions_static(pic_particles* p, int species, float* dst_device, int64_t cap,
                                                     int units, int64_t* n) {
    PIC_REQUIRE(p && n && species >= 0 && species < (int)p->sp.size(), "bad species or NULL argument");
    PIC_REQUIRE(units == 0 || units == 1, "units: 0 cells, 1 domain");
    PIC_REQUIRE(cap >= 0, "negative capacity");
    double sx = 1.0, sy = 1.0;
    if (units == 1) {
        PIC_REQUIRE(p->attached, "domain units need the particles attached to a simulation (its grid spacing)");
        sx = p->attached->g->dx;
        sy = p->attached->g->dy;
    }
    return p->prec == 4 ? export_static_t<float>(p, species, dst_device, cap, sx, sy, n)
                        : export_static_t<double>(p, species, dst_device, cap, sx, sy, n);
}

template <typename T>
static int init_t(pic_particles* p, int s, int64_t n, int nx, int ny, double vth, uint32_t seed, uint32_t case_id,
                  uint32_t id0) {
    SpeciesStore& st = p->sp[s];
    const int par = p->h_cur, b = st.nb;
    int32_t cnt = 0;
    if (int rc = read_count(p, s, par, b, &cnt)) return rc;
    // the kernel appends at the raw device count, so that count must fit as well
    if (cnt < 0 || (int64_t)cnt + n > st.h_cap[b]) {
        set_error("pic_particles_init_maxwellian: not enough free slots (capacity per species exceeded)");
        return PIC_ERR_CAPACITY;
    }
    const size_t off = (size_t)st.h_start[b];
    return launch_init<T>((T*)st.x[par] + off, (T*)st.y[par] + off, (T*)st.vx[par] + off, (T*)st.vy[par] + off,
                          st.id[par] + off, st.d_count[par] + b, (int)n, nx, ny, vth, seed, case_id, (uint32_t)s,
                          id0, lib_stream());
}

template <typename T>
static int init_uniform_t(pic_particles* p, int s, int64_t n, double x0, double x1, double y0, double y1, double v0,
                          double v1, uint32_t seed, uint32_t case_id, uint32_t id0) {
    SpeciesStore& st = p->sp[s];
    const int par = p->h_cur, b = st.nb;
    int32_t cnt = 0;
    if (int rc = read_count(p, s, par, b, &cnt)) return rc;
    if (cnt < 0 || (int64_t)cnt + n > st.h_cap[b]) {
        set_error("pic_particles_init_uniform: not enough free slots (capacity per species exceeded)");
        return PIC_ERR_CAPACITY;
    }
    const size_t off = (size_t)st.h_start[b];
    return launch_init_uniform<T>((T*)st.x[par] + off, (T*)st.y[par] + off, (T*)st.vx[par] + off,
                                  (T*)st.vy[par] + off, st.id[par] + off, st.d_count[par] + b, (int)n, x0, x1, y0, y1,
                                  v0, v1, seed, case_id, (uint32_t)s, id0, lib_stream());
}

extern "C" int pic_particles_init_uniform(pic_particles* p, int species, int64_t n, double x_lo, double x_hi,
                                          double y_lo, double y_hi, double v_lo, double v_hi, uint32_t seed,
                                          uint32_t case_id, uint32_t id0) {
    PIC_REQUIRE(p && species >= 0 && species < (int)p->sp.size(), "bad species");
    PIC_REQUIRE(n >= 0 && n < INT32_MAX, "bad particle count");
    PIC_REQUIRE(x_hi >= x_lo && y_hi >= y_lo && v_hi >= v_lo, "empty range");
    if (n == 0) return 0;
    if (p->attached) {
        if (int rc = sim_sync_impl(p->attached)) return rc;
    }
    int rc = p->prec == 4 ? init_uniform_t<float>(p, species, n, x_lo, x_hi, y_lo, y_hi, v_lo, v_hi, seed, case_id, id0)
                          : init_uniform_t<double>(p, species, n, x_lo, x_hi, y_lo, y_hi, v_lo, v_hi, seed, case_id,
                                                   id0);
    if (rc) return rc;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    return 0;
}

extern "C" int pic_particles_init_maxwellian(pic_particles* p, int species, int64_t n, int nx, int ny, double vth,
                                             uint32_t seed, uint32_t case_id, uint32_t id0) {
    PIC_REQUIRE(p && species >= 0 && species < (int)p->sp.size(), "bad species");
    PIC_REQUIRE(n >= 0 && n < INT32_MAX, "bad particle count");
    PIC_REQUIRE(nx >= 2 && ny >= 2, "grid must have at least 2 nodes per axis");
    if (n == 0) return 0;
    if (p->attached) {
        if (int rc = sim_sync_impl(p->attached)) return rc;
    }
    int rc = p->prec == 4 ? init_t<float>(p, species, n, nx, ny, vth, seed, case_id, id0)
                          : init_t<double>(p, species, n, nx, ny, vth, seed, case_id, id0);
    if (rc) return rc;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    return 0;
}

extern "C" int pic_particles_remove_collided(pic_particles* p) {
    PIC_REQUIRE(p, "particles is NULL");
    return 0;
}

// ============================================================================
// Grid2D
// ============================================================================
extern "C" int pic_grid_create(int m, int n, double X, double Y, pic_grid** out) {
    PIC_REQUIRE(out, "out is NULL");
    PIC_REQUIRE(m >= 3 && n >= 3, "grid needs at least 3 nodes per axis");
    PIC_REQUIRE(X > 0 && Y > 0, "domain size must be positive");
    auto* g = new pic_grid();
    g->m = m;
    g->n = n;
    g->X = X;
    g->Y = Y;
    g->dx = X / (m - 1);
    g->dy = Y / (n - 1);
    const size_t nn = (size_t)m * n;
    if (hipMalloc(&g->rho, nn * 8) != hipSuccess || hipMalloc(&g->phi, nn * 8) != hipSuccess ||
        hipMalloc(&g->E, nn * 16) != hipSuccess || hipMalloc(&g->rhs, nn * 8) != hipSuccess) {
        set_error("pic_grid_create: device allocation failed");
        pic_grid_destroy(g);
        return PIC_ERR_HIP;
    }
    PIC_HIP(hipMemset(g->rho, 0, nn * 8));
    PIC_HIP(hipMemset(g->phi, 0, nn * 8));
    PIC_HIP(hipMemset(g->E, 0, nn * 16));
    PIC_HIP(hipMemset(g->rhs, 0, nn * 8));
    *out = g;
    return 0;
}

extern "C" int pic_grid_destroy(pic_grid* g) {
    if (!g) return 0;
    (void)hipStreamSynchronize(lib_stream());
    (void)hipFree(g->rho);
    (void)hipFree(g->phi);
    (void)hipFree(g->E);
    (void)hipFree(g->rhs);
    delete g;
    return 0;
}

extern "C" int pic_grid_info(const pic_grid* g, int* m, int* n, double* dx, double* dy) {
    PIC_REQUIRE(g, "grid is NULL");
    if (m) *m = g->m;
    if (n) *n = g->n;
    if (dx) *dx = g->dx;
    if (dy) *dy = g->dy;
    return 0;
}

extern "C" int pic_grid_device_ptrs(pic_grid* g, double** rho, double** phi, double** E) {
    PIC_REQUIRE(g, "grid is NULL");
    if (rho) *rho = g->rho;
    if (phi) *phi = g->phi;
    if (E) *E = g->E;
    return 0;
}

extern "C" int pic_grid_download(pic_grid* g, int field, double* out) {
    PIC_REQUIRE(g && out, "NULL argument");
    const size_t nn = (size_t)g->m * g->n;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    switch (field) {
        case PIC_FIELD_RHO: PIC_HIP(hipMemcpy(out, g->rho, nn * 8, hipMemcpyDeviceToHost)); break;
        case PIC_FIELD_PHI: PIC_HIP(hipMemcpy(out, g->phi, nn * 8, hipMemcpyDeviceToHost)); break;
        case PIC_FIELD_RHS: PIC_HIP(hipMemcpy(out, g->rhs, nn * 8, hipMemcpyDeviceToHost)); break;
        case PIC_FIELD_EX:
        case PIC_FIELD_EY: {
            std::vector<double> e(2 * nn);
            PIC_HIP(hipMemcpy(e.data(), g->E, nn * 16, hipMemcpyDeviceToHost));
            const int o = field == PIC_FIELD_EX ? 0 : 1;
            for (size_t i = 0; i < nn; ++i) out[i] = e[2 * i + o];
            break;
        }
        default: set_error("bad field"); return PIC_ERR_ARG;
    }
    return 0;
}

extern "C" int pic_grid_upload(pic_grid* g, int field, const double* in) {
    PIC_REQUIRE(g && in, "NULL argument");
    const size_t nn = (size_t)g->m * g->n;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    switch (field) {
        case PIC_FIELD_RHO: PIC_HIP(hipMemcpy(g->rho, in, nn * 8, hipMemcpyHostToDevice)); break;
        case PIC_FIELD_PHI: PIC_HIP(hipMemcpy(g->phi, in, nn * 8, hipMemcpyHostToDevice)); break;
        case PIC_FIELD_RHS: PIC_HIP(hipMemcpy(g->rhs, in, nn * 8, hipMemcpyHostToDevice)); break;
        default: set_error("field not uploadable"); return PIC_ERR_ARG;
    }
    return 0;
}

extern "C" int pic_grid_compute_E(pic_grid* g) {
    PIC_REQUIRE(g, "grid is NULL");
    if (int rc = launch_efield(g->phi, (double2*)g->E, g->m, g->n, g->dx, g->dy, lib_stream())) return rc;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    return 0;
}

// ============================================================================
// Boundaries2D
// ============================================================================
extern "C" int pic_walls_create(double width, double height, int nx, int ny, pic_walls** out) {
    PIC_REQUIRE(out, "out is NULL");
    auto* w = new pic_walls();
    w->width = width;
    w->height = height;
    w->nx = nx;
    w->ny = ny;
    *out = w;
    return 0;
}

extern "C" int pic_walls_add(pic_walls* w, double x1, double y1, double x2, double y2) {
    PIC_REQUIRE(w, "walls is NULL");
    w->walls.insert(w->walls.end(), {x1, y1, x2, y2});
    w->compiled = false;
    return 0;
}

extern "C" int pic_walls_add_box(pic_walls* w) {
    PIC_REQUIRE(w, "walls is NULL");
    w->box = true;
    return 0;
}

extern "C" int pic_walls_compile(pic_walls* w) {
    PIC_REQUIRE(w, "walls is NULL");
    if (w->d_walls) PIC_HIP(hipFree(w->d_walls));
    w->d_walls = nullptr;
    if (!w->walls.empty()) {
        PIC_HIP(hipMalloc(&w->d_walls, w->walls.size() * sizeof(double)));
        PIC_HIP(hipMemcpy(w->d_walls, w->walls.data(), w->walls.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    w->compiled = true;
    return 0;
}

extern "C" int pic_walls_count(const pic_walls* w, int* n) {
    PIC_REQUIRE(w && n, "NULL argument");
    *n = (int)(w->walls.size() / 4) + (w->box ? 4 : 0);
    return 0;
}

extern "C" int pic_walls_destroy(pic_walls* w) {
    if (!w) return 0;
    (void)hipFree(w->d_walls);
    delete w;
    return 0;
}

// ============================================================================
// multigrid
// ============================================================================
extern "C" int pic_mg_default_params(int preset, pic_mg_params* out) {
    PIC_REQUIRE(out, "out is NULL");
    switch (preset) {
        case PIC_MG_PRESET_REFERENCE: *out = pic_mg_params{0, 3, 3, 20, 1.25, 1e-5, 50, PIC_MG_FP64, 1, 0}; break;
        case PIC_MG_PRESET_STEP: *out = pic_mg_params{0, 2, 2, 1, 1.0, 1e-8, 100, PIC_MG_MIXED, 2, 2}; break;
        case PIC_MG_PRESET_POISSON: *out = pic_mg_params{0, 2, 2, 1, 1.2, 1e-10, 100, PIC_MG_MIXED, 2, 2}; break;
        case PIC_MG_PRESET_STEP_SOR: *out = pic_mg_params{0, 2, 2, 1, 1.2, 1e-8, 100, PIC_MG_MIXED, 2, 2}; break;
        default: set_error("unknown multigrid preset"); return PIC_ERR_ARG;
    }
    return 0;
}

static MGParams to_mg(const pic_mg_params& p) {
    MGParams m{p.levels, p.pre, p.post, p.coarse_sweeps, p.omega, p.tol, p.max_cycles, p.mode};
    m.inner = p.inner;
    m.coarse_direct = p.coarse_direct;
    return m;
}

extern "C" int pic_mg_create(int n, const pic_mg_params* params, pic_mg** out) {
    PIC_REQUIRE(out && params, "NULL argument");
    PIC_REQUIRE(params->mode == PIC_MG_FP64 || params->mode == PIC_MG_MIXED, "bad multigrid mode");
    PIC_REQUIRE(params->inner >= 0 && params->inner <= 4, "inner cycles must be 0..4");
    PIC_REQUIRE(params->coarse_direct >= 0 && params->coarse_direct <= 2, "coarse_direct must be 0, 1 or 2");
    PIC_REQUIRE(lib_stream() != nullptr, "could not create a HIP stream");
    auto* m = new pic_mg();
    m->solver.reset(new MGSolver(n, to_mg(*params), lib_stream()));
    if (int rc = m->solver->init()) {
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

extern "C" int pic_mg_solve(pic_mg* mg, double* phi_d, const double* rhs_d, double h, pic_mg_stats* st) {
    PIC_REQUIRE(mg && phi_d && rhs_d, "NULL argument");
    PIC_REQUIRE(h > 0, "h must be positive");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    auto t0 = std::chrono::steady_clock::now();
    MGStats s{};
    if (int rc = mg->solver->solve(phi_d, rhs_d, h, &s)) return rc;
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->cycles = s.cycles;
        st->rel_residual = s.rel_residual;
        st->rhs_norm = s.rhs_norm;
        st->ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->launches = mg->solver->last_launches();
    }
    return 0;
}

extern "C" int pic_mg_destroy(pic_mg* mg) {
    if (!mg) return 0;
    (void)hipStreamSynchronize(lib_stream());
    delete mg;
    return 0;
}

extern "C" int pic_solve_multigrid(double* phi_d, double* rho_d, int nx, int ny, int levels, double h, double omega) {
    PIC_REQUIRE(phi_d && rho_d, "NULL argument");
    PIC_REQUIRE(nx == ny, "solve_multigrid needs a square grid (Multigrid.cu:84-88)");
    // process-wide solver contexts, like the reference's static pool
    // (Multigrid.cu:187); kept until the process ends
    static std::mutex mu;
    static auto* cache = new std::map<std::tuple<int, int, double>, MGSolver*>();
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(nx, levels, omega);
    auto it = cache->find(key);
    if (it == cache->end()) {
        MGParams p{levels, 3, 3, 20, omega, 1e-5, 50, PIC_MG_FP64};
        auto* s = new MGSolver(nx, p, lib_stream());
        if (int rc = s->init()) {
            delete s;
            return rc;
        }
        it = cache->emplace(key, s).first;
    }
    MGStats st{};
    return it->second->solve(phi_d, rho_d, h, &st);
}

// ============================================================================
// cross sections and MCC
// ============================================================================
extern "C" int pic_read_cross_section(const char* path, int cap, double* energies, double* sigmas, int* n) {
    PIC_REQUIRE(path && n, "NULL argument");
    std::ifstream f(path);
    if (!f.is_open()) {
        set_error(std::string("cannot open cross-section file: ") + path);
        return PIC_ERR_IO;
    }
    std::string line;
    int k = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double e, s;
        if (iss >> e >> s) {
            if (energies && sigmas) {
                if (k >= cap) {
                    set_error("pic_read_cross_section: output capacity too small");
                    return PIC_ERR_CAPACITY;
                }
                energies[k] = e;
                sigmas[k] = s;
            }
            ++k;
        }
    }
    *n = k;
    return 0;
}

extern "C" int pic_xs_create(double gas_density, double gas_temperature_eV, pic_xs** out) {
    PIC_REQUIRE(out, "out is NULL");
    PIC_REQUIRE(gas_density >= 0 && gas_temperature_eV >= 0, "bad gas parameters");
    auto* x = new pic_xs();
    x->density = gas_density;
    x->t_gas = gas_temperature_eV;
    *out = x;
    return 0;
}

extern "C" int pic_xs_add_process(pic_xs* xs, int species, int type, double threshold_eV, int n,
                                  const double* energy_eV, const double* sigma_m2, int product_species) {
    PIC_REQUIRE(xs && energy_eV && sigma_m2 && n >= 2, "need a table of at least 2 points");
    PIC_REQUIRE(type >= 0 && type <= 3, "bad process type");
    for (int i = 1; i < n; ++i) PIC_REQUIRE(energy_eV[i] > energy_eV[i - 1], "energies must increase");
    pic_xs::Proc p{species, type, product_species, threshold_eV, 0.0, {}, {}};
    p.e.assign(energy_eV, energy_eV + n);
    p.s.assign(sigma_m2, sigma_m2 + n);
    xs->procs.push_back(p);
    return 0;
}

extern "C" int pic_xs_add_analytic(pic_xs* xs, int species, int type, double threshold_eV, double peak_m2,
                                   int product_species) {
    PIC_REQUIRE(xs, "xs is NULL");
    PIC_REQUIRE(type >= 0 && type <= 3, "bad process type");
    PIC_REQUIRE(!((type == PIC_PROC_EXCITATION || type == PIC_PROC_IONIZATION) && threshold_eV <= 0),
                "inelastic processes need a positive threshold");
    xs->procs.push_back(pic_xs::Proc{species, type, product_species, threshold_eV, peak_m2, {}, {}});
    return 0;
}

extern "C" int pic_xs_set_table(pic_xs* xs, int points, double emax_eV) {
    PIC_REQUIRE(xs && points >= 2 && emax_eV > 0, "bad table");
    xs->points = points;
    xs->emax = emax_eV;
    return 0;
}

extern "C" int pic_xs_destroy(pic_xs* xs) {
    delete xs;
    return 0;
}

// sigma of a process at energy E (eV): the analytic shape, or the clamped
// linear interpolation of interpolate_cross_sections (kernels.cu:218-263)
static double proc_sigma(const pic_xs::Proc& p, double E) {
    if (p.e.empty()) {
        if (p.type == PIC_PROC_ELASTIC || p.type == PIC_PROC_CHARGE_EXCHANGE) return p.peak;
        const double u = E / p.thr;
        if (u <= 1.0) return 0.0;
        return p.peak * 2.718281828459045 * std::log(u) / u;
    }
    const auto& e = p.e;
    const auto& s = p.s;
    if (E <= e.front()) return s.front();
    if (E >= e.back()) return s.back();
    int lo = 0, hi = (int)e.size() - 2;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (E < e[mid]) hi = mid - 1;
        else if (E >= e[mid + 1]) lo = mid + 1;
        else {
            const double t = (E - e[mid]) / (e[mid + 1] - e[mid]);
            return s[mid] + t * (s[mid + 1] - s[mid]);
        }
    }
    return s.back();
}

struct MccBuild {
    pic_mcc_info info{};
    std::vector<double> tab;   // n_proc * K of n_gas * sigma
};

// null-collision constants of one species: uniform table, and the majorant of
// each energy band (the exact supremum of the interpolated collision
// frequency over the band, x1.001), p_null = 1-exp(-nu_max dt)
static int build_mcc(const pic_xs* xs, const std::vector<SpeciesStore>& sp, int s, double dt, MccBuild& out) {
    std::vector<const pic_xs::Proc*> ps;
    for (auto& p : xs->procs)
        if (p.species == s) ps.push_back(&p);
    PIC_REQUIRE(ps.size() <= PIC_MAX_PROC, "at most 4 processes per species");
    const int K = xs->points;
    const double de = xs->emax / (K - 1);
    pic_mcc_info& I = out.info;
    I = pic_mcc_info{};
    I.n_proc = (int)ps.size();
    I.K = K;
    I.inv_de = 1.0 / de;
    I.efac = 0.5 * sp[s].m / kElementaryCharge;
    I.product_species = -1;
    double partner_mass = sp[s].m;
    for (int q = 0; q < I.n_proc; ++q) {
        I.type[q] = ps[q]->type;
        I.threshold[q] = (ps[q]->type == PIC_PROC_EXCITATION || ps[q]->type == PIC_PROC_IONIZATION) ? ps[q]->thr : 0.0;
        if (ps[q]->type == PIC_PROC_IONIZATION) {
            PIC_REQUIRE(ps[q]->product >= 0 && ps[q]->product < (int)sp.size() && ps[q]->product != s,
                        "ionization needs a product species different from the projectile");
            I.product_species = ps[q]->product;
            partner_mass = sp[ps[q]->product].m;
        }
    }
    I.gas_vth = std::sqrt(xs->t_gas * kElementaryCharge / partner_mass);
    out.tab.assign((size_t)I.n_proc * K, 0.0);
    for (int q = 0; q < I.n_proc; ++q)
        for (int j = 0; j < K; ++j) out.tab[(size_t)q * K + j] = xs->density * proc_sigma(*ps[q], j * de);
    const double m = sp[s].m;
    // collision frequency of the interpolated tables at energy E
    auto nu_at = [&](double E) {
        double x = E / de;
        int j = (int)std::floor(x);
        double t = x - j;
        if (j >= K - 1) {
            j = K - 2;
            t = 1.0;
        }
        const double v = std::sqrt(2.0 * E * kElementaryCharge / m);
        double nu = 0.0;
        for (int q = 0; q < I.n_proc; ++q) {
            if (E < I.threshold[q]) continue;
            const double a = out.tab[(size_t)q * K + j], b = out.tab[(size_t)q * K + j + 1];
            nu += (a + t * (b - a)) * v;
        }
        return nu;
    };
    // Supremum of nu over [lo, hi]. Cut at the table nodes and thresholds in
    // between; on each piece the active processes are fixed and their summed
    // cross section is linear, A + B E, so nu = (A + B E) sqrt(2 E e / m), whose
    // derivative vanishes only at E = -A / (3B) (a maximum when A > 0 > B).
    auto band_max = [&](double lo, double hi) {
        std::vector<double> cut = {lo, hi};
        for (int j = (int)std::ceil(lo / de); j * de < hi && j <= K - 1; ++j)
            if (j * de > lo) cut.push_back(j * de);
        for (int q = 0; q < I.n_proc; ++q)
            if (I.threshold[q] > lo && I.threshold[q] < hi) cut.push_back(I.threshold[q]);
        std::sort(cut.begin(), cut.end());
        double mx = 0.0;
        for (double E : cut) mx = std::max(mx, nu_at(E));
        for (size_t c = 0; c + 1 < cut.size(); ++c) {
            const double a = cut[c], b = cut[c + 1], mid = 0.5 * (a + b);
            if (!(b > a) || mid / de >= K - 1) continue;   // beyond the table nu grows with E: ends suffice
            const int j = (int)std::floor(mid / de);
            double A = 0.0, B = 0.0;
            for (int q = 0; q < I.n_proc; ++q) {
                if (mid < I.threshold[q]) continue;
                const double s0 = out.tab[(size_t)q * K + j], s1 = out.tab[(size_t)q * K + j + 1];
                A += s0 - j * (s1 - s0);
                B += (s1 - s0) / de;
            }
            if (A > 0.0 && B < 0.0) {
                const double Es = -A / (3.0 * B);
                if (Es > a && Es < b) mx = std::max(mx, nu_at(Es));
            }
        }
        return mx;
    };
    const double emax = xs->emax;
    I.emax = emax;
    double numax = 0.0;
    for (int b = 0; b < 16; ++b) {
        const double lo = b == 0 ? 0.0 : std::ldexp(emax, b - 15);
        const double hi = b == 15 ? emax : std::ldexp(emax, b - 14);
        // x1.001 covers the rounding of the kernel's interpolation
        I.band_nu_max[b] = band_max(lo, hi) * 1.001;
        I.band_p_null[b] = -std::expm1(-I.band_nu_max[b] * dt);
        numax = std::max(numax, I.band_nu_max[b]);
    }
    I.nu_max = numax;
    I.p_null = -std::expm1(-I.nu_max * dt);
    return 0;
}

extern "C" int pic_xs_build(const pic_xs* xs, int species, int n_species, const double* masses, double dt,
                            pic_mcc_info* out, int cap, double* table) {
    PIC_REQUIRE(xs && masses && out, "NULL argument");
    PIC_REQUIRE(n_species >= 1 && n_species <= PIC_MAX_SPECIES && species >= 0 && species < n_species,
                "bad species");
    PIC_REQUIRE(dt > 0, "dt must be positive");
    std::vector<SpeciesStore> sp(n_species);
    for (int i = 0; i < n_species; ++i) {
        PIC_REQUIRE(masses[i] > 0, "masses must be positive");
        sp[i].m = masses[i];
    }
    MccBuild mb;
    if (int rc = build_mcc(xs, sp, species, dt, mb)) return rc;
    *out = mb.info;
    if (table) {
        PIC_REQUIRE(cap >= (int)mb.tab.size(), "table capacity too small");
        std::copy(mb.tab.begin(), mb.tab.end(), table);
    }
    return 0;
}

// ============================================================================
// step driver
// ============================================================================
size_t cur_lds(pic_sim* S) {
    return S->p->prec == 4 ? push_lds_bytes<float>(S->bin_w, S->bin_h, S->halo)
                           : push_lds_bytes<double>(S->bin_w, S->bin_h, S->halo);
}

static int rebuild_push_items(pic_sim* S) {
    std::vector<WorkItem> items;
    for (int s = 0; s < (int)S->p->sp.size(); ++s) items_for_species(S->p, s, S->spill_chunk, items);
    static_assert(sizeof(WorkItem) == 16, "work item layout");
    int alloc = S->n_items > 0 ? S->n_items : 0;
    // the sim keeps its own table sized exactly
    if (S->d_items) PIC_HIP(hipFree(S->d_items));
    S->d_items = nullptr;
    alloc = 0;
    if (int rc = upload_items(&S->d_items, &alloc, items)) return rc;
    S->n_items = (int)items.size();
    return 0;
}

// Per-bin candidate walls for the bin tile in C (cell units). A wall is listed
// for a bin when it meets the bin's rectangle grown by S->wall_margin + 1 cells
// (Liang-Barsky clip of the segment); the push kernel uses one bin's list for a
// path whose two ends lie in that bin's rectangle grown by S->wall_margin
// cells, so a wall left out is at least one cell away from every such path.

static bool segment_meets_rect(const double* w, double x0, double y0, double x1, double y1) {
    const double ax = w[0], ay = w[1], dx = w[2] - w[0], dy = w[3] - w[1];
    const double p[4] = {-dx, dx, -dy, dy};
    const double q[4] = {ax - x0, x1 - ax, ay - y0, y1 - ay};
    double t0 = 0.0, t1 = 1.0;
    for (int i = 0; i < 4; ++i) {
        if (p[i] == 0.0) {
            if (q[i] < 0.0) return false;
            continue;
        }
        const double t = q[i] / p[i];
        if (p[i] < 0.0) t0 = std::max(t0, t);
        else t1 = std::min(t1, t);
        if (t0 > t1) return false;
    }
    return true;
}

template <typename T>
static int build_wall_lists(pic_sim* S, CaseDev<T>& C) {
    (void)hipFree(S->d_wall_lists);
    S->d_wall_lists = nullptr;
    S->wall_entries = 0;
    S->wall_max_bin = 0;
    C.wall_off = nullptr;
    C.wall_idx = nullptr;
    C.wall_margin = S->wall_margin;
    if (!S->wall_lists || !S->w || S->w->walls.empty()) return 0;
    const int nw = (int)(S->w->walls.size() / 4);
    const int nb = C.nbx * C.nby;
    std::vector<double> wc(4 * (size_t)nw);
    for (int j = 0; j < nw; ++j) {
        wc[4 * j] = S->w->walls[4 * j] / S->g->dx;
        wc[4 * j + 1] = S->w->walls[4 * j + 1] / S->g->dy;
        wc[4 * j + 2] = S->w->walls[4 * j + 2] / S->g->dx;
        wc[4 * j + 3] = S->w->walls[4 * j + 3] / S->g->dy;
    }
    const int G = S->wall_margin + 1;
    std::vector<std::vector<int>> per(nb);
    for (int j = 0; j < nw; ++j) {
        const double* w = &wc[4 * j];
        const double xmn = std::min(w[0], w[2]), xmx = std::max(w[0], w[2]);
        const double ymn = std::min(w[1], w[3]), ymx = std::max(w[1], w[3]);
        const double lo_x = std::max(-1.0, std::floor((xmn - G) / C.bw) - 1);
        const double hi_x = std::min((double)C.nbx, std::floor((xmx + G) / C.bw) + 1);
        const double lo_y = std::max(-1.0, std::floor((ymn - G) / C.bh) - 1);
        const double hi_y = std::min((double)C.nby, std::floor((ymx + G) / C.bh) + 1);
        for (int by = std::max(0, (int)lo_y); by <= std::min(C.nby - 1, (int)hi_y); ++by)
            for (int bx = std::max(0, (int)lo_x); bx <= std::min(C.nbx - 1, (int)hi_x); ++bx) {
                const double rx0 = bx * C.bw - G, ry0 = by * C.bh - G;
                if (segment_meets_rect(w, rx0, ry0, rx0 + C.bw + 2 * G, ry0 + C.bh + 2 * G))
                    per[by * C.nbx + bx].push_back(j);
            }
    }
    std::vector<int> lists(nb + 1);
    for (int b = 0; b < nb; ++b) {
        lists[b + 1] = lists[b] + (int)per[b].size();
        S->wall_max_bin = std::max(S->wall_max_bin, (int)per[b].size());
    }
    for (int b = 0; b < nb; ++b) lists.insert(lists.end(), per[b].begin(), per[b].end());
    S->wall_entries = lists[nb];
    PIC_HIP(hipMalloc(&S->d_wall_lists, lists.size() * sizeof(int)));
    PIC_HIP(hipMemcpy(S->d_wall_lists, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice));
    C.wall_off = S->d_wall_lists;
    C.wall_idx = S->d_wall_lists + nb + 1;
    return 0;
}

// log2 of a power of two, -1 otherwise
static int log2_exact(int v) {
    if (v <= 0 || (v & (v - 1))) return -1;
    int k = 0;
    while ((1 << k) < v) ++k;
    return k;
}

// dynamic LDS a workgroup may have on the current device
static size_t max_lds_per_block() {
    static size_t v = 0;
    if (!v) {
        int dev = 0, bytes = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && bytes > 0)
            v = (size_t)bytes;
        else
            v = 65536;
    }
    return v;
}

template <typename T>
static int rebin_t(pic_sim* S, int bw, int bh) {
    pic_particles* p = S->p;
    CaseDev<T>& C = hcase<T>(p);
    if (push_lds_bytes<T>(bw, bh, S->halo) > max_lds_per_block()) {
        set_error("bin tile " + std::to_string(bw) + "x" + std::to_string(bh) + " needs " +
                  std::to_string(push_lds_bytes<T>(bw, bh, S->halo)) + " B of LDS per workgroup, more than the " +
                  std::to_string(max_lds_per_block()) + " B the device allows");
        return PIC_ERR_ARG;
    }
    hipStream_t sm = lib_stream();
    const int in = p->h_cur, out = in ^ 1;
    const int nbx = div_up(C.nx - 1, bw), nby = div_up(C.ny - 1, bh);
    const int nb_new = nbx * nby;
    const int ns = (int)p->sp.size();
    std::vector<WorkItem> items;
    std::vector<int> off(ns + 1, 0);
    // spill-region chunks of the sort are larger than the push's: a fresh store
    // is all spill, and the counting kernels pay per workgroup for the bins
    const int sort_chunk = std::max(S->spill_chunk, 32768);
    for (int s = 0; s < ns; ++s) {
        off[s] = (int)items.size();
        items_for_species(p, s, sort_chunk, items);
    }
    off[ns] = (int)items.size();
    if (int rc = upload_items(&p->d_items, &p->items_alloc, items)) return rc;
    if (int rc = ensure_scratch(p, nb_new + 1)) return rc;
    std::vector<std::vector<int32_t>> new_start(ns), new_cap(ns);
    for (int s = 0; s < ns; ++s) {
        SpeciesStore& st = p->sp[s];
        PIC_HIP(hipMemsetAsync(p->d_scratch, 0, (size_t)(nb_new + 1) * sizeof(int32_t), sm));
        if (int rc = launch_hist<T>((CaseDev<T>*)p->d_case, p->d_state, p->d_items + off[s], off[s + 1] - off[s],
                                    sort_chunk, bw, bh, nbx, nb_new, p->d_scratch, sm))
            return rc;
        std::vector<int32_t> hist(nb_new + 1);
        PIC_HIP(hipMemcpyAsync(hist.data(), p->d_scratch, (size_t)(nb_new + 1) * sizeof(int32_t),
                               hipMemcpyDeviceToHost, sm));
        PIC_HIP(hipStreamSynchronize(sm));
        int64_t total = 0;
        for (int b = 0; b < nb_new; ++b) total += hist[b];
        const int64_t C_ = st.capacity;
        int64_t spill_min = std::min<int64_t>(std::max<int64_t>(4096, total / 64), C_ / 8);
        if (total + spill_min > C_) {
            set_error("re-binning: particle count exceeds the species capacity");
            return PIC_ERR_CAPACITY;
        }
        // every bin keeps at least its population in the front part (3/4 of
        // the bin); the slack targets 1.5x the population + 64, 16-aligned
        auto cmin = [](int64_t h) { return (4 * h + 2) / 3 + 4; };
        auto extra = [&](int64_t h) { return std::max<int64_t>(0, h + h / 2 + 64 - cmin(h)); };
        int64_t need = 0, want = 0;
        for (int b = 0; b < nb_new; ++b) {
            need += cmin(hist[b]);
            want += extra(hist[b]) + 15;
        }
        if (need + spill_min > C_) {
            set_error("re-binning: particle count exceeds the species capacity");
            return PIC_ERR_CAPACITY;
        }
        const int64_t avail = C_ - spill_min - need;
        const bool full = want <= avail;
        const double scale = full ? 1.0 : (double)avail / (double)std::max<int64_t>(want, 1);
        auto& ns_ = new_start[s];
        auto& nc_ = new_cap[s];
        ns_.assign(nb_new + 1, 0);
        nc_.assign(nb_new + 1, 0);
        int64_t pos = 0;
        for (int b = 0; b < nb_new; ++b) {
            int64_t c;
            if (full) c = (cmin(hist[b]) + extra(hist[b]) + 15) / 16 * 16;
            else c = cmin(hist[b]) + (int64_t)std::floor(extra(hist[b]) * scale);
            ns_[b] = (int32_t)pos;
            nc_[b] = (int32_t)c;
            pos += c;
        }
        ns_[nb_new] = (int32_t)pos;
        nc_[nb_new] = (int32_t)(C_ - pos);
        // parity `out` gets fresh arrays of the new size (parity `in` is still in use)
        if (int rc = alloc_layout_parity(st, out, nb_new + 1)) return rc;
        PIC_HIP(hipMemcpyAsync(st.d_start[out], ns_.data(), (nb_new + 1) * sizeof(int32_t), hipMemcpyHostToDevice, sm));
        PIC_HIP(hipMemcpyAsync(st.d_cap[out], nc_.data(), (nb_new + 1) * sizeof(int32_t), hipMemcpyHostToDevice, sm));
        PIC_HIP(hipMemsetAsync(st.d_count[out], 0, (nb_new + 1) * sizeof(int32_t), sm));
        SpeciesDev<T>& D = C.sp[s];
        D.start[out] = st.d_start[out];
        D.cap[out] = st.d_cap[out];
        D.count[out] = st.d_count[out];
        D.mcount[out] = st.d_mcount[out];
        D.lay[out] = st.d_count[out];
        D.nb[out] = nb_new;
    }
    if (int rc = upload_case(p)) return rc;
    PIC_HIP(hipMemsetAsync(p->d_ll, 0, sizeof(long long), sm));
    if (int rc = launch_scatter<T>((CaseDev<T>*)p->d_case, p->d_state, p->d_items, (int)items.size(), sort_chunk,
                                   bw, bh, nbx, nb_new, (unsigned long long*)p->d_ll, sm))
        return rc;
    if (int rc = launch_end_step<T>((CaseDev<T>*)p->d_case, p->d_state, 1, 0, sm)) return rc;
    long long lost = 0;
    PIC_HIP(hipMemcpyAsync(&lost, p->d_ll, sizeof(lost), hipMemcpyDeviceToHost, sm));
    PIC_HIP(hipStreamSynchronize(sm));
    p->h_cur = out;
    for (int s = 0; s < ns; ++s) {
        SpeciesStore& st = p->sp[s];
        st.nb = nb_new;
        st.h_start = new_start[s];
        st.h_cap = new_cap[s];
        st.layout_alloc = nb_new + 1;
        if (int rc = alloc_layout_parity(st, in, nb_new + 1)) return rc;
        PIC_HIP(hipMemcpyAsync(st.d_start[in], st.d_start[out], (nb_new + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, sm));
        PIC_HIP(hipMemcpyAsync(st.d_cap[in], st.d_cap[out], (nb_new + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, sm));
        SpeciesDev<T>& D = C.sp[s];
        D.start[in] = st.d_start[in];
        D.cap[in] = st.d_cap[in];
        D.count[in] = st.d_count[in];
        D.mcount[in] = st.d_mcount[in];
        D.lay[in] = st.d_count[in];
        D.nb[in] = nb_new;
    }
    C.bw = bw;
    C.bh = bh;
    C.bw_log2 = log2_exact(bw);
    C.bh_log2 = log2_exact(bh);
    C.nbx = nbx;
    C.nby = nby;
    S->bin_w = bw;
    S->bin_h = bh;
    if (int rc = build_wall_lists<T>(S, C)) return rc;
    if (int rc = upload_case(p)) return rc;
    if (int rc = rebuild_push_items(S)) return rc;
    S->rebins++;
    S->need_rebin = false;
    if (lost) {
        set_error("re-binning lost particles (spill region full)");
        return PIC_ERR_CAPACITY;
    }
    return 0;
}

int sim_rebin(pic_sim* S, int bw, int bh) {
    return S->p->prec == 4 ? rebin_t<float>(S, bw, bh) : rebin_t<double>(S, bw, bh);
}

template <typename T>
static int setup_t(pic_sim* S) {
    pic_particles* p = S->p;
    pic_grid* g = S->g;
    CaseDev<T>& C = hcase<T>(p);
    const int ns = (int)p->sp.size();
    C.nx = g->m;
    C.ny = g->n;
    C.bw = S->bin_w;
    C.bh = S->bin_h;
    C.bw_log2 = log2_exact(S->bin_w);
    C.bh_log2 = log2_exact(S->bin_h);
    C.nbx = div_up(g->m - 1, S->bin_w);
    C.nby = div_up(g->n - 1, S->bin_h);
    C.halo = S->halo;
    C.box = (S->w && S->w->box) ? 1 : 0;
    C.ns = ns;
    C.dx = g->dx;
    C.dy = g->dy;
    C.wall_eps = S->wall_eps;
    C.dtdx = (T)(S->dt / g->dx);
    C.dtdy = (T)(S->dt / g->dy);
    C.phi = g->phi;
    C.n_walls = 0;
    C.walls = nullptr;
    if (S->w && !S->w->walls.empty()) {
        std::vector<double> wc(S->w->walls.size());
        for (size_t i = 0; i < wc.size(); i += 4) {
            wc[i] = S->w->walls[i] / g->dx;
            wc[i + 1] = S->w->walls[i + 1] / g->dy;
            wc[i + 2] = S->w->walls[i + 2] / g->dx;
            wc[i + 3] = S->w->walls[i + 3] / g->dy;
        }
        PIC_HIP(hipMalloc(&S->d_walls_cell, wc.size() * sizeof(double)));
        PIC_HIP(hipMemcpy(S->d_walls_cell, wc.data(), wc.size() * sizeof(double), hipMemcpyHostToDevice));
        C.walls = S->d_walls_cell;
        C.n_walls = (int)(wc.size() / 4);
    }
    if (int rc = build_wall_lists<T>(S, C)) return rc;
    const size_t nn = (size_t)g->m * g->n;
    S->acc.assign(ns, nullptr);
    S->tabs.assign(ns, nullptr);
    S->tab_host.assign(ns, {});
    S->info.assign(ns, pic_mcc_info{});
    S->charge = ChargeArgs{};
    S->charge.ns = ns;
    S->charge.zero = 1;
    S->charge.rho = g->rho;
    S->charge.rhs = g->rhs;
    const double qscale = 1.0 / ((double)(1ull << PIC_QBITS) * (double)(1ull << PIC_QBITS));
    for (int s = 0; s < ns; ++s) {
        fill_species_ptrs<T>(p, s);
        SpeciesDev<T>& D = C.sp[s];
        const SpeciesStore& st = p->sp[s];
        D.kq = (T)(st.q * S->dt / st.m);
        D.efac = (T)(0.5 * st.m / kElementaryCharge);
        PIC_HIP(hipMalloc(&S->acc[s], nn * sizeof(unsigned long long)));
        PIC_HIP(hipMemset(S->acc[s], 0, nn * sizeof(unsigned long long)));
        D.acc = S->acc[s];
        D.n_proc = 0;
        D.tab = nullptr;
        D.K = 2;
        if (S->xs && S->mcc) {
            MccBuild mb;
            if (int rc = build_mcc(S->xs, p->sp, s, S->dt, mb)) return rc;
            S->info[s] = mb.info;
            S->tab_host[s] = mb.tab;
            if (mb.info.n_proc > 0) {
                std::vector<T> tt(mb.tab.size());
                for (size_t i = 0; i < tt.size(); ++i) tt[i] = (T)mb.tab[i];
                PIC_HIP(hipMalloc(&S->tabs[s], tt.size() * sizeof(T)));
                PIC_HIP(hipMemcpy(S->tabs[s], tt.data(), tt.size() * sizeof(T), hipMemcpyHostToDevice));
                D.n_proc = mb.info.n_proc;
                D.K = mb.info.K;
                for (int q = 0; q < PIC_MAX_PROC; ++q) {
                    D.ptype[q] = q < mb.info.n_proc ? mb.info.type[q] : 0;
                    D.thr[q] = (T)(q < mb.info.n_proc ? mb.info.threshold[q] : 0.0);
                }
                D.tab = (const T*)S->tabs[s];
                D.inv_de = (T)mb.info.inv_de;
                D.nu_max = (T)mb.info.nu_max;
                D.p_null = (T)mb.info.p_null;
                D.inv_emax = (T)(1.0 / mb.info.emax);
                for (int b = 0; b < 16; ++b) {
                    D.band_numax[b] = (T)mb.info.band_nu_max[b];
                    D.band_pnull[b] = (T)mb.info.band_p_null[b];
                }
                D.gas_vth = mb.info.gas_vth;
                D.other = mb.info.product_species;
            }
        }
        const double qf = st.q * st.w * qscale / (g->dx * g->dy);
        S->charge.acc[s] = S->acc[s];
        S->charge.qfac[s] = qf;
        S->charge.cfac[s] = -qf / S->eps0;
    }
    if (int rc = upload_case(p)) return rc;
    return 0;
}

extern "C" int pic_sim_default_params(pic_sim_params* out) {
    PIC_REQUIRE(out, "out is NULL");
    *out = pic_sim_params{};
    out->dt = 1e-10;
    out->eps0 = 8.8541878128e-12;
    out->seed = 1;
    out->case_id = 0;
    out->bin_w = 16;
    out->bin_h = 16;
    out->halo = 2;
    out->mcc = 1;
    out->rebin_every = 0;
    out->wall_eps = 1e-10;
    return pic_mg_default_params(PIC_MG_PRESET_STEP_SOR, &out->mg);
}

extern "C" int pic_sim_create(pic_particles* p, pic_grid* g, pic_walls* w, pic_xs* xs, const pic_sim_params* prm,
                              pic_sim** out) {
    PIC_REQUIRE(p && g && prm && out, "NULL argument");
    PIC_REQUIRE(!p->attached, "particles already belong to a pic_sim");
    PIC_REQUIRE(g->m == g->n, "the multigrid solver needs a square grid (Multigrid.cu:84-88)");
    PIC_REQUIRE(std::fabs(g->dx - g->dy) <= 1e-12 * g->dx, "the multigrid solver needs dx == dy");
    PIC_REQUIRE(prm->dt > 0 && prm->eps0 != 0, "bad dt or eps0");
    PIC_REQUIRE(prm->bin_w >= 2 && prm->bin_w <= 64 && prm->bin_h >= 2 && prm->bin_h <= 64, "bin tile must be 2..64 cells");
    PIC_REQUIRE(prm->halo >= 0 && prm->halo <= 8, "halo must be 0..8");
    PIC_REQUIRE(prm->mg.mode == PIC_MG_FP64 || prm->mg.mode == PIC_MG_MIXED, "bad multigrid mode");
    PIC_REQUIRE(prm->mg.inner >= 0 && prm->mg.inner <= 4, "inner cycles must be 0..4");
    PIC_REQUIRE(prm->mg.coarse_direct >= 0 && prm->mg.coarse_direct <= 2, "coarse_direct must be 0, 1 or 2");
    auto* S = new pic_sim();
    S->p = p;
    S->g = g;
    S->w = w;
    S->xs = xs;
    S->dt = prm->dt;
    S->eps0 = prm->eps0;
    S->bin_w = prm->bin_w;
    S->bin_h = prm->bin_h;
    S->halo = prm->halo;
    S->mcc = prm->mcc;
    S->rebin_every = prm->rebin_every;
    S->wall_eps = prm->wall_eps;
    S->mgp = to_mg(prm->mg);
    CaseDev<float>& cf = p->hf;
    CaseDev<double>& cd = p->hd;
    cf.seed = cd.seed = prm->seed;
    cf.case_id = cd.case_id = prm->case_id;
    auto fail = [&](int rc) {
        pic_sim_destroy(S);
        return rc;
    };
    int rc = p->prec == 4 ? setup_t<float>(S) : setup_t<double>(S);
    if (rc) return fail(rc);
    S->mg.reset(new MGSolver(g->m, S->mgp, lib_stream()));
    if ((rc = S->mg->init())) return fail(rc);
    // the spill status is written by end_step_kernel straight into pinned,
    // coherent host memory (read after the step's synchronization)
    if (hipMalloc(&S->phi_prev, (size_t)g->m * g->n * sizeof(double)) != hipSuccess ||
        hipMalloc(&S->d_rebin, sizeof(int)) != hipSuccess || hipMemset(S->d_rebin, 0, sizeof(int)) != hipSuccess ||
        hipHostMalloc(&S->h_status, PIC_MAX_SPECIES * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&S->d_status, S->h_status, 0) != hipSuccess) {
        set_error("status buffers");
        return fail(PIC_ERR_HIP);
    }
    S->mg->set_phi_save(S->phi_prev);
    std::memset(S->h_status, 0, PIC_MAX_SPECIES * sizeof(int));
    for (int k = 0; k < 5; ++k)
        if (hipEventCreate(&S->ev[k]) != hipSuccess) return fail(PIC_ERR_HIP);
    // particles staged in the reference's layout before the simulation existed
    if (p->stage_pending > 0 && (rc = commit_staging(p, p->stage_pending, g->dx, g->dy))) return fail(rc);
    p->attached = S;
    if ((rc = sim_rebin(S, S->bin_w, S->bin_h))) {
        p->attached = nullptr;
        return fail(rc);
    }
    S->rebins = 0;
    *out = S;
    return 0;
}

int mcc_active(pic_sim* S) {
    if (!S->xs || !S->mcc) return 0;
    for (auto& i : S->info)
        if (i.n_proc > 0) return 1;
    return 0;
}

// The charge conversion and end of step that the first residual pass of the
// step's solve runs (the spill status goes to the host's pinned buffer)
static MGStepFuse step_fuse(pic_sim* S) {
    MGStepFuse f{};
    f.charge = S->charge;
    f.cases = S->p->d_case;
    f.states = S->p->d_state;
    f.status = S->d_status;
    f.rebin = S->d_rebin;
    f.fp32 = S->p->prec == 4;
    return f;
}

// The work of one step after the host decisions: push, then the multigrid
// solve of `cycles` cycles whose first residual pass also converts the charge
// and ends the step (spill status to the host), with no host synchronization.
// Graph mode captures exactly this.
template <typename T>
static int enqueue_step_body(pic_sim* S, int flags, size_t lds, int cycles) {
    pic_particles* p = S->p;
    hipStream_t sm = lib_stream();
    if (int rc = launch_push<T>((CaseDev<T>*)p->d_case, p->d_state, S->d_items, S->n_items, S->spill_chunk, flags, -1,
                                lds, sm, S->d_walls_cell != nullptr))
        return rc;
    // the solve's first residual pass also copies phi (its warm start) to phi_prev
    const MGStepFuse f = step_fuse(S);
    return S->mg->enqueue_fixed(cycles, &f);
}

static void drop_graph(pic_sim* S) {
    if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
    S->gexec = nullptr;
}

void sim_drop_graph(pic_sim* S) { drop_graph(S); }

int sim_set_mg(pic_sim* S, const MGParams& prm) {
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    std::unique_ptr<MGSolver> mg(new MGSolver(S->g->m, prm, lib_stream()));
    if (int rc = mg->init()) return rc;
    mg->set_phi_save(S->phi_prev);
    drop_graph(S);
    S->mg = std::move(mg);
    S->mgp = prm;
    S->last_cycles = 0;
    return 0;
}

static int level_params(MGParams& p, int n, const int* pre, const int* post) {
    PIC_REQUIRE(n >= 0 && n <= 16 && (n == 0 || (pre && post)), "level sweeps: bad arguments");
    p.per_level = 1;
    for (int l = 0; l < 16; ++l) {
        p.pre_l[l] = l < n ? pre[l] : p.pre;
        p.post_l[l] = l < n ? post[l] : p.post;
        PIC_REQUIRE(p.pre_l[l] >= 0 && p.pre_l[l] <= 3 && p.post_l[l] >= 0 && p.post_l[l] <= 3,
                    "level sweeps must be in [0,3]");
    }
    return 0;
}

extern "C" int pic_sim_set_level_sweeps(pic_sim* S, int n, const int* pre, const int* post) {
    PIC_REQUIRE(S, "sim is NULL");
    PIC_REQUIRE(!S->ens, "the simulation is in an ensemble");
    MGParams p = S->mgp;
    if (int rc = level_params(p, n, pre, post)) return rc;
    return sim_set_mg(S, p);
}

extern "C" int pic_mg_set_electrodes(pic_mg* m, const uint8_t* mask, const double* phi_bc) {
    PIC_REQUIRE(m, "solver is NULL");
    return m->solver->set_mask(mask, phi_bc);
}

extern "C" int pic_sim_set_electrodes(pic_sim* S, const uint8_t* mask, const double* phi_bc) {
    PIC_REQUIRE(S, "sim is NULL");
    drop_graph(S);
    if (int rc = S->mg->set_mask(mask, phi_bc)) return rc;
    ensemble_invalidate(S);
    return 0;
}

extern "C" int pic_mg_set_level_sweeps(pic_mg* m, int n, const int* pre, const int* post) {
    PIC_REQUIRE(m, "solver is NULL");
    MGParams p = m->solver->params();
    if (int rc = level_params(p, n, pre, post)) return rc;
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    std::unique_ptr<MGSolver> s(new MGSolver(m->solver->n(), p, lib_stream()));
    if (int rc = s->init()) return rc;
    m->solver = std::move(s);
    return 0;
}

namespace pic {
// The spill fill a step leaves (its slot counter, which keeps counting past the
// capacity): above the capacity, particles of the species found no slot and
// were dropped, so the step reports PIC_ERR_CAPACITY.
int spill_status(const SpeciesStore& ss, int fill, int species) {
    const int cap = ss.h_cap[ss.nb];
    if (fill > cap) {
        set_error("spill region full: " + std::to_string((long long)fill - cap) + " particles of species " +
                  std::to_string(species) + " were lost in the last step (raise max_particles or re-bin more often)");
        return PIC_ERR_CAPACITY;
    }
    return 0;
}
}  // namespace pic

// One step as a graph launch.  The graph is captured again when the layout
// (work items, LDS size), the pass flags or the cycle count change.
template <typename T>
static int graph_step(pic_sim* S, int flags, size_t lds, MGStats* st) {
    hipStream_t sm = lib_stream();
    const int cycles = std::max(1, S->last_cycles);
    // A graph with one refinement step (mixed mode: inner V-cycles; FP64: one
    // V-cycle) more than the last solve needed is kept unless the lower count
    // holds for PIC_GRAPH_SHRINK steps: the device flag turns cycles beyond
    // convergence into no-ops, so a count that moves between two values does
    // not re-capture every step, while a count that settled lower re-captures
    // (gated no-op launches are not free).
    const int step = S->mgp.inner_eff();
    S->g_shrink = (S->gexec && cycles < S->g_cycles) ? S->g_shrink + 1 : 0;
    const bool cycles_fit = S->g_cycles >= cycles && S->g_cycles <= cycles + step && S->g_shrink < PIC_GRAPH_SHRINK;
    if (S->gexec && (S->g_items != S->d_items || S->g_n_items != S->n_items || S->g_lds != lds ||
                     S->g_flags != flags || !cycles_fit))
        drop_graph(S);
    if (!S->gexec) {
        if (int rc = S->mg->prepare(S->g->phi, S->g->rhs, S->g->dx)) return rc;   // allocates: not in a capture
        PIC_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_step_body<T>(S, flags, lds, cycles);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(sm, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PIC_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&S->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PIC_HIP(ei);
        S->g_items = S->d_items;
        S->g_n_items = S->n_items;
        S->g_lds = lds;
        S->g_flags = flags;
        S->g_cycles = cycles;
        S->graph_captures++;
    }
    PIC_HIP(hipEventRecord(S->ev[0], sm));
    PIC_HIP(hipGraphLaunch(S->gexec, sm));
    PIC_HIP(hipEventRecord(S->ev[2], sm));
    PIC_HIP(hipEventSynchronize(S->ev[2]));
    S->graph_steps++;
    return S->mg->complete_fixed(st);
}

// Eager steps.  The push of step i + 1 is enqueued before the host has read
// step i's solve and spill status, gated on the device (the solve's stop flag
// set after the cycles enqueued, no species in need of re-binning), so the GPU
// runs it while the host reads the status; the host tests the same two facts
// and enqueues the push again only if the gate held it back.  A step that
// re-bins first (periodic re-binning) or ends the call is not run ahead.
template <typename T>
static int step_t(pic_sim* S, int n) {
    pic_particles* p = S->p;
    hipStream_t sm = lib_stream();
    const int mcc = mcc_active(S);
    const bool walls = S->d_walls_cell != nullptr;
    const int full = PIC_F_DRIFT | PIC_F_DEPOSIT | PIC_F_KICK | (mcc ? PIC_F_MCC : 0);
    bool ahead = false;   // this step's push ran behind the previous step's solve
    int pe = 0;           // this step's push events: ev[pe], ev[pe + 1]
    for (int i = 0; i < n; ++i) {
        if (!ahead &&
            (S->need_rebin || (S->rebin_every > 0 && S->steps > 0 && S->steps % S->rebin_every == 0))) {
            if (int rc = sim_rebin(S, S->bin_w, S->bin_h)) return rc;
        }
        const size_t lds = cur_lds(S);
        const int flags = S->pending ? full : (PIC_F_DRIFT | PIC_F_DEPOSIT);
        if (S->use_graph && S->pending && S->last_cycles > 0) {
            MGStats st{};
            if (int rc = graph_step<T>(S, flags, lds, &st)) return rc;
            S->have_prev = true;
            p->h_cur ^= 1;
            float a = 0;
            (void)hipEventElapsedTime(&a, S->ev[0], S->ev[2]);
            S->ms_graph += a;
            S->steps++;
            S->last_cycles = st.cycles;
            S->last_rel = st.rel_residual;
            S->total_cycles += st.cycles;
            for (int s = 0; s < (int)p->sp.size(); ++s) {
                const SpeciesStore& ss = p->sp[s];
                if (S->h_status[s] > ss.h_cap[ss.nb] / 2) S->need_rebin = true;
                if (int rc = spill_status(ss, S->h_status[s], s)) return rc;
            }
            continue;
        }
        if (!ahead) {
            PIC_HIP(hipEventRecord(S->ev[pe], sm));
            if (int rc = launch_push<T>((CaseDev<T>*)p->d_case, p->d_state, S->d_items, S->n_items, S->spill_chunk,
                                        flags, -1, lds, sm, walls))
                return rc;
            PIC_HIP(hipEventRecord(S->ev[pe + 1], sm));
        }
        p->h_cur ^= 1;
        S->have_prev = true;   // the solve's first residual pass copies phi to phi_prev
        if (int rc = S->mg->prepare(S->g->phi, S->g->rhs, S->g->dx)) return rc;
        // its first residual pass also converts the charge and ends the step
        const MGStepFuse fuse = step_fuse(S);
        if (int rc = S->mg->enqueue_fixed(S->mg->last_cycles(), &fuse)) return rc;
        PIC_HIP(hipEventRecord(S->ev[2], sm));
        // the next step runs ahead unless it re-bins by the period, ends the
        // call or is a graph step
        const int pn = 3 - pe;
        const bool next = i + 1 < n && !S->use_graph &&
                          !(S->rebin_every > 0 && (S->steps + 1) % S->rebin_every == 0);
        if (next) {
            PIC_HIP(hipEventRecord(S->ev[pn], sm));
            if (int rc = launch_push<T>((CaseDev<T>*)p->d_case, p->d_state, S->d_items, S->n_items, S->spill_chunk,
                                        full, -1, lds, sm, walls, S->mg->stop_flag(), S->d_rebin))
                return rc;
            PIC_HIP(hipEventRecord(S->ev[pn + 1], sm));
        }
        PIC_HIP(hipEventSynchronize(S->ev[2]));
        const bool stopped = S->mg->fixed_stopped();
        MGStats st{};
        if (int rc = S->mg->complete_fixed(&st)) return rc;
        if (!stopped) {   // cycles beyond the enqueued ones: time them too
            PIC_HIP(hipEventRecord(S->ev[2], sm));
            PIC_HIP(hipEventSynchronize(S->ev[2]));
        }
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, S->ev[pe], S->ev[pe + 1]);
        (void)hipEventElapsedTime(&b, S->ev[pe + 1], S->ev[2]);
        S->ms_push += a;
        S->ms_field += b;
        S->pending = true;
        S->steps++;
        S->last_cycles = st.cycles;
        S->last_rel = st.rel_residual;
        S->total_cycles += st.cycles;
        bool rebin = false;
        for (int s = 0; s < (int)p->sp.size(); ++s) {
            const SpeciesStore& ss = p->sp[s];
            if (S->h_status[s] > ss.h_cap[ss.nb] / 2) rebin = true;
            if (int rc = spill_status(ss, S->h_status[s], s)) return rc;
        }
        if (rebin) S->need_rebin = true;
        // the gate let the pushed-ahead step run exactly when both hold
        ahead = next && stopped && !rebin;
        if (ahead) {
            pe = pn;
            S->ahead_steps++;
        } else if (next) {
            S->ahead_held++;
        }
    }
    return 0;
}

extern "C" int pic_sim_step(pic_sim* S, int n_steps) {
    PIC_REQUIRE(S, "sim is NULL");
    PIC_REQUIRE(n_steps >= 0, "n_steps must be non-negative");
    auto t0 = std::chrono::steady_clock::now();
    int rc = S->p->prec == 4 ? step_t<float>(S, n_steps) : step_t<double>(S, n_steps);
    S->ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

template <typename T>
static int sync_t(pic_sim* S) {
    pic_particles* p = S->p;
    hipStream_t sm = lib_stream();
    const int flags = PIC_F_KICK | (mcc_active(S) ? PIC_F_MCC : 0);
    if (int rc = launch_push<T>((CaseDev<T>*)p->d_case, p->d_state, S->d_items, S->n_items, S->spill_chunk, flags, -1,
                                cur_lds(S), sm, S->d_walls_cell != nullptr))
        return rc;
    if (int rc = launch_end_step<T>((CaseDev<T>*)p->d_case, p->d_state, 1, 0, sm)) return rc;
    p->h_cur ^= 1;
    PIC_HIP(hipStreamSynchronize(sm));
    S->pending = false;
    return 0;
}

static int sim_sync_impl(pic_sim* S) {
    if (!S->pending) return 0;
    return S->p->prec == 4 ? sync_t<float>(S) : sync_t<double>(S);
}

extern "C" int pic_sim_sync(pic_sim* S) {
    PIC_REQUIRE(S, "sim is NULL");
    return sim_sync_impl(S);
}

// velocity moments of every species (moments_kernel over the species' work
// items, summed on the host in item order) and the field energy and potential
// norm (sums of squares of E and phi over the nodes)
template <typename T>
static int diagnostics_t(pic_sim* S, pic_diagnostics* out) {
    pic_particles* p = S->p;
    pic_grid* g = S->g;
    hipStream_t sm = lib_stream();
    const size_t nn = (size_t)g->m * g->n;
    size_t need = (2 * nn + 255) / 256;
    std::vector<std::vector<WorkItem>> items(p->sp.size());
    for (size_t s = 0; s < p->sp.size(); ++s) {
        items_for_species(p, (int)s, 4096, items[s]);
        need = std::max(need, 4 * items[s].size());
    }
    double* d = nullptr;
    PIC_HIP(hipMalloc(&d, std::max<size_t>(need, 1) * sizeof(double)));
    std::unique_ptr<double, void (*)(double*)> hold(d, [](double* q) { (void)hipFree(q); });
    std::vector<double> h(need);
    for (size_t s = 0; s < p->sp.size(); ++s) {
        const int ni = (int)items[s].size();
        if (int rc = upload_items(&p->d_items, &p->items_alloc, items[s])) return rc;
        if (int rc = launch_moments<T>((CaseDev<T>*)p->d_case, p->d_state, p->d_items, ni, 4096, d, sm)) return rc;
        PIC_HIP(hipMemcpyAsync(h.data(), d, 4 * (size_t)ni * sizeof(double), hipMemcpyDeviceToHost, sm));
        PIC_HIP(hipStreamSynchronize(sm));
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = 0; b < ni; ++b)
            for (int k = 0; k < 4; ++k) a[k] += h[4 * (size_t)b + k];
        const double mw = p->sp[s].m * p->sp[s].w;
        out->count[s] = a[0];
        out->px[s] = mw * a[1];
        out->py[s] = mw * a[2];
        out->kinetic[s] = 0.5 * mw * a[3];
    }
    auto sum_sq = [&](const double* src, size_t n, double* res) -> int {
        const size_t nc = (n + 255) / 256;
        if (int rc = launch_sumsq(src, n, d, sm)) return rc;
        PIC_HIP(hipMemcpyAsync(h.data(), d, nc * sizeof(double), hipMemcpyDeviceToHost, sm));
        PIC_HIP(hipStreamSynchronize(sm));
        double t = 0.0;
        for (size_t c = 0; c < nc; ++c) t += h[c];
        *res = t;
        return 0;
    };
    if (int rc = launch_efield(g->phi, (double2*)g->E, g->m, g->n, g->dx, g->dy, sm)) return rc;
    double e2 = 0.0, p2 = 0.0;
    if (int rc = sum_sq(g->E, 2 * nn, &e2)) return rc;
    if (int rc = sum_sq(g->phi, nn, &p2)) return rc;
    out->field_energy = 0.5 * S->eps0 * e2 * g->dx * g->dy;
    out->phi_l2 = std::sqrt(p2);
    return 0;
}

extern "C" int pic_sim_diagnostics(pic_sim* S, pic_diagnostics* out) {
    PIC_REQUIRE(S && out, "NULL argument");
    if (int rc = sim_sync_impl(S)) return rc;
    *out = pic_diagnostics{};
    out->n_species = (int)S->p->sp.size();
    return S->p->prec == 4 ? diagnostics_t<float>(S, out) : diagnostics_t<double>(S, out);
}

extern "C" int pic_sim_counters(pic_sim* S, int species, uint64_t out[16]) {
    PIC_REQUIRE(S && out, "NULL argument");
    PIC_REQUIRE(species >= 0 && species < (int)S->p->sp.size(), "bad species");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    PIC_HIP(hipMemcpy(out, S->p->sp[species].d_counters, PIC_NCOUNT * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int pic_sim_stats(pic_sim* S, pic_sim_report* out) {
    PIC_REQUIRE(S && out, "NULL argument");
    *out = pic_sim_report{};
    out->steps = S->steps;
    int64_t n = 0;
    if (int rc = pic_particles_last_alive(S->p, -1, &n)) return rc;
    out->alive = n;
    out->last_mg_cycles = S->last_cycles;
    out->last_mg_rel = S->last_rel;
    out->total_mg_cycles = S->total_cycles;
    out->rebins = S->rebins;
    out->ms_push = S->ms_push;
    out->ms_field = S->ms_field;
    out->ms_total = S->ms_total;
    out->graph_steps = S->graph_steps;
    out->graph_captures = S->graph_captures;
    out->ms_graph = S->ms_graph;
    out->ahead_steps = S->ahead_steps;
    out->ahead_held = S->ahead_held;
    return 0;
}

extern "C" int pic_sim_mcc_info(pic_sim* S, int species, pic_mcc_info* out, double* table, int cap) {
    PIC_REQUIRE(S && out, "NULL argument");
    PIC_REQUIRE(species >= 0 && species < (int)S->p->sp.size(), "bad species");
    *out = S->info[species];
    if (table) {
        const auto& t = S->tab_host[species];
        PIC_REQUIRE((int)t.size() <= cap, "table capacity too small");
        std::copy(t.begin(), t.end(), table);
    }
    return 0;
}

extern "C" int pic_sim_rebin(pic_sim* S, int bin_w, int bin_h) {
    PIC_REQUIRE(S, "sim is NULL");
    const int bw = bin_w > 0 ? bin_w : S->bin_w, bh = bin_h > 0 ? bin_h : S->bin_h;
    PIC_REQUIRE(bw >= 2 && bw <= 64 && bh >= 2 && bh <= 64, "bin tile must be 2..64 cells");
    return sim_rebin(S, bw, bh);
}

int ensemble_forget(pic_sim* S);

extern "C" int pic_sim_use_graph(pic_sim* S, int on) {
    PIC_REQUIRE(S, "sim is NULL");
    PIC_REQUIRE(!S->ens, "the simulation is in an ensemble: set graph mode on the ensemble");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    S->use_graph = on != 0;
    if (!S->use_graph) drop_graph(S);
    return 0;
}

// frame file: "PICF", u32 version 1, u64 step, u32 species, u32 units (1:
// domain), f64 width, f64 height, then per species u64 count and count float
// pairs (x, y); little-endian. Written to path + ".tmp" and renamed, so a
// reader never sees a partial frame.
extern "C" int pic_sim_write_frame(pic_sim* S, const char* path) {
    PIC_REQUIRE(S && path, "NULL argument");
    pic_particles* p = S->p;
    const int ns = (int)p->sp.size();
    std::vector<std::vector<float>> xy(ns);
    for (int s = 0; s < ns; ++s) {
        int64_t n = 0;
        if (int rc = pic_particles_export_positions(p, s, nullptr, 0, 1, &n)) return rc;
        if (n > S->frame_cap) {
            (void)hipFree(S->d_frame);
            S->d_frame = nullptr;
            S->frame_cap = 0;
            PIC_HIP(hipMalloc(&S->d_frame, (size_t)n * 2 * sizeof(float)));
            S->frame_cap = n;
        }
        int64_t n2 = 0;
        if (int rc = pic_particles_export_positions(p, s, S->d_frame, S->frame_cap, 1, &n2)) return rc;
        xy[s].resize((size_t)n2 * 2);
        if (n2) PIC_HIP(hipMemcpy(xy[s].data(), S->d_frame, (size_t)n2 * 2 * sizeof(float), hipMemcpyDeviceToHost));
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        set_error(std::string("cannot write frame file: ") + tmp);
        return PIC_ERR_IO;
    }
    const uint32_t version = 1, nsp = (uint32_t)ns, units = 1;
    const uint64_t step = (uint64_t)S->steps;
    const double wh[2] = {S->g->dx * (S->g->m - 1), S->g->dy * (S->g->n - 1)};
    bool ok = std::fwrite("PICF", 1, 4, f) == 4 && std::fwrite(&version, 4, 1, f) == 1 &&
              std::fwrite(&step, 8, 1, f) == 1 && std::fwrite(&nsp, 4, 1, f) == 1 && std::fwrite(&units, 4, 1, f) == 1 &&
              std::fwrite(wh, 8, 2, f) == 2;
    for (int s = 0; s < ns && ok; ++s) {
        const uint64_t c = xy[s].size() / 2;
        ok = std::fwrite(&c, 8, 1, f) == 1 && (c == 0 || std::fwrite(xy[s].data(), 8, c, f) == c);
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        set_error(std::string("cannot write frame file: ") + path);
        return PIC_ERR_IO;
    }
    return 0;
}

extern "C" int pic_sim_set_wall_lists(pic_sim* S, int on, int margin) {
    PIC_REQUIRE(S, "sim is NULL");
    PIC_REQUIRE(margin >= 0 && margin <= 64, "margin must be 0..64 cells");
    PIC_HIP(hipStreamSynchronize(lib_stream()));
    S->wall_lists = on != 0;
    S->wall_margin = margin;
    int rc = S->p->prec == 4 ? build_wall_lists<float>(S, hcase<float>(S->p))
                             : build_wall_lists<double>(S, hcase<double>(S->p));
    if (rc) return rc;
    return upload_case(S->p);
}

extern "C" int pic_sim_wall_list_stats(pic_sim* S, int* n_bins, long long* n_entries, int* max_per_bin) {
    PIC_REQUIRE(S && n_bins && n_entries && max_per_bin, "NULL argument");
    const bool on = S->d_wall_lists != nullptr;
    *n_bins = on ? div_up(S->g->m - 1, S->bin_w) * div_up(S->g->n - 1, S->bin_h) : 0;
    *n_entries = on ? S->wall_entries : 0;
    *max_per_bin = on ? S->wall_max_bin : 0;
    return 0;
}

extern "C" int pic_sim_destroy(pic_sim* S) {
    if (!S) return 0;
    (void)ensemble_forget(S);
    (void)hipStreamSynchronize(lib_stream());
    drop_graph(S);
    if (S->p && S->p->attached == S) {
        (void)sim_sync_impl(S);
        S->p->attached = nullptr;
    }
    for (auto* a : S->acc) (void)hipFree(a);
    for (auto* t : S->tabs) (void)hipFree(t);
    (void)hipFree(S->d_walls_cell);
    (void)hipFree(S->d_wall_lists);
    (void)hipFree(S->d_frame);
    (void)hipFree(S->d_items);
    (void)hipFree(S->phi_prev);
    (void)hipFree(S->d_rebin);
    if (S->h_status) (void)hipHostFree(S->h_status);
    for (auto& e : S->ev)
        if (e) (void)hipEventDestroy(e);
    // the particles keep their data; clear the grid-dependent device pointers
    if (S->p) {
        for (int s = 0; s < (int)S->p->sp.size(); ++s) {
            S->p->hf.sp[s].acc = nullptr;
            S->p->hd.sp[s].acc = nullptr;
            S->p->hf.sp[s].n_proc = 0;
            S->p->hd.sp[s].n_proc = 0;
        }
        S->p->hf.n_walls = S->p->hd.n_walls = 0;
        S->p->hf.walls = S->p->hd.walls = nullptr;
        S->p->hf.wall_off = S->p->hd.wall_off = nullptr;
        S->p->hf.wall_idx = S->p->hd.wall_idx = nullptr;
        (void)upload_case(S->p);
    }
    delete S;
    return 0;
}

// ============================================================================
// elasticCollisions (mcc.cu:114-150)
// ============================================================================
template <typename T>
static int elastic_t(pic_particles* p, int species, double density, double dt, int n, const double* E,
                     const double* sig, uint32_t seed, uint32_t step) {
    pic_xs xs;
    xs.density = density;
    xs.t_gas = 0.0;
    xs.points = 2048;
    xs.emax = E[n - 1];
    xs.procs.push_back(pic_xs::Proc{species, PIC_PROC_ELASTIC, -1, 0.0, 0.0, std::vector<double>(E, E + n),
                                    std::vector<double>(sig, sig + n)});
    MccBuild mb;
    if (int rc = build_mcc(&xs, p->sp, species, dt, mb)) return rc;
    // the table goes to the device before the descriptor changes, so a failure
    // here leaves the store as it was
    std::vector<T> tt(mb.tab.size());
    for (size_t i = 0; i < tt.size(); ++i) tt[i] = (T)mb.tab[i];
    T* dtab = nullptr;
    PIC_HIP(hipMalloc(&dtab, tt.size() * sizeof(T)));
    const hipError_t ec = hipMemcpy(dtab, tt.data(), tt.size() * sizeof(T), hipMemcpyHostToDevice);
    if (ec != hipSuccess) {
        (void)hipFree(dtab);
        set_error(std::string("elastic collisions: table upload: ") + hipGetErrorString(ec));
        return PIC_ERR_HIP;
    }
    CaseDev<T>& C = hcase<T>(p);
    const CaseDev<T> saved = C;
    for (int s = 0; s < (int)p->sp.size(); ++s) C.sp[s].n_proc = 0;
    SpeciesDev<T>& D = C.sp[species];
    D.n_proc = 1;
    D.K = mb.info.K;
    D.ptype[0] = PIC_PROC_ELASTIC;
    D.thr[0] = (T)0;
    D.tab = dtab;
    D.inv_de = (T)mb.info.inv_de;
    D.nu_max = (T)mb.info.nu_max;
    D.p_null = (T)mb.info.p_null;
    D.inv_emax = (T)(1.0 / mb.info.emax);
    for (int b = 0; b < 16; ++b) {
        D.band_numax[b] = (T)mb.info.band_nu_max[b];
        D.band_pnull[b] = (T)mb.info.band_p_null[b];
    }
    D.efac = (T)mb.info.efac;
    D.other = -1;
    C.seed = seed;
    int rc = upload_case(p);
    std::vector<WorkItem> items;
    for (int s = 0; s < (int)p->sp.size() && !rc; ++s) items_for_species(p, s, 4096, items);
    if (!rc) rc = upload_items(&p->d_items, &p->items_alloc, items);
    const size_t lds = push_lds_bytes<T>(C.bw > 0 ? C.bw : 2, C.bh > 0 ? C.bh : 2, C.halo);
    if (!rc)
        rc = launch_push<T>((CaseDev<T>*)p->d_case, p->d_state, p->d_items, (int)items.size(), 4096, PIC_F_MCC,
                            (long long)step, lds, lib_stream(), C.n_walls > 0);
    if (!rc) rc = launch_end_step<T>((CaseDev<T>*)p->d_case, p->d_state, 1, 0, lib_stream());
    if (!rc) p->h_cur ^= 1;
    const hipError_t es = hipStreamSynchronize(lib_stream());
    if (!rc && es != hipSuccess) {
        set_error(std::string("elastic collisions: ") + hipGetErrorString(es));
        rc = PIC_ERR_HIP;
    }
    // the descriptor is restored on every path, then the table freed
    C = saved;
    const int rc2 = upload_case(p);
    (void)hipStreamSynchronize(lib_stream());
    (void)hipFree(dtab);
    return rc ? rc : rc2;
}

extern "C" int pic_elastic_collisions(pic_particles* p, int species, double density, double dt, int n,
                                      const double* energy_eV, const double* sigma_m2, uint32_t seed, uint32_t step) {
    PIC_REQUIRE(p && energy_eV && sigma_m2 && n >= 2, "NULL argument or table shorter than 2");
    PIC_REQUIRE(species >= 0 && species < (int)p->sp.size(), "bad species");
    PIC_REQUIRE(!p->attached, "use the pic_sim MCC for particles that belong to a simulation");
    return p->prec == 4 ? elastic_t<float>(p, species, density, dt, n, energy_eV, sigma_m2, seed, step)
                        : elastic_t<double>(p, species, density, dt, n, energy_eV, sigma_m2, seed, step);
}
