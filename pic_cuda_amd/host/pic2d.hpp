// pic2d.hpp -- C++ host interface of the MI355X PIC step, mirroring the
// reference's classes over the C ABI of libpic_amd.so (include/pic_amd.h).
//
//   reference (dimad285/PIC_CUDA)             here
//   Particles2D            particles.cuh:3-39  pic2d::Particles2D
//   Grid2D                 grid.cuh:7-24       pic2d::Grid2D
//   Boundaries2D           boundaries.h:7-23   pic2d::Boundaries2D
//   solve_multigrid        Multigrid.hpp:3-11  pic2d::solve_multigrid
//   read_cross_section     mcc.hpp:19          pic2d::read_cross_section
//   elasticCollisions      mcc.hpp:21          pic2d::elasticCollisions
//   updateSimulationCUDA   main.cu:37          pic2d::Simulation::step
//
// Errors throw pic2d::Error (the reference exits or prints and continues).
// Positions are in domain units (metres) at this interface, as in the
// reference; they cross the C ABI in cell units.
#pragma once
#include "../../include/pic_amd.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pic2d {

struct Error : std::runtime_error {
    int code;
    Error(int rc, const std::string& what) : std::runtime_error(what), code(rc) {}
};

inline void check(int rc, const char* where) {
    if (rc != PIC_OK) throw Error(rc, std::string(where) + ": " + pic_last_error());
}

class Grid2D {
public:
    Grid2D(int M, int N, double X, double Y) {
        check(pic_grid_create(M, N, X, Y, &h_), "Grid2D");
        check(pic_grid_info(h_, &m, &n, &dx, &dy), "Grid2D");
        x = X;
        y = Y;
        double* e = nullptr;
        check(pic_grid_device_ptrs(h_, &rho, &phi, &e), "Grid2D");
        E = e;
    }
    ~Grid2D() { pic_grid_destroy(h_); }
    Grid2D(const Grid2D&) = delete;
    Grid2D& operator=(const Grid2D&) = delete;

    std::vector<double> download(int field) const {
        std::vector<double> v((size_t)m * n);
        check(pic_grid_download(h_, field, v.data()), "Grid2D::download");
        return v;
    }
    void computeElectricField() { check(pic_grid_compute_E(h_), "computeElectricField"); }
    pic_grid* handle() const { return h_; }

    // public members of the reference (device pointers for rho, phi, E)
    double* rho = nullptr;
    double* phi = nullptr;
    double* E = nullptr;
    int m = 0, n = 0;
    double dx = 0, dy = 0, x = 0, y = 0;

private:
    pic_grid* h_ = nullptr;
};

class Particles2D {
public:
    // staging: allocate the reference's public device arrays r, r_old, v and
    // type for num_particles entries (needed by callers that cudaMemcpy into
    // them, main.cu:232-238; 52 B per entry)
    explicit Particles2D(int64_t num_particles, int precision = PIC_PREC_FP32, bool staging = true) {
        check(pic_particles_create(num_particles, precision, &h_), "Particles2D");
        max_particles = num_particles;
        if (staging)
            check(pic_particles_staging(h_, num_particles, &r, &r_old, &v, &type), "Particles2D staging");
    }
    // particles.cuh:7-20: (x, y) double pairs in domain units and the species
    // index of each entry, filled by the caller; r_old is accepted and unused
    double* r = nullptr;
    double* r_old = nullptr;
    double* v = nullptr;
    int32_t* type = nullptr;
    // particles.cuh:20: `particles.last_alive = n` hands the first n entries
    // of r, v and type to their species (when a Simulation takes the
    // particles, or at once in a running one); reading it counts the live
    // particles of all species
    class LastAlive {
    public:
        explicit LastAlive(Particles2D* o) : o_(o) {}
        LastAlive& operator=(int64_t n) {
            check(pic_particles_set_staged(o_->h_, n), "last_alive");
            return *this;
        }
        operator int64_t() const { return o_->count(); }

    private:
        Particles2D* o_;
    };
    LastAlive last_alive{this};
    ~Particles2D() { pic_particles_destroy(h_); }
    Particles2D(const Particles2D&) = delete;
    Particles2D& operator=(const Particles2D&) = delete;

    void add_species(double m, double q) {
        int idx = 0;
        check(pic_particles_add_species(h_, m, q, &idx), "add_species");
    }
    void change_species(int species, double m, double q) {
        check(pic_particles_change_species(h_, species, m, q), "change_species");
    }
    void set_weight(int species, double w) { check(pic_particles_set_weight(h_, species, w), "set_weight"); }
    // compaction happens inside the step; kept for callers of the reference API
    void removeCollidedParticles() { check(pic_particles_remove_collided(h_), "removeCollidedParticles"); }
    int64_t count(int species = -1) {
        int64_t n = 0;
        check(pic_particles_last_alive(h_, species, &n), "last_alive");
        return n;
    }
    // upload in domain units (x, y in metres, v in m/s)
    void upload(int species, const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& vx,
                const std::vector<double>& vy, const Grid2D& g, const uint32_t* ids = nullptr) {
        std::vector<double> cx(x.size()), cy(y.size());
        for (size_t i = 0; i < x.size(); ++i) {
            cx[i] = x[i] / g.dx;
            cy[i] = y[i] / g.dy;
        }
        check(pic_particles_upload(h_, species, (int64_t)x.size(), cx.data(), cy.data(), vx.data(), vy.data(), ids),
              "upload");
    }
    struct State {
        std::vector<double> x, y, vx, vy;
        std::vector<uint32_t> id;
    };
    State download(int species, const Grid2D& g) {
        int64_t n = 0;
        check(pic_particles_download(h_, species, 0, nullptr, nullptr, nullptr, nullptr, nullptr, &n), "download");
        State s;
        s.x.resize(n);
        s.y.resize(n);
        s.vx.resize(n);
        s.vy.resize(n);
        s.id.resize(n);
        check(pic_particles_download(h_, species, n, s.x.data(), s.y.data(), s.vx.data(), s.vy.data(), s.id.data(), &n),
              "download");
        for (int64_t i = 0; i < n; ++i) {
            s.x[i] *= g.dx;
            s.y[i] *= g.dy;
        }
        return s;
    }
    void init_maxwellian(int species, int64_t n, const Grid2D& g, double vth, uint32_t seed, uint32_t id0 = 0) {
        check(pic_particles_init_maxwellian(h_, species, n, g.m, g.n, vth, seed, 0, id0), "init_maxwellian");
    }
    // uniform positions in [x_lo,x_hi)x[y_lo,y_hi) (domain units) and velocity
    // components uniform in [v_lo,v_hi)
    void init_uniform(int species, int64_t n, const Grid2D& g, double x_lo, double x_hi, double y_lo, double y_hi,
                      double v_lo, double v_hi, uint32_t seed, uint32_t id0 = 0) {
        check(pic_particles_init_uniform(h_, species, n, x_lo / g.dx, x_hi / g.dx, y_lo / g.dy, y_hi / g.dy, v_lo,
                                         v_hi, seed, 0, id0),
              "init_uniform");
    }
    // updateVBO / convertPositionsToFloat (render.h:384-405, kernels.cu:268-273):
    // live positions as float (x, y) pairs into a device buffer (a mapped
    // vertex buffer, say) of room for `cap` pairs; domain units need the
    // particles attached to a Simulation. Returns the live count.
    int64_t export_positions(int species, float* dst_device, int64_t cap, bool domain_units = true) {
        int64_t n = 0;
        check(pic_particles_export_positions(h_, species, dst_device, cap, domain_units ? 1 : 0, &n),
              "export_positions");
        return n;
    }
    pic_particles* handle() const { return h_; }
    int64_t max_particles = 0;

private:
    pic_particles* h_ = nullptr;
};

class Boundaries2D {
public:
    Boundaries2D(double width, double height, int nx, int ny) {
        check(pic_walls_create(width, height, nx, ny, &h_), "Boundaries2D");
    }
    ~Boundaries2D() { pic_walls_destroy(h_); }
    Boundaries2D(const Boundaries2D&) = delete;
    Boundaries2D& operator=(const Boundaries2D&) = delete;
    void add_wall(double x1, double y1, double x2, double y2) { check(pic_walls_add(h_, x1, y1, x2, y2), "add_wall"); }
    void add_box() { check(pic_walls_add_box(h_), "add_box"); }
    void compile_walls() { check(pic_walls_compile(h_), "compile_walls"); }
    int wall_count() const {
        int n = 0;
        check(pic_walls_count(h_, &n), "wall_count");
        return n;
    }
    pic_walls* handle() const { return h_; }

private:
    pic_walls* h_ = nullptr;
};

class CrossSections {
public:
    CrossSections(double gas_density, double gas_temperature_eV) {
        check(pic_xs_create(gas_density, gas_temperature_eV, &h_), "CrossSections");
    }
    ~CrossSections() { pic_xs_destroy(h_); }
    CrossSections(const CrossSections&) = delete;
    CrossSections& operator=(const CrossSections&) = delete;
    void add_process(int species, int type, double threshold, const std::vector<double>& e,
                     const std::vector<double>& s, int product = -1) {
        check(pic_xs_add_process(h_, species, type, threshold, (int)e.size(), e.data(), s.data(), product),
              "add_process");
    }
    void add_analytic(int species, int type, double threshold, double peak, int product = -1) {
        check(pic_xs_add_analytic(h_, species, type, threshold, peak, product), "add_analytic");
    }
    // every block of an LXCat file as a process of `species`; returns the
    // processes added (attachment blocks are skipped)
    int load_lxcat(const std::string& path, int species, int product = -1) {
        int added = 0, skipped = 0;
        check(pic_xs_load_lxcat(h_, path.c_str(), species, product, &added, &skipped), "load_lxcat");
        return added;
    }
    pic_xs* handle() const { return h_; }

private:
    pic_xs* h_ = nullptr;
};

inline std::pair<std::vector<double>, std::vector<double>> read_cross_section(const std::string& filename) {
    int n = 0;
    check(pic_read_cross_section(filename.c_str(), 0, nullptr, nullptr, &n), "read_cross_section");
    std::vector<double> e(n), s(n);
    check(pic_read_cross_section(filename.c_str(), n, e.data(), s.data(), &n), "read_cross_section");
    return {e, s};
}

// initializeParticles(h_positions, h_velocities, num_particles) main.cu:17-32:
// positions uniform in [0.2, 0.8] (domain units) on both axes, velocities
// uniform in [-0.01, 0.01], here seeded and on the device
inline void initializeParticles(Particles2D& p, int species, int64_t n, const Grid2D& g, uint32_t seed) {
    p.init_uniform(species, n, g, 0.2, 0.8, 0.2, 0.8, -0.01, 0.01, seed);
}

// two columns like read_cross_section, numbers also written m*10^{e}
inline std::pair<std::vector<double>, std::vector<double>> read_table(const std::string& filename) {
    int n = 0;
    check(pic_read_table(filename.c_str(), 0, nullptr, nullptr, &n), "read_table");
    std::vector<double> e(n), s(n);
    check(pic_read_table(filename.c_str(), n, e.data(), s.data(), &n), "read_table");
    return {e, s};
}

// solve_multigrid(phi, rho, nx, ny, levels, h, omega): lap(phi) = rho on device arrays
inline void solve_multigrid(double* phi_d, double* rho_d, int nx, int ny, int levels, double h, double omega = 1.25) {
    check(pic_solve_multigrid(phi_d, rho_d, nx, ny, levels, h, omega), "solve_multigrid");
}

inline void elasticCollisions(Particles2D& p, int species, double density, double dt, const std::vector<double>& e,
                              const std::vector<double>& s, uint32_t seed, uint32_t step) {
    check(pic_elastic_collisions(p.handle(), species, density, dt, (int)e.size(), e.data(), s.data(), seed, step),
          "elasticCollisions");
}

class Simulation {
public:
    Simulation(Particles2D& p, Grid2D& g, Boundaries2D* w, CrossSections* xs, const pic_sim_params& prm) {
        check(pic_sim_create(p.handle(), g.handle(), w ? w->handle() : nullptr, xs ? xs->handle() : nullptr, &prm, &h_),
              "Simulation");
    }
    ~Simulation() { pic_sim_destroy(h_); }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;
    // updateSimulationCUDA(particles, grid, boundaries, step, dt, steps, timing)
    void step(int steps) { check(pic_sim_step(h_, steps), "updateSimulation"); }
    void sync() { check(pic_sim_sync(h_), "sync"); }
    pic_sim_report report() {
        pic_sim_report r{};
        check(pic_sim_stats(h_, &r), "report");
        return r;
    }
    // energy and momentum of the synchronized state
    pic_diagnostics diagnostics() {
        pic_diagnostics d{};
        check(pic_sim_diagnostics(h_, &d), "diagnostics");
        return d;
    }
    std::vector<uint64_t> counters(int species) {
        std::vector<uint64_t> c(16);
        check(pic_sim_counters(h_, species, c.data()), "counters");
        return c;
    }
    void write_frame(const std::string& path) { check(pic_sim_write_frame(h_, path.c_str()), "write_frame"); }
    void set_wall_lists(bool on, int margin = 2) { check(pic_sim_set_wall_lists(h_, on ? 1 : 0, margin), "wall_lists"); }
    // embedded electrodes: n*n mask (0 fluid, 1 Dirichlet at phi_bc, 2 solid)
    void set_electrodes(const std::vector<uint8_t>& mask, const std::vector<double>& phi_bc) {
        check(pic_sim_set_electrodes(h_, mask.empty() ? nullptr : mask.data(), phi_bc.empty() ? nullptr : phi_bc.data()),
              "set_electrodes");
    }
    pic_sim* handle() const { return h_; }

private:
    pic_sim* h_ = nullptr;
};

}  // namespace pic2d
