// pic_run.cpp -- command-line step driver (the simulation loop of main.cu:146-351
// without the window): builds a uniform electron/Ar+ plasma in a grounded box,
// runs the step and prints one JSON line with counts and timings.
//
//   pic_run [--cells 128] [--particles 100000] [--steps 100] [--mcc 0|1]
//           [--fp64] [--seed 1] [--te 2.0] [--gas 1e21]
//           [--frames PATH --every K] [--staged]
//
// --staged fills the particles the way main.cu:232-238 does: host arrays of
// positions and velocities copied into the public device arrays r and v,
// species indices into type, then `last_alive = n` (the first n entries are
// electrons, the next n ions; positions and velocities from Weyl sequences,
// x_i = L (0.02 + 0.96 frac(i a)), v_i = vth sqrt(3) (2 frac(i c) - 1)).
//
// With --frames, every K steps the positions are written to the frame file
// PATH (replaced in place), where main.cu hands them to updateVBO every
// UPDATE_STEPS_PER_FRAME steps.
#include "pic2d.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

int main(int argc, char** argv) {
    int cells = 128, steps = 100, mcc = 0, prec = PIC_PREC_FP32;
    long long n = 100000;
    unsigned seed = 1;
    double te = 2.0, gas = 1e21;
    const char* frames = nullptr;
    int every = 0;
    bool staged = false;
    for (int i = 1; i < argc; ++i) {
        auto arg = [&](const char* k) { return std::strcmp(argv[i], k) == 0 && i + 1 < argc; };
        if (arg("--cells")) cells = std::atoi(argv[++i]);
        else if (arg("--particles")) n = std::atoll(argv[++i]);
        else if (arg("--steps")) steps = std::atoi(argv[++i]);
        else if (arg("--mcc")) mcc = std::atoi(argv[++i]);
        else if (arg("--seed")) seed = (unsigned)std::atol(argv[++i]);
        else if (arg("--te")) te = std::atof(argv[++i]);
        else if (arg("--gas")) gas = std::atof(argv[++i]);
        else if (arg("--frames")) frames = argv[++i];
        else if (arg("--every")) every = std::atoi(argv[++i]);
        else if (std::strcmp(argv[i], "--fp64") == 0) prec = PIC_PREC_FP64;
        else if (std::strcmp(argv[i], "--staged") == 0) staged = true;
        else {
            std::fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    const double qe = 1.602176634e-19, me = 9.1093837015e-31, mar = 39.948 * 1.66053906660e-27;
    const double eps0 = 8.8541878128e-12, dens = 1e15;
    const double debye = std::sqrt(eps0 * te / (dens * qe));
    const double dx = 0.5 * debye, L = cells * dx;
    const double dt = 0.2 / std::sqrt(dens * qe * qe / (eps0 * me));
    try {
        pic2d::Grid2D grid(cells + 1, cells + 1, L, L);
        // staged: the reference's MAX_PARTICLES holds both species
        pic2d::Particles2D particles(staged ? 2 * n + 4096 : (int64_t)(n * 1.02) + 4096, prec, staged);
        particles.change_species(0, me, -qe);
        particles.add_species(mar, qe);
        particles.set_weight(0, dens * L * L / n);
        particles.set_weight(1, dens * L * L / n);
        const double vte = std::sqrt(te * qe / me), vti = std::sqrt(0.03 * qe / mar);
        if (staged) {
            std::vector<double> r(4 * n), v(4 * n);
            std::vector<int32_t> type(2 * n);
            for (long long i = 0; i < 2 * n; ++i) {
                const double vth = i < n ? vte : vti;
                r[2 * i] = L * (0.02 + 0.96 * std::fmod((double)i * 0.6180339887498949, 1.0));
                r[2 * i + 1] = L * (0.02 + 0.96 * std::fmod((double)i * 0.7548776662466927, 1.0));
                v[2 * i] = vth * 1.7320508075688772 * (2.0 * std::fmod((double)i * 0.5698402909980532, 1.0) - 1.0);
                v[2 * i + 1] = vth * 1.7320508075688772 * (2.0 * std::fmod((double)i * 0.8191725133961645, 1.0) - 1.0);
                type[i] = i < n ? 0 : 1;
            }
            pic2d::check(pic_host_to_device(particles.r, r.data(), (int64_t)(r.size() * sizeof(double))), "copy r");
            pic2d::check(pic_host_to_device(particles.v, v.data(), (int64_t)(v.size() * sizeof(double))), "copy v");
            pic2d::check(pic_host_to_device(particles.type, type.data(), (int64_t)(type.size() * sizeof(int32_t))),
                         "copy type");
            particles.last_alive = 2 * n;   // handed over when the Simulation takes the particles
        } else {
            particles.init_maxwellian(0, n, grid, vte, seed);
            particles.init_maxwellian(1, n, grid, vti, seed);
        }
        pic2d::Boundaries2D walls(L, L, cells + 1, cells + 1);
        walls.add_box();
        walls.compile_walls();
        pic2d::CrossSections xs(gas, 0.03);
        xs.add_analytic(0, PIC_PROC_ELASTIC, 0.0, 1e-19);
        xs.add_analytic(0, PIC_PROC_EXCITATION, 11.5, 2e-21);
        xs.add_analytic(0, PIC_PROC_IONIZATION, 15.8, 3e-20, 1);
        xs.add_analytic(1, PIC_PROC_CHARGE_EXCHANGE, 0.0, 5e-19);
        pic_sim_params prm;
        pic2d::check(pic_sim_default_params(&prm), "params");
        prm.dt = dt;
        prm.seed = seed;
        prm.mcc = mcc;
        if (prec == PIC_PREC_FP64) prm.mg.mode = PIC_MG_FP64;
        pic2d::Simulation sim(particles, grid, &walls, mcc ? &xs : nullptr, prm);
        if (frames && every <= 0) every = steps;
        int n_frames = 0;
        auto t0 = std::chrono::steady_clock::now();
        if (frames) {
            for (int done = 0; done < steps; done += every) {
                sim.step(std::min(every, steps - done));
                sim.write_frame(frames);
                ++n_frames;
            }
        } else {
            sim.step(steps);
        }
        sim.sync();
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const pic_sim_report r = sim.report();
        const auto ce = sim.counters(0), ci = sim.counters(1);
        std::printf("{\"cells\": %d, \"particles\": %lld, \"steps\": %d, \"seconds\": %.6f, \"alive\": %lld, "
                    "\"absorbed_e\": %llu, \"absorbed_i\": %llu, \"collisions_e\": %llu, \"mg_cycles\": %lld, "
                    "\"particle_steps_per_s\": %.6g, \"frames\": %d}\n",
                    cells, 2 * n, steps, s, (long long)r.alive, (unsigned long long)ce[PIC_CNT_ABSORBED],
                    (unsigned long long)ci[PIC_CNT_ABSORBED],
                    (unsigned long long)(ce[2] + ce[3] + ce[4] + ce[5]), (long long)r.total_mg_cycles,
                    2.0 * n * steps / s, n_frames);
    } catch (const pic2d::Error& e) {
        std::fprintf(stderr, "pic_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
