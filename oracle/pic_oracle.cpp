/*
 * pic_oracle.cpp -- serial CPU restatement of the reference PIC step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See pic_oracle.h
 * for the contract and the list of replacements of nondeterministic reference
 * behaviour.  Built with -ffp-contract=off so that every a*b+c is two roundings
 * unless written as fma(); the HIP kernels are built the same way and use the
 * same explicit fma() forms, which makes per-step results bit-identical.
 */
#include "pic_oracle.h"

#include <cmath>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Replaces the cuRAND    */
/* XORWOW host generator of mcc.cu:123-145.                                   */
/* ------------------------------------------------------------------------ */
extern "C" void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Philox2x32-10 (same paper).  The collision-candidacy draw of every       */
/* particle and step: first word over counter {id, step} with key            */
/* cand_key = first word over {seed, case} with key species.                 */
extern "C" void orc_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], k = key;
    for (int r = 0; r < 10; ++r) {
        if (r) k += 0x9E3779B9u;
        uint64_t p = (uint64_t)0xD256D193u * c0;
        uint32_t hi = (uint32_t)(p >> 32), lo = (uint32_t)p;
        c0 = hi ^ k ^ c1;
        c1 = lo;
    }
    out[0] = c0; out[1] = c1;
}

static uint32_t cand_key(uint32_t seed, uint32_t case_id, uint32_t species) {
    const uint32_t c[2] = {seed, case_id};
    uint32_t o[2];
    orc_philox2x32_10(c, species, o);
    return o[0];
}

template <typename T>
static T cand_uni(uint32_t id, uint32_t step, uint32_t key) {
    const uint32_t c[2] = {id, step};
    uint32_t o[2];
    orc_philox2x32_10(c, key, o);
    return (T)(o[0] >> 8) * (T)0x1p-24;
}

/* Deterministic natural log: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
   log m = 2 atanh(f), f = (m-1)/(m+1), odd series to f^23, Horner with fma. */
extern "C" double orc_det_log(double x) {
    static const double c[11] = {
        0.3333333333333333, 0.2, 0.14285714285714285, 0.1111111111111111,
        0.09090909090909091, 0.07692307692307693, 0.06666666666666667,
        0.058823529411764705, 0.05263157894736842, 0.047619047619047616,
        0.043478260869565216};
    uint64_t u;
    std::memcpy(&u, &x, 8);
    int e = (int)((u >> 52) & 0x7ff) - 1022;
    uint64_t mb = (u & 0x000fffffffffffffull) | (0x3feull << 52);   /* m in [0.5,1) */
    double m;
    std::memcpy(&m, &mb, 8);
    if (m < 0.70710678118654752) { m = m * 2.0; e -= 1; }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
    double p = c[10];
    for (int k = 9; k >= 0; --k) p = std::fma(p, f2, c[k]);
    double s = f2 * p;
    double tf = 2.0 * f;
    double lm = std::fma(tf, s, tf);
    double de = (double)e;
    double r = std::fma(de, 0x1.df473de6af279p-26, lm);
    return std::fma(de, 0x1.62e42f0000000p-1, r);
}

/* ------------------------------------------------------------------------ */
/* E = -grad(phi), kernels.cu:134-191.  Central differences in the interior,  */
/* one-sided differences on edges and corners, applied per axis.              */
/* ------------------------------------------------------------------------ */
static inline double ex_at(const double* phi, int nx, int i, int j, double dx) {
    const double* r = phi + (size_t)j * nx;
    if (i == 0) return -(r[1] - r[0]) / dx;
    if (i == nx - 1) return -(r[i] - r[i - 1]) / dx;
    return -(r[i + 1] - r[i - 1]) / (2.0 * dx);
}
static inline double ey_at(const double* phi, int nx, int ny, int i, int j, double dy) {
    if (j == 0) return -(phi[(size_t)nx + i] - phi[i]) / dy;
    if (j == ny - 1) return -(phi[(size_t)j * nx + i] - phi[(size_t)(j - 1) * nx + i]) / dy;
    return -(phi[(size_t)(j + 1) * nx + i] - phi[(size_t)(j - 1) * nx + i]) / (2.0 * dy);
}

extern "C" void orc_efield(const double* phi, int nx, int ny, double dx, double dy,
                           double* ex, double* ey) {
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            ex[(size_t)j * nx + i] = ex_at(phi, nx, i, j, dx);
            ey[(size_t)j * nx + i] = ey_at(phi, nx, ny, i, j, dy);
        }
}

/* ------------------------------------------------------------------------ */
/* CIC deposit in Q20 fixed point (replaces the FP64 atomics of              */
/* kernels.cu:26-69; same cell test "skip outside [0,m-2]x[0,n-2]").          */
/* ------------------------------------------------------------------------ */
template <typename T>
static inline void deposit_one(T x, T y, int nx, int ny, uint64_t* acc) {
    int ix = (int)std::floor(x), iy = (int)std::floor(y);
    if (ix < 0 || ix >= nx - 1 || iy < 0 || iy >= ny - 1) return;
    T fx = x - (T)ix, fy = y - (T)iy;
    const uint64_t W = 1ull << ORC_QBITS;
    uint64_t qx = (uint64_t)(uint32_t)(fx * (T)W);
    uint64_t qy = (uint64_t)(uint32_t)(fy * (T)W);
    size_t b = (size_t)iy * nx + ix;
    acc[b] += (W - qx) * (W - qy);
    acc[b + 1] += qx * (W - qy);
    acc[b + nx] += (W - qx) * qy;
    acc[b + nx + 1] += qx * qy;
}

extern "C" void orc_deposit(int prec, const void* x, const void* y, int64_t n, int nx, int ny,
                            uint64_t* acc) {
    for (int64_t i = 0; i < n; ++i) {
        if (prec == 4) deposit_one(((const float*)x)[i], ((const float*)y)[i], nx, ny, acc);
        else deposit_one(((const double*)x)[i], ((const double*)y)[i], nx, ny, acc);
    }
}

extern "C" void orc_charge_to_rhs(const uint64_t* const* acc, const double* cfac, int nspec,
                                  int64_t nn, double* rhs) {
    for (int64_t i = 0; i < nn; ++i) {
        double r = 0.0;
        for (int s = 0; s < nspec; ++s) r = std::fma(cfac[s], (double)acc[s][i], r);
        rhs[i] = r;
    }
}

/* ------------------------------------------------------------------------ */
/* Multigrid, Multigrid.cu.                                                   */
/* ------------------------------------------------------------------------ */
/* One colour of red-black SOR, Multigrid.cu:192-227 (same expression form).
   mask (may be null): nodes with a nonzero mask are held fixed, as the
   embedded-boundary smoother of Multigrid_embedded.cu:14-33 holds its
   Dirichlet and solid nodes. */
static void rb_colour(double* phi, const double* rhs, int n, double h2, double omega, int colour,
                      const uint8_t* mask = nullptr) {
    /* the nodes of one colour are independent: any order gives the same values */
#pragma omp parallel for schedule(static) if (n > 256)
    for (int j = 1; j < n - 1; ++j)
        for (int i = 1; i < n - 1; ++i) {
            if (((i + j) & 1) != colour) continue;
            size_t idx = (size_t)j * n + i;
            if (mask && mask[idx]) continue;
            double c = phi[idx];
            double s = phi[idx + n] + phi[idx - n] + phi[idx + 1] + phi[idx - 1];
            double nv = 0.25 * (s - h2 * rhs[idx]);
            phi[idx] = std::fma(omega, nv, (1.0 - omega) * c);
        }
}

/* smooth(): odd colour first, then even (Multigrid.cu:501-504). */
static void smooth_m(double* phi, const double* rhs, int n, double h, double omega, int iters, const uint8_t* mask) {
    double h2 = h * h;
    for (int it = 0; it < iters; ++it) {
        rb_colour(phi, rhs, n, h2, omega, 1, mask);
        rb_colour(phi, rhs, n, h2, omega, 0, mask);
    }
}

extern "C" void orc_mg_smooth(double* phi, const double* rhs, int n, double h, double omega, int iters) {
    smooth_m(phi, rhs, n, h, omega, iters, nullptr);
}

/* residual_2d, Multigrid.cu:269-297 (zero at fixed nodes of a mask, as
   Multigrid_embedded.cu:36-54). */
static void residual_m(double* res, const double* phi, const double* rhs, int n, double h, const uint8_t* mask) {
    double inv_h2 = 1.0 / (h * h);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            size_t idx = (size_t)j * n + i;
            if (i == 0 || i == n - 1 || j == 0 || j == n - 1 || (mask && mask[idx])) { res[idx] = 0.0; continue; }
            double c = phi[idx];
            double lap = (phi[idx - n] + phi[idx + n] + phi[idx - 1] + phi[idx + 1] - 4.0 * c) * inv_h2;
            res[idx] = rhs[idx] - lap;
        }
}

extern "C" void orc_mg_residual(double* res, const double* phi, const double* rhs, int n, double h) {
    residual_m(res, phi, rhs, n, h, nullptr);
}

/* restrict_2d, Multigrid.cu:299-349: full weighting, out-of-range terms dropped. */
extern "C" void orc_mg_restrict(double* coarse, const double* fine, int nf) {
    int nc = (nf + 1) / 2;
    for (int jc = 0; jc < nc; ++jc)
        for (int ic = 0; ic < nc; ++ic) {
            int i = 2 * ic, j = 2 * jc;
            if (i >= nf || j >= nf) { coarse[(size_t)jc * nc + ic] = 0.0; continue; }
            size_t f = (size_t)j * nf + i;
            double s = 4.0 * fine[f];
            if (i > 0) s += 2.0 * fine[f - 1];
            if (i < nf - 1) s += 2.0 * fine[f + 1];
            if (j > 0) s += 2.0 * fine[f - nf];
            if (j < nf - 1) s += 2.0 * fine[f + nf];
            if (i > 0 && j > 0) s += fine[f - nf - 1];
            if (i < nf - 1 && j > 0) s += fine[f - nf + 1];
            if (i > 0 && j < nf - 1) s += fine[f + nf - 1];
            if (i < nf - 1 && j < nf - 1) s += fine[f + nf + 1];
            coarse[(size_t)jc * nc + ic] = s * 0.0625;
        }
}

/* prolong_2d, Multigrid.cu:351-407: fine += bilinear(coarse), clamped at the
   far edge; fixed nodes of a mask take no correction (Multigrid_embedded.cu:89-104). */
static void prolong_m(double* fine, const double* coarse, int nc, const uint8_t* mask) {
    int nf = 2 * (nc - 1) + 1;
    for (int j = 0; j < nf; ++j)
        for (int i = 0; i < nf; ++i) {
            if (mask && mask[(size_t)j * nf + i]) continue;
            int ic = i / 2, jc = j / 2;
            size_t c = (size_t)jc * nc + ic;
            bool io = i & 1, jo = j & 1;
            double v;
            if (!io && !jo) v = coarse[c];
            else if (io && !jo) v = (ic + 1 < nc) ? 0.5 * (coarse[c] + coarse[c + 1]) : coarse[c];
            else if (!io && jo) v = (jc + 1 < nc) ? 0.5 * (coarse[c] + coarse[c + nc]) : coarse[c];
            else v = (ic + 1 < nc && jc + 1 < nc)
                         ? 0.25 * (coarse[c] + coarse[c + 1] + coarse[c + nc] + coarse[c + nc + 1])
                         : coarse[c];
            fine[(size_t)j * nf + i] += v;
        }
}

extern "C" void orc_mg_prolong(double* fine, const double* coarse, int nc) { prolong_m(fine, coarse, nc, nullptr); }

static double sumsq(const double* v, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += v[i] * v[i];
    return s;
}

/* automatic ladder: coarsen while the node count is odd and above 3; the
   coarsest level is then 3x3 or 4x4 (checked by the caller) */
static int mg_levels_for(int n, int requested) {
    if (requested > 0) return requested;
    int L = 1;
    while (n > 3 && (n & 1)) { n = (n + 1) / 2; ++L; }
    return L;
}

/* solve_multigrid, Multigrid.cu:573-642 with v_cycle :539-569.  Differences:
   the relative residual is ||r||/||rhs|| (the reference divides by sum rhs^2),
   and the residual is tested before the first cycle (a converged warm start
   costs no V-cycle).  phi keeps its boundary values (Dirichlet). */
extern "C" int orc_mg_solve(double* phi, const double* rhs, int nx, int ny, double h,
                            const orc_mg_params* p, orc_mg_stats* st) {
    return orc_mg_solve_masked(phi, rhs, nx, ny, h, p, nullptr, nullptr, st);
}

/* Embedded electrodes (Multigrid_embedded.cu:7-190): mask 0 = solved node,
   1 = Dirichlet node held at phi_bc, 2 = solid node held at its value.  The
   mask of a coarse level is that of the fine node it sits on (injection,
   :56-87); fixed nodes are skipped by the smoother, have zero residual and
   take no prolongation.  Restriction and prolongation otherwise keep the
   full-weighting / bilinear forms of Multigrid.cu (the embedded file's
   injection prolongation and its red-black precedence defect, D11, are not
   reproduced). */
extern "C" int orc_mg_solve_masked(double* phi, const double* rhs, int nx, int ny, double h, const orc_mg_params* p,
                                   const uint8_t* mask, const double* phi_bc, orc_mg_stats* st) {
    if (nx != ny || nx < 3) return -1;
    int L = mg_levels_for(nx, p->levels);
    std::vector<int> ns(L);
    ns[0] = nx;
    for (int l = 1; l < L; ++l) {
        if ((ns[l - 1] & 1) == 0) return -2;   /* ladder must stay odd */
        ns[l] = (ns[l - 1] + 1) / 2;
    }
    if (p->levels <= 0 && ns[L - 1] != 3 && ns[L - 1] != 4) return -3;
    std::vector<std::vector<double>> ph(L), rh(L), rs(L);
    for (int l = 0; l < L; ++l) {
        size_t sz = (size_t)ns[l] * ns[l];
        rs[l].assign(sz, 0.0);
        if (l) { ph[l].assign(sz, 0.0); rh[l].assign(sz, 0.0); }
    }
    size_t nn = (size_t)nx * nx;
    /* mask pyramid by injection; Dirichlet values applied to phi */
    std::vector<std::vector<uint8_t>> mk(L);
    if (mask) {
        mk[0].assign(mask, mask + nn);
        for (int l = 1; l < L; ++l) {
            mk[l].assign((size_t)ns[l] * ns[l], 0);
            for (int jc = 0; jc < ns[l]; ++jc)
                for (int ic = 0; ic < ns[l]; ++ic)
                    mk[l][(size_t)jc * ns[l] + ic] = mk[l - 1][(size_t)(2 * jc) * ns[l - 1] + 2 * ic];
        }
        for (size_t k = 0; k < nn; ++k)
            if (mask[k] == 1) phi[k] = phi_bc[k];
    }
    auto M = [&](int l) -> const uint8_t* { return mask ? mk[l].data() : nullptr; };
    double rhs_norm = std::sqrt(sumsq(rhs, nn));
    auto rel_res = [&]() {
        residual_m(rs[0].data(), phi, rhs, nx, h, M(0));
        double r = std::sqrt(sumsq(rs[0].data(), nn));
        return rhs_norm > 0.0 ? r / rhs_norm : r;
    };
    double rel = rel_res();
    int cyc = 0;
    while (!(rel < p->tol) && cyc < p->max_cycles) {
        double hh = h;
        for (int l = 0; l < L - 1; ++l) {
            double* P = l ? ph[l].data() : phi;
            const double* R = l ? rh[l].data() : rhs;
            smooth_m(P, R, ns[l], hh, p->omega, p->per_level ? p->pre_level[l] : p->pre, M(l));
            residual_m(rs[l].data(), P, R, ns[l], hh, M(l));
            orc_mg_restrict(rh[l + 1].data(), rs[l].data(), ns[l]);
            std::fill(ph[l + 1].begin(), ph[l + 1].end(), 0.0);
            hh *= 2.0;
        }
        {
            double* P = (L - 1) ? ph[L - 1].data() : phi;
            const double* R = (L - 1) ? rh[L - 1].data() : rhs;
            /* a 4x4 coarsest level of an automatic ladder gets at least 32 sweeps */
            const int cs = (p->levels <= 0 && ns[L - 1] == 4 && p->coarse_sweeps < 32) ? 32 : p->coarse_sweeps;
            smooth_m(P, R, ns[L - 1], hh, p->omega, cs, M(L - 1));
        }
        for (int l = L - 2; l >= 0; --l) {
            hh /= 2.0;
            double* P = l ? ph[l].data() : phi;
            const double* R = l ? rh[l].data() : rhs;
            prolong_m(P, ph[l + 1].data(), ns[l + 1], M(l));
            smooth_m(P, R, ns[l], hh, p->omega, p->per_level ? p->post_level[l] : p->post, M(l));
        }
        ++cyc;
        rel = rel_res();
    }
    if (st) { st->cycles = cyc; st->rel_residual = rel; st->rhs_norm = rhs_norm; }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Wall test, collisions.cu:7-54 (Bourke u_a/u_b, inclusive, parallel cut).   */
/* ------------------------------------------------------------------------ */
extern "C" int orc_segment_hits(double x1, double y1, double x2, double y2,
                                const double* walls, int n_walls, double eps) {
    for (int j = 0; j < n_walls; ++j) {
        double x3 = walls[4 * j], y3 = walls[4 * j + 1], x4 = walls[4 * j + 2], y4 = walls[4 * j + 3];
        double den = (y4 - y3) * (x2 - x1) - (x4 - x3) * (y2 - y1);
        if (std::fabs(den) < eps) continue;
        double ua = ((x4 - x3) * (y1 - y3) - (y4 - y3) * (x1 - x3)) / den;
        double ub = ((x2 - x1) * (y1 - y3) - (y2 - y1) * (x1 - x3)) / den;
        if (ua >= 0.0 && ua <= 1.0 && ub >= 0.0 && ub <= 1.0) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-particle random stream: Philox counter {id, step, call, case},        */
/* key {seed, species}.  Uniforms carry 24 bits: u = (r >> 8) * 2^-24.        */
/* ------------------------------------------------------------------------ */
struct Stream {
    uint32_t ctr[4], key[2], buf[4];
    int idx;
    Stream(uint32_t id, uint32_t step, uint32_t case_id, uint32_t seed, uint32_t species) {
        ctr[0] = id; ctr[1] = step; ctr[2] = 0; ctr[3] = case_id;
        key[0] = seed; key[1] = species;
        orc_philox4x32_10(ctr, key, buf);
        idx = 0;
    }
    uint32_t u32() {
        if (idx == 4) { ctr[2] += 1; orc_philox4x32_10(ctr, key, buf); idx = 0; }
        return buf[idx++];
    }
    template <typename T> T uni() { return (T)(u32() >> 8) * (T)0x1p-24; }
};

/* isotropic 2D direction by rejection from the square (no trigonometry) */
template <typename T>
static bool iso_dir(Stream& rs, T& c, T& s) {
    for (int t = 0; t < 16; ++t) {
        T a = (T)2 * rs.uni<T>() - (T)1;
        T b = (T)2 * rs.uni<T>() - (T)1;
        T q = std::fma(a, a, b * b);
        if (q > (T)0 && q <= (T)1) {
            c = std::fma(a, a, -(b * b)) / q;
            s = ((T)2 * a) * b / q;
            return true;
        }
    }
    c = (T)1; s = (T)0;
    return false;
}

/* Gaussian pair by the Marsaglia polar method with det_log, scaled by sigma */
template <typename T>
static bool gauss_pair(Stream& rs, double sigma, T& g1, T& g2) {
    for (int t = 0; t < 16; ++t) {
        T a = (T)2 * rs.uni<T>() - (T)1;
        T b = (T)2 * rs.uni<T>() - (T)1;
        T q = std::fma(a, a, b * b);
        if (q > (T)0 && q < (T)1) {
            double qd = (double)q;
            double f = std::sqrt(-2.0 * orc_det_log(qd) / qd);
            g1 = (T)(sigma * ((double)a * f));
            g2 = (T)(sigma * ((double)b * f));
            return true;
        }
    }
    g1 = (T)0; g2 = (T)0;
    return false;
}

/* ------------------------------------------------------------------------ */
/* Fused particle step for one species (main.cu:37-140 stage order with the   */
/* kick of step n-1 moved to the front of step n):                            */
/*   kick (updateParticleVelocities, kernels.cu:71-131)                       */
/*   MCC  (null-collision form of collision_kernel, mcc.cu:56-85)             */
/*   drift (update_coordinates, kernels.cu:8-23)                              */
/*   walls (detect_intersections, collisions.cu:7-54; box fast path)          */
/*   compaction (removeCollidedParticles, particles.cu:217-261; set semantics)*/
/*   deposit (updateDensity, kernels.cu:26-69; fixed point)                   */
/* ------------------------------------------------------------------------ */
template <typename T>
struct PP { T x, y, vx, vy; uint32_t id; };

template <typename T>
static inline bool absorbed(const orc_push_params* p, T x0, T y0, T x1, T y1) {
    if (p->box_walls) {
        if (x1 <= (T)0 || x1 >= (T)(p->nx - 1) || y1 <= (T)0 || y1 >= (T)(p->ny - 1)) return true;
    }
    if (p->n_walls > 0)
        return orc_segment_hits((double)x0, (double)y0, (double)x1, (double)y1, p->walls, p->n_walls,
                                p->wall_eps) != 0;
    return false;
}

/* energy band of the null-collision majorant (same operations as the kernels) */
template <typename T>
static inline int mcc_band(T E, T inv_emax) {
    const T r = E * inv_emax;
    if (!(r >= (T)0x1p-14)) return 0;
    int e = 0;
    (void)std::frexp(r, &e);
    return std::min(14 + e, 15);
}

/* One particle of a pass: kick, collisions, drift, walls, deposit.  emit(kind,
   particle) receives the survivor (kind 0), a born particle of the same species
   (1) and a born particle of the other species (2), in that order. */
template <typename T, typename Emit>
static inline void push_one(const orc_push_params* p, const double* phi, const T* tab, PP<T> q,
                            uint64_t* acc_self, uint64_t* acc_other, uint64_t* cnt, Emit&& emit) {
    const int nx = p->nx, ny = p->ny;
    const T kq = (T)p->kq, dtdx = (T)p->dtdx, dtdy = (T)p->dtdy;
    const T efac = (T)p->efac, inv_de = (T)p->inv_de;
    const int K = p->K;
    T x = q.x, y = q.y, vx = q.vx, vy = q.vy;
    const uint32_t id = q.id;
    if (p->do_kick) {
        int ix = (int)std::floor(x), iy = (int)std::floor(y);
        if (!(ix < 0 || ix >= nx - 1 || iy < 0 || iy >= ny - 1)) {
            T fx = x - (T)ix, fy = y - (T)iy;
            T gx = (T)1 - fx, gy = (T)1 - fy;
            T w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
            T e[4][2];
            const int ci[4] = {ix, ix + 1, ix, ix + 1}, cj[4] = {iy, iy, iy + 1, iy + 1};
            for (int c = 0; c < 4; ++c) {
                e[c][0] = (T)ex_at(phi, nx, ci[c], cj[c], p->dx);
                e[c][1] = (T)ey_at(phi, nx, ny, ci[c], cj[c], p->dy);
            }
            T Ex = std::fma(w11, e[3][0], std::fma(w01, e[2][0], std::fma(w10, e[1][0], w00 * e[0][0])));
            T Ey = std::fma(w11, e[3][1], std::fma(w01, e[2][1], std::fma(w10, e[1][1], w00 * e[0][1])));
            vx = std::fma(kq, Ex, vx);
            vy = std::fma(kq, Ey, vy);
        }
    }
    bool have_born = false;
    PP<T> be{}, bi{};
    if (p->do_mcc && p->n_proc > 0) {
        T u0 = cand_uni<T>(id, p->step, cand_key(p->seed, p->case_id, p->species));
        Stream rs(id, p->step, p->case_id, p->seed, p->species);
        T u1 = rs.uni<T>();
        T v2 = std::fma(vx, vx, vy * vy);
        T E = efac * v2;
        const int band = mcc_band(E, (T)p->inv_emax);
        if (u0 < (T)p->band_p_null[band]) {
            const T nu_max = (T)p->band_nu_max[band];
            cnt[ORC_C_NULL_CAND]++;
            T v = std::sqrt(v2);
            T s = E * inv_de;
            int j;
            T t;
            if (s < (T)(K - 1)) { j = (int)s; t = s - (T)j; }
            else { j = K - 2; t = (T)1; }
            T R = u1 * nu_max;
            T cum = (T)0;
            int sel = -1;
            for (int qq = 0; qq < p->n_proc; ++qq) {
                if (E < (T)p->proc_thr[qq]) continue;
                T a = tab[(size_t)qq * K + j], b = tab[(size_t)qq * K + j + 1];
                T sig = std::fma(t, b - a, a);
                cum = std::fma(sig, v, cum);
                if (sel < 0 && R < cum) sel = qq;
            }
            if (cum > nu_max) cnt[ORC_C_NUMAX_EXCEEDED]++;
            if (sel >= 0) {
                cnt[ORC_C_PROC0 + sel]++;
                int ty = p->proc_type[sel];
                T thr = (T)p->proc_thr[sel];
                T c, sn;
                if (ty == ORC_ELASTIC) {
                    if (!iso_dir(rs, c, sn)) cnt[ORC_C_REJECT_FALLBACK]++;
                    vx = v * c; vy = v * sn;
                } else if (ty == ORC_EXCITATION) {
                    T e1 = E - thr;
                    T v1 = std::sqrt(e1 / efac);
                    if (!iso_dir(rs, c, sn)) cnt[ORC_C_REJECT_FALLBACK]++;
                    vx = v1 * c; vy = v1 * sn;
                } else if (ty == ORC_IONIZATION) {
                    T er = E - thr;
                    T us = rs.uni<T>();
                    T e1 = er * us;
                    T e2 = er - e1;
                    T v1 = std::sqrt(e1 / efac), v2s = std::sqrt(e2 / efac);
                    if (!iso_dir(rs, c, sn)) cnt[ORC_C_REJECT_FALLBACK]++;
                    vx = v1 * c; vy = v1 * sn;
                    if (!iso_dir(rs, c, sn)) cnt[ORC_C_REJECT_FALLBACK]++;
                    be.x = x; be.y = y; be.vx = v2s * c; be.vy = v2s * sn;
                    T g1, g2;
                    if (!gauss_pair(rs, p->gas_vth, g1, g2)) cnt[ORC_C_REJECT_FALLBACK]++;
                    bi.x = x; bi.y = y; bi.vx = g1; bi.vy = g2;
                    be.id = rs.u32();
                    bi.id = rs.u32();
                    have_born = true;
                } else { /* charge exchange: take a gas-atom velocity */
                    T g1, g2;
                    if (!gauss_pair(rs, p->gas_vth, g1, g2)) cnt[ORC_C_REJECT_FALLBACK]++;
                    vx = g1; vy = g2;
                }
            }
        }
    }
    T x0 = x, y0 = y;
    if (p->do_drift) {
        x = std::fma(vx, dtdx, x);
        y = std::fma(vy, dtdy, y);
    }
    if (absorbed(p, x0, y0, x, y)) {
        cnt[ORC_C_ABSORBED]++;
    } else {
        deposit_one(x, y, nx, ny, acc_self);
        emit(0, PP<T>{x, y, vx, vy, id});
    }
    if (have_born) {
        /* born particles drift from the collision point when the pass drifts */
        PP<T> b[2] = {be, bi};
        for (int w = 0; w < 2; ++w) {
            T bx = p->do_drift ? std::fma(b[w].vx, dtdx, b[w].x) : b[w].x;
            T by = p->do_drift ? std::fma(b[w].vy, dtdy, b[w].y) : b[w].y;
            if (absorbed(p, b[w].x, b[w].y, bx, by)) { cnt[ORC_C_BORN_ABSORBED]++; continue; }
            b[w].x = bx; b[w].y = by;
            if (w == 0) {
                deposit_one(bx, by, nx, ny, acc_self);
                emit(1, b[w]);
                cnt[ORC_C_BORN_SELF]++;
            } else {
                deposit_one(bx, by, nx, ny, acc_other);
                emit(2, b[w]);
                cnt[ORC_C_BORN_OTHER]++;
            }
        }
    }
}

/* A pass over one species.  Survivors keep their order and are followed by the
   born particles of this species in the order of creation; born particles of
   the other species go to the *O arrays.  The OpenMP build splits the input into
   one contiguous range per thread, collects each range's results privately and
   concatenates them in range order, so its output equals the serial one
   (fixed-point charge is summed exactly). */
template <typename T>
static int64_t push_t(const orc_push_params* p, const double* phi,
                      T* X, T* Y, T* VX, T* VY, uint32_t* ID, int64_t n, int64_t cap,
                      uint64_t* acc_self, T* XO, T* YO, T* VXO, T* VYO, uint32_t* IDO,
                      int64_t cap_o, int64_t* n_other, uint64_t* acc_other, uint64_t* cnt) {
    const int K = p->K;
    std::vector<T> tab;
    if (p->do_mcc && p->n_proc > 0) {
        tab.resize((size_t)p->n_proc * K);
        for (size_t i = 0; i < tab.size(); ++i) tab[i] = (T)p->tab[i];
    }
    int64_t no = 0, k = 0;
    auto put_other = [&](const PP<T>& b) {
        if (no < cap_o) { XO[no] = b.x; YO[no] = b.y; VXO[no] = b.vx; VYO[no] = b.vy; IDO[no] = b.id; }
        ++no;
    };
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
    if (n < 65536) nt = 1;
#endif
    if (nt == 1) {
        std::vector<PP<T>> born_self;
        for (int64_t i = 0; i < n; ++i) {
            push_one(p, phi, tab.data(), PP<T>{X[i], Y[i], VX[i], VY[i], ID[i]}, acc_self, acc_other, cnt,
                     [&](int kind, const PP<T>& r) {
                         if (kind == 0) { X[k] = r.x; Y[k] = r.y; VX[k] = r.vx; VY[k] = r.vy; ID[k] = r.id; ++k; }
                         else if (kind == 1) born_self.push_back(r);
                         else put_other(r);
                     });
        }
        for (auto& b : born_self) {
            if (k < cap) { X[k] = b.x; Y[k] = b.y; VX[k] = b.vx; VY[k] = b.vy; ID[k] = b.id; }
            ++k;
        }
        if (n_other) *n_other = no;
        return k;
    }
#ifdef _OPENMP
    const size_t nn = (size_t)p->nx * p->ny;
    struct Part {
        std::vector<PP<T>> out[3];
        std::vector<uint64_t> acc[2];
        uint64_t cnt[ORC_NCOUNT];
    };
    std::vector<Part> parts(nt);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        Part& pt = parts[t];
        pt.acc[0].assign(nn, 0);
        if (acc_other) pt.acc[1].assign(nn, 0);
        std::fill(pt.cnt, pt.cnt + ORC_NCOUNT, 0ull);
        const int64_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
        pt.out[0].reserve(i1 - i0);
        for (int64_t i = i0; i < i1; ++i)
            push_one(p, phi, tab.data(), PP<T>{X[i], Y[i], VX[i], VY[i], ID[i]}, pt.acc[0].data(),
                     acc_other ? pt.acc[1].data() : nullptr, pt.cnt,
                     [&](int kind, const PP<T>& r) { pt.out[kind].push_back(r); });
    }
    for (int kind = 0; kind < 2; ++kind)
        for (int t = 0; t < nt; ++t)
            for (const auto& r : parts[t].out[kind]) {
                if (k < cap) { X[k] = r.x; Y[k] = r.y; VX[k] = r.vx; VY[k] = r.vy; ID[k] = r.id; }
                ++k;
            }
    for (int t = 0; t < nt; ++t)
        for (const auto& r : parts[t].out[2]) put_other(r);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nn; ++i) {
        uint64_t a = 0, b = 0;
        for (int t = 0; t < nt; ++t) {
            a += parts[t].acc[0][i];
            if (acc_other) b += parts[t].acc[1][i];
        }
        acc_self[i] += a;
        if (acc_other) acc_other[i] += b;
    }
    for (int t = 0; t < nt; ++t)
        for (int c = 0; c < ORC_NCOUNT; ++c) cnt[c] += parts[t].cnt[c];
#endif
    if (n_other) *n_other = no;
    return k;
}

extern "C" int64_t orc_push(const orc_push_params* p, const double* phi,
                            void* x, void* y, void* vx, void* vy, uint32_t* id, int64_t n, int64_t cap,
                            uint64_t* acc_self,
                            void* xo, void* yo, void* vxo, void* vyo, uint32_t* ido, int64_t cap_o,
                            int64_t* n_other, uint64_t* acc_other, uint64_t* counters) {
    if (p->prec == 4)
        return push_t<float>(p, phi, (float*)x, (float*)y, (float*)vx, (float*)vy, id, n, cap, acc_self,
                             (float*)xo, (float*)yo, (float*)vxo, (float*)vyo, ido, cap_o, n_other,
                             acc_other, counters);
    return push_t<double>(p, phi, (double*)x, (double*)y, (double*)vx, (double*)vy, id, n, cap, acc_self,
                          (double*)xo, (double*)yo, (double*)vxo, (double*)vyo, ido, cap_o, n_other,
                          acc_other, counters);
}

/* ------------------------------------------------------------------------ */
/* Seeded initial state (replaces main.cu:17-32, which used random_device).  */
/* ------------------------------------------------------------------------ */
template <typename T>
static void init_t(int64_t n, int nx, int ny, double vth, uint32_t seed, uint32_t case_id,
                   uint32_t species, uint32_t id0, T* x, T* y, T* vx, T* vy, uint32_t* id) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t pid = id0 + (uint32_t)i;
        Stream rs(pid, 0xFFFFFFFFu, case_id, seed, species);
        uint32_t rx = rs.u32(), ry = rs.u32();
        x[i] = (T)(((double)rx + 0.5) * 0x1p-32 * (double)(nx - 1));
        y[i] = (T)(((double)ry + 0.5) * 0x1p-32 * (double)(ny - 1));
        double g1 = 0.0, g2 = 0.0;
        gauss_pair<double>(rs, vth, g1, g2);
        vx[i] = (T)g1;
        vy[i] = (T)g2;
        id[i] = pid;
    }
}

/* Uniform positions in a box and uniform velocities in a range, the
   distributions of initializeParticles (main.cu:17-32) made reproducible:
   Philox counter {i, 0xFFFFFFFE, c, case_id}, key {seed, species}; the four
   words of the first block give x, y, vx, vy. */
template <typename T>
static void init_uniform_t(int64_t n, double x0, double x1, double y0, double y1, double v0, double v1,
                           uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0, T* x, T* y, T* vx,
                           T* vy, uint32_t* id) {
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t pid = id0 + (uint32_t)i;
        Stream rs(pid, 0xFFFFFFFEu, case_id, seed, species);
        const double ux = ((double)rs.u32() + 0.5) * 0x1p-32, uy = ((double)rs.u32() + 0.5) * 0x1p-32;
        const double uvx = ((double)rs.u32() + 0.5) * 0x1p-32, uvy = ((double)rs.u32() + 0.5) * 0x1p-32;
        x[i] = (T)(x0 + ux * (x1 - x0));
        y[i] = (T)(y0 + uy * (y1 - y0));
        vx[i] = (T)(v0 + uvx * (v1 - v0));
        vy[i] = (T)(v0 + uvy * (v1 - v0));
        id[i] = pid;
    }
}

extern "C" void orc_init_uniform(int prec, int64_t n, double x0, double x1, double y0, double y1, double v0,
                                 double v1, uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0,
                                 void* x, void* y, void* vx, void* vy, uint32_t* id) {
    if (prec == 4)
        init_uniform_t<float>(n, x0, x1, y0, y1, v0, v1, seed, case_id, species, id0, (float*)x, (float*)y,
                              (float*)vx, (float*)vy, id);
    else
        init_uniform_t<double>(n, x0, x1, y0, y1, v0, v1, seed, case_id, species, id0, (double*)x, (double*)y,
                               (double*)vx, (double*)vy, id);
}

extern "C" void orc_init_particles(int prec, int64_t n, int nx, int ny, double vth,
                                   uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0,
                                   void* x, void* y, void* vx, void* vy, uint32_t* id) {
    if (prec == 4) init_t<float>(n, nx, ny, vth, seed, case_id, species, id0, (float*)x, (float*)y,
                                 (float*)vx, (float*)vy, id);
    else init_t<double>(n, nx, ny, vth, seed, case_id, species, id0, (double*)x, (double*)y,
                        (double*)vx, (double*)vy, id);
}
