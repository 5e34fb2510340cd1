// push.hip -- particle kernels for gfx950.
//
//  push_kernel    one fused pass per step: CIC gather of E from an LDS tile (E is
//                 computed from phi with the edge rules of kernels.cu:134-191),
//                 kick (kernels.cu:71-131), null-collision MCC (replaces
//                 mcc.cu:56-85), drift (kernels.cu:8-23), wall test
//                 (collisions.cu:7-54), Q20 fixed-point CIC deposit into an LDS
//                 tile flushed with one 64-bit atomic per node (replaces the FP64
//                 atomics of kernels.cu:26-69) and wave-ballot compaction of the
//                 survivors into the bin of their new cell (replaces
//                 removeCollidedParticles, particles.cu:217-261).
//  init_kernel    seeded particles: uniform positions with Maxwellian velocities,
//  init_uniform_kernel  or uniform in a box and velocity range (replace main.cu:17-32).
//  hist/scatter   re-binning of a species store (counting sort by tile).
//  gather_kernel  flattening of a store for download.
//  export_kernel  live positions as a packed float2 stream (replaces
//                 convertPositionsToFloat, kernels.cu:268-273).
//  end_step       buffer-parity flip and step counter, one block per case.
//
// Bin layout.  Bin b of a buffer is [start, start+cap): survivors that stay in
// their bin fill the front part [start, start+cap-cap/4) from slots handed out
// by a counter in LDS (only the block that owns the bin writes there); particles
// that move in from other bins, and born particles, fill the back part downwards
// from start+cap-1 with one global atomic per destination bin per wave.  Movers
// are staged in LDS and written in a flush phase, so no global atomic sits on
// the path of a particle that stays.  Overflow of either part goes to the spill
// region (index nb).
#include "pic_device.hpp"
#include "field.hpp"
#include "step_device.hpp"
#include "push.hpp"
#include <type_traits>

namespace pic {

__device__ inline float dsqrt(float v) { return sqrtf(v); }
__device__ inline double dsqrt(double v) { return sqrt(v); }
__device__ inline float dfloor(float v) { return floorf(v); }
__device__ inline double dfloor(double v) { return floor(v); }
__device__ inline float dfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double dfma(double a, double b, double c) { return fma(a, b, c); }

// E = -grad(phi) at a node, kernels.cu:134-191 (per-axis one-sided at edges)
__device__ inline double ex_node(const double* __restrict__ phi_, int nx, int i, int j, double dx) {
    const gptr<const double> r = G(phi_) + (size_t)j * nx;
    if (i == 0) return -(r[1] - r[0]) / dx;
    if (i == nx - 1) return -(r[i] - r[i - 1]) / dx;
    return -(r[i + 1] - r[i - 1]) / (2.0 * dx);
}
__device__ inline double ey_node(const double* __restrict__ phi_, int nx, int ny, int i, int j, double dy) {
    const gptr<const double> phi = G(phi_);
    if (j == 0) return -(phi[(size_t)nx + i] - phi[i]) / dy;
    if (j == ny - 1) return -(phi[(size_t)j * nx + i] - phi[(size_t)(j - 1) * nx + i]) / dy;
    return -(phi[(size_t)(j + 1) * nx + i] - phi[(size_t)(j - 1) * nx + i]) / (2.0 * dy);
}

// a species' arrays of one parity (pointers into its packed blocks)
template <typename T>
struct PartsOf {
    T* x;
    T* y;
    T* vx;
    T* vy;
    uint32_t* id;
};
template <typename T>
__device__ inline PartsOf<T> parts_of(const SpeciesDev<T>& S, int par) {
    T* b = (T*)S.blk[par];
    const size_t c = S.pcap;
    return {b, b + c, b + 2 * c, b + 3 * c, (uint32_t*)(b + 4 * c)};
}
struct LayOf {
    int32_t* count;
    int32_t* mcount;
    const int32_t* start;
    const int32_t* cap;
};
template <typename T>
__device__ inline LayOf lay_of(const SpeciesDev<T>& S, int par) {
    int32_t* b = S.lay[par];
    const int e = S.nb[par] + 1;
    return {b, b + e, b + 2 * e, b + 3 * e};
}
// element i of an array: a 32-bit byte offset from the array's base (which is
// uniform and stays in scalar registers); every array holds under 2^32 bytes
template <typename T>
__device__ inline gptr<T> el(T* a, uint32_t i) {
    return (gptr<T>)((gptr<char>)G((char*)a) + i * (uint32_t)sizeof(T));
}

__device__ inline uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline int front_cap(int cap) { return cap - (cap >> 2); }
__device__ inline int back_cap(int cap) { return cap >> 2; }

// Wave-aggregated slot allocation in the output buffer.  back == false: the
// front part of bin `key` (re-binning); back == true: the back part (movers and
// born particles).  Overflow goes to the spill region; -1 when that is full too.
// key == nb addresses the spill region directly.  Every lane of the wave must
// call it.  The lanes are grouped by key first (a scalar loop over the distinct
// keys, no memory access); then the lowest lane of every group adds the group's
// size to its bin's counter, all groups in one atomic instruction, so a call
// waits for one atomic round trip however many bins its lanes go to.
template <typename T>
__device__ inline int alloc_slot(const SpeciesDev<T>& S, int out, bool want, int key, bool back) {
    const int lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    const int nb = S.nb[out];
    const LayOf L = lay_of(S, out);
    const bool binned = want && key != nb;
    // this lane's group: the lanes that want the same bin
    uint64_t group = 0;
    uint64_t act = __ballot(binned);
    while (act) {
        const int first = __ffsll((unsigned long long)act) - 1;
        const int k = __shfl(key, first);
        const uint64_t m = __ballot(binned && key == k);
        if (binned && key == k) group = m;
        act &= ~m;
    }
    int cap = 0, start = 0;
    if (binned) {
        cap = *el(L.cap, key);
        start = *el(L.start, key);
    }
    const int leader = binned ? __ffsll((unsigned long long)group) - 1 : lane;
    int base = 0;
    if (binned && leader == lane)
        base = gatomic_add(back ? L.mcount + key : L.count + key, (int)__popcll(group));
    base = __shfl(base, leader);
    int dst = -1;
    bool spill = want && !binned;
    if (binned) {
        const int slot = base + (int)__popcll(group & lt);
        if (back) {
            if (slot < back_cap(cap)) dst = start + cap - 1 - slot;
            else spill = true;
        } else {
            if (slot < front_cap(cap)) dst = start + slot;
            else spill = true;
        }
    }
    const uint64_t om = __ballot(spill);
    if (om) {
        const int sl = __ffsll((unsigned long long)om) - 1;
        int sbase = 0;
        if (lane == sl) sbase = gatomic_add(L.count + nb, (int)__popcll(om));
        sbase = __shfl(sbase, sl);
        if (spill) {
            const int s2 = sbase + __popcll(om & lt);
            dst = (s2 < *el(L.cap, nb)) ? *el(L.start, nb) + s2 : -1;
        }
    }
    return dst;
}

template <typename T>
__device__ inline int bin_key(const CaseDev<T>& C, int nb, T x, T y) {
    if (nb == 0) return 0;
    int ix = (int)dfloor(x), iy = (int)dfloor(y);
    ix = min(max(ix, 0), C.nx - 2);
    iy = min(max(iy, 0), C.ny - 2);
    // ix, iy >= 0: a shift is the division for power-of-two tiles
    const int bx = C.bw_log2 >= 0 ? ix >> C.bw_log2 : ix / C.bw;
    const int by = C.bh_log2 >= 0 ? iy >> C.bh_log2 : iy / C.bh;
    return by * C.nbx + bx;
}

// segment-segment intersection of collisions.cu:34-41
__device__ inline bool segment_hits(const double* w, double X1, double Y1, double X2, double Y2, double eps) {
    const double x3 = w[0], y3 = w[1], x4 = w[2], y4 = w[3];
    const double den = (y4 - y3) * (X2 - X1) - (x4 - x3) * (Y2 - Y1);
    if (fabs(den) < eps) return false;
    const double ua = ((x4 - x3) * (Y1 - y3) - (y4 - y3) * (X1 - x3)) / den;
    const double ub = ((X2 - X1) * (Y1 - y3) - (Y2 - Y1) * (X1 - x3)) / den;
    return ua >= 0.0 && ua <= 1.0 && ub >= 0.0 && ub <= 1.0;
}

// Absorption test of one path (box edges, then the segment test of
// detect_intersections, collisions.cu:7-54). The walls come from the per-bin
// candidate lists, each holding every wall that meets its bin grown by
// wall_margin + 1 cells:
//  * a path with both ends within wall_margin cells of bin `bin` (cell origin
//    cx0, cy0; the common case, uniform over the workgroup) tests that bin's
//    list;
//  * any other path tests the lists of every bin its bounding box meets
//    (every point of the path lies in one of them), at most 16 bins;
//  * a longer path, one outside the binned area, or no lists: every wall.
// All three give the full scan's answer. WALLS false compiles the box test
// alone, for launches in which no case has internal walls (the segment code
// costs the particle loop registers even when it does not run).
template <typename T, bool WALLS>
__device__ inline bool hits_walls(const CaseDev<T>& C, int bin, int cx0, int cy0, T x0, T y0, T x1, T y1) {
    if (C.box) {
        if (x1 <= (T)0 || x1 >= (T)(C.nx - 1) || y1 <= (T)0 || y1 >= (T)(C.ny - 1)) return true;
    }
    if (WALLS && C.n_walls > 0) {
        const double X1 = x0, Y1 = y0, X2 = x1, Y2 = y1;
        const double rx0 = cx0 - C.wall_margin, ry0 = cy0 - C.wall_margin;
        const double rx1 = cx0 + C.bw + C.wall_margin, ry1 = cy0 + C.bh + C.wall_margin;
        const bool inside = bin >= 0 && C.wall_off && X1 >= rx0 && X1 <= rx1 && Y1 >= ry0 && Y1 <= ry1 &&
                            X2 >= rx0 && X2 <= rx1 && Y2 >= ry0 && Y2 <= ry1;
        bool hit = false;
        if (inside) {
            const int o0 = G(C.wall_off)[bin], o1 = G(C.wall_off)[bin + 1];
            for (int k = o0; k < o1; ++k)
                hit |= segment_hits(C.walls + 4 * G(C.wall_idx)[k], X1, Y1, X2, Y2, C.wall_eps);
            return hit;
        }
        const double lx = fmin(X1, X2), hx = fmax(X1, X2), ly = fmin(Y1, Y2), hy = fmax(Y1, Y2);
        if (C.wall_off && lx >= 0.0 && ly >= 0.0 && hx < (double)(C.nbx * C.bw) && hy < (double)(C.nby * C.bh)) {
            const int bx0 = (int)(lx / C.bw), bx1 = (int)(hx / C.bw);
            const int by0 = (int)(ly / C.bh), by1 = (int)(hy / C.bh);
            if ((bx1 - bx0 + 1) * (by1 - by0 + 1) <= 16) {
                for (int by = by0; by <= by1; ++by)
                    for (int bx = bx0; bx <= bx1; ++bx) {
                        const int b = by * C.nbx + bx;
                        const int o0 = G(C.wall_off)[b], o1 = G(C.wall_off)[b + 1];
                        for (int k = o0; k < o1; ++k)
                            hit |= segment_hits(C.walls + 4 * G(C.wall_idx)[k], X1, Y1, X2, Y2, C.wall_eps);
                    }
                return hit;
            }
        }
        for (int j = 0; j < C.n_walls; ++j) hit |= segment_hits(C.walls + 4 * j, X1, Y1, X2, Y2, C.wall_eps);
        return hit;
    }
    return false;
}

// Q20 CIC deposit (cell test of kernels.cu:53).  LDS window if the four nodes
// fall inside it, global 64-bit atomics otherwise.
template <typename T>
__device__ inline void deposit(T x, T y, int nx, int ny, bool use_lds, int wx0, int wy0, int aw, int ah,
                               unsigned long long* lds, unsigned long long* gacc, unsigned* lcnt) {
    const int ix = (int)dfloor(x), iy = (int)dfloor(y);
    if (ix < 0 || ix >= nx - 1 || iy < 0 || iy >= ny - 1) return;
    const T fx = x - (T)ix, fy = y - (T)iy;
    const unsigned W = 1u << PIC_QBITS;
    // fx, fy in [0, 1): qx, qy and W - qx, W - qy fit in 21 bits (the masks
    // change no value and let the products use the 24-bit multiplier)
    const unsigned qx = (unsigned)(fx * (T)W) & 0x1FFFFFu, qy = (unsigned)(fy * (T)W) & 0x1FFFFFu;
    const unsigned px = (W - qx) & 0x1FFFFFu, py = (W - qy) & 0x1FFFFFu;
    const unsigned long long w00 = (unsigned long long)px * py, w10 = (unsigned long long)qx * py;
    const unsigned long long w01 = (unsigned long long)px * qy, w11 = (unsigned long long)qx * qy;
    const int lx = ix - wx0, ly = iy - wy0;
    if (use_lds && lx >= 0 && lx < aw - 1 && ly >= 0 && ly < ah - 1) {
        unsigned long long* b = lds + ly * aw + lx;
        atomicAdd(b, w00);
        atomicAdd(b + 1, w10);
        atomicAdd(b + aw, w01);
        atomicAdd(b + aw + 1, w11);
    } else {
        unsigned long long* b = gacc + (size_t)iy * nx + ix;
        gatomic_add(b, w00);
        gatomic_add(b + 1, w10);
        gatomic_add(b + nx, w01);
        gatomic_add(b + nx + 1, w11);
        atomicAdd(lcnt + PIC_C_GLOBAL_DEPOSIT, 1u);
    }
}

template <typename T>
__device__ inline bool iso_dir(Stream& rs, T& c, T& s) {
    for (int t = 0; t < 16; ++t) {
        const T a = (T)2 * rs.uni<T>() - (T)1;
        const T b = (T)2 * rs.uni<T>() - (T)1;
        const T q = dfma(a, a, b * b);
        if (q > (T)0 && q <= (T)1) {
            c = dfma(a, a, -(b * b)) / q;
            s = ((T)2 * a) * b / q;
            return true;
        }
    }
    c = (T)1;
    s = (T)0;
    return false;
}

template <typename T>
__device__ inline bool gauss_pair(Stream& rs, double sigma, T& g1, T& g2) {
    for (int t = 0; t < 16; ++t) {
        const T a = (T)2 * rs.uni<T>() - (T)1;
        const T b = (T)2 * rs.uni<T>() - (T)1;
        const T q = dfma(a, a, b * b);
        if (q > (T)0 && q < (T)1) {
            const double qd = (double)q;
            const double f = sqrt(-2.0 * det_log(qd) / qd);
            g1 = (T)(sigma * ((double)a * f));
            g2 = (T)(sigma * ((double)b * f));
            return true;
        }
    }
    g1 = (T)0;
    g2 = (T)0;
    return false;
}

template <typename T>
struct Born {
    T ex, ey, evx, evy, ix, iy, ivx, ivy;
    uint32_t eid, iid;
};

__device__ inline float dfrexp(float v, int* e) { return frexpf(v, e); }
__device__ inline double dfrexp(double v, int* e) { return frexp(v, e); }

// energy band of the null-collision majorant: 0 below emax*2^-14 (and NaN),
// b = 15 + floor(log2(E/emax)) above, 15 from emax/2 on
template <typename T>
__device__ inline int mcc_band(T E, T inv_emax) {
    const T r = E * inv_emax;
    if (!(r >= (T)0x1p-14)) return 0;
    int e = 0;
    (void)dfrexp(r, &e);
    return min(14 + e, 15);
}

// The candidacy uniform on the device.  The key passes through an empty asm
// statement inside the particle loop, so its ten round keys are formed there
// (one scalar add each) instead of being hoisted into ten scalar registers
// for the whole loop (they were spilled to vector lanes and read back one by
// one).
template <typename T>
__device__ inline T cand_uni_dev(uint32_t id, uint32_t step, uint32_t key) {
    asm volatile("" : "+s"(key));
    return cand_uni<T>(id, step, key);
}

// The null-collision candidacy test of mcc_collide alone: the candidacy draw
// against the candidate probability of the particle's energy band.  ckey:
// cand_key(seed, case, species).
template <typename T>
__device__ inline bool mcc_candidate(const SpeciesDev<T>& S, const T* band_lds, uint32_t id, uint32_t step,
                                     uint32_t ckey, T vx, T vy) {
    const T u0 = cand_uni_dev<T>(id, step, ckey);
    const T v2 = dfma(vx, vx, vy * vy);
    const T E = S.efac * v2;
    return u0 < band_lds[mcc_band(E, S.inv_emax)];
}

// Null-collision MCC for one particle (replaces elasticCollisions,
// mcc.cu:114-150, and its kernels kernels.cu:196-263, mcc.cu:56-85).  The
// majorant is that of the particle's energy band (band values in LDS).
template <typename T>
__device__ inline bool mcc_collide(const SpeciesDev<T>& S, const T* band_lds, uint32_t id, uint32_t step,
                                   uint32_t case_id, uint32_t seed, uint32_t ckey, T x, T y, T& vx, T& vy,
                                   unsigned* lcnt, Born<T>& b) {
    const T u0 = cand_uni_dev<T>(id, step, ckey);
    Stream rs(id, step, case_id, seed, S.key);
    const T u1 = rs.uni<T>();
    const T v2 = dfma(vx, vx, vy * vy);
    const T E = S.efac * v2;
    const int band = mcc_band(E, S.inv_emax);
    if (!(u0 < band_lds[band])) return false;
    atomicAdd(lcnt + PIC_C_NULL_CAND, 1u);
    const T nu_max = band_lds[16 + band];
    const T v = dsqrt(v2);
    const T s = E * S.inv_de;
    const int K = S.K;
    int j;
    T t;
    if (s < (T)(K - 1)) { j = (int)s; t = s - (T)j; }   // false for NaN as well
    else { j = K - 2; t = (T)1; }
    const T R = u1 * nu_max;
    T cum = (T)0;
    int sel = -1;
    for (int q = 0; q < S.n_proc; ++q) {
        if (E < S.thr[q]) continue;
        const T a = G(S.tab)[(size_t)q * K + j], bb = G(S.tab)[(size_t)q * K + j + 1];
        const T sig = dfma(t, bb - a, a);
        cum = dfma(sig, v, cum);
        if (sel < 0 && R < cum) sel = q;
    }
    if (cum > nu_max) atomicAdd(lcnt + PIC_C_NUMAX_EXCEEDED, 1u);
    if (sel < 0) return false;
    atomicAdd(lcnt + PIC_C_PROC0 + sel, 1u);
    const int ty = S.ptype[sel];
    const T thr = S.thr[sel];
    T c, sn;
    if (ty == PIC_ELASTIC) {
        if (!iso_dir(rs, c, sn)) atomicAdd(lcnt + PIC_C_REJECT_FALLBACK, 1u);
        vx = v * c;
        vy = v * sn;
    } else if (ty == PIC_EXCITATION) {
        const T e1 = E - thr;
        const T v1 = dsqrt(e1 / S.efac);
        if (!iso_dir(rs, c, sn)) atomicAdd(lcnt + PIC_C_REJECT_FALLBACK, 1u);
        vx = v1 * c;
        vy = v1 * sn;
    } else if (ty == PIC_IONIZATION) {
        const T er = E - thr;
        const T us = rs.uni<T>();
        const T e1 = er * us;
        const T e2 = er - e1;
        const T v1 = dsqrt(e1 / S.efac), v2s = dsqrt(e2 / S.efac);
        if (!iso_dir(rs, c, sn)) atomicAdd(lcnt + PIC_C_REJECT_FALLBACK, 1u);
        vx = v1 * c;
        vy = v1 * sn;
        if (!iso_dir(rs, c, sn)) atomicAdd(lcnt + PIC_C_REJECT_FALLBACK, 1u);
        b.ex = x; b.ey = y; b.evx = v2s * c; b.evy = v2s * sn;
        T g1, g2;
        if (!gauss_pair(rs, S.gas_vth, g1, g2)) atomicAdd(lcnt + PIC_C_REJECT_FALLBACK, 1u);
        b.ix = x; b.iy = y; b.ivx = g1; b.ivy = g2;
        b.eid = rs.u32();
        b.iid = rs.u32();
        return true;
    } else {  // charge exchange
        T g1, g2;
        if (!gauss_pair(rs, S.gas_vth, g1, g2)) atomicAdd(lcnt + PIC_C_REJECT_FALLBACK, 1u);
        vx = g1;
        vy = g2;
    }
    return false;
}

// The input of a work item: bin b is its front part [start, start+n1) and its
// back part [start+cap-n2, start+cap); a spill chunk is one range.
struct Range {
    int b1, n1, b2, n2;
    __device__ inline int n() const { return n1 + n2; }
    __device__ inline int at(int t) const { return t < n1 ? b1 + t : b2 + (t - n1); }
};

template <typename T>
__device__ inline Range item_range(const SpeciesDev<T>& S, int in, const WorkItem& it, int chunk) {
    const int nb = S.nb[in];
    const LayOf L = lay_of(S, in);
    Range r;
    if (it.bin < nb) {
        const int start = *el(L.start, it.bin), cap = *el(L.cap, it.bin);
        r.n1 = min(*el(L.count, it.bin), front_cap(cap));
        r.n2 = min(*el(L.mcount, it.bin), back_cap(cap));
        r.b1 = start;
        r.b2 = start + cap - r.n2;
    } else {
        const int tot = min(*el(L.count, nb), *el(L.cap, nb));
        r.b1 = *el(L.start, nb) + it.chunk * chunk;
        r.n1 = max(0, min(chunk, tot - it.chunk * chunk));
        r.b2 = 0;
        r.n2 = 0;
    }
    return r;
}

// LDS staging of the particles that leave their bin (and of born particles of
// the same species) until a flush writes them to their destination bins

template <typename T>
struct Stage {
    T* x;
    T* y;
    T* vx;
    T* vy;
    uint32_t* id;
    int* key;
};

// LDS accesses of one wavefront complete in order; the fences keep the
// compiler from moving LDS accesses across this point
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// write a wavefront's staged particles to their destination bins (the whole
// wavefront calls it; nm is uniform)
template <typename T>
__device__ inline void flush_stage(const SpeciesDev<T>& S, int out, const Stage<T>& sg, int nm, unsigned* lcnt) {
    const int lane = __lane_id();
    wave_lds_sync();
    for (int base = 0; base < nm; base += 64) {
        const int e = base + lane;
        const bool want = e < nm;
        const int key = want ? sg.key[e] : 0;
        const int dst = alloc_slot(S, out, want, key, true);
        if (want) {
            if (dst >= 0) {
                const PartsOf<T> P = parts_of(S, out);
                *el(P.x, dst) = sg.x[e];
                *el(P.y, dst) = sg.y[e];
                *el(P.vx, dst) = sg.vx[e];
                *el(P.vy, dst) = sg.vy[e];
                *el(P.id, dst) = sg.id[e];
                if (dst >= *el(lay_of(S, out).start, S.nb[out])) atomicAdd(lcnt + PIC_C_SPILLED, 1u);
            } else {
                atomicAdd(lcnt + PIC_C_LOST, 1u);
            }
        }
    }
    wave_lds_sync();
}

// append the lanes with `want` to the wavefront's stage; nm is the
// wavefront's (uniform) fill
template <typename T>
__device__ inline void stage_put(const Stage<T>& sg, int& nm, bool want, T x, T y, T vx, T vy, uint32_t id, int key) {
    const uint64_t m = __ballot(want);
    if (!m) return;
    const int base = nm;
    nm += __popcll(m);
    if (want) {
        const int e = base + __popcll(m & lanemask_lt());
        sg.x[e] = x;
        sg.y[e] = y;
        sg.vx[e] = vx;
        sg.vy[e] = vy;
        sg.id[e] = id;
        sg.key[e] = key;
    }
}

// ---------------------------------------------------------------------------
// the fused particle kernel
// ---------------------------------------------------------------------------
// gate_conv / gate_rebin (null: no gate): a push enqueued ahead of the host's
// check of the previous step does nothing unless *gate_conv (the solve's
// convergence flag) is set and *gate_rebin (no species needs re-binning) is
// clear; both are read with the item's set-up loads and tested before the
// first global write.
template <typename T, bool WALLS>
__global__ __launch_bounds__(256) void push_kernel(const CaseDev<T>* __restrict__ cases,
                                                   const CaseState* __restrict__ states,
                                                   const WorkItem* __restrict__ items, int spill_chunk,
                                                   int flags_arg, long long mcc_key,
                                                   const int* __restrict__ gate_conv,
                                                   const int* __restrict__ gate_rebin) {
    using T2 = typename Vec2<T>::type;
    extern __shared__ unsigned long long smem_u64[];
    __shared__ int sh_stay;
    const int go = (gate_conv ? *gate_conv : 1) && !(gate_rebin ? *gate_rebin : 0);
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const CaseState st = states[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = st.cur, out = in ^ 1;
    const Range rg = item_range(S, in, it, spill_chunk);
    const int n = rg.n();
    const bool is_bin = it.bin < S.nb[in];
    // an empty bin still publishes its (zero) stay count; nothing else to do
    if (n <= 0) return;
    if (!go) return;   // uniform over the block: a gated push whose step is not ready writes nothing

    const int nx = C.nx, ny = C.ny;
    const int H = C.halo;
    const int bx = is_bin ? it.bin % C.nbx : 0, by = is_bin ? it.bin / C.nbx : 0;
    const int cx0 = bx * C.bw, cy0 = by * C.bh;
    const int ew = C.bw + 1, eh = C.bh + 1;                  // E window nodes
    const int aw = C.bw + 1 + 2 * H, ah = C.bh + 1 + 2 * H;  // deposit window nodes
    const int wx0 = cx0 - H, wy0 = cy0 - H;
    // LDS: deposit window (u64), E window (T2), stage (4 T + id + key), counters,
    // energy-band constants of the collisions
    unsigned long long* accw = smem_u64;
    // the E window starts on a 16-byte boundary (an even count of u64 entries):
    // double2 loads from LDS need it
    T2* ew_lds = reinterpret_cast<T2*>(smem_u64 + (is_bin ? ((aw * ah + 1) & ~1) : 0));
    T* stage_base = reinterpret_cast<T*>(ew_lds + (is_bin ? ew * eh : 0));
    // the stage is split into one part per wavefront (PIC_MSTAGE / 4 entries)
    constexpr int WST = PIC_MSTAGE / 4;
    const int wv = threadIdx.x >> 6;
    Stage<T> sg;
    sg.x = stage_base + wv * WST;
    sg.y = stage_base + PIC_MSTAGE + wv * WST;
    sg.vx = stage_base + 2 * PIC_MSTAGE + wv * WST;
    sg.vy = stage_base + 3 * PIC_MSTAGE + wv * WST;
    sg.id = reinterpret_cast<uint32_t*>(stage_base + 4 * PIC_MSTAGE) + wv * WST;
    sg.key = reinterpret_cast<int*>(stage_base + 4 * PIC_MSTAGE) + PIC_MSTAGE + wv * WST;
    unsigned* lcnt = reinterpret_cast<unsigned*>(reinterpret_cast<int*>(stage_base + 4 * PIC_MSTAGE) + 2 * PIC_MSTAGE);
    int nmw = 0;   // this wavefront's stage fill
    // candidate probability and majorant of each energy band (32 values)
    T* band_lds = reinterpret_cast<T*>(lcnt + PIC_NCOUNT);
    // collision candidates deferred by each wavefront (PIC_MQUEUE entries per
    // wavefront: x, y, vx, vy after the kick, and the id)
    T* queue_base = band_lds + 32;
    T* qx = queue_base + wv * PIC_MQUEUE;
    T* qy = queue_base + 4 * PIC_MQUEUE + wv * PIC_MQUEUE;
    T* qvx = queue_base + 8 * PIC_MQUEUE + wv * PIC_MQUEUE;
    T* qvy = queue_base + 12 * PIC_MQUEUE + wv * PIC_MQUEUE;
    uint32_t* qid = reinterpret_cast<uint32_t*>(queue_base + 16 * PIC_MQUEUE) + wv * PIC_MQUEUE;
    int nq = 0;    // this wavefront's queue fill
    const int flags = flags_arg >= 0 ? flags_arg : st.flags;
    const bool kick = flags & PIC_F_KICK, mcc = (flags & PIC_F_MCC) && S.n_proc > 0, drift = flags & PIC_F_DRIFT,
               dep = flags & PIC_F_DEPOSIT;
    // the kick and collisions of a pass complete step st.step - 1 (the drift
    // starts step st.step), so collisions are keyed on st.step - 1
    const uint32_t mcc_step = mcc_key >= 0 ? (uint32_t)mcc_key : st.step - 1u;
    const uint32_t ckey = cand_key(C.seed, C.case_id, S.key);

    for (int k = threadIdx.x; k < PIC_NCOUNT; k += blockDim.x) lcnt[k] = 0u;
    if (threadIdx.x < 32) band_lds[threadIdx.x] = threadIdx.x < 16 ? S.band_pnull[threadIdx.x] : S.band_numax[threadIdx.x - 16];
    if (threadIdx.x == 0) sh_stay = 0;
    if (is_bin) {
        for (int k = threadIdx.x; k < aw * ah; k += blockDim.x) accw[k] = 0ull;
        if (kick) {
            for (int k = threadIdx.x; k < ew * eh; k += blockDim.x) {
                const int i = cx0 + k % ew, j = cy0 + k / ew;
                T2 e;
                if (i < nx && j < ny) {
                    e.x = (T)ex_node(C.phi, nx, i, j, C.dx);
                    e.y = (T)ey_node(C.phi, nx, ny, i, j, C.dy);
                } else {
                    e.x = (T)0;
                    e.y = (T)0;
                }
                ew_lds[k] = e;
            }
        }
    }
    __syncthreads();

    const T kq = S.kq, dtdx = C.dtdx, dtdy = C.dtdy;
    const SpeciesDev<T>* O = (S.other >= 0) ? &C.sp[S.other] : nullptr;
    const int wall_bin = is_bin ? it.bin : -1;
    const LayOf Lo = lay_of(S, out);
    const int home_start = is_bin ? *el(Lo.start, it.bin) : 0;
    const int home_front = is_bin ? front_cap(*el(Lo.cap, it.bin)) : 0;
    const int lane = __lane_id();

    // Drift, walls, deposit and compaction of one particle per lane after the
    // kick. FULL: the particle is a collision candidate and collides first
    // (born particles included); otherwise it does not collide. The whole
    // wavefront calls it (ballots inside).
    auto finish = [&](auto full_tag, bool valid, T x, T y, T vx, T vy, uint32_t id) {
        constexpr bool FULL = decltype(full_tag)::value;
        bool alive = false, born = false, be_alive = false, bi_alive = false;
        int key = 0, key_e = 0, key_i = 0;
        Born<T> b;
        if (valid) {
            if (FULL) born = mcc_collide(S, band_lds, id, mcc_step, C.case_id, C.seed, ckey, x, y, vx, vy, lcnt, b);
            const T x0 = x, y0 = y;
            if (drift) {
                x = dfma(vx, dtdx, x);
                y = dfma(vy, dtdy, y);
            }
            if (hits_walls<T, WALLS>(C, wall_bin, cx0, cy0, x0, y0, x, y)) {
                atomicAdd(lcnt + PIC_C_ABSORBED, 1u);
            } else {
                alive = true;
                if (dep) deposit(x, y, nx, ny, is_bin, wx0, wy0, aw, ah, accw, S.acc, lcnt);
                key = bin_key(C, S.nb[out], x, y);
            }
            if (FULL && born) {
                // born particles drift from the collision point when the pass drifts
                const T ex1 = drift ? dfma(b.evx, dtdx, b.ex) : b.ex, ey1 = drift ? dfma(b.evy, dtdy, b.ey) : b.ey;
                if (hits_walls<T, WALLS>(C, wall_bin, cx0, cy0, b.ex, b.ey, ex1, ey1)) {
                    atomicAdd(lcnt + PIC_C_BORN_ABSORBED, 1u);
                } else {
                    be_alive = true;
                    b.ex = ex1;
                    b.ey = ey1;
                    if (dep) deposit(ex1, ey1, nx, ny, is_bin, wx0, wy0, aw, ah, accw, S.acc, lcnt);
                    key_e = bin_key(C, S.nb[out], ex1, ey1);
                    atomicAdd(lcnt + PIC_C_BORN_SELF, 1u);
                }
                const T ix1 = drift ? dfma(b.ivx, dtdx, b.ix) : b.ix, iy1 = drift ? dfma(b.ivy, dtdy, b.iy) : b.iy;
                if (hits_walls<T, WALLS>(C, wall_bin, cx0, cy0, b.ix, b.iy, ix1, iy1)) {
                    atomicAdd(lcnt + PIC_C_BORN_ABSORBED, 1u);
                } else if (O) {
                    bi_alive = true;
                    b.ix = ix1;
                    b.iy = iy1;
                    if (dep) deposit(ix1, iy1, nx, ny, false, 0, 0, 0, 0, accw, O->acc, lcnt);
                    key_i = bin_key(C, O->nb[out], ix1, iy1);
                    atomicAdd(lcnt + PIC_C_BORN_OTHER, 1u);
                }
            }
        }
        // survivors that stay in this bin: front slots from the LDS counter
        bool stay = alive && is_bin && key == it.bin;
        const uint64_t sm = __ballot(stay);
        if (sm) {
            const int leader = __ffsll((unsigned long long)sm) - 1;
            int sbase = 0;
            if (lane == leader) sbase = atomicAdd(&sh_stay, __popcll(sm));
            sbase = __shfl(sbase, leader);
            if (stay) {
                const int slot = sbase + __popcll(sm & lanemask_lt());
                if (slot < home_front) {
                    const int dst = home_start + slot;
                    const PartsOf<T> P = parts_of(S, out);
                    *el(P.x, dst) = x;
                    *el(P.y, dst) = y;
                    *el(P.vx, dst) = vx;
                    *el(P.vy, dst) = vy;
                    *el(P.id, dst) = id;
                } else {
                    stay = false;   // front part full: goes with the movers
                }
            }
        }
        // movers and born particles of this species: staged
        stage_put(sg, nmw, alive && !stay, x, y, vx, vy, id, key);
        if (FULL) {
            stage_put(sg, nmw, be_alive, b.ex, b.ey, b.evx, b.evy, b.eid, key_e);
            // born ions of the other species (rare): written directly
            if (O && __ballot(bi_alive)) {
                const int di = alloc_slot(*O, out, bi_alive, key_i, true);
                if (bi_alive) {
                    if (di >= 0) {
                        const PartsOf<T> P = parts_of(*O, out);
                        *el(P.x, di) = b.ix;
                        *el(P.y, di) = b.iy;
                        *el(P.vx, di) = b.ivx;
                        *el(P.vy, di) = b.ivy;
                        *el(P.id, di) = b.iid;
                    } else {
                        atomicAdd(lcnt + PIC_C_LOST, 1u);
                    }
                }
            }
        }
        // a call adds at most 128 entries per wavefront (a mover and a born
        // electron per lane)
        if (nmw > WST - 128) {
            flush_stage(S, out, sg, nmw, lcnt);
            nmw = 0;
        }
    };
    // the deferred candidates of this wavefront, one per lane
    auto flush_queue = [&]() {
        wave_lds_sync();
        const bool v = lane < nq;
        T x = (T)0, y = (T)0, vx = (T)0, vy = (T)0;
        uint32_t id = 0;
        if (v) {
            x = qx[lane];
            y = qy[lane];
            vx = qvx[lane];
            vy = qvy[lane];
            id = qid[lane];
        }
        wave_lds_sync();
        finish(std::true_type{}, v, x, y, vx, vy, id);
        nq = 0;
    };

    // The particles of the next pass are loaded before this pass's stores:
    // the memory counter completes in order, so waiting for them never waits
    // for the stores, and no default value or branch precedes the loads (every
    // lane loads; lanes past the end load the last particle, n > 0).
    T px, py, pvx, pvy;
    uint32_t pid;
    {
        const int src = rg.at(min((int)threadIdx.x, n - 1));
        const PartsOf<T> P = parts_of(S, in);
        px = *el(P.x, src);
        py = *el(P.y, src);
        pvx = *el(P.vx, src);
        pvy = *el(P.vy, src);
        pid = *el(P.id, src);
    }
    for (int base = 0; base < n; base += blockDim.x) {
        const int t = base + threadIdx.x;
        const bool valid = t < n;
        T x = px, y = py, vx = pvx, vy = pvy;
        uint32_t id = pid;
        if (base + (int)blockDim.x < n) {
            const int src = rg.at(min(t + (int)blockDim.x, n - 1));
            const PartsOf<T> P = parts_of(S, in);
            px = *el(P.x, src);
            py = *el(P.y, src);
            pvx = *el(P.vx, src);
            pvy = *el(P.vy, src);
            pid = *el(P.id, src);
        }
        if (valid && kick) {
            const int ix = (int)dfloor(x), iy = (int)dfloor(y);
            if (!(ix < 0 || ix >= nx - 1 || iy < 0 || iy >= ny - 1)) {
                const T fx = x - (T)ix, fy = y - (T)iy;
                const T gx = (T)1 - fx, gy = (T)1 - fy;
                const T w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
                T2 e00, e10, e01, e11;
                const int lx = ix - cx0, ly = iy - cy0;
                if (is_bin && lx >= 0 && lx < C.bw && ly >= 0 && ly < C.bh) {
                    const T2* p = ew_lds + ly * ew + lx;
                    e00 = p[0];
                    e10 = p[1];
                    e01 = p[ew];
                    e11 = p[ew + 1];
                } else {
                    e00.x = (T)ex_node(C.phi, nx, ix, iy, C.dx);
                    e00.y = (T)ey_node(C.phi, nx, ny, ix, iy, C.dy);
                    e10.x = (T)ex_node(C.phi, nx, ix + 1, iy, C.dx);
                    e10.y = (T)ey_node(C.phi, nx, ny, ix + 1, iy, C.dy);
                    e01.x = (T)ex_node(C.phi, nx, ix, iy + 1, C.dx);
                    e01.y = (T)ey_node(C.phi, nx, ny, ix, iy + 1, C.dy);
                    e11.x = (T)ex_node(C.phi, nx, ix + 1, iy + 1, C.dx);
                    e11.y = (T)ey_node(C.phi, nx, ny, ix + 1, iy + 1, C.dy);
                }
                const T Ex = dfma(w11, e11.x, dfma(w01, e01.x, dfma(w10, e10.x, w00 * e00.x)));
                const T Ey = dfma(w11, e11.y, dfma(w01, e01.y, dfma(w10, e10.y, w00 * e00.y)));
                vx = dfma(kq, Ex, vx);
                vy = dfma(kq, Ey, vy);
            }
        }
        // collision candidates leave the batch for this wavefront's queue,
        // which is processed (one candidate per lane) whenever it would overflow
        bool cand = false;
        if (mcc) {
            cand = valid && mcc_candidate(S, band_lds, id, mcc_step, ckey, vx, vy);
            const uint64_t cmask = __ballot(cand);
            if (cmask) {
                if (nq + __popcll(cmask) > PIC_MQUEUE) flush_queue();
                if (cand) {
                    const int q = nq + __popcll(cmask & lanemask_lt());
                    qx[q] = x;
                    qy[q] = y;
                    qvx[q] = vx;
                    qvy[q] = vy;
                    qid[q] = id;
                }
                nq += __popcll(cmask);
            }
        }
        finish(std::false_type{}, valid && !cand, x, y, vx, vy, id);
    }
    if (nq > 0) flush_queue();
    if (nmw > 0) flush_stage(S, out, sg, nmw, lcnt);
    // every wavefront's stay count, deposits and counters are in LDS now
    __syncthreads();
    if (is_bin && threadIdx.x == 0) *el(Lo.count, it.bin) = min(sh_stay, home_front);
    if (is_bin) {
        for (int k = threadIdx.x; k < aw * ah; k += blockDim.x) {
            const unsigned long long v = accw[k];
            if (v) {
                const int i = wx0 + k % aw, j = wy0 + k / aw;
                if (i >= 0 && i < nx && j >= 0 && j < ny) gatomic_add(S.acc + (size_t)j * nx + i, v);
            }
        }
    }
    for (int k = threadIdx.x; k < PIC_NCOUNT; k += blockDim.x)
        if (lcnt[k]) gatomic_add(S.counters + k, (unsigned long long)lcnt[k]);
}

// ---------------------------------------------------------------------------
// seeded initial particles (uniform in the open box, Maxwellian velocities)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void init_kernel(T* x, T* y, T* vx, T* vy, uint32_t* idv, const int32_t* base_ptr, int n, int nx,
                            int ny, double vth, uint32_t seed, uint32_t case_id, uint32_t species,
                            uint32_t id0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int dst = *base_ptr + i;
    const uint32_t pid = id0 + (uint32_t)i;
    Stream rs(pid, 0xFFFFFFFFu, case_id, seed, species);
    const uint32_t rx = rs.u32(), ry = rs.u32();
    x[dst] = (T)(((double)rx + 0.5) * 0x1p-32 * (double)(nx - 1));
    y[dst] = (T)(((double)ry + 0.5) * 0x1p-32 * (double)(ny - 1));
    double g1 = 0.0, g2 = 0.0;
    gauss_pair<double>(rs, vth, g1, g2);
    vx[dst] = (T)g1;
    vy[dst] = (T)g2;
    idv[dst] = pid;
}

// uniform positions in a box and uniform velocities (initializeParticles,
// main.cu:17-32, made reproducible): the four words of the first Philox block
// of counter {id, 0xFFFFFFFE, 0, case} give x, y, vx, vy
template <typename T>
__global__ void init_uniform_kernel(T* x, T* y, T* vx, T* vy, uint32_t* idv, const int32_t* base_ptr, int n,
                                    double x0, double x1, double y0, double y1, double v0, double v1, uint32_t seed,
                                    uint32_t case_id, uint32_t species, uint32_t id0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int dst = *base_ptr + i;
    const uint32_t pid = id0 + (uint32_t)i;
    Stream rs(pid, 0xFFFFFFFEu, case_id, seed, species);
    const double ux = ((double)rs.u32() + 0.5) * 0x1p-32, uy = ((double)rs.u32() + 0.5) * 0x1p-32;
    const double uvx = ((double)rs.u32() + 0.5) * 0x1p-32, uvy = ((double)rs.u32() + 0.5) * 0x1p-32;
    x[dst] = (T)(x0 + ux * (x1 - x0));
    y[dst] = (T)(y0 + uy * (y1 - y0));
    vx[dst] = (T)(v0 + uvx * (v1 - v0));
    vy[dst] = (T)(v0 + uvy * (v1 - v0));
    idv[dst] = pid;
}

__global__ void add_count_kernel(int32_t* c, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c += n;
}

// ---------------------------------------------------------------------------
// particles staged in the reference's layout (Particles2D::r, v, type of
// particles.cuh:7-20: (x, y) double pairs in domain units, an int species
// index per particle) handed to their species' spill regions
// ---------------------------------------------------------------------------
// counts[s]: staged entries of species s among the first n; counts[PIC_MAX_SPECIES]:
// entries whose type is not a species
__global__ void staged_count_kernel(const int32_t* __restrict__ type, long long n, int n_species,
                                    int32_t* counts) {
    __shared__ int c[PIC_MAX_SPECIES + 1];
    if (threadIdx.x <= PIC_MAX_SPECIES) c[threadIdx.x] = 0;
    __syncthreads();
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int t = type[i];
        atomicAdd(&c[(t >= 0 && t < n_species) ? t : PIC_MAX_SPECIES], 1);
    }
    __syncthreads();
    if (threadIdx.x <= PIC_MAX_SPECIES && c[threadIdx.x]) atomicAdd(&counts[threadIdx.x], c[threadIdx.x]);
}

// entry i goes to the next free spill slot of its species (one atomic per
// species present in the wavefront), with id id0 + i; positions divided by the
// cell size into cell units
template <typename T>
__global__ void staged_append_kernel(const double* __restrict__ r, const double* __restrict__ v,
                                     const int32_t* __restrict__ type, long long n, double dx, double dy,
                                     int n_species, StagedDst<T> d, uint32_t id0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const int s = live ? type[i] : -1;   // every type was checked by the count pass
    const int lane = __lane_id();
    int slot = -1;
    for (int q = 0; q < n_species; ++q) {
        const uint64_t m = __ballot(s == q);
        if (!m) continue;
        const int leader = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(d.fill + q, __popcll(m));
        base = __shfl(base, leader);
        if (s == q) slot = base + __popcll(m & lanemask_lt());
    }
    if (!live) return;
    d.x[s][slot] = (T)(r[2 * i] / dx);
    d.y[s][slot] = (T)(r[2 * i + 1] / dy);
    d.vx[s][slot] = (T)v[2 * i];
    d.vy[s][slot] = (T)v[2 * i + 1];
    d.id[s][slot] = id0 + (uint32_t)i;
}

// ---------------------------------------------------------------------------
// re-binning: histogram of new bin keys, then scatter into the new layout
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline int key_for(int nxm2, int nym2, int bw, int bh, int nbx, int nb, T x, T y) {
    if (nb == 0) return 0;
    int ix = (int)dfloor(x), iy = (int)dfloor(y);
    ix = min(max(ix, 0), nxm2);
    iy = min(max(iy, 0), nym2);
    return (iy / bh) * nbx + (ix / bw);
}

template <typename T>
__global__ __launch_bounds__(256) void hist_kernel(const CaseDev<T>* __restrict__ cases,
                                                   const CaseState* __restrict__ states,
                                                   const WorkItem* __restrict__ items, int chunk, int bw, int bh,
                                                   int nbx, int nb_new, int32_t* hist) {
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    if (n <= 0) return;
    const int lane = __lane_id();
    for (int base = 0; base < n; base += blockDim.x) {
        const int t = base + threadIdx.x;
        const bool want = t < n;
        int key = 0;
        if (want) {
            const int s = rg.at(t);
            key = key_for(C.nx - 2, C.ny - 2, bw, bh, nbx, nb_new, S.x[in][s], S.y[in][s]);
        }
        uint64_t act = __ballot(want);
        while (act) {
            const int leader = __ffsll((unsigned long long)act) - 1;
            const int k = __shfl(key, leader);
            const uint64_t m = __ballot(want && key == k) & act;
            if (lane == leader) atomicAdd(hist + k, __popcll(m));
            act &= ~m;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void scatter_kernel(const CaseDev<T>* __restrict__ cases,
                                                      const CaseState* __restrict__ states,
                                                      const WorkItem* __restrict__ items, int chunk, int bw,
                                                      int bh, int nbx, unsigned long long* lost) {
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur, out = in ^ 1;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    if (n <= 0) return;
    for (int base = 0; base < n; base += blockDim.x) {
        const int t = base + threadIdx.x;
        const bool want = t < n;
        int key = 0;
        T x = 0, y = 0, vx = 0, vy = 0;
        uint32_t id = 0;
        if (want) {
            const int s = rg.at(t);
            x = S.x[in][s];
            y = S.y[in][s];
            vx = S.vx[in][s];
            vy = S.vy[in][s];
            id = S.id[in][s];
            key = key_for(C.nx - 2, C.ny - 2, bw, bh, nbx, S.nb[out], x, y);
        }
        const int dst = alloc_slot(S, out, want, key, false);
        if (want) {
            if (dst >= 0) {
                S.x[out][dst] = x;
                S.y[out][dst] = y;
                S.vx[out][dst] = vx;
                S.vy[out][dst] = vy;
                S.id[out][dst] = id;
            } else {
                atomicAdd(lost, 1ull);
            }
        }
    }
}

// Counting sort of the re-binning in LDS (layouts of up to PIC_SORT_LDS_BINS
// bins): a workgroup counts its particles per destination bin in LDS and adds
// the nonzero counts to the global histogram, one atomic per (workgroup, bin);
// the scatter counts the same way, reserves each bin's range with one global
// atomic per (workgroup, bin), then takes each particle's slot from an LDS
// atomic on that range.  The wavefront-aggregated kernels above remain for
// larger layouts.
template <typename T>
__global__ __launch_bounds__(256) void hist_lds_kernel(const CaseDev<T>* __restrict__ cases,
                                                       const CaseState* __restrict__ states,
                                                       const WorkItem* __restrict__ items, int chunk, int bw, int bh,
                                                       int nbx, int nb_new, int32_t* hist) {
    extern __shared__ int lh[];
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    if (n <= 0) return;
    const PartsOf<T> P = parts_of(S, in);
    for (int k = threadIdx.x; k < nb_new; k += blockDim.x) lh[k] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int src = rg.at(t);
        atomicAdd(&lh[key_for(C.nx - 2, C.ny - 2, bw, bh, nbx, nb_new, *el(P.x, src), *el(P.y, src))], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nb_new; k += blockDim.x) {
        const int v = lh[k];
        if (v) gatomic_add(hist + k, v);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void scatter_lds_kernel(const CaseDev<T>* __restrict__ cases,
                                                          const CaseState* __restrict__ states,
                                                          const WorkItem* __restrict__ items, int chunk, int bw,
                                                          int bh, int nbx, unsigned long long* lost) {
    extern __shared__ int lh[];   // per bin: next output position, then the end of the bin's front part
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur, out = in ^ 1;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    if (n <= 0) return;
    const int nb = S.nb[out];
    int* pos = lh;
    int* lim = lh + nb;
    const PartsOf<T> Pi = parts_of(S, in), Po = parts_of(S, out);
    const LayOf Lo = lay_of(S, out);
    for (int k = threadIdx.x; k < nb; k += blockDim.x) pos[k] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int src = rg.at(t);
        atomicAdd(&pos[key_for(C.nx - 2, C.ny - 2, bw, bh, nbx, nb, *el(Pi.x, src), *el(Pi.y, src))], 1);
    }
    __syncthreads();
    // this workgroup's range in the front part of every bin it fills, as
    // absolute positions in the output buffer
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        const int v = pos[k];
        if (v) {
            const int start = *el(Lo.start, k);
            pos[k] = start + gatomic_add(Lo.count + k, v);
            lim[k] = start + front_cap(*el(Lo.cap, k));
        }
    }
    __syncthreads();
    // the next pass's particles are loaded before this pass's stores
    T px, py, pvx, pvy;
    uint32_t pid;
    {
        const int src = rg.at(min((int)threadIdx.x, n - 1));
        px = *el(Pi.x, src);
        py = *el(Pi.y, src);
        pvx = *el(Pi.vx, src);
        pvy = *el(Pi.vy, src);
        pid = *el(Pi.id, src);
    }
    for (int base = 0; base < n; base += blockDim.x) {
        const int t = base + threadIdx.x;
        const T x = px, y = py, vx = pvx, vy = pvy;
        const uint32_t id = pid;
        if (base + (int)blockDim.x < n) {
            const int src = rg.at(min(t + (int)blockDim.x, n - 1));
            px = *el(Pi.x, src);
            py = *el(Pi.y, src);
            pvx = *el(Pi.vx, src);
            pvy = *el(Pi.vy, src);
            pid = *el(Pi.id, src);
        }
        if (t >= n) continue;
        const int key = key_for(C.nx - 2, C.ny - 2, bw, bh, nbx, nb, x, y);
        int dst = atomicAdd(&pos[key], 1);
        if (dst >= lim[key]) {
            // front part full (a layout smaller than the population): the spill region
            const int s2 = gatomic_add(Lo.count + nb, 1);
            dst = s2 < *el(Lo.cap, nb) ? *el(Lo.start, nb) + s2 : -1;
        }
        if (dst >= 0) {
            *el(Po.x, dst) = x;
            *el(Po.y, dst) = y;
            *el(Po.vx, dst) = vx;
            *el(Po.vy, dst) = vy;
            *el(Po.id, dst) = id;
        } else {
            atomicAdd(lost, 1ull);
        }
    }
}

// flatten the input layout of one species into buffer `out` from index 0
template <typename T>
__global__ __launch_bounds__(256) void gather_kernel(const CaseDev<T>* __restrict__ cases,
                                                     const CaseState* __restrict__ states,
                                                     const WorkItem* __restrict__ items, int chunk,
                                                     int32_t* counter) {
    __shared__ int sh_base;
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur, out = in ^ 1;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    if (n <= 0) return;
    // one atomic per item: its particles stay contiguous
    if (threadIdx.x == 0) sh_base = atomicAdd(counter, n);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int s = rg.at(t), d = sh_base + t;
        S.x[out][d] = S.x[in][s];
        S.y[out][d] = S.y[in][s];
        S.vx[out][d] = S.vx[in][s];
        S.vy[out][d] = S.vy[in][s];
        S.id[out][d] = S.id[in][s];
    }
}

// number of particles held by one species of one case (both parts of every
// bin plus the spill region), for last_alive
template <typename T>
__global__ __launch_bounds__(256) void count_kernel(const CaseDev<T>* __restrict__ cases,
                                                    const CaseState* __restrict__ states, int c, int s,
                                                    long long* out) {
    __shared__ long long sh[256];
    const CaseDev<T>& C = cases[c];
    const int cur = states[c].cur;
    const SpeciesDev<T>& S = C.sp[s];
    const int nb = S.nb[cur];
    long long acc = 0;
    for (int b = threadIdx.x; b <= nb; b += 256) {
        const int cap = S.cap[cur][b];
        if (b < nb) acc += min(S.count[cur][b], front_cap(cap)) + min(S.mcount[cur][b], back_cap(cap));
        else acc += min(S.count[cur][b], cap);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

template <typename T>
__global__ __launch_bounds__(256) void end_step_kernel(const CaseDev<T>* __restrict__ cases, CaseState* states,
                                                       int advance_step, int* __restrict__ status_out) {
    end_step_block(cases, states, (int)blockIdx.x, advance_step, status_out, nullptr);
}

// Undo the bin counts a push left in the output parity (trial pushes of the
// tuner): the state of every case stays what it was before the push.
template <typename T>
__global__ __launch_bounds__(256) void clear_out_kernel(const CaseDev<T>* __restrict__ cases,
                                                        const CaseState* __restrict__ states) {
    const int c = blockIdx.x;
    const CaseDev<T>& C = cases[c];
    const int out = states[c].cur ^ 1;
    for (int s = 0; s < C.ns; ++s) {
        const SpeciesDev<T>& S = C.sp[s];
        const int nb = S.nb[out];
        for (int k = threadIdx.x; k <= nb; k += blockDim.x) {
            G(S.count[out])[k] = 0;
            G(S.mcount[out])[k] = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
int launch_push(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items,
                int spill_chunk, int flags, long long mcc_key, size_t lds, hipStream_t s, bool walls,
                const int* gate_conv, const int* gate_rebin) {
    if (n_items <= 0) return 0;
    if (walls)
        hipLaunchKernelGGL((push_kernel<T, true>), dim3(n_items), dim3(256), lds, s, cases, states, items,
                           spill_chunk, flags, mcc_key, gate_conv, gate_rebin);
    else
        hipLaunchKernelGGL((push_kernel<T, false>), dim3(n_items), dim3(256), lds, s, cases, states, items,
                           spill_chunk, flags, mcc_key, gate_conv, gate_rebin);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_init(T* x, T* y, T* vx, T* vy, uint32_t* id, int32_t* count_ptr, int n, int nx, int ny, double vth,
                uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(init_kernel<T>, dim3(div_up(n, 256)), dim3(256), 0, s, x, y, vx, vy, id, count_ptr, n, nx,
                       ny, vth, seed, case_id, species, id0);
    hipLaunchKernelGGL(add_count_kernel, dim3(1), dim3(64), 0, s, count_ptr, n);
    PIC_HIP(hipGetLastError());
    return 0;
}

int launch_staged_count(const int32_t* type, long long n, int n_species, int32_t* counts, hipStream_t s) {
    if (n <= 0) return 0;
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(staged_count_kernel, dim3((unsigned)blocks), dim3(256), 0, s, type, n, n_species, counts);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_staged_append(const double* r, const double* v, const int32_t* type, long long n, double dx, double dy,
                         int n_species, const StagedDst<T>& d, uint32_t id0, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(staged_append_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, r, v, type, n,
                       dx, dy, n_species, d, id0);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_init_uniform(T* x, T* y, T* vx, T* vy, uint32_t* id, int32_t* count_ptr, int n, double x0, double x1,
                        double y0, double y1, double v0, double v1, uint32_t seed, uint32_t case_id,
                        uint32_t species, uint32_t id0, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(init_uniform_kernel<T>, dim3(div_up(n, 256)), dim3(256), 0, s, x, y, vx, vy, id, count_ptr, n,
                       x0, x1, y0, y1, v0, v1, seed, case_id, species, id0);
    hipLaunchKernelGGL(add_count_kernel, dim3(1), dim3(64), 0, s, count_ptr, n);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_hist(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                int bw, int bh, int nbx, int nb_new, int32_t* hist, hipStream_t s) {
    if (n_items <= 0) return 0;
    if (nb_new <= PIC_SORT_LDS_BINS)
        hipLaunchKernelGGL(hist_lds_kernel<T>, dim3(n_items), dim3(256), (size_t)nb_new * sizeof(int), s, cases, states,
                           items, chunk, bw, bh, nbx, nb_new, hist);
    else
        hipLaunchKernelGGL(hist_kernel<T>, dim3(n_items), dim3(256), 0, s, cases, states, items, chunk, bw, bh, nbx,
                           nb_new, hist);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_scatter(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items,
                   int chunk, int bw, int bh, int nbx, int nb_new, unsigned long long* lost, hipStream_t s) {
    if (n_items <= 0) return 0;
    if (nb_new <= PIC_SORT_LDS_BINS)
        hipLaunchKernelGGL(scatter_lds_kernel<T>, dim3(n_items), dim3(256), 2 * (size_t)nb_new * sizeof(int), s, cases,
                           states, items, chunk, bw, bh, nbx, lost);
    else
        hipLaunchKernelGGL(scatter_kernel<T>, dim3(n_items), dim3(256), 0, s, cases, states, items, chunk, bw, bh, nbx,
                           lost);
    PIC_HIP(hipGetLastError());
    return 0;
}

// Positions of a store's live particles as (x*sx, y*sy) float pairs, packed
// from dst[0]: each block takes its item's range with one atomic (so items
// land in an unspecified order, each item's particles contiguous); pairs past
// cap are counted, not written. The store is only read.
template <typename T>
__global__ __launch_bounds__(256) void export_kernel(const CaseDev<T>* __restrict__ cases,
                                                     const CaseState* __restrict__ states,
                                                     const WorkItem* __restrict__ items, int chunk,
                                                     int32_t* counter, float2* __restrict__ dst, int64_t cap,
                                                     double sx, double sy) {
    __shared__ int sh_base;
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    if (n <= 0) return;
    if (threadIdx.x == 0) sh_base = atomicAdd(counter, n);
    __syncthreads();
    const int64_t base = sh_base;
    const gptr<const T> X = G((const T*)S.x[in]), Y = G((const T*)S.y[in]);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int64_t d = base + t;
        if (d < cap) {
            const int s = rg.at(t);
            float2 v;
            v.x = (float)((double)X[s] * sx);
            v.y = (float)((double)Y[s] * sy);
            dst[d] = v;
        }
    }
}

// Velocity moments of the particles of each work item, in FP64: out[4 b + k]
// for item b holds the sums of 1, vx, vy and vx^2 + vy^2 (fixed-order block
// reduction; the host adds the items in order).
template <typename T>
__global__ __launch_bounds__(256) void moments_kernel(const CaseDev<T>* __restrict__ cases,
                                                      const CaseState* __restrict__ states,
                                                      const WorkItem* __restrict__ items, int chunk,
                                                      double* __restrict__ out) {
    __shared__ double sh[4][256];
    const WorkItem it = items[blockIdx.x];
    const CaseDev<T>& C = cases[it.case_idx];
    const SpeciesDev<T>& S = C.sp[it.species];
    const int in = states[it.case_idx].cur;
    const Range rg = item_range(S, in, it, chunk);
    const int n = rg.n();
    const gptr<const T> VX = G((const T*)S.vx[in]), VY = G((const T*)S.vy[in]);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int s = rg.at(t);
        const double vx = (double)VX[s], vy = (double)VY[s];
        a0 += 1.0;
        a1 += vx;
        a2 += vy;
        a3 += vx * vx + vy * vy;
    }
    const int tid = threadIdx.x;
    sh[0][tid] = a0;
    sh[1][tid] = a1;
    sh[2][tid] = a2;
    sh[3][tid] = a3;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k)
            for (int q = 0; q < 4; ++q) sh[q][tid] = sh[q][tid] + sh[q][tid + k];
        __syncthreads();
    }
    if (tid < 4) G(out)[4 * (size_t)blockIdx.x + tid] = sh[tid][0];
}

template <typename T>
int launch_moments(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                   double* out, hipStream_t s) {
    if (n_items <= 0) return 0;
    hipLaunchKernelGGL(moments_kernel<T>, dim3(n_items), dim3(256), 0, s, cases, states, items, chunk, out);
    PIC_HIP(hipGetLastError());
    return 0;
}

// the work items of one species of a store (case 0): its nb bins, then
// `chunks` chunks of its spill region; thread 0 also clears the counter
__global__ __launch_bounds__(256) void species_items_kernel(WorkItem* __restrict__ items, int32_t* counter,
                                                            int species, int nb, int chunks) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k == 0) *G(counter) = 0;
    if (k >= nb + chunks) return;
    const gptr<int> w = G(&items[k].case_idx);
    w[0] = 0;                       // case_idx
    w[1] = species;
    w[2] = k < nb ? k : nb;         // bin (nb: the spill region)
    w[3] = k < nb ? 0 : k - nb;     // chunk
}

int launch_species_items(WorkItem* items, int32_t* counter, int species, int nb, int chunks, hipStream_t s) {
    const int n = nb + chunks;
    hipLaunchKernelGGL(species_items_kernel, dim3((n + 255) / 256), dim3(256), 0, s, items, counter, species, nb,
                       chunks);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_export(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                  int32_t* counter, float2* dst, int64_t cap, double sx, double sy, hipStream_t s) {
    if (n_items <= 0) return 0;
    hipLaunchKernelGGL(export_kernel<T>, dim3(n_items), dim3(256), 0, s, cases, states, items, chunk, counter, dst,
                       cap, sx, sy);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_gather(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                  int32_t* counter, hipStream_t s) {
    if (n_items <= 0) return 0;
    hipLaunchKernelGGL(gather_kernel<T>, dim3(n_items), dim3(256), 0, s, cases, states, items, chunk, counter);
    PIC_HIP(hipGetLastError());
    return 0;
}


template <typename T>
int launch_end_step(const CaseDev<T>* cases, CaseState* states, int n_cases, int advance, hipStream_t s,
                    int* status_out) {
    hipLaunchKernelGGL(end_step_kernel<T>, dim3(n_cases), dim3(256), 0, s, cases, states, advance, status_out);
    PIC_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int launch_clear_out(const CaseDev<T>* cases, const CaseState* states, int n_cases, hipStream_t s) {
    hipLaunchKernelGGL(clear_out_kernel<T>, dim3(n_cases), dim3(256), 0, s, cases, states);
    PIC_HIP(hipGetLastError());
    return 0;
}


template <typename T>
int launch_counts(const CaseDev<T>* cases, const CaseState* states, int c, int sp, long long* out, hipStream_t s) {
    hipLaunchKernelGGL(count_kernel<T>, dim3(1), dim3(256), 0, s, cases, states, c, sp, out);
    PIC_HIP(hipGetLastError());
    return 0;
}

#define PIC_LAUNCH_INST(T)                                                                                       \
    template int launch_push<T>(const CaseDev<T>*, const CaseState*, const WorkItem*, int, int, int, long long,  \
                                size_t, hipStream_t, bool, const int*, const int*);                              \
    template int launch_counts<T>(const CaseDev<T>*, const CaseState*, int, int, long long*, hipStream_t);       \
    template int launch_staged_append<T>(const double*, const double*, const int32_t*, long long, double, double, \
                                         int, const StagedDst<T>&, uint32_t, hipStream_t);                         \
    template int launch_init<T>(T*, T*, T*, T*, uint32_t*, int32_t*, int, int, int, double, uint32_t, uint32_t,  \
                                uint32_t, uint32_t, hipStream_t);                                                \
    template int launch_init_uniform<T>(T*, T*, T*, T*, uint32_t*, int32_t*, int, double, double, double, double, \
                                        double, double, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t);     \
    template int launch_hist<T>(const CaseDev<T>*, const CaseState*, const WorkItem*, int, int, int, int, int,   \
                                int, int32_t*, hipStream_t);                                                     \
    template int launch_scatter<T>(const CaseDev<T>*, const CaseState*, const WorkItem*, int, int, int, int, int, \
                                   int, unsigned long long*, hipStream_t);                                       \
    template int launch_gather<T>(const CaseDev<T>*, const CaseState*, const WorkItem*, int, int, int32_t*,       \
                                  hipStream_t);                                                                  \
    template int launch_moments<T>(const CaseDev<T>*, const CaseState*, const WorkItem*, int, int, double*,       \
                                   hipStream_t);                                                                 \
    template int launch_export<T>(const CaseDev<T>*, const CaseState*, const WorkItem*, int, int, int32_t*,       \
                                  float2*, int64_t, double, double, hipStream_t);                                \
    template int launch_end_step<T>(const CaseDev<T>*, CaseState*, int, int, hipStream_t, int*);                 \
    template int launch_clear_out<T>(const CaseDev<T>*, const CaseState*, int, hipStream_t);
PIC_LAUNCH_INST(float)
PIC_LAUNCH_INST(double)

}  // namespace pic
