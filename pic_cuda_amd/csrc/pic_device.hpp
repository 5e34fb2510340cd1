// pic_device.hpp -- device-resident descriptors of cases (independent plasmas)
// and species stores, shared by the kernels and the host objects.
#pragma once
#include "pic_common.hpp"

namespace pic {

// One species store: SoA particles in two buffers (parity 0/1).  Each buffer
// is cut into bins (one per tile of bw x bh cells) plus a spill region with
// index nb.  count[p][b] counts the slots requested in bin b of buffer p;
// the filled part is min(count, cap).
template <typename T>
struct SpeciesDev {
    T* x[2];
    T* y[2];
    T* vx[2];
    T* vy[2];
    uint32_t* id[2];
    const int32_t* start[2];   // nb[p]+1 entries
    const int32_t* cap[2];
    int32_t* count[2];         // front part of each bin (survivors that stayed); spill fill at nb
    int32_t* mcount[2];        // back part of each bin (particles that moved in)
    int nb[2];
    // the same arrays packed: x, y, vx, vy, id of parity p in one block of
    // pcap slots per array (blk[p]); count, mcount, start, cap of parity p in
    // one block of 4 x (nb[p] + 1) entries (lay[p])
    char* blk[2];
    int32_t* lay[2];
    uint32_t pcap;
    // physics
    T kq;                      // q dt / m
    T efac;                    // E[eV] = efac * v^2
    T p_null;                  // 1 - exp(-nu_max dt)
    T nu_max;
    T inv_de;                  // 1 / table energy step
    T inv_emax;                // 1 / table end: energy bands of the majorant
    T band_pnull[16];          // candidate probability per energy band
    T band_numax[16];          // majorant per energy band
    int K;                     // table points
    int n_proc;
    int ptype[PIC_MAX_PROC];
    T thr[PIC_MAX_PROC];       // eV
    const T* tab;              // n_proc * K: n_gas * sigma_k(j dE)
    double gas_vth;            // sqrt(k T_gas / M) of the neutral partner species
    uint32_t key;              // Philox key word (species index)
    int other;                 // species that receives born ions (-1: none)
    unsigned long long* acc;   // Q20 charge accumulator (nx*ny)
    unsigned long long* counters;
};

template <typename T>
struct CaseDev {
    int nx, ny;                // nodes
    int bw, bh, nbx, nby;      // bin tile (cells) and bins per axis
    int bw_log2, bh_log2;      // log2 of the tile sides, -1 when not a power of two
    int halo;                  // deposit window halo (nodes)
    int box;                   // absorbing box walls on the domain edges
    int n_walls;
    int ns;
    double dx, dy;
    double wall_eps;
    T dtdx, dtdy;
    const double* phi;
    const double* walls;       // cell units x1,y1,x2,y2
    // per-bin candidate walls (CSR over the bins; null: every particle tests
    // every wall): the walls whose bounding box comes within wall_margin + 1
    // cells of the bin, for particles whose path stays within wall_margin cells
    const int* wall_off;       // nbx*nby + 1
    const int* wall_idx;
    int wall_margin;
    int wall_pad;
    uint32_t seed, case_id;
    SpeciesDev<T> sp[PIC_MAX_SPECIES];
};

// mutable per-case state, advanced on the device by end_step
struct CaseState {
    uint32_t step;
    int cur;                   // input buffer parity
    int flags;                 // bit0 kick, bit1 mcc, bit2 drift
    int pad;
};

enum { PIC_F_KICK = 1, PIC_F_MCC = 2, PIC_F_DRIFT = 4, PIC_F_DEPOSIT = 8 };

// entries of the LDS stage of particles leaving their bin (a pass of 256
// particles adds at most 512: 256 movers and 256 born particles)
#define PIC_MSTAGE 768
// entries of each wavefront's queue of deferred collision candidates
#define PIC_MQUEUE 64
// layouts of up to this many bins are sorted in LDS when re-binning (two int
// per bin: 64 KB)
#define PIC_SORT_LDS_BINS 8192

// work item of the particle kernels: a bin of a species of a case, or a chunk
// of its spill region (bin == nb)
struct WorkItem {
    int case_idx;
    int species;
    int bin;
    int chunk;
};

}  // namespace pic
