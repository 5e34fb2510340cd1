// push.hpp -- launchers of the particle kernels in push.hip.
#pragma once
#include "pic_device.hpp"
#include "field.hpp"

namespace pic {

// walls: some case of the launch has internal wall segments (false launches
// the instantiation without the segment test); gate_conv / gate_rebin (device
// ints, null: no gate): the push runs only if *gate_conv != 0 and
// *gate_rebin == 0 when it starts, and otherwise writes nothing
template <typename T>
int launch_push(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items,
                int spill_chunk, int flags, long long mcc_key, size_t lds, hipStream_t s, bool walls,
                const int* gate_conv = nullptr, const int* gate_rebin = nullptr);
template <typename T>
int launch_counts(const CaseDev<T>* cases, const CaseState* states, int c, int sp, long long* out, hipStream_t s);
template <typename T>
int launch_init(T* x, T* y, T* vx, T* vy, uint32_t* id, int32_t* count_ptr, int n, int nx, int ny, double vth,
                uint32_t seed, uint32_t case_id, uint32_t species, uint32_t id0, hipStream_t s);
template <typename T>
int launch_init_uniform(T* x, T* y, T* vx, T* vy, uint32_t* id, int32_t* count_ptr, int n, double x0, double x1,
                        double y0, double y1, double v0, double v1, uint32_t seed, uint32_t case_id,
                        uint32_t species, uint32_t id0, hipStream_t s);
// destinations of staged particles: per species, the five arrays at its first
// free spill slot, and the slots taken so far (zeroed before the launch)
template <typename T>
struct StagedDst {
    T* x[PIC_MAX_SPECIES];
    T* y[PIC_MAX_SPECIES];
    T* vx[PIC_MAX_SPECIES];
    T* vy[PIC_MAX_SPECIES];
    uint32_t* id[PIC_MAX_SPECIES];
    int32_t* fill;
};
// per-species counts of type[0..n) into counts[0..PIC_MAX_SPECIES) and, at
// counts[PIC_MAX_SPECIES], the entries that are not a species (counts zeroed before)
int launch_staged_count(const int32_t* type, long long n, int n_species, int32_t* counts, hipStream_t s);
// entry i (r, v: (x, y) double pairs in domain units) to species type[i], id id0 + i
template <typename T>
int launch_staged_append(const double* r, const double* v, const int32_t* type, long long n, double dx, double dy,
                         int n_species, const StagedDst<T>& d, uint32_t id0, hipStream_t s);
template <typename T>
int launch_hist(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                int bw, int bh, int nbx, int nb_new, int32_t* hist, hipStream_t s);
template <typename T>
int launch_scatter(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items,
                   int chunk, int bw, int bh, int nbx, int nb_new, unsigned long long* lost, hipStream_t s);
template <typename T>
int launch_gather(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                  int32_t* counter, hipStream_t s);
// per work item: sums of 1, vx, vy, vx^2 + vy^2 (FP64) into out[4 * item + k]
template <typename T>
int launch_moments(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                   double* out, hipStream_t s);
template <typename T>
int launch_export(const CaseDev<T>* cases, const CaseState* states, const WorkItem* items, int n_items, int chunk,
                  int32_t* counter, float2* dst, int64_t cap, double sx, double sy, hipStream_t s);
// the work items of a store's species (nb bins, then `chunks` spill chunks)
// written on the device, and the counter at `counter` cleared
int launch_species_items(WorkItem* items, int32_t* counter, int species, int nb, int chunks, hipStream_t s);
// end of step: clear the next output's counts, flip the parity, advance the
// step; status_out (PIC_MAX_SPECIES ints per case, may be null) receives the
// spill fill of every species
template <typename T>
int launch_end_step(const CaseDev<T>* cases, CaseState* states, int n_cases, int advance, hipStream_t s,
                    int* status_out = nullptr);
// clear the bin counts of the output parity (after a trial push)
template <typename T>
int launch_clear_out(const CaseDev<T>* cases, const CaseState* states, int n_cases, hipStream_t s);

// dynamic LDS of push_kernel for a bin tile
template <typename T>
inline size_t push_lds_bytes(int bw, int bh, int halo) {
    const size_t aw = bw + 1 + 2 * halo, ah = bh + 1 + 2 * halo;
    return ((aw * ah + 1) & ~(size_t)1) * 8 + (size_t)(bw + 1) * (bh + 1) * 2 * sizeof(T) + (size_t)PIC_MSTAGE * (4 * sizeof(T) + 8) +
           PIC_NCOUNT * 4 + 32 * sizeof(T) + (size_t)4 * PIC_MQUEUE * (4 * sizeof(T) + 4);
}

}  // namespace pic
