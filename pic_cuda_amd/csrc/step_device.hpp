// step_device.hpp -- the end of a particle step on the device, shared by the
// particle kernels (push.hip) and the first residual pass of a step's solve
// (multigrid.hip), which can run it in the same launch.
#pragma once
#include "pic_device.hpp"

namespace pic {

// one block per case: clear the counts of the buffer that becomes the next
// output, flip the parity and advance the step counter; with status_out, also
// the spill fill of every species of the new state (read back by the host
// after the Poisson solve to decide on re-binning)
template <typename T>
__device__ inline void end_step_block(const CaseDev<T>* __restrict__ cases, CaseState* states, int c,
                                      int advance_step, int* __restrict__ status_out, int* __restrict__ rebin_out) {
    const CaseDev<T>& C = cases[c];
    const int in = states[c].cur;
    __syncthreads();
    for (int s = 0; s < C.ns; ++s) {
        const SpeciesDev<T>& S = C.sp[s];
        const int nb = S.nb[in];
        for (int k = threadIdx.x; k <= nb; k += blockDim.x) {
            S.count[in][k] = 0;
            S.mcount[in][k] = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        states[c].cur = in ^ 1;
        if (advance_step) states[c].step += 1;
        if (status_out)
            for (int s = 0; s < PIC_MAX_SPECIES; ++s) {
                int v = 0;
                if (s < C.ns) {
                    const SpeciesDev<T>& S = C.sp[s];
                    v = S.count[in ^ 1][S.nb[in ^ 1]];
                }
                status_out[c * PIC_MAX_SPECIES + s] = v;
            }
        // the host's re-binning rule (a spill region more than half full), on
        // the device for a push enqueued before the host has read the status
        if (rebin_out) {
            int need = 0;
            for (int s = 0; s < C.ns; ++s) {
                const SpeciesDev<T>& S = C.sp[s];
                const int nb = S.nb[in ^ 1];
                if (S.count[in ^ 1][nb] > S.cap[in ^ 1][nb] / 2) need = 1;
            }
            rebin_out[c] = need;
        }
    }
}

}  // namespace pic
