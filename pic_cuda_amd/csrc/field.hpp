// field.hpp -- launchers of the grid kernels in field.hip.
#pragma once
#include "pic_common.hpp"

namespace pic {

struct ChargeArgs {
    unsigned long long* acc[PIC_MAX_SPECIES];
    double cfac[PIC_MAX_SPECIES];   // rhs factor: -q w / (dx dy eps0 2^(2Q))
    double qfac[PIC_MAX_SPECIES];   // rho factor:  q w / (dx dy 2^(2Q))
    int ns;
    int zero;                       // clear the accumulators after reading
    double* rho;                    // may be null
    double* rhs;
};

// node i of the charge conversion: rhs (and rho) from the Q20 accumulators,
// which are cleared when a.zero is set; returns the rhs value written
__device__ inline double charge_node(const ChargeArgs& a, size_t i) {
    double r = 0.0, q = 0.0;
    for (int s = 0; s < a.ns; ++s) {
        const double v = (double)a.acc[s][i];
        r = fma(a.cfac[s], v, r);
        q = fma(a.qfac[s], v, q);
        if (a.zero) a.acc[s][i] = 0ull;
    }
    if (a.rho) a.rho[i] = q;
    a.rhs[i] = r;
    return r;
}


// batched: one ChargeArgs per case, items (case, 256-node chunk) in `chunks`
// (int4 {case, chunk, 0, 0}); nn per case in `nns`
int launch_charge_batch(const ChargeArgs* args, const long long* nns, const int4* chunks, int n_chunks,
                        hipStream_t s);
int launch_efield(const double* phi, double2* E, int nx, int ny, double dx, double dy, hipStream_t s);
// zero the interior nodes of an n x n field (the boundary keeps its values)
int launch_zero_interior(double* f, int n, hipStream_t s);
// partial[c] = sum of a[i]^2 over the 256-value chunk c, c < ceil(n / 256)
int launch_sumsq(const double* a, size_t n, double* partial, hipStream_t s);

}  // namespace pic
