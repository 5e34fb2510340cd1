// pic_common.hpp -- shared definitions for the gfx950 PIC kernels and the host
// objects behind the C ABI.
//
// Built with -ffp-contract=off: every fused multiply-add in a kernel is written
// as an explicit fma(), so the arithmetic of a step is fixed by the source and
// does not depend on the compiler's contraction choices.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>

#define PIC_QBITS 20          // fixed-point fraction bits of CIC weights
#define PIC_MAX_SPECIES 4
#define PIC_MAX_PROC 4
#define PIC_NCOUNT 16

enum { PIC_ELASTIC = 0, PIC_EXCITATION = 1, PIC_IONIZATION = 2, PIC_CHARGE_EXCHANGE = 3 };

// per-species event counters (device, uint64)
enum {
    PIC_C_ABSORBED = 0, PIC_C_NULL_CAND = 1, PIC_C_PROC0 = 2,
    PIC_C_BORN_SELF = 6, PIC_C_BORN_OTHER = 7, PIC_C_BORN_ABSORBED = 8,
    PIC_C_NUMAX_EXCEEDED = 9, PIC_C_REJECT_FALLBACK = 10, PIC_C_SPILLED = 11,
    PIC_C_LOST = 12, PIC_C_GLOBAL_DEPOSIT = 13
};

namespace pic {

// thread-local last error, reported through pic_last_error()
void set_error(const std::string& msg);
const char* last_error();

#define PIC_HIP(expr)                                                             \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::pic::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));  \
            return -2;                                                            \
        }                                                                         \
    } while (0)

#define PIC_REQUIRE(cond, msg)                  \
    do {                                        \
        if (!(cond)) {                          \
            ::pic::set_error(msg);              \
            return -1;                          \
        }                                       \
    } while (0)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter {id, step, call, case},
// key {seed, species}.  Same arithmetic on host and device.
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
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

// Philox2x32-10 (same paper): counter {c0, c1}, key k; one 32-bit product per
// round against two for the 4x32 form.
__host__ __device__ inline void philox2x32_10(uint32_t c0, uint32_t c1, uint32_t k, uint32_t out[2]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) k += 0x9E3779B9u;
        const uint64_t p = (uint64_t)0xD256D193u * c0;
        const uint32_t hi = (uint32_t)(p >> 32), lo = (uint32_t)p;
        c0 = hi ^ k ^ c1;
        c1 = lo;
    }
    out[0] = c0;
    out[1] = c1;
}

// The collision-candidacy draw, made for every particle at every step: the
// first word of Philox2x32-10 over counter {id, step}, keyed by cand_key of
// (seed, case, species) (itself the first word of Philox2x32-10 over {seed,
// case} with key species).  The collision itself draws from Stream.
__host__ __device__ inline uint32_t cand_key(uint32_t seed, uint32_t case_id, uint32_t species) {
    uint32_t o[2];
    philox2x32_10(seed, case_id, species, o);
    return o[0];
}
template <typename T>
__host__ __device__ inline T cand_uni(uint32_t id, uint32_t step, uint32_t key) {
    uint32_t o[2];
    philox2x32_10(id, step, key, o);
    return (T)(o[0] >> 8) * (T)0x1p-24;
}

// Natural log with a fixed evaluation order: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh((m-1)/(m+1)) by the odd series to degree 23 (Horner, fma).
__host__ __device__ inline double det_log(double x) {
    uint64_t u = (uint64_t)__builtin_bit_cast(unsigned long long, x);
    int e = (int)((u >> 52) & 0x7ff) - 1022;
    uint64_t mb = (u & 0x000fffffffffffffull) | (0x3feull << 52);
    double m = __builtin_bit_cast(double, (unsigned long long)mb);
    if (m < 0.70710678118654752) { m = m * 2.0; e -= 1; }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
    double p = 0.043478260869565216;
    p = fma(p, f2, 0.047619047619047616);
    p = fma(p, f2, 0.05263157894736842);
    p = fma(p, f2, 0.058823529411764705);
    p = fma(p, f2, 0.06666666666666667);
    p = fma(p, f2, 0.07692307692307693);
    p = fma(p, f2, 0.09090909090909091);
    p = fma(p, f2, 0.1111111111111111);
    p = fma(p, f2, 0.14285714285714285);
    p = fma(p, f2, 0.2);
    p = fma(p, f2, 0.3333333333333333);
    double s = f2 * p;
    double tf = 2.0 * f;
    double lm = fma(tf, s, tf);
    double de = (double)e;
    double r = fma(de, 0x1.df473de6af279p-26, lm);
    return fma(de, 0x1.62e42f0000000p-1, r);
}

// Per-particle random stream; uniforms carry 24 bits: (r >> 8) * 2^-24.
// The four words live in named registers (a dynamically indexed array would be
// placed in scratch memory on the GPU).
struct Stream {
    uint32_t c0, c1, c2, c3, k0, k1;
    uint32_t b0, b1, b2, b3;
    int idx;
    __host__ __device__ inline void refill() {
        uint32_t o[4];
        philox4x32_10(c0, c1, c2, c3, k0, k1, o);
        b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
        idx = 0;
    }
    __host__ __device__ inline Stream(uint32_t id, uint32_t step, uint32_t case_id, uint32_t seed,
                                      uint32_t species)
        : c0(id), c1(step), c2(0), c3(case_id), k0(seed), k1(species), idx(0) {
        refill();
    }
    __host__ __device__ inline uint32_t u32() {
        if (idx == 4) {
            c2 += 1;
            refill();
        }
        uint32_t v = idx == 0 ? b0 : idx == 1 ? b1 : idx == 2 ? b2 : b3;
        ++idx;
        return v;
    }
    template <typename T>
    __host__ __device__ inline T uni() { return (T)(u32() >> 8) * (T)0x1p-24; }
};

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

__host__ __device__ inline int div_up(int a, int b) { return (a + b - 1) / b; }

// Explicit global-address-space access.  A pointer read out of a struct is
// generic to the compiler, which then emits flat memory instructions; flat
// loads count against the LDS wait counter too, so every wait for an LDS access
// would also wait for the loads in flight.  Device buffers here are always
// global memory.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ inline gptr<T> G(T* p) { return (gptr<T>)p; }
// relaxed device-scope atomic add (the semantics of atomicAdd) on global memory
template <typename T>
__device__ inline T gatomic_add(T* p, T v) {
    return __hip_atomic_fetch_add(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace pic
