// common.hpp -- shared device types and helpers for the gfx950 e-ph kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/lelph_hip.h"

// The reference's truncated pi (src/common/constants.h:11) enters every
// 2*pi phase of the path; twiddles of the FFT use the exact value.
#define LELPH_RPI 3.1415927
#define LELPH_EPS 1e-5
#define LELPH_EPS6 1e-6
#define LELPH_E2 2.0

template <typename T>
struct cpx
{
    T x, y;
};
using z64 = cpx<double>;
using c32 = cpx<float>;

template <typename T>
__host__ __device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ inline cpx<T> cadd(cpx<T> a, cpx<T> b)
{
    return {a.x + b.x, a.y + b.y};
}
template <typename T>
__host__ __device__ inline cpx<T> csub(cpx<T> a, cpx<T> b)
{
    return {a.x - b.x, a.y - b.y};
}
template <typename T>
__host__ __device__ inline cpx<T> cconj(cpx<T> a)
{
    return {a.x, -a.y};
}
template <typename T>
__host__ __device__ inline cpx<T> cscale(cpx<T> a, T s)
{
    return {a.x * s, a.y * s};
}

// e^{sgn 2 pi i G0.r} at grid point p = (ix, iy, iz) (r = (ix/Nx, iy/Ny,
// iz/Nz) in crystal units): G0.r reduced exactly in integers, the fraction
// in FP64, rounded once to T
template <typename T>
__device__ inline cpx<T> box_phase(int64_t p, int Nx, int Ny, int Nz, int g0, int g1, int g2,
                                   int sgn)
{
    const int ix = (int)(p / ((int64_t)Ny * Nz)), iy = (int)((p / Nz) % Ny), iz = (int)(p % Nz);
    auto fr = [](int g, int i, int n) {
        int64_t m = ((int64_t)g * i) % n;
        if (m < 0) m += n;
        return (double)m / n;
    };
    double f = fr(g0, ix, Nx) + fr(g1, iy, Ny) + fr(g2, iz, Nz);
    f -= floor(f);
    double s, c;
    sincospi(2.0 * f * sgn, &s, &c);
    return {(T)c, (T)s};
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// accesses (lgkmcnt) and synchronises, but unlike __syncthreads it does not
// wait for outstanding global loads, so register prefetches stay in flight
// across it.  Global -> LDS copies are waited for explicitly (vmcnt) before
// it.  The memory clobber keeps the compiler from moving LDS accesses
// across the barrier.
__device__ inline void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

#define HIP_TRY(expr)                                                   \
    do                                                                  \
    {                                                                   \
        hipError_t _e = (expr);                                         \
        if (_e != hipSuccess)                                           \
        {                                                               \
            fprintf(stderr, "lelph: HIP error %s at %s:%d\n",           \
                    hipGetErrorString(_e), __FILE__, __LINE__);         \
            return LELPH_ERR_HIP;                                       \
        }                                                               \
    } while (0)

#define LELPH_TRY(expr)          \
    do                           \
    {                            \
        int _s = (expr);         \
        if (_s != LELPH_OK)      \
            return _s;           \
    } while (0)

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
