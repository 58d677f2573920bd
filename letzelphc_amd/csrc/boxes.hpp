// boxes.hpp -- pruned sphere -> real-space transform of the FP32 psi boxes.
//
// Same result as scatter + three full FFT passes (the inverse transform of
// src/fft/bwdfft.c:13-120), with the work restricted to where the sphere
// lives, as the reference does with its z sticks (bwdfft.c:63-110):
//   1. the rotated sphere's G vectors are grouped by x plane (counting sort
//      on the box index; integer atomics, the order inside a plane has no
//      effect on the result);
//   2. k_plane_z: one block per (occupied x plane, box): the plane's
//      coefficients (SU(2) rotation, translation phase, time-reversal
//      conjugate as k_scatter) are scattered into an LDS plane, transformed
//      along z there and written out (the box needs no zeroing);
//   3. the y pass runs only on occupied planes;
//   4. the x pass treats the unoccupied planes as zeros without reading them.
// HBM traffic per band drops from ~7 box volumes to ~2.
#pragma once

#include "fft.hpp"
#include "wfc.hpp"

// per-plane counts: a block histogram in LDS, then one global atomic per
// touched plane (nx <= 1024)
__global__ __launch_bounds__(256) void k_plane_count(const int* __restrict__ idx, int64_t npw,
                                                     int nyz, int nx, int* __restrict__ cnt)
{
    __shared__ int h[1024];
    for (int t = threadIdx.x; t < nx; t += 256) h[t] = 0;
    __syncthreads();
    const int64_t g0 = (int64_t)blockIdx.x * 4096;
    for (int64_t g = g0 + threadIdx.x; g < npw && g < g0 + 4096; g += 256)
        atomicAdd(&h[idx[g] / nyz], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < nx; t += 256)
        if (h[t]) atomicAdd(&cnt[t], h[t]);
}

// exclusive scan of cnt (nx <= 1024) into off and cur (one block)
__global__ void k_plane_scan(const int* __restrict__ cnt, int nx, int* __restrict__ off,
                             int* __restrict__ cur)
{
    __shared__ int s[1024];
    const int t = threadIdx.x;
    s[t] = t < nx ? cnt[t] : 0;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1)
    {
        const int v = t >= d ? s[t - d] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    if (t < nx)
    {
        const int e = s[t] - cnt[t];
        off[t] = e;
        cur[t] = e;
    }
}

// G order grouped by plane: each block reserves a range per plane with one
// global atomic, then places its G vectors with LDS cursors
__global__ __launch_bounds__(256) void k_plane_fill(const int* __restrict__ idx, int64_t npw,
                                                    int nyz, int nx, int* __restrict__ cur,
                                                    int* __restrict__ perm)
{
    __shared__ int h[1024];
    for (int t = threadIdx.x; t < nx; t += 256) h[t] = 0;
    __syncthreads();
    const int64_t g0 = (int64_t)blockIdx.x * 4096;
    for (int64_t g = g0 + threadIdx.x; g < npw && g < g0 + 4096; g += 256)
        atomicAdd(&h[idx[g] / nyz], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < nx; t += 256)
        if (h[t]) h[t] = atomicAdd(&cur[t], h[t]);
    __syncthreads();
    for (int64_t g = g0 + threadIdx.x; g < npw && g < g0 + 4096; g += 256)
        perm[atomicAdd(&h[idx[g] / nyz], 1)] = (int)g;
}

// Stockham stages on tl lines of length N in LDS, line stride LS >= N
// (runtime line count; same butterflies as FftStagesC).
template <typename T, int SIGN, int N, int S, int NT, int LS = N>
struct FftStagesP
{
    static __device__ inline void run(cpx<T>*& in, cpx<T>*& out, const cpx<T>* W, int tid,
                                      int tl)
    {
        constexpr int R = fft_radix_c(N, S);
        if constexpr (R != 0)
        {
            constexpr int Ns = fft_ns_c(N, S);
            constexpr int NBF = N / R;
            constexpr int TWS = N / (Ns * R);
            for (int w = tid; w < tl * NBF; w += NT)
            {
                const int t = w / NBF, j = w - t * NBF;
                const int k = j % Ns;
                cpx<T> v[R];
                const cpx<T>* src = in + t * LS + j;
                v[0] = src[0];
#pragma unroll
                for (int r = 1; r < R; ++r) v[r] = cmul(src[r * NBF], tw<T, SIGN>(W, r * k * TWS));
                if constexpr (R <= 5)
                    dft_small<T, SIGN>(v, R, W, N);
                else
                {
                    cpx<T> o[R];
#pragma unroll
                    for (int s = 0; s < R; ++s)
                    {
                        cpx<T> acc = {0, 0};
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            acc = cadd(acc, cmul(v[r], tw<T, SIGN>(W, ((r * s) % R) * (N / R))));
                        o[s] = acc;
                    }
#pragma unroll
                    for (int s = 0; s < R; ++s) v[s] = o[s];
                }
                cpx<T>* dst = out + t * LS + (j / Ns) * Ns * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
            }
            __syncthreads();
            cpx<T>* tmp = in;
            in = out;
            out = tmp;
            FftStagesP<T, SIGN, N, S + 1, NT, LS>::run(in, out, W, tid, tl);
        }
    }
};

// In-place Stockham stages on TLC lines of length N in one LDS buffer,
// element j of line t at t * LS + j * ES: each thread loads its
// butterflies' inputs and computes them in registers, a barrier, then it
// stores the outputs over the same buffer.  TF: consecutive threads take
// consecutive lines (for lines adjacent in LDS, ES > 1), else consecutive
// butterflies of one line.  Same butterflies as FftStagesP.
template <typename T, int SIGN, int N, int S, int NT, int LS, int ES, int TLC, bool TF>
struct FftStagesIP
{
    static __device__ inline void run(cpx<T>* buf, const cpx<T>* W, int tid)
    {
        constexpr int R = fft_radix_c(N, S);
        if constexpr (R != 0)
        {
            constexpr int Ns = fft_ns_c(N, S);
            constexpr int NBF = N / R;
            constexpr int TWS = N / (Ns * R);
            constexpr int NW = TLC * NBF;
            constexpr int PER = (NW + NT - 1) / NT;
            cpx<T> v[PER][R];
#pragma unroll
            for (int u = 0; u < PER; ++u)
            {
                const int w = tid + u * NT;
                if (NW % NT == 0 || w < NW)
                {
                    const int t = TF ? w % TLC : w / NBF, j = TF ? w / TLC : w % NBF;
                    const int k = j % Ns;
                    const cpx<T>* src = buf + t * LS + j * ES;
                    v[u][0] = src[0];
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        v[u][r] = cmul(src[r * NBF * ES], tw<T, SIGN>(W, r * k * TWS));
                    if constexpr (R <= 5)
                        dft_small<T, SIGN>(v[u], R, W, N);
                    else
                    {
                        cpx<T> o[R];
#pragma unroll
                        for (int s = 0; s < R; ++s)
                        {
                            cpx<T> acc = {0, 0};
#pragma unroll
                            for (int r = 0; r < R; ++r)
                                acc = cadd(acc,
                                           cmul(v[u][r], tw<T, SIGN>(W, ((r * s) % R) * (N / R))));
                            o[s] = acc;
                        }
#pragma unroll
                        for (int s = 0; s < R; ++s) v[u][s] = o[s];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < PER; ++u)
            {
                const int w = tid + u * NT;
                if (NW % NT == 0 || w < NW)
                {
                    const int t = TF ? w % TLC : w / NBF, j = TF ? w / TLC : w % NBF;
                    const int k = j % Ns;
                    cpx<T>* dst = buf + t * LS + ((j / Ns) * Ns * R + k) * ES;
#pragma unroll
                    for (int r = 0; r < R; ++r) dst[r * Ns * ES] = v[u][r];
                }
            }
            __syncthreads();
            FftStagesIP<T, SIGN, N, S + 1, NT, LS, ES, TLC, TF>::run(buf, W, tid);
        }
    }
};

// LDS bytes of k_plane_z for a (ny, nz) plane (z twiddles; with the y pass
// fused, the y twiddles and the [z][y] layout's padded line stride ny + 1)
static inline size_t plane_lds_bytes(int ny, int nz, bool fuse_y = false)
{
    return sizeof(c32) * (fuse_y ? (size_t)2 * (ny + 1) * nz + nz + ny : (size_t)2 * ny * nz + nz);
}

// One block per (x plane, box b = set * nspinor + sigma): gather the plane's
// coefficients into LDS, backward FFT along z (e^{+i}, unnormalised), write
// the plane.  Empty planes are skipped (never written).  NY > 0 (= ny): the
// y pass too, in LDS: the z-transformed plane transposed to [z][y], FFT of
// its NZ lines of length NY, written back in [y][z] order, so the plane
// leaves the kernel transformed along y and z.  A plane's buffers take most
// of the LDS (one block per CU at 96 x 96), so the block is 16 waves.
#define PZ_THREADS 1024
template <int NZ, int NY = 0>
__global__ __launch_bounds__(PZ_THREADS) void k_plane_z(
    const z64* __restrict__ wfc, int nspinor, int64_t npw, const int* __restrict__ idx,
    const z64* __restrict__ ph, SideGeom geo, const int* __restrict__ perm,
    const int* __restrict__ off, const int* __restrict__ cnt, int ny, int64_t nf,
    const c32* __restrict__ Wg, const c32* __restrict__ Wgy, c32* __restrict__ box)
{
    extern __shared__ __attribute__((aligned(16))) char pz_smem[];
    const int x = blockIdx.x;
    const int n = cnt[x];
    if (n == 0) return;
    const int b = blockIdx.y;
    const int64_t s = b / nspinor;
    const int sig = b - (int)s * nspinor;
    const int nyz = ny * NZ;
    // buffers hold the [y][z] plane, or with the fused y pass the [z][y] one
    // with line stride NY + 1 (a stride of NY lands every lane of a
    // transpose on one bank)
    constexpr int LY = NY > 0 ? NY + 1 : 0;
    const int cap = NY > 0 ? NZ * LY : nyz;
    c32* p0 = reinterpret_cast<c32*>(pz_smem);
    c32* p1 = p0 + cap;
    c32* W = p1 + cap;
    c32* Wy = W + NZ;
    const int tid = threadIdx.x;
    for (int t = tid; t < NZ; t += PZ_THREADS) W[t] = Wg[t];
    if constexpr (NY > 0)
        for (int t = tid; t < NY; t += PZ_THREADS) Wy[t] = Wgy[t];
    for (int t = tid; t < nyz; t += PZ_THREADS) p0[t] = c32{0.f, 0.f};
    __syncthreads();
    const z64* w = wfc + s * nspinor * npw;
    const int o = off[x];
    for (int k = tid; k < n; k += PZ_THREADS)
    {
        const int g = perm[o + k];
        const z64 p = ph[g];
        z64 c;
        if (nspinor == 1)
            c = cmul(w[g], p);
        else
        {
            const z64 u = w[g], d = w[npw + g];
            c = sig == 0 ? cadd(cmul(u, geo.su2[0]), cmul(d, geo.su2[1]))
                         : cadd(cmul(u, geo.su2[2]), cmul(d, geo.su2[3]));
            c = cmul(c, p);
        }
        if (geo.conj) c.y = -c.y;
        p0[idx[g] - x * nyz] = c32{(float)c.x, (float)c.y};
    }
    __syncthreads();
    c32* in = p0;
    c32* out = p1;
    FftStagesP<float, 1, NZ, 0, PZ_THREADS>::run(in, out, W, tid, ny);
    c32* dst = box + (int64_t)b * nf + (int64_t)x * nyz;
    if constexpr (NY > 0)
    {
        for (int t = tid; t < nyz; t += PZ_THREADS) out[(t % NZ) * LY + t / NZ] = in[t];
        __syncthreads();
        c32* in2 = out;
        c32* out2 = in;
        FftStagesP<float, 1, NY, 0, PZ_THREADS, LY>::run(in2, out2, Wy, tid, NZ);
        for (int t = tid; t < nyz; t += PZ_THREADS) dst[t] = in2[(t % NZ) * LY + t / NZ];
    }
    else
        for (int t = tid; t < nyz; t += PZ_THREADS) dst[t] = in[t];
}

// LDS bytes of k_plane_zy: one [y][z] plane and the twiddles of both axes
static inline size_t plane_zy_lds_bytes(int n)
{
    return sizeof(c32) * ((size_t)n * n + 2 * (size_t)n);
}

// k_plane_z with the fused y pass on square y-z planes (N x N) in one LDS
// buffer, [y][z]: gather, z stages in place (lines contiguous), y stages in
// place on the same layout (lines N apart in y, consecutive threads on
// consecutive z), the plane written as it lies.  About 74 KB of LDS per
// block instead of 147 KB and no transposes, so two blocks share a CU and
// one block's gather overlaps the other's transforms.  Per line the
// arithmetic is that of k_plane_z<N, N>.
template <int N>
__global__ __launch_bounds__(PZ_THREADS, 8) void k_plane_zy(
    const z64* __restrict__ wfc, int nspinor, int64_t npw, const int* __restrict__ idx,
    const z64* __restrict__ ph, SideGeom geo, const int* __restrict__ perm,
    const int* __restrict__ off, const int* __restrict__ cnt, int64_t nf,
    const c32* __restrict__ Wg, const c32* __restrict__ Wgy, c32* __restrict__ box)
{
    extern __shared__ __attribute__((aligned(16))) char pz_smem[];
    const int x = blockIdx.x;
    const int n = cnt[x];
    if (n == 0) return;
    const int b = blockIdx.y;
    const int64_t s = b / nspinor;
    const int sig = b - (int)s * nspinor;
    constexpr int NYZ = N * N;
    c32* buf = reinterpret_cast<c32*>(pz_smem);
    c32* W = buf + NYZ;
    c32* Wy = W + N;
    const int tid = threadIdx.x;
    for (int t = tid; t < N; t += PZ_THREADS)
    {
        W[t] = Wg[t];
        Wy[t] = Wgy[t];
    }
    for (int t = tid; t < NYZ; t += PZ_THREADS) buf[t] = c32{0.f, 0.f};
    __syncthreads();
    const z64* w = wfc + s * nspinor * npw;
    const int o = off[x];
    for (int k = tid; k < n; k += PZ_THREADS)
    {
        const int g = perm[o + k];
        const z64 p = ph[g];
        z64 c;
        if (nspinor == 1)
            c = cmul(w[g], p);
        else
        {
            const z64 u = w[g], d = w[npw + g];
            c = sig == 0 ? cadd(cmul(u, geo.su2[0]), cmul(d, geo.su2[1]))
                         : cadd(cmul(u, geo.su2[2]), cmul(d, geo.su2[3]));
            c = cmul(c, p);
        }
        if (geo.conj) c.y = -c.y;
        buf[idx[g] - x * NYZ] = c32{(float)c.x, (float)c.y};
    }
    __syncthreads();
    // z: line y at y * N, elements contiguous
    FftStagesIP<float, 1, N, 0, PZ_THREADS, N, 1, N, false>::run(buf, W, tid);
    // y: line z at z, elements N apart
    FftStagesIP<float, 1, N, 0, PZ_THREADS, 1, N, N, true>::run(buf, Wy, tid);
    c32* dst = box + (int64_t)b * nf + (int64_t)x * NYZ;
    for (int t = tid; t < NYZ; t += PZ_THREADS) dst[t] = buf[t];
}

// launch k_plane_z for a compile-time NZ; false if NZ is not compiled.
// Wy not NULL asks for the fused y pass (cubic y-z planes, ny == nz); false
// if that form is not compiled for this size or does not fit in LDS.
static bool plane_z_launch(int nz, int nx, int nbox, size_t lds, const z64* wfc, int nspinor,
                           int64_t npw, const int* idx, const z64* ph, const SideGeom& geo,
                           const int* perm, const int* off, const int* cnt, int ny, int64_t nf,
                           const c32* W, const c32* Wy, c32* box, hipStream_t st)
{
    if (Wy && (ny != nz || lds > 160 * 1024)) return false;
#define PZ_CASE(NZV)                                                                              \
    case NZV:                                                                                     \
        if (Wy && plane_zy_lds_bytes(NZV) <= 80 * 1024)                                           \
        {                                                                                         \
            const int lz = (int)plane_zy_lds_bytes(NZV);                                          \
            if (hipFuncSetAttribute((const void*)k_plane_zy<NZV>,                                 \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lz) != hipSuccess) \
                return false;                                                                     \
            hipLaunchKernelGGL((k_plane_zy<NZV>), dim3((unsigned)nx, (unsigned)nbox),             \
                               dim3(PZ_THREADS), lz, st, wfc, nspinor, npw, idx, ph, geo, perm,   \
                               off, cnt, nf, W, Wy, box);                                         \
            return true;                                                                          \
        }                                                                                         \
        if (Wy)                                                                                   \
        {                                                                                         \
            if (hipFuncSetAttribute((const void*)k_plane_z<NZV, NZV>,                             \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=      \
                hipSuccess)                                                                       \
                return false;                                                                     \
            hipLaunchKernelGGL((k_plane_z<NZV, NZV>), dim3((unsigned)nx, (unsigned)nbox),         \
                               dim3(PZ_THREADS), lds, st, wfc, nspinor, npw, idx, ph, geo, perm,  \
                               off, cnt, ny, nf, W, Wy, box);                                     \
            return true;                                                                          \
        }                                                                                         \
        if (hipFuncSetAttribute((const void*)k_plane_z<NZV>,                                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) \
            return false;                                                                         \
        hipLaunchKernelGGL(k_plane_z<NZV>, dim3((unsigned)nx, (unsigned)nbox), dim3(PZ_THREADS), lds, st, \
                           wfc, nspinor, npw, idx, ph, geo, perm, off, cnt, ny, nf, W, nullptr, box); \
        return true;
    switch (nz)
    {
        PZ_CASE(12)
        PZ_CASE(24)
        PZ_CASE(32)
        PZ_CASE(36)
        PZ_CASE(45)
        PZ_CASE(48)
        PZ_CASE(60)
        PZ_CASE(64)
        PZ_CASE(72)
        PZ_CASE(96)
        PZ_CASE(108)
        PZ_CASE(216)
        default: return false;
    }
#undef PZ_CASE
}
