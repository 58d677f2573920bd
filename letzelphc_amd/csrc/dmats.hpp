// dmats.hpp -- electronic representation matrices of a symmetry R:
// electronic_reps (src/symmetries/elec_reps.c:16-413),
//     D[s][b][a] = < S2 k2, b | U(R) U(S1) | k1, a >
// with k_BZ = S1 k1 and R k_BZ = S2 k2 (+ G).  The reference sorts both
// rotated G lists and matches them with bsearch (gsort.c:49-95); here the
// R S1 G list of k1 is entered into a box-index map (G mod N is a bijection
// on the sphere), the S2 G + G0 list of k2 looks its partners up in that map,
// and the sandwich is one complex GEMM per spin over the k1 sphere
// (k_zgemm_nt, FP64 MFMA).  Missing partners contribute zero, as in
// elec_reps.c:334-341.
#pragma once

#include "common.hpp"

// host-computed FP64 geometry of one D-matrix (elec_reps.c:83-176)
struct DmGeom
{
    double M1[9];     // cartesian G (no 2pi) -> crystal S1 G
    double M2[9];     // -> crystal R S1 G
    double M3[9];     // -> crystal S2 G
    double ulmc[3];   // crystal G0 = S2 k2 - R S1 k1
    double kcrys[3];  // crystal S1 k1
    double Rkcrys[3]; // crystal R S1 k1
    double tau1c[3], tauRc[3], tau2c[3]; // crystal fractional translations
    int tr1, trR, tr2;
    z64 su1[4], suR[4], su2[4];
};

__device__ inline z64 dm_phase(const double* k, const double* g, const double* tau)
{
    // e^{-i 2 pi (k + G).tau}, trans_wfc.c:15-52 (ELPH_PI)
    const double d = (k[0] + g[0]) * tau[0] + (k[1] + g[1]) * tau[1] + (k[2] + g[2]) * tau[2];
    double s, c;
    sincos(-2.0 * LELPH_RPI * d, &s, &c);
    return {c, s};
}

__device__ inline void dm_rot(const double* M, const double* G, double* out)
{
    for (int i = 0; i < 3; ++i) out[i] = G[0] * M[3 * i] + G[1] * M[3 * i + 1] + G[2] * M[3 * i + 2];
}

__device__ inline int dm_boxidx(const double* c, int Nx, int Ny, int Nz)
{
    const int N[3] = {Nx, Ny, Nz};
    int b[3];
    for (int i = 0; i < 3; ++i)
    {
        int v = (int)rint(c[i]);
        v = v < 0 ? v + N[i] : v;
        if (v < 0 || v >= N[i]) return -1;
        b[i] = v;
    }
    return (b[0] * Ny + b[1]) * Nz + b[2];
}

// k1 side, per G: the map entry of R S1 G and the three phases
// (tau1 with S1 G, tauR and tau2 with R S1 G)
__global__ void k_dm_side1(const double* __restrict__ gvec, int64_t npw, DmGeom geo, int Nx,
                           int Ny, int Nz, int* __restrict__ map, z64* __restrict__ ph,
                           int* __restrict__ err)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npw) return;
    const double G[3] = {gvec[3 * g], gvec[3 * g + 1], gvec[3 * g + 2]};
    double g1[3], g2[3];
    dm_rot(geo.M1, G, g1);
    dm_rot(geo.M2, G, g2);
    const int b = dm_boxidx(g2, Nx, Ny, Nz);
    if (b < 0)
        atomicOr(err, 1);
    else
        map[b] = (int)g;
    ph[3 * g] = dm_phase(geo.kcrys, g1, geo.tau1c);
    ph[3 * g + 1] = dm_phase(geo.Rkcrys, g2, geo.tauRc);
    ph[3 * g + 2] = dm_phase(geo.Rkcrys, g2, geo.tau2c);
}

// k2 side, per G: partner position in the k1 list (-1 if none)
__global__ void k_dm_side2(const double* __restrict__ gvec, int64_t npw, DmGeom geo, int Nx,
                           int Ny, int Nz, const int* __restrict__ map, int* __restrict__ pos)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npw) return;
    const double G[3] = {gvec[3 * g], gvec[3 * g + 1], gvec[3 * g + 2]};
    double g3[3];
    dm_rot(geo.M3, G, g3);
    for (int i = 0; i < 3; ++i) g3[i] += geo.ulmc[i];
    const int b = dm_boxidx(g3, Nx, Ny, Nz);
    pos[g] = b < 0 ? -1 : map[b];
}

__device__ inline void dm_su2(const z64* su, z64& u, z64& d)
{
    const z64 u2 = cadd(cmul(u, su[0]), cmul(d, su[1]));
    const z64 d2 = cadd(cmul(u, su[2]), cmul(d, su[3]));
    u = u2;
    d = d2;
}

// wRS[set][sp][g]: U(R) U(S1) applied to the k1 coefficients (elec_reps.c:310-328)
__global__ void k_dm_rot1(const z64* __restrict__ w1, int64_t nsets, int nspinor, int64_t npw,
                          const z64* __restrict__ ph, DmGeom geo, z64* __restrict__ wRS)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsets * npw) return;
    const int64_t s = t / npw, g = t - s * npw;
    const z64* w = w1 + s * nspinor * npw;
    z64 u = w[g], d = nspinor == 2 ? w[npw + g] : z64{0, 0};
    if (nspinor == 2) dm_su2(geo.su1, u, d);
    u = cmul(u, ph[3 * g]);
    d = cmul(d, ph[3 * g]);
    if (geo.tr1)
    {
        u.y = -u.y;
        d.y = -d.y;
    }
    if (nspinor == 2) dm_su2(geo.suR, u, d);
    u = cmul(u, ph[3 * g + 1]);
    d = cmul(d, ph[3 * g + 1]);
    if (geo.trR)
    {
        u.y = -u.y;
        d.y = -d.y;
    }
    z64* o = wRS + s * nspinor * npw;
    o[g] = u;
    if (nspinor == 2) o[npw + g] = d;
}

// wS2[set][sp][pos]: the k2 coefficients moved to their k1 partners, U(S2),
// translation, and the sandwich conjugate (elec_reps.c:329-373); the
// destination is zeroed beforehand
__global__ void k_dm_rot2(const z64* __restrict__ w2, int64_t nsets, int nspinor, int64_t npw2,
                          int64_t npw1, const int* __restrict__ pos, const z64* __restrict__ ph,
                          DmGeom geo, z64* __restrict__ wS2)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsets * npw2) return;
    const int64_t s = t / npw2, g = t - s * npw2;
    const int p = pos[g];
    if (p < 0) return;
    const z64* w = w2 + s * nspinor * npw2;
    z64 u = w[g], d = nspinor == 2 ? w[npw2 + g] : z64{0, 0};
    if (nspinor == 2) dm_su2(geo.su2, u, d);
    const z64 f = ph[3 * p + 2];
    u = cmul(u, f);
    d = cmul(d, f);
    if (!geo.tr2)
    {
        u.y = -u.y;
        d.y = -d.y;
    }
    z64* o = wS2 + s * nspinor * npw1;
    o[p] = u;
    if (nspinor == 2) o[npw1 + p] = d;
}

// Batched complex FP64 GEMM for the q-star rotation (elph_q_rotate,
// src/symmetries/elphq_expand.c:16-73): C[b] = opA(A[b]) opB(B[b]) with
// opA in {A, conj(A)}, opB in {B, conj(B), B^T}; square n x n matrices, batch
// strides sa/sb/sc (elements; 0 = the same matrix for every batch entry),
// 32 x 32 tiles through LDS.
template <bool CONJA, bool CONJB, bool TRANSB>
__global__ __launch_bounds__(256) void k_zgemm_bat(const z64* __restrict__ A, int64_t sa,
                                                   const z64* __restrict__ B, int64_t sb,
                                                   z64* __restrict__ Cm, int64_t sc, int n)
{
    __shared__ z64 ta[32][33];
    __shared__ z64 tb[32][33];
    const int tiles = (n + 31) / 32;
    const int bi = blockIdx.x / tiles, bj = blockIdx.x % tiles;
    const int64_t bt = blockIdx.y;
    const z64* a = A + bt * sa;
    const z64* b = B + bt * sb;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    z64 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int k0 = 0; k0 < n; k0 += 32)
    {
        for (int r = ty; r < 32; r += 8)
        {
            const int i = bi * 32 + r, k = k0 + tx;
            z64 v = (i < n && k < n) ? a[(int64_t)i * n + k] : z64{0, 0};
            if (CONJA) v.y = -v.y;
            ta[r][tx] = v;
            const int kk = k0 + r, j = bj * 32 + tx;
            z64 w = {0, 0};
            if (kk < n && j < n) w = TRANSB ? b[(int64_t)j * n + kk] : b[(int64_t)kk * n + j];
            if (CONJB) w.y = -w.y;
            tb[r][tx] = w;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < 32; ++k)
        {
            const z64 bv = tb[k][tx];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = cadd(acc[q], cmul(ta[ty + 8 * q][k], bv));
        }
        __syncthreads();
    }
    z64* c = Cm + bt * sc;
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const int i = bi * 32 + ty + 8 * q, j = bj * 32 + tx;
        if (i < n && j < n) c[(int64_t)i * n + j] = acc[q];
    }
}
