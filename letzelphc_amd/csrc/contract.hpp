// contract.hpp -- local e-ph contraction over the real-space grid on MFMA.
//
// The reference computes, per mode nu and band n, the forward FFT of
// dV_nu(r) psi_n,k(r), keeps the k+q sphere times 1/N and contracts it with
// psi_m,k+q in G space (src/fft/convolution.c:13-189,
// src/dvloc/compute_Vbare_loc.c:286-342).  By Parseval on the same box this
// equals
//     g[nu][n][m] = (1/N) sum_{sigma,r} conj(psi_kq,m,sigma(r))
//                                    (dV_nu psi_k,n)_sigma(r)
// where psi_kq(r) is the backward transform of the (shifted, rotated) k+q
// sphere -- exactly, since both sides use the same box.  That is one complex
// GEMM per (nu, spin): rows (nu, n), columns m, reduction over (sigma, r).
// The A operand dV.psi_k is formed while staging tiles into LDS, so it never
// touches HBM.  The reduction is split over K; partials are reduced in a
// fixed order in FP64 (no float atomics: deterministic).
//
// MFMA lane layouts used (16x16x4, wave64): A: lane l holds A[l%16][l/16];
// B: lane l holds B[l/16][l%16]; D: lane l, register v holds D[row][l%16]
// with row = Mfma<T>::row(l, v) (FP64: l/16 + 4v, measured on gfx950;
// FP32: 4(l/16) + v).
#pragma once

#include "common.hpp"

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double>
{
    typedef dbl4 acc;
    static __device__ inline acc op(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float>
{
    typedef flt4 acc;
    static __device__ inline acc op(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

#define CT_KC 16
#define CT_LD (CT_KC + 1)
#define CT_VB 2  // modes per block (wide tile; its tile count sets the K splits)
#define CT_NB 32 // k bands (rows) per block
#define CT_MB 64 // k+q bands (columns) per block
#define CT_FLUSH 4

// Block tile shapes: VB modes x NB rows x MB columns over 4 waves, wave w
// taking mode w / NCG and column group w % NCG, each wave RB x CB blocks of
// 16 x 16.  Every shape issues, for each g element, the same MFMAs on the
// same operands in the same order (per 4-point step the first products of
// all the wave's blocks, then the second ones), so with the same K splits
// all shapes give bit-identical partials; they differ only in how much of a
// tile is padding when nb is not a multiple of the tile.
template <int VB_, int RB_, int CB_>
struct CtShape
{
    static constexpr int VB = VB_, RB = RB_, CB = CB_;
    static constexpr int NCG = 4 / VB_;
    static constexpr int NB = 16 * RB_;
    static constexpr int MB = 16 * CB_ * NCG;
    static_assert(VB_ == 1 || VB_ == 2 || VB_ == 4, "4 waves per block");
};
typedef CtShape<2, 2, 2> CtWide; // 2 modes x 32 x 64
typedef CtShape<4, 2, 2> Ct32;   // 4 modes x 32 x 32
typedef CtShape<4, 1, 3> Ct16x48;   // 4 modes x 16 x 48
typedef CtShape<4, 1, 1> Ct16;   // 4 modes x 16 x 16

// Loads A[n][r] = (dV_nu psi_k,n)_sigma(r) for the block's modes.
template <typename T>
__device__ inline cpx<T> load_dvpsi(const cpx<T>* __restrict__ psik,
                                    const cpx<T>* __restrict__ dv, int v, int is, int n,
                                    int sp, int64_t r, int nb, int nspinor, int nmag,
                                    int64_t N)
{
    const cpx<T>* pk = psik + ((int64_t)(is * nb + n) * nspinor) * N;
    if (nmag == 4 && nspinor == 2)
    {
        // 2x2 spinor potential (src/fft/convolution.c:322-352)
        const cpx<T>* V = dv + (int64_t)v * 4 * N;
        return cadd(cmul(V[(int64_t)(2 * sp) * N + r], pk[r]),
                    cmul(V[(int64_t)(2 * sp + 1) * N + r], pk[N + r]));
    }
    const cpx<T>* V = dv + ((int64_t)v * nmag + (nmag == 2 ? is : 0)) * N;
    return cmul(V[r], pk[(int64_t)sp * N + r]);
}

// One tile of the contraction: tile index bx (mode group, row tile, column
// tile), K split `split`, spin `is`.  The kernel k_contract runs one tile
// per block; the work-queue kernel k_contract_queue runs tiles of many
// (k, k+q) pairs and cells from a list.
template <typename T, typename S>
__device__ inline void contract_tile(const cpx<T>* __restrict__ psik,
                                     const cpx<T>* __restrict__ psikq,
                                     const cpx<T>* __restrict__ dv, z64* __restrict__ part,
                                     int nm, int nb, int nspin, int nspinor, int nmag, int64_t N,
                                     int64_t chunks_per_split, int bx, int split, int is)
{
    typedef typename Mfma<T>::acc acc_t;
    constexpr int VB = S::VB, NB = S::NB, MB = S::MB, RB = S::RB, CB = S::CB;
    __shared__ cpx<T> sA[VB][NB][CT_LD];
    __shared__ cpx<T> sB[MB][CT_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntn = (nb + NB - 1) / NB, ntm = (nb + MB - 1) / MB;
    const int tm = bx % ntm;
    bx /= ntm;
    const int tn = bx % ntn;
    bx /= ntn;
    const int v0 = bx * VB, n0 = tn * NB, m0 = tm * MB;
    const int64_t nchunks = (N + CT_KC - 1) / CT_KC;
    const int64_t c0 = (int64_t)split * chunks_per_split;
    const int64_t c1 = (c0 + chunks_per_split < nchunks) ? c0 + chunks_per_split : nchunks;
    const int wv = wave / S::NCG, wm = wave % S::NCG;

    // FP32: the MFMA accumulators are added into FP64 every CT_FLUSH chunks,
    // so no FP32 summation chain is longer than CT_FLUSH * CT_KC points.
    constexpr bool flush = sizeof(T) == 4;
    acc_t accr[RB][CB], acci[RB][CB];
    double dr[RB][CB][4], di[RB][CB][4];
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
        for (int b = 0; b < CB; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                accr[a][b][v] = 0;
                acci[a][b][v] = 0;
                dr[a][b][v] = 0;
                di[a][b][v] = 0;
            }

    int nfl = 0;
    for (int sp = 0; sp < nspinor; ++sp)
    {
        for (int64_t c = c0; c < c1; ++c)
        {
            if (flush && ++nfl > CT_FLUSH)
            {
                nfl = 1;
#pragma unroll
                for (int a = 0; a < RB; ++a)
#pragma unroll
                    for (int b = 0; b < CB; ++b)
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                        {
                            dr[a][b][e] += (double)accr[a][b][e];
                            di[a][b][e] += (double)acci[a][b][e];
                            accr[a][b][e] = 0;
                            acci[a][b][e] = 0;
                        }
            }
            const int64_t r0 = c * CT_KC;
            for (int f = tid; f < VB * NB * CT_KC; f += 256)
            {
                const int vv = f / (NB * CT_KC);
                const int rem = f - vv * (NB * CT_KC);
                const int nn = rem / CT_KC, rr = rem - nn * CT_KC;
                const int v = v0 + vv, n = n0 + nn;
                const int64_t r = r0 + rr;
                // padding reads a valid element and stores zero (no branch
                // around the load)
                const bool in = v < nm && n < nb && r < N;
                cpx<T> a = load_dvpsi<T>(psik, dv, in ? v : 0, is, in ? n : 0, sp, in ? r : 0, nb,
                                         nspinor, nmag, N);
                if (!in) a = {0, 0};
                sA[vv][nn][rr] = a;
            }
            for (int f = tid; f < MB * CT_KC; f += 256)
            {
                const int mm = f / CT_KC, rr = f - mm * CT_KC;
                const int m = m0 + mm;
                const int64_t r = r0 + rr;
                const bool in = m < nb && r < N;
                cpx<T> b = psikq[((int64_t)(is * nb + (in ? m : 0)) * nspinor + sp) * N + (in ? r : 0)];
                if (!in) b = {0, 0};
                sB[mm][rr] = b;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < CT_KC / 4; ++kk)
            {
                const int i = lane & 15, kq = kk * 4 + (lane >> 4);
                cpx<T> a[RB], b[CB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) a[rb] = sA[wv][rb * 16 + i][kq];
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) b[cb] = sB[(wm * CB + cb) * 16 + i][kq];
                // C += a conj(b): re = ar br + ai bi, im = ai br - ar bi; the
                // first products of all 8 accumulators go out before the
                // second ones, so no MFMA waits on the one just issued
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
                    {
                        accr[rb][cb] = Mfma<T>::op(a[rb].x, b[cb].x, accr[rb][cb]);
                        acci[rb][cb] = Mfma<T>::op(a[rb].y, b[cb].x, acci[rb][cb]);
                    }
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
                    {
                        accr[rb][cb] = Mfma<T>::op(a[rb].y, b[cb].y, accr[rb][cb]);
                        acci[rb][cb] = Mfma<T>::op(-a[rb].x, b[cb].y, acci[rb][cb]);
                    }
            }
            __syncthreads();
        }
    }
    const int v = v0 + wv;
    if (v >= nm) return;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                const int row = n0 + rb * 16 + Mfma<T>::row(lane, e);
                const int col = m0 + (wm * CB + cb) * 16 + (lane & 15);
                if (row < nb && col < nb)
                {
                    const int64_t o = ((((int64_t)split * nm + v) * nspin + is) * nb + row) * nb + col;
                    part[o] = {dr[rb][cb][e] + (double)accr[rb][cb][e],
                               di[rb][cb][e] + (double)acci[rb][cb][e]};
                }
            }
}

// k_contract with dV and psi_k staged separately and their product formed
// as the A fragments are read (A = dV psi_k in the same order as
// load_dvpsi): the block's 2 modes share the psi_k rows and its 32 rows the
// dV values, so a chunk moves 0.5 + 8 KB of A-side data instead of 32 KB
// of element-wise loads.  Not for the 2x2 potential of spinors (nmag 4,
// nspinor 2).  Used for FP32; in FP64 the products formed beside the FP64
// MFMAs cost more than the loads saved.
template <typename T, typename S>
__device__ inline void contract_sep_tile(const cpx<T>* __restrict__ psik,
                                         const cpx<T>* __restrict__ psikq,
                                         const cpx<T>* __restrict__ dv, z64* __restrict__ part,
                                         int nm, int nb, int nspin, int nspinor, int nmag,
                                         int64_t N, int64_t chunks_per_split, int bx, int split,
                                         int is)
{
    typedef typename Mfma<T>::acc acc_t;
    constexpr int VB = S::VB, NB = S::NB, MB = S::MB, RB = S::RB, CB = S::CB;
    __shared__ cpx<T> sV[VB][CT_LD]; // dV of the block's modes
    __shared__ cpx<T> sP[NB][CT_LD]; // psi_k of the block's rows
    __shared__ cpx<T> sB[MB][CT_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntn = (nb + NB - 1) / NB, ntm = (nb + MB - 1) / MB;
    const int tm = bx % ntm;
    bx /= ntm;
    const int tn = bx % ntn;
    bx /= ntn;
    const int v0 = bx * VB, n0 = tn * NB, m0 = tm * MB;
    const int64_t nchunks = (N + CT_KC - 1) / CT_KC;
    const int64_t c0 = (int64_t)split * chunks_per_split;
    const int64_t c1 = (c0 + chunks_per_split < nchunks) ? c0 + chunks_per_split : nchunks;
    const int wv = wave / S::NCG, wm = wave % S::NCG;

    // FP32: the MFMA accumulators are added into FP64 every CT_FLUSH chunks,
    // so no FP32 summation chain is longer than CT_FLUSH * CT_KC points.
    constexpr bool flush = sizeof(T) == 4;
    acc_t accr[RB][CB], acci[RB][CB];
    double dr[RB][CB][4], di[RB][CB][4];
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
        for (int b = 0; b < CB; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                accr[a][b][v] = 0;
                acci[a][b][v] = 0;
                dr[a][b][v] = 0;
                di[a][b][v] = 0;
            }

    int nfl = 0;
    for (int sp = 0; sp < nspinor; ++sp)
    {
        for (int64_t c = c0; c < c1; ++c)
        {
            if (flush && ++nfl > CT_FLUSH)
            {
                nfl = 1;
#pragma unroll
                for (int a = 0; a < RB; ++a)
#pragma unroll
                    for (int b = 0; b < CB; ++b)
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                        {
                            dr[a][b][e] += (double)accr[a][b][e];
                            di[a][b][e] += (double)acci[a][b][e];
                            accr[a][b][e] = 0;
                            acci[a][b][e] = 0;
                        }
            }
            const int64_t r0 = c * CT_KC;
            if (tid < VB * CT_KC)
            {
                const int vv = tid / CT_KC, rr = tid - vv * CT_KC;
                const int v = v0 + vv;
                const int64_t r = r0 + rr;
                const bool in = v < nm && r < N;
                cpx<T> x = dv[((int64_t)(in ? v : 0) * nmag + (nmag == 2 ? is : 0)) * N + (in ? r : 0)];
                if (!in) x = {0, 0};
                sV[vv][rr] = x;
            }
            for (int f = tid; f < NB * CT_KC; f += 256)
            {
                const int nn = f / CT_KC, rr = f - nn * CT_KC;
                const int n = n0 + nn;
                const int64_t r = r0 + rr;
                const bool in = n < nb && r < N;
                cpx<T> x = psik[((int64_t)(is * nb + (in ? n : 0)) * nspinor + sp) * N + (in ? r : 0)];
                if (!in) x = {0, 0};
                sP[nn][rr] = x;
            }
            for (int f = tid; f < MB * CT_KC; f += 256)
            {
                const int mm = f / CT_KC, rr = f - mm * CT_KC;
                const int m = m0 + mm;
                const int64_t r = r0 + rr;
                const bool in = m < nb && r < N;
                cpx<T> b = psikq[((int64_t)(is * nb + (in ? m : 0)) * nspinor + sp) * N + (in ? r : 0)];
                if (!in) b = {0, 0};
                sB[mm][rr] = b;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < CT_KC / 4; ++kk)
            {
                const int i = lane & 15, kq = kk * 4 + (lane >> 4);
                cpx<T> a[RB], b[CB];
                const cpx<T> vq = sV[wv][kq];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) a[rb] = cmul(vq, sP[rb * 16 + i][kq]);
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) b[cb] = sB[(wm * CB + cb) * 16 + i][kq];
                // C += a conj(b): re = ar br + ai bi, im = ai br - ar bi; the
                // first products of all 8 accumulators go out before the
                // second ones, so no MFMA waits on the one just issued
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
                    {
                        accr[rb][cb] = Mfma<T>::op(a[rb].x, b[cb].x, accr[rb][cb]);
                        acci[rb][cb] = Mfma<T>::op(a[rb].y, b[cb].x, acci[rb][cb]);
                    }
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
                    {
                        accr[rb][cb] = Mfma<T>::op(a[rb].y, b[cb].y, accr[rb][cb]);
                        acci[rb][cb] = Mfma<T>::op(-a[rb].x, b[cb].y, acci[rb][cb]);
                    }
            }
            __syncthreads();
        }
    }
    const int v = v0 + wv;
    if (v >= nm) return;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                const int row = n0 + rb * 16 + Mfma<T>::row(lane, e);
                const int col = m0 + (wm * CB + cb) * 16 + (lane & 15);
                if (row < nb && col < nb)
                {
                    const int64_t o = ((((int64_t)split * nm + v) * nspin + is) * nb + row) * nb + col;
                    part[o] = {dr[rb][cb][e] + (double)accr[rb][cb][e],
                               di[rb][cb][e] + (double)acci[rb][cb][e]};
                }
            }
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_contract(const cpx<T>* __restrict__ psik,
                                                  const cpx<T>* __restrict__ psikq,
                                                  const cpx<T>* __restrict__ dv,
                                                  z64* __restrict__ part, int nm, int nb,
                                                  int nspin, int nspinor, int nmag,
                                                  int64_t N, int64_t chunks_per_split)
{
    contract_tile<T, S>(psik, psikq, dv, part, nm, nb, nspin, nspinor, nmag, N, chunks_per_split,
                        blockIdx.x, blockIdx.y, blockIdx.z);
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_contract_sep(const cpx<T>* __restrict__ psik,
                                                      const cpx<T>* __restrict__ psikq,
                                                      const cpx<T>* __restrict__ dv,
                                                      z64* __restrict__ part, int nm, int nb,
                                                      int nspin, int nspinor, int nmag,
                                                      int64_t N, int64_t chunks_per_split)
{
    contract_sep_tile<T, S>(psik, psikq, dv, part, nm, nb, nspin, nspinor, nmag, N,
                            chunks_per_split, blockIdx.x, blockIdx.y, blockIdx.z);
}

// A (k, k+q) pair of a work-queue launch: its boxes, potential, partials
// and shape (the arguments of k_contract / k_contract_sep).
template <typename T>
struct QueuePair
{
    const cpx<T>* psik;
    const cpx<T>* psikq;
    const cpx<T>* dv;
    z64* part;
    int nm, nb, nspin, nspinor, nmag, sep;
    int64_t N, chunks_per_split;
};

// One work item: a tile of one pair (pair index, tile, K split, spin).
struct QueueItem
{
    int pair, tile, split, spin;
};

// Persistent work-queue contraction over the tiles of many pairs (of one
// precision, possibly of different cells, grids, band and mode counts):
// each block takes the next item from an atomic counter (zeroed by the
// caller with a kernel before the launch) until the list is exhausted, so
// every block reaches the exit.  The items are listed largest first.  A
// tile computes exactly what k_contract / k_contract_sep computes for it,
// so the partials, and g after the fixed-order reduction, are bit-identical
// to one launch per pair.  SEP: the launch may contain separate-staging
// pairs (FP32 without a 2x2 spinor potential).  S: the tile shape of every
// pair of the launch.
template <typename T, bool SEP, typename S>
__global__ __launch_bounds__(256) void k_contract_queue(const QueuePair<T>* __restrict__ pairs,
                                                        const QueueItem* __restrict__ items,
                                                        int nitems, int* __restrict__ counter)
{
    __shared__ int s_next;
    for (;;)
    {
        if (threadIdx.x == 0) s_next = atomicAdd(counter, 1);
        __syncthreads();
        const int w = s_next;
        __syncthreads(); // every thread has read s_next before it changes
        if (w >= nitems) break;
        const QueueItem it = items[w];
        const QueuePair<T> p = pairs[it.pair];
        if (SEP && p.sep)
            contract_sep_tile<T, S>(p.psik, p.psikq, p.dv, p.part, p.nm, p.nb, p.nspin,
                                    p.nspinor, p.nmag, p.N, p.chunks_per_split, it.tile,
                                    it.split, it.spin);
        else
            contract_tile<T, S>(p.psik, p.psikq, p.dv, p.part, p.nm, p.nb, p.nspin, p.nspinor,
                                p.nmag, p.N, p.chunks_per_split, it.tile, it.split, it.spin);
    }
}

// zeroes n 16-byte units (a kernel rather than a memset node, so that it
// can sit in a captured graph); grid-stride
__global__ void k_zero_u4(uint4* __restrict__ p, int64_t n)
{
    const uint4 z = {0u, 0u, 0u, 0u};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z;
}

// zeroes n 32-bit words
__global__ void k_zero_words(unsigned* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// g[i] = scale * sum_s part[s][i], fixed order.
__global__ void k_reduce_parts(const z64* __restrict__ part, int nsplit, int64_t count,
                               double scale, z64* __restrict__ g)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double re = 0, im = 0;
    // eight loads in flight, then their adds in split order (the same sum)
    int s = 0;
    for (; s + 8 <= nsplit; s += 8)
    {
        z64 p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = part[(int64_t)(s + u) * count + i];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            re += p[u].x;
            im += p[u].y;
        }
    }
    for (; s < nsplit; ++s)
    {
        const z64 p = part[(int64_t)s * count + i];
        re += p.x;
        im += p.y;
    }
    g[i] = {re * scale, im * scale};
}

// dst[i] = (float) src[i]
__global__ void k_z64_to_c32(const z64* __restrict__ s, c32* __restrict__ d, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = {(float)s[i].x, (float)s[i].y};
}
