// split.hpp -- local e-ph contraction with split-product MFMA (bf16).
//
// Same contraction as contract.hpp (real-space form of elphLocal,
// src/dvloc/compute_Vbare_loc.c:286-342 via Parseval), with FP32 operands
// split into NP bf16 pieces x = x0 + x1 (+ x2), each the round-to-nearest of
// the remaining residual, and each real product x*y replaced by the NPROD
// piece products with pa + pb < NP (NP = 2: 3 products; NP = 3: 6 products).
// Products run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; every
// SP_FLUSH chunks the FP32 accumulators are added into FP64 registers, so no
// FP32 summation chain spans more than SP_FLUSH * 32 grid points.
//
// Data flow per (k, k+q) pair:
//   k_presplit_B : psi_k+q (FP32 box) -> bf16 pieces in an MFMA-ready,
//                  chunk-major layout Bp[spin][spinor][chunk][piece][re/im]
//                  [m (padded to 128)][32 points]; written once per pair.
//   k_contract_split : one block = one mode nu x 128 k bands (rows) x 64 k+q
//                  bands (columns) x one range of grid points (split-K).
//                  8 waves, each 16 rows x 64 columns.  Per chunk of 32 points
//                  the block stages A = dV_nu psi_k (FP32, formed while
//                  staging) and the B pieces into LDS, double buffered, with
//                  the next chunk prefetched into registers during the MFMAs;
//                  each wave splits its A fragment once per chunk and reuses
//                  it across its 4 column blocks.
//   Blocks are ordered so that all tiles of one K range run on one XCD
//   (block b runs on XCD b % 8): the psi tiles of that range are fetched
//   into that XCD's L2 once.
//
// 16x16x32 fragment layout (wave64): A: lane l holds A[l%16][8(l/16)+0..7];
// B: lane l holds B[8(l/16)+0..7][l%16]; D: lane l, v holds D[4(l/16)+v][l%16].
#pragma once

#include "common.hpp"
#include "contract.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4v __attribute__((ext_vector_type(4)));

#define SP_KC 32      // grid points per chunk
#define SP_TR 128     // rows (k bands) of a block tile
#define SP_TC 64      // columns (k+q bands) of a block tile
#define SP_PAD 128    // band padding of the presplit layout
#define SP_AL 34      // LDS row stride of A (complex), 16 B pad
#define SP_FLUSH 4    // chunks per FP32 accumulation chain
#define SP_THREADS 512

__device__ inline bf16x8 neg_bf16x8(bf16x8 x)
{
    typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
    u16x8 u = __builtin_bit_cast(u16x8, x);
    u ^= (unsigned short)0x8000;
    return __builtin_bit_cast(bf16x8, u);
}

template <int NP>
__device__ inline void split_bf16(float x, __bf16* p)
{
    float r = x;
#pragma unroll
    for (int i = 0; i < NP; ++i)
    {
        p[i] = (__bf16)r;
        r = r - (float)p[i];
    }
}

// Bp layout size in bf16 elements
static inline int64_t presplit_elems(int nspin, int nspinor, int64_t nchunks, int NP, int nbp)
{
    return (int64_t)nspin * nspinor * nchunks * NP * 2 * nbp * SP_KC;
}

// psi_k+q box (nspin, nb, nspinor, N) FP32 -> Bp pieces.  Grid: x over the
// padded points (nchunks * 32), y over (spin, spinor, row m < nbp); rows
// m >= nb and points >= N are zero.
template <int NP>
__global__ void k_presplit_B(const c32* __restrict__ box, int nb, int nbp, int nspinor,
                             int64_t N, int64_t nchunks, __bf16* __restrict__ Bp)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nchunks * SP_KC) return;
    const int64_t c = r >> 5;
    const int kk = (int)(r & 31);
    const int y = blockIdx.y;
    const int m = y % nbp;
    const int sp = (y / nbp) % nspinor;
    const int is = y / (nbp * nspinor);
    c32 v = {0, 0};
    if (m < nb && r < N) v = box[((int64_t)(is * nb + m) * nspinor + sp) * N + r];
    __bf16 pr[NP], pi[NP];
    split_bf16<NP>(v.x, pr);
    split_bf16<NP>(v.y, pi);
    const int64_t base = (((int64_t)(is * nspinor + sp) * nchunks + c) * NP) * 2;
#pragma unroll
    for (int p = 0; p < NP; ++p)
    {
        Bp[((base + 2 * p + 0) * nbp + m) * SP_KC + kk] = pr[p];
        Bp[((base + 2 * p + 1) * nbp + m) * SP_KC + kk] = pi[p];
    }
}

struct SplitArgs
{
    int nm, nb, nbp, nspin, nspinor, nmag;
    int ntn, ntm; // row / column tiles per mode
    int nsplit;   // K splits; nspin * nsplit is a multiple of 8
    int64_t N, nchunks, chunks_per_split;
};

template <int NP>
struct SplitStage
{
    static constexpr int B16 = NP * 2 * SP_TC * SP_KC * 2 / 16 / SP_THREADS; // uint4 per thread
    uint4 psi[4]; // 8 complex: psi_k (or, for 2x2 spinor potentials, dV psi_k)
    c32 dvv;      // dV_nu at point tid of the chunk (tid < SP_KC), else unused
    uint4 b[B16 > 0 ? B16 : 1];
};

// Block stage of chunk c: psi_k rows [n0, n0+128) x 32 points (thread ->
// row tid/4, 8 points), the 32 dV values, and the B pieces.
// FAST (one spinor, diagonal potential, N a multiple of 32): 16-byte loads,
// no bounds tests; rows n >= nb read row nb-1 (their results are never
// stored) and the product dV psi is formed when the wave builds its A
// fragment.  Otherwise the product is staged element by element and dV = 1.
template <int NP, bool FAST>
__device__ inline void stage_load(SplitStage<NP>& st, const c32* __restrict__ psik,
                                  const __bf16* __restrict__ Bp, const c32* __restrict__ dv,
                                  const SplitArgs& a, int v, int is, int sp, int n0, int m0,
                                  int64_t c, int tid)
{
    const int arow = tid >> 2, aoff = (tid & 3) * 8;
    const int n = n0 + arow;
    const int64_t r0 = c * SP_KC + aoff;
    const int64_t N = a.N;
    if constexpr (FAST)
    {
        const int nn = n < a.nb ? n : a.nb - 1;
        const uint4* pk = reinterpret_cast<const uint4*>(psik + (int64_t)(is * a.nb + nn) * N + r0);
#pragma unroll
        for (int k = 0; k < 4; ++k) st.psi[k] = pk[k];
        if (tid < SP_KC)
        {
            const c32* V = dv + ((int64_t)v * a.nmag + (a.nmag == 2 ? is : 0)) * N;
            st.dvv = V[c * SP_KC + tid];
        }
    }
    else
    {
        c32 vals[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            const int64_t r = r0 + k;
            c32 val = {0, 0};
            if (n < a.nb && r < N)
                val = load_dvpsi<float>(psik, dv, v, is, n, sp, r, a.nb, a.nspinor, a.nmag, N);
            vals[k] = val;
        }
        __builtin_memcpy(st.psi, vals, sizeof(vals));
        st.dvv = {1.f, 0.f};
    }
    const uint4* src = reinterpret_cast<const uint4*>(
        Bp + ((((int64_t)(is * a.nspinor + sp) * a.nchunks + c) * NP * 2) * a.nbp + m0) * SP_KC);
    constexpr int PLANE16 = SP_TC * SP_KC * 2 / 16; // uint4 per plane of the tile
#pragma unroll
    for (int q = 0; q < SplitStage<NP>::B16; ++q)
    {
        const int idx = tid + q * SP_THREADS;
        const int plane = idx / PLANE16, within = idx % PLANE16;
        st.b[q] = src[(int64_t)plane * (a.nbp * SP_KC * 2 / 16) + within];
    }
}

template <int NP>
__device__ inline void stage_store(const SplitStage<NP>& st, c32* sA, __bf16* sB, c32* sD,
                                   int tid)
{
    const int arow = tid >> 2, aoff = (tid & 3) * 8;
    uint4* d = reinterpret_cast<uint4*>(sA + arow * SP_AL + aoff);
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = st.psi[k];
    if (tid < SP_KC) sD[tid] = st.dvv;
    uint4* db = reinterpret_cast<uint4*>(sB);
#pragma unroll
    for (int q = 0; q < SplitStage<NP>::B16; ++q) db[tid + q * SP_THREADS] = st.b[q];
}

typedef __attribute__((address_space(3))) char lds_char;

// FAST staging of chunk c straight from global memory into LDS
// (global_load_lds_dwordx4: each wave instruction fills 1 KB of LDS, lane l
// its bytes [16 l, 16 l + 16)).  A: the padded tile, 128 rows x 272 bytes,
// as 34 slices; pad slots reload the row's first 16 bytes.  B: NP*2 planes of
// 64 rows x 64 bytes.  D: the 32 dV values (wave 0, lanes 0..15).  The
// caller waits (s_waitcnt) and synchronises before reading the buffer.
template <int NP>
__device__ inline void stage_async(const c32* __restrict__ psik, const __bf16* __restrict__ Bp,
                                   const c32* __restrict__ dv, const SplitArgs& a, int v, int is,
                                   int n0, int m0, int64_t c, lds_char* LA, lds_char* LB,
                                   lds_char* LD, int wave, int lane)
{
    const int64_t N = a.N;
    constexpr int RB = SP_AL * 8;             // bytes per padded LDS row of A
    constexpr int ASL = SP_TR * RB / 1024;    // 1 KB slices of the A tile
    static_assert(SP_TR * RB % 1024 == 0, "A tile must be whole 1 KB slices");
#pragma unroll
    for (int i = wave; i < ASL; i += 8)
    {
        const int o = i * 1024 + lane * 16;
        const int row = o / RB;
        int col = o - row * RB;
        if (col >= SP_KC * 8) col = 0;
        const int n = n0 + row;
        const int nn = n < a.nb ? n : a.nb - 1;
        const char* src =
            reinterpret_cast<const char*>(psik + (int64_t)(is * a.nb + nn) * N + c * SP_KC) + col;
        __builtin_amdgcn_global_load_lds(src, LA + i * 1024, 16, 0, 0);
    }
    const __bf16* bsrc =
        Bp + (((int64_t)is * a.nchunks + c) * NP * 2 * a.nbp + m0) * SP_KC;
#pragma unroll
    for (int i = wave; i < NP * 2 * 4; i += 8)
    {
        const int plane = i >> 2, part = i & 3;
        const char* src = reinterpret_cast<const char*>(bsrc + (int64_t)plane * a.nbp * SP_KC) +
                          part * 1024 + lane * 16;
        __builtin_amdgcn_global_load_lds(src, LB + i * 1024, 16, 0, 0);
    }
    if (wave == 0 && lane < SP_KC / 2)
    {
        const c32* V = dv + ((int64_t)v * a.nmag + (a.nmag == 2 ? is : 0)) * N + c * SP_KC;
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const char*>(V) + lane * 16, LD, 16, 0,
                                         0);
    }
}

template <int NP, int NPROD, bool FAST>
__global__ __launch_bounds__(SP_THREADS, 2) void k_contract_split(
    const c32* __restrict__ psik, const __bf16* __restrict__ Bp, const c32* __restrict__ dv,
    z64* __restrict__ part, SplitArgs a)
{
    constexpr int BELEMS = NP * 2 * SP_TC * SP_KC; // bf16 per B stage
    extern __shared__ __attribute__((aligned(16))) char sp_smem[];
    c32* sA0 = reinterpret_cast<c32*>(sp_smem);
    c32* sA1 = sA0 + SP_TR * SP_AL;
    __bf16* sB0 = reinterpret_cast<__bf16*>(sA1 + SP_TR * SP_AL);
    __bf16* sB1 = sB0 + BELEMS;
    c32* sD0 = reinterpret_cast<c32*>(sB1 + BELEMS);
    c32* sD1 = sD0 + SP_KC;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // XCD-aware decode: group g = (spin, split) runs on XCD g % 8
    const int per_group = a.nm * a.ntn * a.ntm;
    const int b = blockIdx.x;
    const int x = b & 7, j = b >> 3;
    const int g = x + 8 * (j / per_group);
    const int tile = j % per_group;
    const int v = tile / (a.ntn * a.ntm);
    const int tn = (tile / a.ntm) % a.ntn, tm = tile % a.ntm;
    const int is = g / a.nsplit, split = g % a.nsplit;
    if (is >= a.nspin) return;
    const int n0 = tn * SP_TR, m0 = tm * SP_TC;
    const int64_t c0 = (int64_t)split * a.chunks_per_split;
    const int64_t c1 = (c0 + a.chunks_per_split < a.nchunks) ? c0 + a.chunks_per_split : a.nchunks;

    flt4v accr[4], acci[4];
    double dr[4][4], di[4][4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            accr[cb][e] = 0.f;
            acci[cb][e] = 0.f;
            dr[cb][e] = 0.0;
            di[cb][e] = 0.0;
        }

    SplitStage<NP> st;
    lds_char* L = (lds_char*)sp_smem;
    constexpr int A_BYTES = SP_TR * SP_AL * 8;
    constexpr int B_BYTES = BELEMS * 2;
    int nfl = 0;
    for (int sp = 0; sp < a.nspinor; ++sp)
    {
        if (c0 >= c1) break;
        if constexpr (FAST)
        {
            stage_async<NP>(psik, Bp, dv, a, v, is, n0, m0, c0, L, L + 2 * A_BYTES,
                            L + 2 * A_BYTES + 2 * B_BYTES, wave, lane);
            __builtin_amdgcn_s_waitcnt(0);
        }
        else
        {
            stage_load<NP, FAST>(st, psik, Bp, dv, a, v, is, sp, n0, m0, c0, tid);
            stage_store<NP>(st, sA0, sB0, sD0, tid);
        }
        __syncthreads();
        int cur = 0;
        for (int64_t c = c0; c < c1; ++c)
        {
            const bool more = (c + 1 < c1);
            if (more)
            {
                if constexpr (FAST)
                    stage_async<NP>(psik, Bp, dv, a, v, is, n0, m0, c + 1,
                                    L + (cur ? 0 : A_BYTES), L + 2 * A_BYTES + (cur ? 0 : B_BYTES),
                                    L + 2 * A_BYTES + 2 * B_BYTES + (cur ? 0 : SP_KC * 8), wave,
                                    lane);
                else
                    stage_load<NP, FAST>(st, psik, Bp, dv, a, v, is, sp, n0, m0, c + 1, tid);
            }
            const c32* sA = cur ? sA1 : sA0;
            const __bf16* sb = cur ? sB1 : sB0;
            const c32* sD = cur ? sD1 : sD0;

            // A fragment: rows wave*16 + lane%16, points 8(lane/16)..+7,
            // A = dV_nu psi_k formed here in FP32, then split
            bf16x8 are[NP], aim[NP];
            {
                const c32* s = sA + (wave * 16 + (lane & 15)) * SP_AL + 8 * (lane >> 4);
                const c32* d = sD + 8 * (lane >> 4);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                {
                    const c32 val = cmul(d[k], s[k]);
                    __bf16 pr[NP], pi[NP];
                    split_bf16<NP>(val.x, pr);
                    split_bf16<NP>(val.y, pi);
#pragma unroll
                    for (int p = 0; p < NP; ++p)
                    {
                        are[p][k] = pr[p];
                        aim[p][k] = pi[p];
                    }
                }
            }
            const int ko = 8 * (lane >> 4);
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
            {
                bf16x8 br[NP], bi[NP];
                const int ro = cb * 16 + (lane & 15);
#pragma unroll
                for (int p = 0; p < NP; ++p)
                {
                    br[p] = *reinterpret_cast<const bf16x8*>(sb + ((2 * p) * SP_TC + ro) * SP_KC + ko);
                    bi[p] = *reinterpret_cast<const bf16x8*>(sb + ((2 * p + 1) * SP_TC + ro) * SP_KC + ko);
                }
#pragma unroll
                for (int p1 = 0; p1 < NP; ++p1)
#pragma unroll
                    for (int p2 = 0; p2 < NP; ++p2)
                    {
                        if (p1 + p2 >= NP) continue;
                        if (NPROD == 3 && p1 + p2 > 1) continue;
                        // C += a conj(b): re = ar br + ai bi, im = ai br - ar bi
                        accr[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(are[p1], br[p2], accr[cb], 0, 0, 0);
                        acci[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aim[p1], br[p2], acci[cb], 0, 0, 0);
                        accr[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aim[p1], bi[p2], accr[cb], 0, 0, 0);
                        acci[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(neg_bf16x8(are[p1]), bi[p2], acci[cb], 0, 0, 0);
                    }
            }
            if (++nfl == SP_FLUSH)
            {
                nfl = 0;
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                    {
                        dr[cb][e] += (double)accr[cb][e];
                        di[cb][e] += (double)acci[cb][e];
                        accr[cb][e] = 0.f;
                        acci[cb][e] = 0.f;
                    }
            }
            if (more)
            {
                // the other buffer was last read before the previous barrier
                if constexpr (FAST)
                    __builtin_amdgcn_s_waitcnt(0);
                else
                    stage_store<NP>(st, cur ? sA0 : sA1, cur ? sB0 : sB1, cur ? sD0 : sD1, tid);
                __syncthreads();
                cur ^= 1;
            }
        }
        __syncthreads();
    }
    const int row = n0 + wave * 16;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int rr = row + 4 * (lane >> 4) + e;
            const int cc = m0 + cb * 16 + (lane & 15);
            if (rr < a.nb && cc < a.nb)
            {
                const int64_t o = ((((int64_t)split * a.nm + v) * a.nspin + is) * a.nb + rr) * a.nb + cc;
                part[o] = {dr[cb][e] + (double)accr[cb][e], di[cb][e] + (double)acci[cb][e]};
            }
        }
}

template <int NP>
static inline size_t split_lds_bytes()
{
    return 2 * (sizeof(c32) * (SP_TR * SP_AL + SP_KC) + sizeof(__bf16) * (size_t)NP * 2 * SP_TC * SP_KC);
}
