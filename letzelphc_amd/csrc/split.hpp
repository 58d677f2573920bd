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
//   k_presplit_B : psi_k+q (FP32 box) -> bf16 pieces, chunk-major
//                  Bp[spin][spinor][chunk][piece][re/im][kq][m][8]: for one
//                  chunk of 32 points, kq = point / 8 and 8 consecutive points
//                  per 16-byte unit; m padded to a multiple of 128.
//   k_contract_split : one block = one mode nu x 128 k bands (rows) x 64 k+q
//                  bands (columns) x one range of grid points (split-K).
//                  8 waves, each 16 rows x 64 columns.  Per chunk of 32 points
//                  the B pieces (64 columns) and the 32 dV values are staged
//                  into LDS, double buffered; each wave forms its A fragment
//                  dV_nu psi_k in FP32, splits it once and reuses it across
//                  its 4 column blocks.
//   Aligned fast path (one spinor, diagonal potential, N a multiple of 32):
//   the B pieces and dV go global -> LDS directly (global_load_lds), and the
//   A fragment (each wave's own 16 rows) goes global -> registers, one chunk
//   ahead; LDS then holds only B and dV.  Otherwise A = dV psi_k is staged
//   element by element through registers into LDS.
//   Blocks are ordered so that all tiles of one K range run on one XCD
//   (block b runs on XCD b % 8): the psi tiles of that range are fetched
//   into that XCD's L2 once.
//
// 16x16x32 fragment layout (wave64): A: lane l holds A[l%16][8(l/16)+0..7];
// B: lane l holds B[8(l/16)+0..7][l%16]; D: lane l, v holds D[4(l/16)+v][l%16].
// The k-group-major B layout makes the 16 lanes of one k group read 256
// consecutive LDS bytes: the B fragment reads are free of bank conflicts.
#pragma once

#include "common.hpp"
#include "contract.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) char lds_char;

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

// psi_k+q box (nspin, nb, nspinor, N) FP32 -> Bp pieces.  Block (chunk c,
// rows 64 y .. 64 y + 63, spin * nspinor + spinor): the chunk's 32 points of
// the 64 rows are read row by row (32 threads per 256-byte row segment) into
// LDS, rows 33 elements apart so the column reads below do not share banks;
// then thread t takes row m = t % 64 and k group kq = t / 64, splits points
// 8 kq .. 8 kq + 7 and writes one 16-byte unit per piece plane, consecutive
// threads on consecutive rows (1 KB contiguous per wave and plane).  Rows
// m >= nb and points >= N are zero.  G0 = (g0x, g0y, g0z) not zero: the box
// holds psi_k+q without its umklapp vector, and every point is multiplied by
// e^{+2 pi i G0.r} (the shift theorem, as k_box_shift) before the split.
#define PS_ROWS 64
template <int NP>
__global__ __launch_bounds__(256) void k_presplit_B(const c32* __restrict__ box, int nb, int nbp,
                                                    int nspinor, int64_t N, int64_t nchunks,
                                                    __bf16* __restrict__ Bp, int Nx, int Ny,
                                                    int Nz, int g0x, int g0y, int g0z)
{
    __shared__ c32 tile[PS_ROWS * (SP_KC + 1)];
    __shared__ c32 eph[SP_KC];
    const int tid = threadIdx.x;
    const int64_t c = blockIdx.x;
    const int m0 = blockIdx.y * PS_ROWS;
    const int ss = blockIdx.z; // spin * nspinor + spinor
    const int is = ss / nspinor, sp = ss % nspinor;
    const int64_t r0 = c * SP_KC;
    const bool shift = (g0x | g0y | g0z) != 0;
    if (shift)
    {
        if (tid < SP_KC)
            eph[tid] = r0 + tid < N ? box_phase<float>(r0 + tid, Nx, Ny, Nz, g0x, g0y, g0z, 1)
                                    : c32{0, 0};
        __syncthreads();
    }
    for (int f = tid; f < PS_ROWS * SP_KC; f += 256)
    {
        const int m = f / SP_KC, j = f % SP_KC;
        c32 v = {0, 0};
        if (m0 + m < nb && r0 + j < N)
            v = box[((int64_t)(is * nb + m0 + m) * nspinor + sp) * N + r0 + j];
        if (shift) v = cmul(v, eph[j]);
        tile[m * (SP_KC + 1) + j] = v;
    }
    __syncthreads();
    const int m = tid % PS_ROWS, kq = tid / PS_ROWS;
    bf16x8 pr[NP], pi[NP];
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        const c32 v = tile[m * (SP_KC + 1) + kq * 8 + j];
        __bf16 a[NP], b[NP];
        split_bf16<NP>(v.x, a);
        split_bf16<NP>(v.y, b);
#pragma unroll
        for (int p = 0; p < NP; ++p)
        {
            pr[p][j] = a[p];
            pi[p][j] = b[p];
        }
    }
    const int64_t plane = (int64_t)nbp * SP_KC; // bf16 per plane
    __bf16* base = Bp + ((int64_t)ss * nchunks + c) * NP * 2 * plane + (int64_t)kq * nbp * 8 +
                   (int64_t)(m0 + m) * 8;
#pragma unroll
    for (int p = 0; p < NP; ++p)
    {
        *reinterpret_cast<bf16x8*>(base + (2 * p) * plane) = pr[p];
        *reinterpret_cast<bf16x8*>(base + (2 * p + 1) * plane) = pi[p];
    }
}

struct SplitArgs
{
    int nm, nb, nbp, nspin, nspinor, nmag;
    int nmt;      // mode tiles (k_contract_split_pc: modes in groups; otherwise nm)
    int ntn, ntm; // row / column tiles per mode tile
    int nsplit;   // K splits; nspin * nsplit is a multiple of 8
    int64_t N, nchunks, chunks_per_split;
};

template <int NP>
struct SplitStage
{
    static constexpr int B16 = NP * 2 * SP_TC * SP_KC * 2 / 16 / SP_THREADS; // uint4 per thread
    uint4 psi[4]; // 8 complex: dV psi_k (generic path)
    c32 dvv;      // 1 (the product is already formed)
    uint4 b[B16 > 0 ? B16 : 1];
};

// bf16 offset of the B tile (piece plane 0, kq 0, column m0) of chunk c
__device__ inline int64_t btile_base(const SplitArgs& a, int NP, int is, int sp, int m0, int64_t c)
{
    return (((int64_t)(is * a.nspinor + sp) * a.nchunks + c) * NP * 2) * a.nbp * SP_KC +
           (int64_t)m0 * 8;
}

// Generic-path block stage of chunk c: A = dV psi_k for rows [n0, n0+128)
// x 32 points (thread -> row tid/4, 8 points), element by element with
// bounds tests, and the B pieces (LDS slot s of a plane = kq * 64 + row).
template <int NP>
__device__ inline void stage_load(SplitStage<NP>& st, const c32* __restrict__ psik,
                                  const __bf16* __restrict__ Bp, const c32* __restrict__ dv,
                                  const SplitArgs& a, int v, int is, int sp, int n0, int m0,
                                  int64_t c, int tid)
{
    const int arow = tid >> 2, aoff = (tid & 3) * 8;
    const int n = n0 + arow;
    const int64_t r0 = c * SP_KC + aoff;
    const int64_t N = a.N;
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
    const __bf16* bt = Bp + btile_base(a, NP, is, sp, m0, c);
#pragma unroll
    for (int q = 0; q < SplitStage<NP>::B16; ++q)
    {
        const int idx = tid + q * SP_THREADS;
        const int plane = idx >> 8, within = idx & 255;
        const int kq = within >> 6, row = within & 63;
        st.b[q] = *reinterpret_cast<const uint4*>(bt + (int64_t)plane * a.nbp * SP_KC +
                                                  (int64_t)kq * a.nbp * 8 + row * 8);
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

// Fast-path staging of chunk c straight from global memory into LDS
// (global_load_lds_dwordx4: each wave instruction fills 1 KB of LDS, lane l
// its bytes [16 l, 16 l + 16)): NP*2 planes x 4 k groups of 64 columns x
// 16 B, and the 32 dV values (wave 0, lanes 0..15).  The caller waits
// (s_waitcnt) and synchronises before reading the buffer.
template <int NP>
__device__ inline void stage_async(const __bf16* __restrict__ Bp, const c32* __restrict__ dv,
                                   const SplitArgs& a, int v, int is, int m0, int64_t c,
                                   lds_char* LB, lds_char* LD, int wave, int lane)
{
    const __bf16* bt = Bp + btile_base(a, NP, is, 0, m0, c);
    static_assert(NP * 2 * 4 % 8 == 0, "B planes must spread evenly over 8 waves");
#pragma unroll
    for (int k = 0; k < NP * 2 * 4 / 8; ++k)
    {
        const int i = wave + 8 * k;
        const int plane = i >> 2, kq = i & 3;
        const __bf16* src = bt + (int64_t)plane * a.nbp * SP_KC + (int64_t)kq * a.nbp * 8 + lane * 8;
        __builtin_amdgcn_global_load_lds(src, LB + i * 1024, 16, 0, 0);
    }
    if (wave == 0 && lane < SP_KC / 2)
    {
        const c32* V = dv + ((int64_t)v * a.nmag + (a.nmag == 2 ? is : 0)) * a.N + c * SP_KC;
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const char*>(V) + lane * 16, LD, 16, 0,
                                         0);
    }
}

template <int NP, bool FAST>
__host__ __device__ constexpr int split_a_bytes()
{
    return FAST ? 0 : SP_TR * SP_AL * 8;
}
template <int NP>
__host__ __device__ constexpr int split_b_bytes()
{
    return NP * 2 * SP_TC * SP_KC * 2;
}

template <int NP, int NPROD, bool FAST>
__global__ __launch_bounds__(SP_THREADS, 2) void k_contract_split(
    const c32* __restrict__ psik, const __bf16* __restrict__ Bp, const c32* __restrict__ dv,
    z64* __restrict__ part, SplitArgs a)
{
    constexpr int A_BYTES = split_a_bytes<NP, FAST>();
    constexpr int B_BYTES = split_b_bytes<NP>();
    extern __shared__ __attribute__((aligned(16))) char sp_smem[];
    lds_char* L = (lds_char*)sp_smem;
    c32* sA0 = reinterpret_cast<c32*>(sp_smem);
    c32* sA1 = reinterpret_cast<c32*>(sp_smem + A_BYTES);
    __bf16* sB0 = reinterpret_cast<__bf16*>(sp_smem + 2 * A_BYTES);
    __bf16* sB1 = reinterpret_cast<__bf16*>(sp_smem + 2 * A_BYTES + B_BYTES);
    c32* sD0 = reinterpret_cast<c32*>(sp_smem + 2 * A_BYTES + 2 * B_BYTES);
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

    // fast path: this lane's A row (rows >= nb read row nb-1; their results
    // are never stored) and 8 points 8(lane/16)..+7 of each chunk
    const int arow = min(n0 + wave * 16 + (lane & 15), a.nb - 1);
    const uint4* arow_p = reinterpret_cast<const uint4*>(
        psik + (int64_t)(is * a.nb + arow) * a.N + 8 * (lane >> 4));
    uint4 areg[4], anext[4];

    SplitStage<NP> st;
    int nfl = 0;
    for (int sp = 0; sp < a.nspinor; ++sp)
    {
        if (c0 >= c1) break;
        if constexpr (FAST)
        {
            stage_async<NP>(Bp, dv, a, v, is, m0, c0, L + 2 * A_BYTES, L + 2 * A_BYTES + 2 * B_BYTES,
                            wave, lane);
#pragma unroll
            for (int k = 0; k < 4; ++k) areg[k] = arow_p[c0 * (SP_KC / 2) + k];
            __builtin_amdgcn_s_waitcnt(0);
        }
        else
        {
            stage_load<NP>(st, psik, Bp, dv, a, v, is, sp, n0, m0, c0, tid);
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
                {
                    stage_async<NP>(Bp, dv, a, v, is, m0, c + 1,
                                    L + 2 * A_BYTES + (cur ? 0 : B_BYTES),
                                    L + 2 * A_BYTES + 2 * B_BYTES + (cur ? 0 : SP_KC * 8), wave,
                                    lane);
#pragma unroll
                    for (int k = 0; k < 4; ++k) anext[k] = arow_p[(c + 1) * (SP_KC / 2) + k];
                }
                else
                    stage_load<NP>(st, psik, Bp, dv, a, v, is, sp, n0, m0, c + 1, tid);
            }
            const __bf16* sb = cur ? sB1 : sB0;
            const c32* sD = cur ? sD1 : sD0;

            // A fragment: rows wave*16 + lane%16, points 8(lane/16)..+7,
            // A = dV_nu psi_k formed here in FP32, then split
            bf16x8 are[NP], aim[NP];
            {
                c32 av[8];
                if constexpr (FAST)
                    __builtin_memcpy(av, areg, sizeof(av));
                else
                {
                    const c32* s = (cur ? sA1 : sA0) + (wave * 16 + (lane & 15)) * SP_AL + 8 * (lane >> 4);
#pragma unroll
                    for (int k = 0; k < 8; ++k) av[k] = s[k];
                }
                const c32* d = sD + 8 * (lane >> 4);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                {
                    const c32 val = cmul(d[k], av[k]);
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
            bf16x8 nre[NP];
#pragma unroll
            for (int p = 0; p < NP; ++p) nre[p] = neg_bf16x8(are[p]);
            // B fragment: LDS slot kq * 64 + column of each piece plane
            const __bf16* bq = sb + ((lane >> 4) * SP_TC + (lane & 15)) * 8;
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
            {
                bf16x8 br[NP], bi[NP];
#pragma unroll
                for (int p = 0; p < NP; ++p)
                {
                    br[p] = *reinterpret_cast<const bf16x8*>(bq + (2 * p) * SP_TC * SP_KC + cb * 128);
                    bi[p] = *reinterpret_cast<const bf16x8*>(bq + (2 * p + 1) * SP_TC * SP_KC + cb * 128);
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
                        acci[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nre[p1], bi[p2], acci[cb], 0, 0, 0);
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
                {
                    __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
                    for (int k = 0; k < 4; ++k) areg[k] = anext[k];
                }
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

template <int NP, bool FAST>
static inline size_t split_lds_bytes()
{
    return 2 * ((size_t)split_a_bytes<NP, FAST>() + split_b_bytes<NP>() + sizeof(c32) * SP_KC);
}
