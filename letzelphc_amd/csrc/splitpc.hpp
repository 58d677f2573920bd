// splitpc.hpp -- producer/consumer form of the bf16x3 split contraction.
//
// Same arithmetic as k_contract_split (split.hpp): A = dV_nu psi_k split into
// 2 bf16 pieces, B = psi_k+q pieces (presplit), 3 piece products per real
// product on v_mfma_f32_16x16x32_bf16, FP32 chains of PC_FLUSH * 32 points
// flushed into FP64.  The difference is the division of labour inside a
// block, which lets the VALU work of the split overlap the MFMAs:
//
//   block tile: PC_MG modes x PC_TR / PC_MG k bands (rows) x 128 k+q bands
//   (columns) x one K range, i.e. four A row blocks of 16: with PC_MG = 4
//   one per mode over the same 16 rows (psi_k is loaded once for four
//   modes), with PC_MG = 1 four row blocks of one mode.  8 waves: 4
//   consumers, 2 producers, 2 copy waves (wave w runs on SIMD w % 4).
//   producer p (0, 1): loads psi_k and dV of the A row blocks 2p, 2p+1
//     straight into registers (three chunks ahead), forms dV psi_k in
//     FP32, splits it and writes their A fragments into LDS in MFMA lane
//     order.
//   copy waves: global -> LDS copies (global_load_lds) of the B pieces of
//     the chunks ahead.
//   consumer p: reads the A fragments of row block p and the B fragments of
//     the 8 column blocks from LDS and runs the 96 MFMAs of the chunk.
//   One barrier per chunk; A pieces are double buffered (chunk c+1 is
//   produced while chunk c is consumed) and B sits in a ring of three, its
//   copies started two chunks ahead, so global latency hides behind two
//   chunks of MFMA work.
//
// LDS per buffer: A pieces [piece][re/im][row block 4][64 lanes x 16 B]
// (8 KB per piece) and B [piece][re/im][kq 4][column 128][16 B] (16 KB per
// piece); both are read as 1 KB contiguous per wave instruction.
// Requirements (checked by the host): one spinor, diagonal potential, N a
// multiple of 32, bands padded to 128 in the presplit layout.
#pragma once

#include "split.hpp"

#define PC_TR 64  // A rows of a block tile (modes x bands)
#ifndef PC_MG
#define PC_MG 4   // modes per block tile (1 or 4)
#endif
static_assert(PC_MG == 1 || PC_MG == 4, "PC_MG is 1 or 4");
#ifndef PC_BLOCKS
#define PC_BLOCKS 768 // target block count (sets the number of K splits)
#endif
#define PC_FLUSH 16 // chunks per FP32 accumulation chain (512 points)
#define PC_TC 128 // columns of a block tile
#ifndef PC_BRING
#define PC_BRING 3 // B buffers in the ring (copies run PC_BRING - 1 chunks ahead)
#endif

template <int NP>
__host__ __device__ constexpr int pc_a_bytes()
{
    return NP * 2 * 4 * 1024;
}
template <int NP>
__host__ __device__ constexpr int pc_b_bytes()
{
    return NP * 2 * 4 * 2048;
}
// A pieces double buffered, B pieces in a ring of PC_BRING (copies run
// PC_BRING - 1 chunks ahead of their use)
template <int NP>
static inline size_t pc_lds_bytes()
{
    return 2 * (size_t)pc_a_bytes<NP>() + PC_BRING * (size_t)pc_b_bytes<NP>();
}

// one chunk of this lane's A row and dV: 8 complex each
struct PcRaw
{
    uint4 a[4];
    uint4 d[4];
};

__device__ inline void pc_load_raw(PcRaw& r, const uint4* arow, const uint4* drow, int64_t c)
{
#pragma unroll
    for (int k = 0; k < 4; ++k) r.a[k] = arow[c * (SP_KC / 2) + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = drow[c * (SP_KC / 2) + k];
}

// dV psi_k -> NP bf16 pieces (re, im) -> LDS A fragment slots of row block rb
template <int NP>
__device__ inline void pc_produce(const PcRaw& r, lds_char* LA, int rb, int lane)
{
    c32 av[8], dv[8];
    __builtin_memcpy(av, r.a, sizeof(av));
    __builtin_memcpy(dv, r.d, sizeof(dv));
    bf16x8 are[NP], aim[NP];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        const c32 val = cmul(dv[k], av[k]);
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
    const int slot = rb * 1024 + lane * 16;
#pragma unroll
    for (int p = 0; p < NP; ++p)
    {
        *reinterpret_cast<__attribute__((address_space(3))) bf16x8*>(
            LA + (2 * p) * 4096 + slot) = are[p];
        *reinterpret_cast<__attribute__((address_space(3))) bf16x8*>(
            LA + (2 * p + 1) * 4096 + slot) = aim[p];
    }
}

// B pieces of chunk c, columns [m0, m0+128): NP*2 planes x 4 k groups x 2
// halves of 64 columns, one 1 KB global_load_lds each, spread over the 2
// copy waves (pw = 0, 1): PC_BCOPY instructions per wave.
#define PC_BCOPY(NP) ((NP) * 2 * 8 / 2)

// s_waitcnt immediate (gfx9 encoding): vmcnt(vm), expcnt not waited, lgkmcnt(0)
__host__ __device__ constexpr int waitcnt_vm(int vm)
{
    return (vm & 15) | ((vm >> 4) << 14) | (7 << 4);
}
template <int NP>
__device__ inline void pc_stage_B(const __bf16* __restrict__ Bp, const SplitArgs& a, int is,
                                  int m0, int64_t c, lds_char* LB, int pw, int lane)
{
    const __bf16* bt = Bp + btile_base(a, NP, is, 0, m0, c);
#pragma unroll
    for (int k = 0; k < PC_BCOPY(NP); ++k)
    {
        const int i = pw + 2 * k;
        const int plane = i >> 3, kq = (i >> 1) & 3, half = i & 1;
        const __bf16* src = bt + (int64_t)plane * a.nbp * SP_KC + (int64_t)kq * a.nbp * 8 +
                            (half * 64 + lane) * 8;
        __builtin_amdgcn_global_load_lds(src, LB + i * 1024, 16, 0, 0);
    }
}


// one chunk of this lane's two A row blocks: two psi_k rows and one dV
// (PC_MG = 1), or one psi_k row and the dV of two modes (PC_MG = 4)
struct PcRaw2
{
#if PC_MG == 1
    uint4 a[2][4];
    uint4 d[1][4];
#else
    uint4 a[1][4];
    uint4 d[2][4];
#endif
};

__device__ inline void pc_load_raw2(PcRaw2& r, const uint4* const* ap, const uint4* const* dp,
                                    int i)
{
    constexpr int NA = PC_MG == 1 ? 2 : 1;
#pragma unroll
    for (int h = 0; h < NA; ++h)
#pragma unroll
        for (int k = 0; k < 4; ++k) r.a[h][k] = ap[h][i * (SP_KC / 2) + k];
#pragma unroll
    for (int h = 0; h < 3 - NA; ++h)
#pragma unroll
        for (int k = 0; k < 4; ++k) r.d[h][k] = dp[h][i * (SP_KC / 2) + k];
}

template <int NP>
__device__ inline void pc_produce2(const PcRaw2& r, lds_char* LA, int pw, int lane)
{
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        PcRaw one;
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            one.a[k] = r.a[PC_MG == 1 ? h : 0][k];
            one.d[k] = r.d[PC_MG == 1 ? 0 : h][k];
        }
        pc_produce<NP>(one, LA, 2 * pw + h, lane);
    }
}

// one split-wave step i: split raw(i+1) into A buffer (i+1)%2, load raw(i+3),
// barrier (its lgkmcnt wait covers the LDS writes)
template <int NP>
__device__ __forceinline__ void pc_split_step(int i, int n, const PcRaw2& rp, PcRaw2& rl,
                                              const uint4* const* ap, const uint4* const* dp,
                                              lds_char* L, int pw, int lane)
{
    constexpr int A_BYTES = pc_a_bytes<NP>();
    if (i < n && i + 1 < n) pc_produce2<NP>(rp, L + ((i + 1) & 1) * A_BYTES, pw, lane);
    pc_load_raw2(rl, ap, dp, i + 3 < n ? i + 3 : n - 1);
    if (i < n) lds_barrier();
}

// one chunk of the consumer: 8 column blocks of 16, 12 MFMAs each
template <int NP>
__device__ __forceinline__ void pc_chunk16(const lds_char* LA, const lds_char* LB, int w4,
                                           int lane, flt4v* accr, flt4v* acci)
{
    bf16x8 are[NP], aim[NP], nre[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p)
    {
        are[p] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8*>(
            LA + ((2 * p) * 4 + w4) * 1024 + lane * 16);
        aim[p] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8*>(
            LA + ((2 * p + 1) * 4 + w4) * 1024 + lane * 16);
        nre[p] = neg_bf16x8(are[p]);
    }
    // B fragment of column block cb: plane, kq = lane/16, column cb*16 + lane%16
    const lds_char* bq = LB + ((lane >> 4) * PC_TC + (lane & 15)) * 16;
#pragma unroll
    for (int cb = 0; cb < 8; ++cb)
    {
        bf16x8 br[NP], bi[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p)
        {
            br[p] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8*>(
                bq + (2 * p) * 8192 + cb * 256);
            bi[p] = *reinterpret_cast<const __attribute__((address_space(3))) bf16x8*>(
                bq + (2 * p + 1) * 8192 + cb * 256);
        }
#pragma unroll
        for (int p1 = 0; p1 < NP; ++p1)
#pragma unroll
            for (int p2 = 0; p2 < NP; ++p2)
            {
                if (p1 + p2 > 1) continue; // the 3 products of the 2-piece split
                // C += a conj(b): re = ar br + ai bi, im = ai br - ar bi
                accr[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(are[p1], br[p2], accr[cb], 0, 0, 0);
                acci[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aim[p1], br[p2], acci[cb], 0, 0, 0);
                accr[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aim[p1], bi[p2], accr[cb], 0, 0, 0);
                acci[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nre[p1], bi[p2], acci[cb], 0, 0, 0);
            }
    }
}

__device__ __forceinline__ void pc_flush_block(flt4v& ar, flt4v& ai, double* dr, double* di)
{
#pragma unroll
    for (int e = 0; e < 4; ++e)
    {
        dr[e] += (double)ar[e];
        di[e] += (double)ai[e];
        ar[e] = 0.f;
        ai[e] = 0.f;
    }
}

// partials of an empty K range (more splits than chunks): zeros, so the
// fixed-order reduction over all splits reads defined values
__device__ inline void pc_store_zero(const SplitArgs& a, z64* __restrict__ part, int split, int v,
                                     int is, int row, int m0, int lane)
{
    if (v >= a.nm) return;
#pragma unroll
    for (int cb = 0; cb < 8; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int rr = row + 4 * (lane >> 4) + e;
            const int cc = m0 + cb * 16 + (lane & 15);
            if (rr < a.nb && cc < a.nb)
            {
                const int64_t o = ((((int64_t)split * a.nm + v) * a.nspin + is) * a.nb + rr) * a.nb + cc;
                part[o] = {0.0, 0.0};
            }
        }
}

// consumer: A row block w4 = mode v, rows row .. row + 15, columns m0 ..
// m0 + 127 (8 blocks of 16); FP32 chains of PC_FLUSH chunks flushed into
// FP64; modes v >= nm (padding of the last mode group) are not stored
template <int NP>
__device__ __forceinline__ void pc_consume16(const lds_char* L, int n, int w4, int lane,
                                             const SplitArgs& a, z64* __restrict__ part,
                                             int split, int v, int is, int row, int m0)
{
    constexpr int A_BYTES = pc_a_bytes<NP>();
    constexpr int B_BYTES = pc_b_bytes<NP>();
    flt4v accr[8], acci[8];
    double dr[8][4], di[8][4];
#pragma unroll
    for (int cb = 0; cb < 8; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            accr[cb][e] = 0.f;
            acci[cb][e] = 0.f;
            dr[cb][e] = 0.0;
            di[cb][e] = 0.0;
        }
    lds_barrier();
    int u = 0, u3 = 0, nfl = 0;
    for (int i = 0; i < n; ++i)
    {
        pc_chunk16<NP>(L + u * A_BYTES, L + 2 * A_BYTES + u3 * B_BYTES, w4, lane, accr, acci);
        if (++nfl == PC_FLUSH)
        {
            nfl = 0;
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) pc_flush_block(accr[cb], acci[cb], dr[cb], di[cb]);
        }
        lds_barrier();
        u ^= 1;
        u3 = (u3 == PC_BRING - 1) ? 0 : u3 + 1;
    }
    if (v >= a.nm) return;
#pragma unroll
    for (int cb = 0; cb < 8; ++cb)
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
__global__ __launch_bounds__(SP_THREADS, 2) void k_contract_split_pc(
    const c32* __restrict__ psik, const __bf16* __restrict__ Bp, const c32* __restrict__ dv,
    z64* __restrict__ part, SplitArgs a)
{
    constexpr int A_BYTES = pc_a_bytes<NP>();
    constexpr int B_BYTES = pc_b_bytes<NP>();
    extern __shared__ __attribute__((aligned(16))) char pc_smem[];
    lds_char* L = (lds_char*)pc_smem;
    // A buffer u at L + u*A_BYTES (u = 0, 1), B buffer t at
    // L + 2*A_BYTES + t*B_BYTES (t = 0 .. PC_BRING - 1)

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool producer = wave >= 4;
    const int w4 = wave & 3;
    // XCD-aware decode: group g = (spin, split) runs on XCD g % 8
    const int per_group = a.nmt * a.ntn * a.ntm;
    const int b = blockIdx.x;
    const int x = b & 7, j = b >> 3;
    const int g = x + 8 * (j / per_group);
    const int tile = j % per_group;
    const int vt = tile / (a.ntn * a.ntm); // mode tile
    const int tn = (tile / a.ntm) % a.ntn, tm = tile % a.ntm;
    const int is = g / a.nsplit, split = g % a.nsplit;
    if (is >= a.nspin) return;
    const int n0 = tn * (PC_TR / PC_MG), m0 = tm * PC_TC;
    // mode and first row of consumer w4
    const int vc = PC_MG == 1 ? vt : PC_MG * vt + w4;
    const int rowc = PC_MG == 1 ? n0 + 16 * w4 : n0;
    const int64_t c0 = (int64_t)split * a.chunks_per_split;
    const int64_t c1 = (c0 + a.chunks_per_split < a.nchunks) ? c0 + a.chunks_per_split : a.nchunks;
    if (c0 >= c1) // uniform over the block; no barrier is reached
    {
        if (wave < 4) pc_store_zero(a, part, split, vc, is, rowc, m0, lane);
        return;
    }

    const int n = (int)(c1 - c0); // chunks of this block
    if (wave >= 6)
    {
        // copy waves: B(i + PC_BRING - 1) is started in step i, B(i+1) must
        // have landed before the barrier that ends step i; a wave waits only
        // for its own copies, PC_BCOPY per chunk
        static_assert(PC_BRING == 3 || PC_BRING == 4, "B ring of 3 or 4");
        const int pw = wave - 6;
        lds_char* LB = L + 2 * A_BYTES;
        for (int k = 0; k < PC_BRING - 1 && k < n; ++k)
            pc_stage_B<NP>(Bp, a, is, m0, c0 + k, LB + k * B_BYTES, pw, lane);
        // B(0) landed: later prologue copies may stay in flight
        if (PC_BRING == 4 && n > 2)
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(2 * PC_BCOPY(NP)));
        else if (n > 1)
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(PC_BCOPY(NP)));
        else
            __builtin_amdgcn_s_waitcnt(0);
        lds_barrier();
        int t3 = PC_BRING - 1;
        for (int i = 0; i < n; ++i)
        {
            if (i + PC_BRING - 1 < n)
            {
                pc_stage_B<NP>(Bp, a, is, m0, c0 + i + PC_BRING - 1, LB + t3 * B_BYTES, pw,
                               lane);
                // B(i+1) landed, the PC_BRING - 2 later chunks may be in flight
                __builtin_amdgcn_s_waitcnt(waitcnt_vm((PC_BRING - 2) * PC_BCOPY(NP)));
            }
            else if (PC_BRING == 4 && i + 3 == n)
                __builtin_amdgcn_s_waitcnt(waitcnt_vm(PC_BCOPY(NP))); // B(i+2) may fly
            else
                __builtin_amdgcn_s_waitcnt(0);
            lds_barrier();
            t3 = (t3 == PC_BRING - 1) ? 0 : t3 + 1;
        }
        return;
    }
    if (producer)
    {
        // split waves: A row blocks 2 pw, 2 pw + 1 (rows n0 + 32 pw + 16 h
        // of mode vt, or rows n0 of modes 4 vt + 2 pw + h), row lane%16,
        // points 8(lane/16)..+7 of every chunk; raw chunk i+3 loads while
        // i+1 is split (three register sets, loads clamped to the last chunk
        // so every set keeps fixed registers); padding modes read the last
        // mode and padding rows the last band
        const int pw = wave - 4;
        const uint4* ap[2];
        const uint4* dp[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
            const int arow =
                min(n0 + (PC_MG == 1 ? 32 * pw + 16 * h : 0) + (lane & 15), a.nb - 1);
            const int vm = min(PC_MG == 1 ? vt : PC_MG * vt + 2 * pw + h, a.nm - 1);
            ap[h] = reinterpret_cast<const uint4*>(psik + (int64_t)(is * a.nb + arow) * a.N +
                                                   8 * (lane >> 4)) + c0 * (SP_KC / 2);
            dp[h] = reinterpret_cast<const uint4*>(
                        dv + ((int64_t)vm * a.nmag + (a.nmag == 2 ? is : 0)) * a.N +
                        8 * (lane >> 4)) + c0 * (SP_KC / 2);
        }
        PcRaw2 R0, R1, R2;
        pc_load_raw2(R0, ap, dp, 0);
        pc_produce2<NP>(R0, L, pw, lane);
        pc_load_raw2(R1, ap, dp, n > 1 ? 1 : n - 1);
        pc_load_raw2(R2, ap, dp, n > 2 ? 2 : n - 1);
        lds_barrier();
        for (int i = 0; i < n; i += 3)
        {
            pc_split_step<NP>(i, n, R1, R0, ap, dp, L, pw, lane);
            pc_split_step<NP>(i + 1, n, R2, R1, ap, dp, L, pw, lane);
            pc_split_step<NP>(i + 2, n, R0, R2, ap, dp, L, pw, lane);
        }
        __builtin_amdgcn_s_waitcnt(0);
        return;
    }

    pc_consume16<NP>(L, n, w4, lane, a, part, split, vc, is, rowc, m0);
}
