// learned.hip -- a small network that stands in for the Fourier interpolation
// of dV_q onto fine q points (src/interpolation/interpolation_driver.c:
// 813-856, fft_R2q_dvscf src/interpolation/interpolation_utilities.c:
// 165-238), trained on the GPU by the kernels of this file.
//
// What it predicts.  The short-range, lattice-periodic part of the
// potential in the cartesian basis, u_q(r) = e^{-i 2 pi q.x} (dV_q(r) -
// dV_q^LR(r)) (nf complex fields, nout = 2 nf real channels, x the
// fractional grid coordinate), at a fine q from the coarse grid: the input
// of a sample is, per grid point, the periodic parts of the 8 coarse q
// around it (each with its own q' unwrapped next to q, so that no
// e^{iGr} jump enters), their trilinear blend and sin/cos features of the
// position of q in its coarse cell; the network's output is added to the
// trilinear blend (a residual).
//
// The network.  L = depth + 2 3-D convolutions, 3x3x3 taps, periodic
// padding on the FFT grid, channels NN_CIN -> NN_W -> ... -> NN_W ->
// nout (padded to NN_W), SiLU between layers, linear output.  Operands in
// BF16, MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, FP32
// master weights, Adam.  Each convolution is an implicit GEMM (k_nn_conv:
// the A fragments read from the 27 neighbours' channel rows); its backward
// is the input gradient as the same implicit GEMM over the mirrored taps
// and the weight gradient as one GEMM (split over the points) on transposed
// columns built through LDS (k_nn_im2col_t).  Every reduction has a fixed
// order, so a run is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"

typedef __bf16 nn_bf16x8 __attribute__((ext_vector_type(8)));
typedef float nn_flt4 __attribute__((ext_vector_type(4)));

#define NN_CIN 128  // input channels of the first layer (features, zero padded)
#define NN_W 64     // hidden width; the output layer is padded to it
#define NN_TAPS 27  // 3x3x3
#define NN_MAXL 16  // layers at most
#define NN_FEAT_NB 8

namespace {

struct NBuf
{
    void* p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= n) return LELPH_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess)
        {
            p = nullptr;
            return LELPH_ERR_ALLOC;
        }
        n = bytes;
        return LELPH_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename X>
    X* as()
    {
        return reinterpret_cast<X*>(p);
    }
};

// grid geometry: the periodic neighbour of point p (sample-major index
// s * N + (ix * Ny + iy) * Nz + iz) at tap t = (dx+1) * 9 + (dy+1) * 3 + (dz+1)
struct NnGrid
{
    int nx, ny, nz;
    int64_t n;
};

__device__ inline int64_t nn_nbr(const NnGrid& g, int64_t p, int t, int sgn)
{
    const int64_t s = p / g.n;
    int64_t r = p - s * g.n;
    const int iz = (int)(r % g.nz);
    r /= g.nz;
    const int iy = (int)(r % g.ny);
    const int ix = (int)(r / g.ny);
    const int dx = sgn * (t / 9 - 1), dy = sgn * ((t / 3) % 3 - 1), dz = sgn * (t % 3 - 1);
    const int jx = (ix + dx + g.nx) % g.nx, jy = (iy + dy + g.ny) % g.ny, jz = (iz + dz + g.nz) % g.nz;
    return s * g.n + ((int64_t)jx * g.ny + jy) * g.nz + jz;
}

// colT[t * C + c][p] = A[nbr(p, t)][c] (C = 64 or 128).  Block: 64 points
// x the C channels of one tap through LDS: the neighbours' channel rows are
// read as 16-byte units (a point's channels are contiguous), stored with a
// row stride of C + 2 bf16 (65 or 33 words, so the column reads below fall
// on distinct banks), then every thread writes 8 consecutive points of one
// channel (16 bytes; 8 threads cover a channel's 128-byte row segment).
// P is a multiple of 64 (the host requires grids of 64k points).
#define NN_T_PTS 64
__global__ __launch_bounds__(256) void k_nn_im2col_t(const __bf16* __restrict__ A, int C,
                                                     NnGrid g, int64_t P,
                                                     __bf16* __restrict__ colT)
{
    __shared__ uint32_t tile[NN_T_PTS * (NN_CIN + 2) / 2];
    const int64_t p0 = (int64_t)blockIdx.x * NN_T_PTS;
    const int t = blockIdx.y;
    const int cu = C / 8, rw = (C + 2) / 2; // 16-byte units per point, words per LDS row
    for (int u = threadIdx.x; u < NN_T_PTS * cu; u += blockDim.x)
    {
        const int r = u / cu, k = u % cu;
        const int64_t q = nn_nbr(g, p0 + r, t, 1);
        const uint4 v = *reinterpret_cast<const uint4*>(A + q * C + 8 * k);
        uint32_t* d = tile + r * rw + 4 * k;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    __syncthreads();
    const __bf16* t16 = reinterpret_cast<const __bf16*>(tile);
    for (int v = threadIdx.x; v < C * (NN_T_PTS / 8); v += blockDim.x)
    {
        const int c = v / (NN_T_PTS / 8), j = v % (NN_T_PTS / 8);
        nn_bf16x8 o;
#pragma unroll
        for (int i = 0; i < 8; ++i) o[i] = t16[(8 * j + i) * (2 * rw) + c];
        *reinterpret_cast<nn_bf16x8*>(colT + ((int64_t)t * C + c) * P + p0 + 8 * j) = o;
    }
}

// C[M][N] (or partial split s of it) = sum_k A[m][k] B[n][k], A and B BF16
// with k contiguous, FP32 accumulation on MFMA.  Block 64 x 64 of C, 4 waves
// of 32 x 32 (2 x 2 tiles of 16 x 16), k in steps of 32.  Fragment layout of
// v_mfma_f32_16x16x32_bf16 (wave64): A: lane l holds A[l % 16][8 (l / 16) +
// 0..7]; B: lane l holds B[8 (l / 16) + 0..7][l % 16]; D: lane l, v holds
// D[4 (l / 16) + v][l % 16].  M and N multiples of 64, K of 32 (checked by
// the host).  Split s covers k in [s * kper, min(K, (s + 1) * kper)).
__global__ __launch_bounds__(256) void k_nn_gemm(const __bf16* __restrict__ A,
                                                 const __bf16* __restrict__ B, int64_t M,
                                                 int64_t N, int64_t K, int64_t kper,
                                                 float* __restrict__ C)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t ntn = N / 64;
    const int64_t tm = blockIdx.x / ntn, tn = blockIdx.x % ntn;
    const int64_t m0 = tm * 64 + (wave >> 1) * 32, n0 = tn * 64 + (wave & 1) * 32;
    const int64_t split = blockIdx.y;
    const int64_t k0 = split * kper;
    const int64_t k1 = (k0 + kper < K) ? k0 + kper : K;
    nn_flt4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (nn_flt4){0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, kq = 8 * (lane >> 4);
    const __bf16* pa0 = A + (m0 + r) * K + kq;
    const __bf16* pa1 = A + (m0 + 16 + r) * K + kq;
    const __bf16* pb0 = B + (n0 + r) * K + kq;
    const __bf16* pb1 = B + (n0 + 16 + r) * K + kq;
    for (int64_t k = k0; k < k1; k += 32)
    {
        const nn_bf16x8 a0 = *reinterpret_cast<const nn_bf16x8*>(pa0 + k);
        const nn_bf16x8 a1 = *reinterpret_cast<const nn_bf16x8*>(pa1 + k);
        const nn_bf16x8 b0 = *reinterpret_cast<const nn_bf16x8*>(pb0 + k);
        const nn_bf16x8 b1 = *reinterpret_cast<const nn_bf16x8*>(pb1 + k);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
    }
    float* out = C + split * M * N;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int64_t m = m0 + rb * 16 + 4 * (lane >> 4) + v;
                const int64_t n = n0 + cb * 16 + r;
                out[m * N + n] = acc[rb][cb][v];
            }
}

// out[i] = sum_s part[s][i], fixed order
__global__ void k_nn_sum_splits(const float* __restrict__ part, int ns, int64_t count,
                                float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float a = 0.f;
    for (int s = 0; s < ns; ++s) a += part[(int64_t)s * count + i];
    out[i] = a;
}

// Implicit-GEMM convolution: out[p][n] = sum_t sum_c X[nbr(p, sgn t)][c]
// B[n * bn + t * bt + c] over the 27 taps t and CR channels c (a multiple
// of 32), with the tiling, fragment layout and k order (t, then c) of
// k_nn_gemm, the A fragments read straight from the neighbours' channel rows
// (16 bytes of one row per lane) instead of from a column buffer.
//   forward: X = the layer input (CR = C), sgn = +1, B = Wt (bn = 27 C, bt =
//            C): the same products in the same order as k_nn_gemm on
//            explicit columns col[p][t C + c] = X[nbr(p, t)][c] would issue;
//   input gradient: X = dZ (CR = 64), sgn = -1, B = Wk (bn = 64, bt = 64 C):
//            dA[p][ci] = sum_t sum_c dZ[nbr(p, -t)][c] W[t C + ci][c].
// P a multiple of 64, N of 64 (checked by the host).
__global__ __launch_bounds__(256) void k_nn_conv(const __bf16* __restrict__ X, int CR, int sgn,
                                                 NnGrid g, const __bf16* __restrict__ B,
                                                 int64_t bn, int64_t bt, int64_t N,
                                                 float* __restrict__ out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t ntn = N / 64;
    const int64_t tm = blockIdx.x / ntn, tn = blockIdx.x % ntn;
    const int64_t m0 = tm * 64 + (wave >> 1) * 32, n0 = tn * 64 + (wave & 1) * 32;
    nn_flt4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (nn_flt4){0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, kq = 8 * (lane >> 4);
    const int64_t p0 = m0 + r, p1 = m0 + 16 + r;
    for (int t = 0; t < NN_TAPS; ++t)
    {
        const __bf16* xa0 = X + nn_nbr(g, p0, t, sgn) * CR + kq;
        const __bf16* xa1 = X + nn_nbr(g, p1, t, sgn) * CR + kq;
        const __bf16* pb0 = B + (n0 + r) * bn + (int64_t)t * bt + kq;
        const __bf16* pb1 = B + (n0 + 16 + r) * bn + (int64_t)t * bt + kq;
        for (int c = 0; c < CR; c += 32)
        {
            const nn_bf16x8 a0 = *reinterpret_cast<const nn_bf16x8*>(xa0 + c);
            const nn_bf16x8 a1 = *reinterpret_cast<const nn_bf16x8*>(xa1 + c);
            const nn_bf16x8 b0 = *reinterpret_cast<const nn_bf16x8*>(pb0 + c);
            const nn_bf16x8 b1 = *reinterpret_cast<const nn_bf16x8*>(pb1 + c);
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int64_t m = m0 + rb * 16 + 4 * (lane >> 4) + v;
                const int64_t n = n0 + cb * 16 + r;
                out[m * N + n] = acc[rb][cb][v];
            }
}

__device__ inline float nn_sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

// Z[p][c] += b[c]; hidden layers also write A = SiLU(Z) in BF16
__global__ void k_nn_bias_act(float* __restrict__ Z, const float* __restrict__ b, int64_t P,
                              int act, __bf16* __restrict__ A)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * NN_W) return;
    const float z = Z[i] + b[i % NN_W];
    Z[i] = z;
    if (act) A[i] = (__bf16)(z * nn_sigmoid(z));
}

// output layer: r = base + Z - target on the nout channels; dZ = 2 r / cnt
// (zero on the padding); per-block sums of r^2 in lossp[blockIdx.x]
__global__ void k_nn_loss(const float* __restrict__ Z, const float* __restrict__ base,
                          const float* __restrict__ target, int64_t P, int nout, float inv_cnt,
                          float* __restrict__ dZ, float* __restrict__ lossp)
{
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float rr = 0.f;
    if (i < P * NN_W)
    {
        const int64_t p = i / NN_W;
        const int c = (int)(i % NN_W);
        float d = 0.f;
        if (c < nout)
        {
            const float res = base[p * nout + c] + Z[i] - target[p * nout + c];
            rr = res * res;
            d = 2.f * res * inv_cnt;
        }
        dZ[i] = d;
    }
    red[threadIdx.x] = rr;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// hidden layers: dZ = dA * SiLU'(Z) in place of dA (FP32)
__global__ void k_nn_silu_bwd(float* __restrict__ d, const float* __restrict__ Z, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i];
    const float s = nn_sigmoid(z);
    d[i] *= s * (1.f + z * (1.f - s));
}

// dZ (FP32, [P][NN_W]) -> BF16 [P][NN_W] and transposed [NN_W][P], through a
// 64 x 64 LDS tile
__global__ void k_nn_dz_pack(const float* __restrict__ dZf, int64_t P, __bf16* __restrict__ dZ,
                             __bf16* __restrict__ dZT, float* __restrict__ bpart)
{
    __shared__ float tile[64][NN_W + 1];
    const int64_t p0 = (int64_t)blockIdx.x * 64;
    for (int f = threadIdx.x; f < 64 * NN_W; f += blockDim.x)
    {
        const int pp = f / NN_W, c = f % NN_W;
        const float v = dZf[(p0 + pp) * NN_W + c];
        tile[pp][c] = v;
        dZ[(p0 + pp) * NN_W + c] = (__bf16)v;
    }
    __syncthreads();
    // the bias gradient's partial of these 64 points, in point order
    if (threadIdx.x < NN_W)
    {
        float s = 0.f;
        for (int pp = 0; pp < 64; ++pp) s += tile[pp][threadIdx.x];
        bpart[blockIdx.x * NN_W + threadIdx.x] = s;
    }
    for (int f = threadIdx.x; f < 64 * NN_W; f += blockDim.x)
    {
        const int c = f / 64, pp = f % 64;
        dZT[(int64_t)c * P + p0 + pp] = (__bf16)tile[pp][c];
    }
}

// gb[c] = sum over the 64-point partials of k_nn_dz_pack, one block of 256
// threads: 4 groups of 64 (one channel per thread, coalesced rows), each
// summing every 4th partial in order, then the 4 groups in order
__global__ __launch_bounds__(256) void k_nn_bias_grad(const float* __restrict__ bpart,
                                                      int64_t nblk, float* __restrict__ gb)
{
    __shared__ float red[4][NN_W];
    const int c = threadIdx.x % NN_W, grp = threadIdx.x / NN_W;
    float a = 0.f;
    for (int64_t b = grp; b < nblk; b += 4) a += bpart[b * NN_W + c];
    red[grp][c] = a;
    __syncthreads();
    if (threadIdx.x < NN_W) gb[c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

// Adam on FP32 masters (mask: entries held at zero, the padding)
__global__ void k_nn_adam(float* __restrict__ w, const float* __restrict__ gr,
                          float* __restrict__ m, float* __restrict__ v,
                          const unsigned char* __restrict__ mask, int64_t n, float lr, float b1,
                          float b2, float c1, float c2, float eps)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[i]) return;
    const float gg = gr[i];
    const float mm = b1 * m[i] + (1.f - b1) * gg;
    const float vv = b2 * v[i] + (1.f - b2) * gg * gg;
    m[i] = mm;
    v[i] = vv;
    w[i] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 copies of W[K][NN_W]: Wt[NN_W][K] (forward B operand) and Wk[K][NN_W]
// (column-gradient B operand)
__global__ void k_nn_pack_w(const float* __restrict__ W, int64_t K, __bf16* __restrict__ Wt,
                            __bf16* __restrict__ Wk)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * NN_W) return;
    const int64_t k = i / NN_W;
    const int c = (int)(i % NN_W);
    const __bf16 b = (__bf16)W[i];
    Wk[i] = b;
    Wt[(int64_t)c * K + k] = b;
}

// one sample's neighbours: coarse index, unwrapped q' (crystal), trilinear
// weight; and the features of q's position in its coarse cell
struct NnSample
{
    int iqc[NN_FEAT_NB];
    double qp[NN_FEAT_NB][3];
    float w[NN_FEAT_NB];
    float off[6];
};

// X0[s][r][c] (BF16, NN_CIN channels) and base[s][r][2 nf] (FP32) of the
// samples from the coarse short-range fields coarse[iqc][f][r] (complex
// FP32): u_j = e^{-i 2 pi q'_j . x} coarse[iqc_j] / scale per neighbour
// (channels 2 (j nf + f) + {0, 1}), their trilinear blend (channels
// 16 nf + 2 f + {0, 1}, and base), the offset features, zeros after.
__global__ void k_nn_features(const c32* __restrict__ coarse, int nf, NnGrid g,
                              const NnSample* __restrict__ smp, float inv_scale,
                              __bf16* __restrict__ X0, float* __restrict__ base)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (r >= g.n) return;
    const NnSample& sm = smp[s];
    const int iz = (int)(r % g.nz), iy = (int)((r / g.nz) % g.ny), ix = (int)(r / ((int64_t)g.nz * g.ny));
    const double x[3] = {(double)ix / g.nx, (double)iy / g.ny, (double)iz / g.nz};
    __bf16* xo = X0 + ((int64_t)s * g.n + r) * NN_CIN;
    float* bo = base + ((int64_t)s * g.n + r) * 2 * nf;
    for (int f = 0; f < 2 * nf; ++f) bo[f] = 0.f;
    for (int j = 0; j < NN_FEAT_NB; ++j)
    {
        const double ph = -2.0 * (sm.qp[j][0] * x[0] + sm.qp[j][1] * x[1] + sm.qp[j][2] * x[2]);
        double sn, cs;
        sincospi(ph, &sn, &cs);
        const c32 e = {(float)cs, (float)sn};
        for (int f = 0; f < nf; ++f)
        {
            c32 u = cmul(e, coarse[((int64_t)sm.iqc[j] * nf + f) * g.n + r]);
            u.x *= inv_scale;
            u.y *= inv_scale;
            xo[2 * (j * nf + f)] = (__bf16)u.x;
            xo[2 * (j * nf + f) + 1] = (__bf16)u.y;
            bo[2 * f] += sm.w[j] * u.x;
            bo[2 * f + 1] += sm.w[j] * u.y;
        }
    }
    const int cb = 2 * nf * NN_FEAT_NB;
    for (int f = 0; f < 2 * nf; ++f) xo[cb + f] = (__bf16)bo[f];
    for (int k = 0; k < 6; ++k) xo[cb + 2 * nf + k] = (__bf16)sm.off[k];
    for (int c = cb + 2 * nf + 6; c < NN_CIN; ++c) xo[c] = (__bf16)0.f;
}

// out[s][r][c] = (base + Z[.][c]) * scale for the nout channels
__global__ void k_nn_out(const float* __restrict__ Z, const float* __restrict__ base, int64_t P,
                         int nout, float scale, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * nout) return;
    const int64_t p = i / nout;
    const int c = (int)(i % nout);
    out[i] = (base[i] + Z[p * NN_W + c]) * scale;
}

__global__ void k_nn_f32_to_bf16(const float* __restrict__ a, int64_t n, __bf16* __restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (__bf16)a[i];
}

// splitmix64 stream for the initial weights
inline uint64_t nn_mix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double nn_unif(uint64_t& s) { return ((nn_mix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

} // namespace

struct lelph_nn
{
    hipStream_t stream = nullptr;
    NnGrid g{};
    int nf = 0, nout = 0, depth = 0, L = 0;
    int cin[NN_MAXL] = {};
    int64_t K[NN_MAXL] = {};
    NBuf W[NN_MAXL], b[NN_MAXL], mW[NN_MAXL], vW[NN_MAXL], mb[NN_MAXL], vb[NN_MAXL];
    NBuf Wt[NN_MAXL], Wk[NN_MAXL], gW[NN_MAXL], gb[NN_MAXL], maskW[NN_MAXL], maskb[NN_MAXL];
    int64_t step = 0;
    // coarse short-range fields
    int64_t qgrid[3] = {0, 0, 0};
    int nqc = 0;
    NBuf coarse;
    float scale = 1.f;
    // training targets (nq, N, nout) FP32
    NBuf targets;
    int64_t ntarg = 0;
    // batch work space, sized for smax samples
    int smax = 0;
    NBuf smp, X0, base, tgt, A[NN_MAXL], Z[NN_MAXL], colT, dZf, dZ, dZT, dA, part, lossp, bpart;
    std::vector<float> lossh;
};

namespace {

static int nn_launch_gemm(lelph_nn* m, const __bf16* A, const __bf16* B, int64_t M, int64_t N,
                          int64_t K, int ksplit, float* C, float* part)
{
    if (M % 64 || N % 64 || K % 32 || ksplit < 1) return LELPH_ERR_ARG;
    const int64_t kper = cdiv(cdiv(K, 32), ksplit) * 32;
    const int64_t ns = cdiv(K, kper);
    const dim3 grid((unsigned)((M / 64) * (N / 64)), (unsigned)ns);
    hipLaunchKernelGGL(k_nn_gemm, grid, dim3(256), 0, m->stream, A, B, M, N, K, kper,
                       ns > 1 ? part : C);
    HIP_TRY(hipGetLastError());
    if (ns > 1)
    {
        hipLaunchKernelGGL(k_nn_sum_splits, dim3((unsigned)cdiv(M * N, 256)), dim3(256), 0,
                           m->stream, part, (int)ns, M * N, C);
        HIP_TRY(hipGetLastError());
    }
    return LELPH_OK;
}

static int nn_ensure_batch(lelph_nn* m, int S)
{
    if (S <= m->smax) return LELPH_OK;
    const int64_t P = (int64_t)S * m->g.n;
    const int64_t Kmax = (int64_t)NN_TAPS * NN_CIN;
    LELPH_TRY(m->smp.ensure(sizeof(NnSample) * S));
    LELPH_TRY(m->X0.ensure(sizeof(__bf16) * P * NN_CIN));
    LELPH_TRY(m->base.ensure(sizeof(float) * P * m->nout));
    LELPH_TRY(m->tgt.ensure(sizeof(float) * P * m->nout));
    for (int l = 0; l < m->L; ++l)
    {
        LELPH_TRY(m->A[l].ensure(sizeof(__bf16) * P * NN_W));
        LELPH_TRY(m->Z[l].ensure(sizeof(float) * P * NN_W));
    }
    LELPH_TRY(m->colT.ensure(sizeof(__bf16) * P * Kmax));
    LELPH_TRY(m->dZf.ensure(sizeof(float) * P * NN_W));
    LELPH_TRY(m->dZ.ensure(sizeof(__bf16) * P * NN_W));
    LELPH_TRY(m->dZT.ensure(sizeof(__bf16) * P * NN_W));
    LELPH_TRY(m->dA.ensure(sizeof(float) * P * NN_W));
    LELPH_TRY(m->part.ensure(sizeof(float) * 32 * Kmax * NN_W));
    LELPH_TRY(m->lossp.ensure(sizeof(float) * cdiv(P * NN_W, 256)));
    LELPH_TRY(m->bpart.ensure(sizeof(float) * (P / 64) * NN_W));
    m->smax = S;
    return LELPH_OK;
}

// forward over the batch in X0 (S samples); layer l's input is A[l-1] (or
// X0), its pre-activation Z[l], its output A[l]
static int nn_forward(lelph_nn* m, int S)
{
    const int64_t P = (int64_t)S * m->g.n;
    for (int l = 0; l < m->L; ++l)
    {
        const __bf16* in = l == 0 ? m->X0.as<__bf16>() : m->A[l - 1].as<__bf16>();
        const int C = m->cin[l];
        hipLaunchKernelGGL(k_nn_conv, dim3((unsigned)(P / 64)), dim3(256), 0, m->stream, in, C, 1,
                           m->g, m->Wt[l].as<__bf16>(), m->K[l], (int64_t)C, (int64_t)NN_W,
                           m->Z[l].as<float>());
        HIP_TRY(hipGetLastError());
        const int act = l + 1 < m->L;
        hipLaunchKernelGGL(k_nn_bias_act, dim3((unsigned)cdiv(P * NN_W, 256)), dim3(256), 0,
                           m->stream, m->Z[l].as<float>(), m->b[l].as<float>(), P, act,
                           m->A[l].as<__bf16>());
        HIP_TRY(hipGetLastError());
    }
    return LELPH_OK;
}

// loss of the batch (base + output vs tgt) and every gradient into gW, gb
static int nn_backward(lelph_nn* m, int S, double* loss)
{
    const int64_t P = (int64_t)S * m->g.n;
    const int last = m->L - 1;
    const int64_t nblk = cdiv(P * NN_W, 256);
    const float inv = 1.f / (float)(P * m->nout);
    hipLaunchKernelGGL(k_nn_loss, dim3((unsigned)nblk), dim3(256), 0, m->stream,
                       m->Z[last].as<float>(), m->base.as<float>(), m->tgt.as<float>(), P, m->nout,
                       inv, m->dZf.as<float>(), m->lossp.as<float>());
    HIP_TRY(hipGetLastError());
    for (int l = last; l >= 0; --l)
    {
        if (l < last)
        {
            // dZf holds dA of this layer's output
            hipLaunchKernelGGL(k_nn_silu_bwd, dim3((unsigned)cdiv(P * NN_W, 256)), dim3(256), 0,
                               m->stream, m->dZf.as<float>(), m->Z[l].as<float>(), P * NN_W);
            HIP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_nn_dz_pack, dim3((unsigned)(P / 64)), dim3(256), 0, m->stream,
                           m->dZf.as<float>(), P, m->dZ.as<__bf16>(), m->dZT.as<__bf16>(),
                           m->bpart.as<float>());
        hipLaunchKernelGGL(k_nn_bias_grad, dim3(1), dim3(256), 0, m->stream, m->bpart.as<float>(),
                           P / 64, m->gb[l].as<float>());
        HIP_TRY(hipGetLastError());
        const __bf16* in = l == 0 ? m->X0.as<__bf16>() : m->A[l - 1].as<__bf16>();
        const int C = m->cin[l];
        hipLaunchKernelGGL(k_nn_im2col_t, dim3((unsigned)(P / NN_T_PTS), NN_TAPS), dim3(256), 0,
                           m->stream, in, C, m->g, P, m->colT.as<__bf16>());
        HIP_TRY(hipGetLastError());
        // dW[k][c] = sum_p colT[k][p] dZT[c][p], split over the points
        const int ks = (int)std::min<int64_t>(32, std::max<int64_t>(1, P / 4096));
        LELPH_TRY(nn_launch_gemm(m, m->colT.as<__bf16>(), m->dZT.as<__bf16>(), m->K[l], NN_W, P,
                                 ks, m->gW[l].as<float>(), m->part.as<float>()));
        if (l > 0)
        {
            // dA[p][ci] = sum_t sum_c dZ[nbr(p, -t)][c] Wk[t C + ci][c]
            hipLaunchKernelGGL(k_nn_conv, dim3((unsigned)((P / 64) * (C / 64))), dim3(256), 0,
                               m->stream, m->dZ.as<__bf16>(), NN_W, -1, m->g,
                               m->Wk[l].as<__bf16>(), (int64_t)NN_W, (int64_t)C * NN_W,
                               (int64_t)C, m->dZf.as<float>());
            HIP_TRY(hipGetLastError());
        }
    }
    if (loss)
    {
        std::vector<float> lp((size_t)nblk);
        HIP_TRY(hipMemcpyAsync(lp.data(), m->lossp.p, sizeof(float) * nblk, hipMemcpyDeviceToHost,
                               m->stream));
        HIP_TRY(hipStreamSynchronize(m->stream));
        double a = 0;
        for (float v : lp) a += v;
        *loss = a / (double)(P * m->nout);
    }
    return LELPH_OK;
}

static int nn_pack_all(lelph_nn* m)
{
    for (int l = 0; l < m->L; ++l)
    {
        hipLaunchKernelGGL(k_nn_pack_w, dim3((unsigned)cdiv(m->K[l] * NN_W, 256)), dim3(256), 0,
                           m->stream, m->W[l].as<float>(), m->K[l], m->Wt[l].as<__bf16>(),
                           m->Wk[l].as<__bf16>());
        HIP_TRY(hipGetLastError());
    }
    return LELPH_OK;
}

static int nn_adam_step(lelph_nn* m, float lr)
{
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
    for (int l = 0; l < m->L; ++l)
    {
        const int64_t nw = m->K[l] * NN_W;
        hipLaunchKernelGGL(k_nn_adam, dim3((unsigned)cdiv(nw, 256)), dim3(256), 0, m->stream,
                           m->W[l].as<float>(), m->gW[l].as<float>(), m->mW[l].as<float>(),
                           m->vW[l].as<float>(), m->maskW[l].as<unsigned char>(), nw, lr, b1, b2,
                           c1, c2, 1e-8f);
        hipLaunchKernelGGL(k_nn_adam, dim3(1), dim3(NN_W), 0, m->stream, m->b[l].as<float>(),
                           m->gb[l].as<float>(), m->mb[l].as<float>(), m->vb[l].as<float>(),
                           m->maskb[l].as<unsigned char>(), (int64_t)NN_W, lr, b1, b2, c1, c2,
                           1e-8f);
        HIP_TRY(hipGetLastError());
    }
    return nn_pack_all(m);
}

// host side of a sample: the 8 coarse points around q and the features
static NnSample nn_sample(const lelph_nn* m, const double* q)
{
    NnSample s{};
    int i0[3];
    double f[3];
    for (int d = 0; d < 3; ++d)
    {
        const double x = q[d] * (double)m->qgrid[d];
        i0[d] = (int)std::floor(x);
        f[d] = x - (double)i0[d];
        s.off[2 * d] = (float)std::sin(M_PI * f[d]);
        s.off[2 * d + 1] = (float)std::cos(M_PI * f[d]);
    }
    for (int j = 0; j < NN_FEAT_NB; ++j)
    {
        int id[3];
        double w = 1.0;
        for (int d = 0; d < 3; ++d)
        {
            const int bit = (j >> (2 - d)) & 1;
            const int ii = i0[d] + bit;
            s.qp[j][d] = (double)ii / (double)m->qgrid[d];
            id[d] = ((ii % (int)m->qgrid[d]) + (int)m->qgrid[d]) % (int)m->qgrid[d];
            w *= bit ? f[d] : 1.0 - f[d];
        }
        s.iqc[j] = (int)(((int64_t)id[0] * m->qgrid[1] + id[1]) * m->qgrid[2] + id[2]);
        s.w[j] = (float)w;
    }
    return s;
}

static int nn_features(lelph_nn* m, int S, const double* qpts)
{
    std::vector<NnSample> hs((size_t)S);
    for (int s = 0; s < S; ++s) hs[s] = nn_sample(m, qpts + 3 * s);
    HIP_TRY(hipMemcpyAsync(m->smp.p, hs.data(), sizeof(NnSample) * S, hipMemcpyHostToDevice,
                           m->stream));
    hipLaunchKernelGGL(k_nn_features, dim3((unsigned)cdiv(m->g.n, 128), (unsigned)S), dim3(128),
                       0, m->stream, m->coarse.as<c32>(), m->nf, m->g, m->smp.as<NnSample>(),
                       1.f / m->scale, m->X0.as<__bf16>(), m->base.as<float>());
    HIP_TRY(hipGetLastError());
    // the host arrays must outlive the copy
    HIP_TRY(hipStreamSynchronize(m->stream));
    return LELPH_OK;
}

static void nn_free(lelph_nn* m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    NBuf* all[] = {&m->coarse, &m->targets, &m->smp, &m->X0, &m->base, &m->tgt, &m->colT,
                   &m->dZf, &m->dZ, &m->dZT, &m->dA, &m->part, &m->lossp, &m->bpart};
    for (NBuf* b : all) b->release();
    for (int l = 0; l < NN_MAXL; ++l)
    {
        NBuf* per[] = {&m->W[l], &m->b[l], &m->mW[l], &m->vW[l], &m->mb[l], &m->vb[l],
                       &m->Wt[l], &m->Wk[l], &m->gW[l], &m->gb[l], &m->maskW[l], &m->maskb[l],
                       &m->A[l], &m->Z[l]};
        for (NBuf* b : per) b->release();
    }
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

} // namespace

extern "C" {

int lelph_nn_create(const int64_t grid[3], int nfields, int depth, uint64_t seed, lelph_nn** out)
{
    if (!grid || !out || nfields < 1 || 2 * nfields > NN_W || depth < 0 || depth + 2 > NN_MAXL)
        return LELPH_ERR_ARG;
    if (2 * nfields * NN_FEAT_NB + 2 * nfields + 6 > NN_CIN) return LELPH_ERR_UNSUPPORTED;
    if (grid[0] < 3 || grid[1] < 3 || grid[2] < 3) return LELPH_ERR_ARG;
    const int64_t n = grid[0] * grid[1] * grid[2];
    if (n % 64) return LELPH_ERR_UNSUPPORTED; // GEMM row tiles of 64 points
    *out = nullptr;
    lelph_nn* m = new (std::nothrow) lelph_nn();
    if (!m) return LELPH_ERR_ALLOC;
    auto fail = [&](int st) {
        nn_free(m);
        return st;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(LELPH_ERR_HIP);
    m->g = NnGrid{(int)grid[0], (int)grid[1], (int)grid[2], n};
    m->nf = nfields;
    m->nout = 2 * nfields;
    m->depth = depth;
    m->L = depth + 2;
    uint64_t rs = seed;
    for (int l = 0; l < m->L; ++l)
    {
        m->cin[l] = l == 0 ? NN_CIN : NN_W;
        m->K[l] = (int64_t)NN_TAPS * m->cin[l];
        const int64_t nw = m->K[l] * NN_W;
        // He-normal initial weights on the real fan-in; the output layer
        // starts at zero, so the untrained model is the trilinear blend
        const int fan_in = NN_TAPS * (l == 0 ? 2 * nfields * NN_FEAT_NB + 2 * nfields + 6 : NN_W);
        const double sd = std::sqrt(2.0 / (double)fan_in);
        std::vector<float> w((size_t)nw, 0.f);
        std::vector<unsigned char> mask((size_t)nw, 1), maskb(NN_W, 1);
        for (int64_t k = 0; k < m->K[l]; ++k)
            for (int c = 0; c < NN_W; ++c)
            {
                const int ci = (int)(k % m->cin[l]);
                const bool pad_in = l == 0 && ci >= 2 * nfields * NN_FEAT_NB + 2 * nfields + 6;
                const bool pad_out = l + 1 == m->L && c >= m->nout;
                if (pad_in || pad_out)
                {
                    mask[k * NN_W + c] = 0;
                    continue;
                }
                if (l + 1 == m->L) continue;
                const double u1 = nn_unif(rs), u2 = nn_unif(rs);
                w[k * NN_W + c] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) *
                                          std::cos(2.0 * M_PI * u2));
            }
        if (l + 1 == m->L)
            for (int c = m->nout; c < NN_W; ++c) maskb[c] = 0;
        if (m->W[l].ensure(sizeof(float) * nw) || m->mW[l].ensure(sizeof(float) * nw) ||
            m->vW[l].ensure(sizeof(float) * nw) || m->gW[l].ensure(sizeof(float) * nw) ||
            m->Wt[l].ensure(sizeof(__bf16) * nw) || m->Wk[l].ensure(sizeof(__bf16) * nw) ||
            m->maskW[l].ensure(nw) || m->b[l].ensure(sizeof(float) * NN_W) ||
            m->mb[l].ensure(sizeof(float) * NN_W) || m->vb[l].ensure(sizeof(float) * NN_W) ||
            m->gb[l].ensure(sizeof(float) * NN_W) || m->maskb[l].ensure(NN_W))
            return fail(LELPH_ERR_ALLOC);
        if (hipMemcpy(m->W[l].p, w.data(), sizeof(float) * nw, hipMemcpyHostToDevice) ||
            hipMemset(m->mW[l].p, 0, sizeof(float) * nw) ||
            hipMemset(m->vW[l].p, 0, sizeof(float) * nw) ||
            hipMemcpy(m->maskW[l].p, mask.data(), nw, hipMemcpyHostToDevice) ||
            hipMemset(m->b[l].p, 0, sizeof(float) * NN_W) ||
            hipMemset(m->mb[l].p, 0, sizeof(float) * NN_W) ||
            hipMemset(m->vb[l].p, 0, sizeof(float) * NN_W) ||
            hipMemcpy(m->maskb[l].p, maskb.data(), NN_W, hipMemcpyHostToDevice))
            return fail(LELPH_ERR_HIP);
    }
    int st = nn_pack_all(m);
    if (st == LELPH_OK && hipStreamSynchronize(m->stream) != hipSuccess) st = LELPH_ERR_HIP;
    if (st != LELPH_OK) return fail(st);
    *out = m;
    return LELPH_OK;
}

void lelph_nn_destroy(lelph_nn* m) { nn_free(m); }

int lelph_nn_nparams(lelph_nn* m, int64_t* n)
{
    if (!m || !n) return LELPH_ERR_ARG;
    int64_t a = 0;
    for (int l = 0; l < m->L; ++l) a += m->K[l] * NN_W + NN_W;
    *n = a;
    return LELPH_OK;
}

int lelph_nn_get_params(lelph_nn* m, float* out)
{
    if (!m || !out) return LELPH_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(m->stream));
    for (int l = 0; l < m->L; ++l)
    {
        HIP_TRY(hipMemcpy(out, m->W[l].p, sizeof(float) * m->K[l] * NN_W, hipMemcpyDeviceToHost));
        out += m->K[l] * NN_W;
        HIP_TRY(hipMemcpy(out, m->b[l].p, sizeof(float) * NN_W, hipMemcpyDeviceToHost));
        out += NN_W;
    }
    return LELPH_OK;
}

int lelph_nn_set_params(lelph_nn* m, const float* in)
{
    if (!m || !in) return LELPH_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(m->stream));
    for (int l = 0; l < m->L; ++l)
    {
        HIP_TRY(hipMemcpy(m->W[l].p, in, sizeof(float) * m->K[l] * NN_W, hipMemcpyHostToDevice));
        in += m->K[l] * NN_W;
        HIP_TRY(hipMemcpy(m->b[l].p, in, sizeof(float) * NN_W, hipMemcpyHostToDevice));
        in += NN_W;
    }
    LELPH_TRY(nn_pack_all(m));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return LELPH_OK;
}

int lelph_nn_set_coarse(lelph_nn* m, const int64_t qgrid[3], const double* fields, double scale)
{
    if (!m || !qgrid || !fields || !(scale > 0)) return LELPH_ERR_ARG;
    if (qgrid[0] < 1 || qgrid[1] < 1 || qgrid[2] < 1) return LELPH_ERR_ARG;
    const int64_t nqc = qgrid[0] * qgrid[1] * qgrid[2];
    const int64_t cnt = nqc * m->nf * m->g.n;
    std::vector<c32> h((size_t)cnt);
    for (int64_t i = 0; i < cnt; ++i) h[i] = {(float)fields[2 * i], (float)fields[2 * i + 1]};
    LELPH_TRY(m->coarse.ensure(sizeof(c32) * cnt));
    HIP_TRY(hipMemcpy(m->coarse.p, h.data(), sizeof(c32) * cnt, hipMemcpyHostToDevice));
    std::memcpy(m->qgrid, qgrid, sizeof(m->qgrid));
    m->nqc = (int)nqc;
    m->scale = (float)scale;
    return LELPH_OK;
}

int lelph_nn_loss_grad(lelph_nn* m, int S, const float* X0, const float* base,
                       const float* target, double* loss, float* grad)
{
    if (!m || S < 1 || !X0 || !base || !target || !loss) return LELPH_ERR_ARG;
    LELPH_TRY(nn_ensure_batch(m, S));
    const int64_t P = (int64_t)S * m->g.n;
    std::vector<__bf16> xb((size_t)(P * NN_CIN));
    for (int64_t i = 0; i < P * NN_CIN; ++i) xb[i] = (__bf16)X0[i];
    HIP_TRY(hipMemcpy(m->X0.p, xb.data(), sizeof(__bf16) * P * NN_CIN, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->base.p, base, sizeof(float) * P * m->nout, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->tgt.p, target, sizeof(float) * P * m->nout, hipMemcpyHostToDevice));
    LELPH_TRY(nn_forward(m, S));
    LELPH_TRY(nn_backward(m, S, loss));
    if (grad)
    {
        HIP_TRY(hipStreamSynchronize(m->stream));
        for (int l = 0; l < m->L; ++l)
        {
            HIP_TRY(hipMemcpy(grad, m->gW[l].p, sizeof(float) * m->K[l] * NN_W,
                              hipMemcpyDeviceToHost));
            grad += m->K[l] * NN_W;
            HIP_TRY(hipMemcpy(grad, m->gb[l].p, sizeof(float) * NN_W, hipMemcpyDeviceToHost));
            grad += NN_W;
        }
    }
    return LELPH_OK;
}

int lelph_nn_forward(lelph_nn* m, int S, const float* X0, const float* base, float* out)
{
    if (!m || S < 1 || !X0 || !base || !out) return LELPH_ERR_ARG;
    LELPH_TRY(nn_ensure_batch(m, S));
    const int64_t P = (int64_t)S * m->g.n;
    std::vector<__bf16> xb((size_t)(P * NN_CIN));
    for (int64_t i = 0; i < P * NN_CIN; ++i) xb[i] = (__bf16)X0[i];
    HIP_TRY(hipMemcpy(m->X0.p, xb.data(), sizeof(__bf16) * P * NN_CIN, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->base.p, base, sizeof(float) * P * m->nout, hipMemcpyHostToDevice));
    LELPH_TRY(nn_forward(m, S));
    hipLaunchKernelGGL(k_nn_out, dim3((unsigned)cdiv(P * m->nout, 256)), dim3(256), 0, m->stream,
                       m->Z[m->L - 1].as<float>(), m->base.as<float>(), P, m->nout, 1.f,
                       m->tgt.as<float>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, m->tgt.p, sizeof(float) * P * m->nout, hipMemcpyDeviceToHost,
                           m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return LELPH_OK;
}

int lelph_nn_features(lelph_nn* m, int S, const double* qpts, float* X0, float* base)
{
    if (!m || S < 1 || !qpts || !m->nqc) return LELPH_ERR_ARG;
    LELPH_TRY(nn_ensure_batch(m, S));
    LELPH_TRY(nn_features(m, S, qpts));
    const int64_t P = (int64_t)S * m->g.n;
    if (X0)
    {
        std::vector<__bf16> xb((size_t)(P * NN_CIN));
        HIP_TRY(hipMemcpy(xb.data(), m->X0.p, sizeof(__bf16) * P * NN_CIN, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < P * NN_CIN; ++i) X0[i] = (float)xb[i];
    }
    if (base)
        HIP_TRY(hipMemcpy(base, m->base.p, sizeof(float) * P * m->nout, hipMemcpyDeviceToHost));
    return LELPH_OK;
}

int lelph_nn_set_targets(lelph_nn* m, int64_t nq, const float* targets)
{
    if (!m || nq < 1 || !targets) return LELPH_ERR_ARG;
    const size_t bytes = sizeof(float) * (size_t)(nq * m->g.n * m->nout);
    LELPH_TRY(m->targets.ensure(bytes));
    std::vector<float> h((size_t)(nq * m->g.n * m->nout));
    const float inv = 1.f / m->scale;
    for (size_t i = 0; i < h.size(); ++i) h[i] = targets[i] * inv;
    HIP_TRY(hipMemcpy(m->targets.p, h.data(), bytes, hipMemcpyHostToDevice));
    m->ntarg = nq;
    return LELPH_OK;
}

int lelph_nn_train(lelph_nn* m, int64_t nq, const double* qpts, int epochs, int batch, double lr,
                   uint64_t seed, double* loss_hist)
{
    if (!m || nq < 1 || !qpts || epochs < 1 || batch < 1 || !(lr > 0)) return LELPH_ERR_ARG;
    if (nq != m->ntarg || !m->nqc) return LELPH_ERR_STATE;
    LELPH_TRY(nn_ensure_batch(m, batch));
    std::vector<int64_t> order((size_t)nq);
    for (int64_t i = 0; i < nq; ++i) order[i] = i;
    uint64_t rs = seed;
    std::vector<double> bq((size_t)(3 * batch));
    const size_t one = sizeof(float) * (size_t)(m->g.n * m->nout);
    for (int e = 0; e < epochs; ++e)
    {
        for (int64_t i = nq - 1; i > 0; --i) std::swap(order[i], order[nn_mix(rs) % (uint64_t)(i + 1)]);
        double acc = 0;
        int64_t nb = 0;
        for (int64_t s0 = 0; s0 + batch <= nq; s0 += batch)
        {
            for (int s = 0; s < batch; ++s)
            {
                const int64_t iq = order[s0 + s];
                std::memcpy(bq.data() + 3 * s, qpts + 3 * iq, 3 * sizeof(double));
                HIP_TRY(hipMemcpyAsync(m->tgt.as<char>() + one * s,
                                       m->targets.as<char>() + one * iq, one,
                                       hipMemcpyDeviceToDevice, m->stream));
            }
            LELPH_TRY(nn_features(m, batch, bq.data()));
            LELPH_TRY(nn_forward(m, batch));
            double l = 0;
            LELPH_TRY(nn_backward(m, batch, &l));
            LELPH_TRY(nn_adam_step(m, (float)lr));
            acc += l;
            ++nb;
        }
        if (loss_hist) loss_hist[e] = nb ? acc / nb : 0.0;
    }
    HIP_TRY(hipStreamSynchronize(m->stream));
    return LELPH_OK;
}

int lelph_nn_predict(lelph_nn* m, int64_t nq, const double* qpts, float* out)
{
    if (!m || nq < 1 || !qpts || !out || !m->nqc) return LELPH_ERR_ARG;
    const int S = (int)std::min<int64_t>(nq, std::max(1, m->smax > 0 ? m->smax : 8));
    LELPH_TRY(nn_ensure_batch(m, S));
    for (int64_t q0 = 0; q0 < nq; q0 += S)
    {
        const int s = (int)std::min<int64_t>(S, nq - q0);
        LELPH_TRY(nn_features(m, s, qpts + 3 * q0));
        LELPH_TRY(nn_forward(m, s));
        const int64_t P = (int64_t)s * m->g.n;
        hipLaunchKernelGGL(k_nn_out, dim3((unsigned)cdiv(P * m->nout, 256)), dim3(256), 0,
                           m->stream, m->Z[m->L - 1].as<float>(), m->base.as<float>(), P, m->nout,
                           m->scale, m->tgt.as<float>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out + q0 * m->g.n * m->nout, m->tgt.p,
                               sizeof(float) * P * m->nout, hipMemcpyDeviceToHost, m->stream));
        HIP_TRY(hipStreamSynchronize(m->stream));
    }
    return LELPH_OK;
}

} // extern "C"
