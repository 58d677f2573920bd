// nonloc.hpp -- Kleinman-Bylander part of g: add_elphNonLocal
// (src/nonloc/Vbare_nloc.c:35-577) and sum_VNL_KKp
// (src/nonloc/sum_VNL_KKp.c:8-58).
//
// Per side (K = S(k+q+G) and K' = S(k+G), cartesian without 2pi):
//   1. k_nl_geom   : K, real Y_lm(K) (FP64), the rotated coefficients
//                    (su2, translation phase e^{-i 2pi K.tau}).
//   2. k_nl_beta   : beta[a][p][x][G] = F_l(K) e^{-/+ i 2pi K.tau_a} Y_lm(K)
//                    {1, +/- sign K_x,y,z}  (FP64 phases and tables).
//   3. k_zgemm_nt  : <beta|psi> projector GEMM on MFMA (conjugation per side
//                    and time reversal, Vbare_nloc.c:397-418).
//   4. k_nl_kkp    : the 4-term K/K' sum per atom (sum_VNL_KKp).
//   5. k_nl_mode   : cart -> mode with prefactor -2 pi i e^2 and accumulate
//                    onto g (Vbare_nloc.c:541-558).
#pragma once

#include "common.hpp"
#include "contract.hpp"

__device__ inline double d_fact(int n)
{
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return n < 0 ? 0.0 : f;
}

// legendre (src/common/numerical_func.c:29-94) for m >= 0, as Ylm uses it
__device__ inline double d_legendre(int l, int m, double x)
{
    if (l == 0 && m == 0) return 1.0;
    if (m > l) return 0.0;
    double p = 1.0;
    const double s = sqrt(1 - x * x);
    for (int i = 1; i <= m; ++i) p *= (2 * i - 1) * s;
    if (l == m) return p;
    double pmm = p;
    p *= (2 * m + 1) * x;
    if (l == m + 1) return p;
    for (int i = m + 2; i <= l; ++i)
    {
        double t = p;
        p = (2 * i - 1) * x * p - (i + m - 1) * pmm;
        p = p / (i - m);
        pmm = t;
    }
    return p;
}

// Ylm (src/common/numerical_func.c:97-152): real harmonics, Blanco et al.
__device__ inline double d_Ylm(int l, int m, const double* v)
{
    const double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nrm < LELPH_EPS6) return l != 0 ? 0.0 : 1.0 / sqrt(4 * LELPH_RPI);
    const double cost = v[2] / nrm;
    const double phi = atan2(v[1], v[0]);
    double sh;
    if (m > 0)
        sh = sqrt(2.0) * cos(m * phi);
    else if (m < 0)
    {
        m = -m;
        sh = sqrt(2.0) * sin(m * phi);
    }
    else
        sh = 1;
    sh *= d_legendre(l, m, cost);
    sh *= sqrt(d_fact(l - m) * (2 * l + 1) / (4 * LELPH_RPI * d_fact(l + m)));
    return sh;
}

struct NlSide
{
    double S[9];   // cartesian symmetry
    double k[3];   // iBZ k, cartesian no 2pi
    double tau[3]; // fractional translation (cart)
    int conj_w;    // conjugate the wavefunction in the projector GEMM
    z64 su2[4];
};

// K, Y_lm(K) (nlm rows), rotated coefficients wr[(s,b)][sigma][G]
// blockIdx.y: sets spy * y .. spy * y + spy - 1 of the rotation (K and Y_lm
// written by y = 0).  TW: the rotated coefficients rounded once to the
// precision of the projector GEMM that reads them.
template <typename TW>
__global__ void k_nl_geom(const double* __restrict__ gvec, int64_t npw, NlSide sd, int lmax,
                          const z64* __restrict__ wfc, int64_t nsets, int nspinor,
                          double* __restrict__ Kc, double* __restrict__ Y,
                          cpx<TW>* __restrict__ wr, int spy)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npw) return;
    const double t[3] = {sd.k[0] + gvec[3 * g], sd.k[1] + gvec[3 * g + 1],
                         sd.k[2] + gvec[3 * g + 2]};
    double K[3];
    for (int i = 0; i < 3; ++i) K[i] = sd.S[3 * i] * t[0] + sd.S[3 * i + 1] * t[1] + sd.S[3 * i + 2] * t[2];
    if (blockIdx.y == 0)
    {
        Kc[3 * g] = K[0];
        Kc[3 * g + 1] = K[1];
        Kc[3 * g + 2] = K[2];
        for (int l = 0; l <= lmax; ++l)
            for (int im = 0; im <= 2 * l; ++im)
                Y[(int64_t)(l * l + im) * npw + g] = d_Ylm(l, im - l, K);
    }
    double s, c;
    sincos(-2.0 * LELPH_RPI * (K[0] * sd.tau[0] + K[1] * sd.tau[1] + K[2] * sd.tau[2]), &s, &c);
    const z64 ph = {c, s};
    const int64_t s0 = (int64_t)blockIdx.y * spy;
    const int64_t s1 = s0 + spy < nsets ? s0 + spy : nsets;
    for (int64_t st = s0; st < s1; ++st)
    {
        const z64* w = wfc + st * nspinor * npw;
        cpx<TW>* o = wr + st * nspinor * npw;
        if (nspinor == 1)
        {
            const z64 v = cmul(w[g], ph);
            o[g] = {(TW)v.x, (TW)v.y};
        }
        else
        {
            z64 u = w[g], d = w[npw + g];
            const z64 a = cmul(cadd(cmul(u, sd.su2[0]), cmul(d, sd.su2[1])), ph);
            const z64 b = cmul(cadd(cmul(u, sd.su2[2]), cmul(d, sd.su2[3])), ph);
            o[g] = {(TW)a.x, (TW)a.y};
            o[npw + g] = {(TW)b.x, (TW)b.y};
        }
    }
}

// <beta_K'|psi> of a (k point, symmetry) key from its <beta_K|psi>: the
// K' projectors are the complex conjugates of the K ones (phase
// e^{+2 pi i K.tau} for e^{-2 pi i K.tau}), the three K-weighted rows with
// the opposite sign, and psi enters the other side's product conjugated
// the other way ('T' for 'C'), so bbKp = diag(1, -1, -1, -1) conj(bbK) row
// by row of four; exact (sign flips and conjugation only).
__global__ void k_nl_mirror(const z64* __restrict__ src, int64_t rows, int64_t ncol,
                            z64* __restrict__ dst)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * ncol) return;
    const int64_t r = t / ncol;
    const z64 v = src[t];
    dst[t] = (r % 4 == 0) ? z64{v.x, -v.y} : z64{-v.x, v.y};
}

// projector list of an atom: entry p -> (lidx, l, m index)
struct NlProj
{
    int natom, pmax, ntype;
};

// beta[(a*pmax + p)*4 + x][G]; side = -1 for K (k+q), +1 for K' (k)
template <typename T>
__global__ void k_nl_beta(const double* __restrict__ Kc, const double* __restrict__ Y,
                          const double* __restrict__ Fk, int64_t npw, NlProj pj,
                          const int* __restrict__ plist, // (natom, pmax, 3): lidx,l,lm or -1
                          const int* __restrict__ atype, const double* __restrict__ apos,
                          const double* __restrict__ Fsign, int side, cpx<T>* __restrict__ beta)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ap = blockIdx.y; // a*pmax + p
    if (g >= npw) return;
    const int a = ap / pj.pmax;
    const int* e = plist + 3 * ap;
    cpx<T>* b = beta + (int64_t)ap * 4 * npw;
    if (e[0] < 0)
    {
        for (int x = 0; x < 4; ++x) b[(int64_t)x * npw + g] = {0, 0};
        return;
    }
    const int lidx = e[0], lm = e[2];
    const int t = atype[a];
    const double* K = Kc + 3 * g;
    const double* tau = apos + 3 * a;
    const double td = K[0] * tau[0] + K[1] * tau[1] + K[2] * tau[2];
    double s, c;
    sincos(side * 2.0 * LELPH_RPI * td, &s, &c);
    const double f = Fk[((int64_t)lidx * pj.ntype + t) * npw + g] * Y[(int64_t)lm * npw + g];
    const z64 b0 = {f * c, f * s};
    const double sg = Fsign[lidx * pj.ntype + t] * (side < 0 ? 1.0 : -1.0);
    b[g] = {(T)b0.x, (T)b0.y};
    for (int x = 0; x < 3; ++x)
    {
        const double kx = K[x] * sg;
        b[(int64_t)(x + 1) * npw + g] = {(T)(b0.x * kx), (T)(b0.y * kx)};
    }
}

// C[M][Nc] (+)= A[M][K] op(B[Nc][K])^T, complex, split over K into partials
// part[split][M][Nc].  op = conj when conjB.
#define GZ_MB 64
#define GZ_NB 64
template <typename T>
__global__ __launch_bounds__(256) void k_zgemm_nt(const cpx<T>* __restrict__ A,
                                                  const cpx<T>* __restrict__ B, int M, int Nc,
                                                  int64_t K, int conjB, int64_t chunks_per_split,
                                                  z64* __restrict__ part)
{
    typedef typename Mfma<T>::acc acc_t;
    __shared__ cpx<T> sA[GZ_MB][CT_LD];
    __shared__ cpx<T> sB[GZ_NB][CT_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntn = (Nc + GZ_NB - 1) / GZ_NB;
    const int tm = blockIdx.x / ntn, tn = blockIdx.x % ntn;
    const int m0 = tm * GZ_MB, n0 = tn * GZ_NB;
    const int split = blockIdx.y;
    const int64_t nchunks = (K + CT_KC - 1) / CT_KC;
    const int64_t c0 = (int64_t)split * chunks_per_split;
    const int64_t c1 = (c0 + chunks_per_split < nchunks) ? c0 + chunks_per_split : nchunks;
    const int wr_ = wave >> 1, wc = wave & 1;
    acc_t accr[2][2], acci[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v)
            {
                accr[a][b][v] = 0;
                acci[a][b][v] = 0;
            }
    const T sgn = conjB ? (T)-1 : (T)1;
    for (int64_t c = c0; c < c1; ++c)
    {
        const int64_t k0 = c * CT_KC;
        for (int f = tid; f < GZ_MB * CT_KC; f += 256)
        {
            const int mm = f / CT_KC, kk = f - mm * CT_KC;
            const int64_t k = k0 + kk;
            cpx<T> v = {0, 0};
            if (m0 + mm < M && k < K) v = A[(int64_t)(m0 + mm) * K + k];
            sA[mm][kk] = v;
        }
        for (int f = tid; f < GZ_NB * CT_KC; f += 256)
        {
            const int nn = f / CT_KC, kk = f - nn * CT_KC;
            const int64_t k = k0 + kk;
            cpx<T> v = {0, 0};
            if (n0 + nn < Nc && k < K) v = B[(int64_t)(n0 + nn) * K + k];
            v.y *= sgn;
            sB[nn][kk] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CT_KC / 4; ++kk)
        {
            const int i = lane & 15, kq = kk * 4 + (lane >> 4);
            cpx<T> a[2], b[2];
            a[0] = sA[wr_ * 32 + i][kq];
            a[1] = sA[wr_ * 32 + 16 + i][kq];
            b[0] = sB[wc * 32 + i][kq];
            b[1] = sB[wc * 32 + 16 + i][kq];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb)
                {
                    // C += a b: re = ar br - ai bi, im = ar bi + ai br
                    accr[rb][cb] = Mfma<T>::op(a[rb].x, b[cb].x, accr[rb][cb]);
                    accr[rb][cb] = Mfma<T>::op(-a[rb].y, b[cb].y, accr[rb][cb]);
                    acci[rb][cb] = Mfma<T>::op(a[rb].x, b[cb].y, acci[rb][cb]);
                    acci[rb][cb] = Mfma<T>::op(a[rb].y, b[cb].x, acci[rb][cb]);
                }
        }
        __syncthreads();
    }
    for (int rb = 0; rb < 2; ++rb)
        for (int cb = 0; cb < 2; ++cb)
            for (int e = 0; e < 4; ++e)
            {
                const int row = m0 + wr_ * 32 + rb * 16 + Mfma<T>::row(lane, e);
                const int col = n0 + wc * 32 + cb * 16 + (lane & 15);
                if (row < M && col < Nc)
                    part[((int64_t)split * M + row) * Nc + col] = {(double)accr[rb][cb][e],
                                                                   (double)acci[rb][cb][e]};
            }
}

// Reduced-precision projector GEMMs (the FP32 and BF16-split forms of the
// <beta|psi> products, src/nonloc/Vbare_nloc.c:409-418).  beta (phases and
// radial / Y_lm tables in FP64) and the rotated wavefunctions are split into
// NP bf16 pieces per real component, x = x0 + x1 (+ x2), each the
// round-to-nearest of the remaining FP32 residual, stored as planes
// P[piece][re/im][row][k] with k contiguous and zero padded to Kp (a
// multiple of 32) and rows to Rp (a multiple of 64); negim flips the sign of
// the imaginary planes (the conjugate operand).
template <int NP>
__global__ void k_nl_presplit(const z64* __restrict__ X, int R, int64_t K, int Rp, int64_t Kp,
                              int negim, __bf16* __restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (k >= Kp || r >= Rp) return;
    z64 v = {0, 0};
    if (r < R && k < K) v = X[(int64_t)r * K + k];
    float re = (float)v.x, im = (float)(negim ? -v.y : v.y);
    const int64_t plane = (int64_t)Rp * Kp;
#pragma unroll
    for (int p = 0; p < NP; ++p)
    {
        const __bf16 a = (__bf16)re, b = (__bf16)im;
        out[(2 * p) * plane + (int64_t)r * Kp + k] = a;
        out[(2 * p + 1) * plane + (int64_t)r * Kp + k] = b;
        re -= (float)a;
        im -= (float)b;
    }
}

typedef __bf16 nl_bf16x8 __attribute__((ext_vector_type(8)));
typedef float nl_flt4 __attribute__((ext_vector_type(4)));

__device__ inline nl_bf16x8 nl_neg(nl_bf16x8 x)
{
    typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
    u16x8 u = __builtin_bit_cast(u16x8, x);
    u ^= (unsigned short)0x8000;
    return __builtin_bit_cast(nl_bf16x8, u);
}

// C[M][N] (partial of split s) = sum_k A[m][k] B[n][k] over the split's k
// range, complex, from the piece planes of k_nl_presplit (B's imaginary
// planes already carry the conjugation sign): Cr = Ar Br - Ai Bi, Ci = Ar Bi
// + Ai Br, each real product as the piece products p1 + p2 < NP (NPROD 3:
// p1 + p2 <= 1) on v_mfma_f32_16x16x32_bf16, FP32 accumulation within the
// split, partials in FP64.  Block 64 x 64, 4 waves of 32 x 32; fragments are
// read straight from global memory (16 bytes per lane: A lane l holds row
// l % 16, k 8 (l / 16) .. + 7; B the same for column l % 16).
template <int NP, int NPROD>
__global__ __launch_bounds__(256) void k_zgemm_nt_bf(const __bf16* __restrict__ Ap,
                                                     const __bf16* __restrict__ Bp, int M,
                                                     int Nc, int Mp, int Np, int64_t Kp,
                                                     int64_t chunks_per_split,
                                                     z64* __restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntn = Np / 64;
    const int tm = blockIdx.x / ntn, tn = blockIdx.x % ntn;
    const int m0 = tm * 64 + (wave >> 1) * 32, n0 = tn * 64 + (wave & 1) * 32;
    const int split = blockIdx.y;
    const int64_t nch = Kp / 32;
    const int64_t c0 = (int64_t)split * chunks_per_split;
    const int64_t c1 = (c0 + chunks_per_split < nch) ? c0 + chunks_per_split : nch;
    const int64_t pa = (int64_t)Mp * Kp, pb = (int64_t)Np * Kp;
    nl_flt4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
        {
            accr[a][b] = (nl_flt4){0.f, 0.f, 0.f, 0.f};
            acci[a][b] = (nl_flt4){0.f, 0.f, 0.f, 0.f};
        }
    const int r = lane & 15, kq = 8 * (lane >> 4);
    for (int64_t c = c0; c < c1; ++c)
    {
        const int64_t k = c * 32 + kq;
        nl_bf16x8 ar[NP][2], ai[NP][2], nai[NP][2], br[NP][2], bi[NP][2];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int h = 0; h < 2; ++h)
            {
                const int64_t ra = (int64_t)(m0 + 16 * h + r) * Kp + k;
                const int64_t rb = (int64_t)(n0 + 16 * h + r) * Kp + k;
                ar[p][h] = *reinterpret_cast<const nl_bf16x8*>(Ap + (2 * p) * pa + ra);
                ai[p][h] = *reinterpret_cast<const nl_bf16x8*>(Ap + (2 * p + 1) * pa + ra);
                nai[p][h] = nl_neg(ai[p][h]);
                br[p][h] = *reinterpret_cast<const nl_bf16x8*>(Bp + (2 * p) * pb + rb);
                bi[p][h] = *reinterpret_cast<const nl_bf16x8*>(Bp + (2 * p + 1) * pb + rb);
            }
#pragma unroll
        for (int p1 = 0; p1 < NP; ++p1)
#pragma unroll
            for (int p2 = 0; p2 < NP; ++p2)
            {
                if (p1 + p2 >= NP) continue;
                if (NPROD == 3 && p1 + p2 > 1) continue;
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                    {
                        accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[p1][a], br[p2][b],
                                                                            accr[a][b], 0, 0, 0);
                        accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nai[p1][a], bi[p2][b],
                                                                            accr[a][b], 0, 0, 0);
                        acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[p1][a], bi[p2][b],
                                                                            acci[a][b], 0, 0, 0);
                        acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[p1][a], br[p2][b],
                                                                            acci[a][b], 0, 0, 0);
                    }
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int row = m0 + 16 * a + 4 * (lane >> 4) + v;
                const int col = n0 + 16 * b + r;
                if (row < M && col < Nc)
                    part[((int64_t)split * M + row) * Nc + col] = {(double)accr[a][b][v],
                                                                   (double)acci[a][b][v]};
            }
}

// ebuf[a][x][s][n][m] = sum f (K3_x[m] Kp0[n] + K0[m] Kp3_x[n])
// bbK/bbKp rows (a*pmax+p)*4+x, columns (s*nb + b)*nspinor + sigma.
__global__ void k_nl_kkp(const z64* __restrict__ bbK, const z64* __restrict__ bbKp,
                         NlProj pj, const int* __restrict__ plist, const int* __restrict__ atype,
                         const double* __restrict__ socfac, // (nltimesj, ntype)
                         const z64* __restrict__ fco,       // packed f coefficients or null
                         const int64_t* __restrict__ fco_off, // (nltimesj*ntype) offsets or -1
                         int nspin, int nb, int nspinor, z64* __restrict__ ebuf)
{
    const int64_t per_a = (int64_t)3 * nspin * nb * nb;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)pj.natom * per_a) return;
    const int a = (int)(t / per_a);
    int64_t r = t - a * per_a;
    const int x = (int)(r / ((int64_t)nspin * nb * nb));
    r -= (int64_t)x * nspin * nb * nb;
    const int s = (int)(r / ((int64_t)nb * nb));
    r -= (int64_t)s * nb * nb;
    const int n = (int)(r / nb), m = (int)(r % nb);
    const int ty = atype[a];
    const int64_t ncol = (int64_t)nspin * nb * nspinor;
    z64 acc = {0, 0};
    for (int p1 = 0; p1 < pj.pmax; ++p1)
    {
        const int* e1 = plist + 3 * (a * pj.pmax + p1);
        if (e1[0] < 0) continue;
        const z64* K0 = bbK + ((int64_t)(a * pj.pmax + p1) * 4 + 0) * ncol;
        const z64* K3 = bbK + ((int64_t)(a * pj.pmax + p1) * 4 + 1 + x) * ncol;
        for (int p2 = 0; p2 < pj.pmax; ++p2)
        {
            const int* e2 = plist + 3 * (a * pj.pmax + p2);
            if (e2[0] != e1[0]) continue;
            const int lidx = e1[0], l = e1[1];
            const int64_t off = fco_off ? fco_off[lidx * pj.ntype + ty] : -1;
            if (off < 0 && p1 != p2) continue;
            const z64* Kp0 = bbKp + ((int64_t)(a * pj.pmax + p2) * 4 + 0) * ncol;
            const z64* Kp3 = bbKp + ((int64_t)(a * pj.pmax + p2) * 4 + 1 + x) * ncol;
            const int m1 = e1[2] - l * l, m2 = e2[2] - l * l;
            for (int s1 = 0; s1 < nspinor; ++s1)
                for (int s2 = 0; s2 < nspinor; ++s2)
                {
                    z64 f;
                    if (off >= 0)
                        f = fco[off + ((int64_t)(2 * l + 1) * m1 + m2) * nspinor * nspinor +
                                nspinor * s1 + s2];
                    else
                        f = {s1 == s2 ? socfac[lidx * pj.ntype + ty] : 0.0, 0.0};
                    if (f.x == 0.0 && f.y == 0.0) continue;
                    const int64_t cm = ((int64_t)s * nb + m) * nspinor + s1;
                    const int64_t cn = ((int64_t)s * nb + n) * nspinor + s2;
                    z64 term = cadd(cmul(K3[cm], Kp0[cn]), cmul(K0[cm], Kp3[cn]));
                    acc = cadd(acc, cmul(f, term));
                }
        }
    }
    ebuf[t] = acc;
}

// g[v][s][n][m] += pre * sum_j eig[v][j] ebuf[j][s][n][m], j = (a, x)
__global__ void k_nl_mode(const z64* __restrict__ ebuf, const z64* __restrict__ eig, int nm,
                          int n3a, int64_t es, z64 pre, z64* __restrict__ g)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nm * es) return;
    const int v = (int)(t / es);
    const int64_t i = t - (int64_t)v * es;
    z64 acc = {0, 0};
    for (int j = 0; j < n3a; ++j) acc = cadd(acc, cmul(eig[(int64_t)v * n3a + j], ebuf[(int64_t)j * es + i]));
    g[t] = cadd(g[t], cmul(pre, acc));
}
