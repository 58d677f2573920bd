// fft.hpp -- batched mixed-radix 3-D FFT for gfx950.
//
// The box layout is the reference's (Nx, Ny, Nz) row major with z fastest
// (src/fft/bwdfft.c:136-219).  A 3-D transform is three passes, one per
// axis.  Each pass is one kernel: a workgroup stages `tl` lines of one axis
// through LDS with coalesced loads (lines adjacent in memory are taken
// together so every row of the tile is one contiguous read), runs the
// Stockham autosort stages in LDS (radix 2/3/4/5 butterflies, any other
// prime by direct DFT) and writes the lines back.  Twiddles are a table of
// e^{+2 pi i t/n} computed on the host in FP64 (rounded to FP32 for the
// FP32 path).  Sign and normalisation follow FFTW: BACKWARD (+1) is e^{+i},
// FORWARD (-1) is e^{-i}, both unnormalised.
#pragma once

#include <cmath>
#include <vector>

#include "common.hpp"

#define FFT_MAX_STAGES 16
#define FFT_MAX_RADIX 16

struct FftStages
{
    int nst;
    int radix[FFT_MAX_STAGES];
};

struct FftAxis
{
    int n = 0;
    FftStages st{};
    z64* w64 = nullptr; // device (n) e^{+2 pi i t / n}
    c32* w32 = nullptr;
};

static inline int fft_factor(int n, FftStages* st)
{
    st->nst = 0;
    int m = n;
    while (m % 4 == 0 && st->nst < FFT_MAX_STAGES)
    {
        st->radix[st->nst++] = 4;
        m /= 4;
    }
    for (int p = 2; m > 1 && st->nst < FFT_MAX_STAGES;)
    {
        if (m % p == 0)
        {
            if (p > FFT_MAX_RADIX) return LELPH_ERR_UNSUPPORTED;
            st->radix[st->nst++] = p;
            m /= p;
        }
        else
            ++p;
    }
    return m == 1 ? LELPH_OK : LELPH_ERR_UNSUPPORTED;
}

static inline int fft_axis_init(FftAxis* ax, int n)
{
    ax->n = n;
    LELPH_TRY(fft_factor(n, &ax->st));
    std::vector<z64> w(n);
    std::vector<c32> wf(n);
    for (int t = 0; t < n; ++t)
    {
        double a = 2.0 * M_PI * (double)t / (double)n;
        w[t] = {cos(a), sin(a)};
        wf[t] = {(float)w[t].x, (float)w[t].y};
    }
    HIP_TRY(hipMalloc(&ax->w64, sizeof(z64) * n));
    HIP_TRY(hipMalloc(&ax->w32, sizeof(c32) * n));
    HIP_TRY(hipMemcpy(ax->w64, w.data(), sizeof(z64) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ax->w32, wf.data(), sizeof(c32) * n, hipMemcpyHostToDevice));
    return LELPH_OK;
}

static inline void fft_axis_free(FftAxis* ax)
{
    if (ax->w64) (void)hipFree(ax->w64);
    if (ax->w32) (void)hipFree(ax->w32);
    ax->w64 = nullptr;
    ax->w32 = nullptr;
}

// twiddle w^(SIGN): the table holds e^{+i}; forward uses the conjugate.
template <typename T, int SIGN>
__device__ inline cpx<T> tw(const cpx<T>* W, int t)
{
    cpx<T> w = W[t];
    if (SIGN < 0) w.y = -w.y;
    return w;
}

// multiply by SIGN * i
template <typename T, int SIGN>
__device__ inline cpx<T> mul_si(cpx<T> a)
{
    return SIGN > 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

template <typename T, int SIGN>
__device__ inline void dft_small(cpx<T>* v, int R, const cpx<T>* W, int n)
{
    if (R == 2)
    {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
    else if (R == 3)
    {
        const T c = (T)-0.5, s = (T)0.86602540378443864676;
        cpx<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        cpx<T> m = {v[0].x + c * t1.x, v[0].y + c * t1.y};
        cpx<T> is = mul_si<T, SIGN>(cpx<T>{s * t2.x, s * t2.y});
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, is);
        v[2] = csub(m, is);
    }
    else if (R == 4)
    {
        cpx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx<T> c = cadd(v[1], v[3]), d = mul_si<T, SIGN>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
    else if (R == 5)
    {
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        cpx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        cpx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        cpx<T> m1 = {v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
        cpx<T> m2 = {v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
        cpx<T> s_1 = mul_si<T, SIGN>(cpx<T>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
        cpx<T> s_2 = mul_si<T, SIGN>(cpx<T>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(m1, s_1);
        v[4] = csub(m1, s_1);
        v[2] = cadd(m2, s_2);
        v[3] = csub(m2, s_2);
    }
}

// One Stockham stage of radix R on nl lines of n points held in LDS
// (Govindaraju et al., SC'08, "FftIteration"): butterfly j reads
// in[j + r n/R], twiddles by w^(r (j mod Ns) n/(Ns R)), and writes
// out[(j/Ns) Ns R + (j mod Ns) + r Ns].
template <typename T, int SIGN, int R>
__device__ inline void fft_stage(const cpx<T>* in, cpx<T>* out, const cpx<T>* W, int n,
                                 int nl, int Ns, int tid, int bd)
{
    const int nbf = n / R;
    const int tws = n / (Ns * R);
    for (int w = tid; w < nl * nbf; w += bd)
    {
        const int t = w / nbf, j = w - t * nbf;
        const int k = j % Ns;
        cpx<T> v[R];
        const cpx<T>* src = in + t * n + j;
        v[0] = src[0];
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(src[r * nbf], tw<T, SIGN>(W, r * k * tws));
        dft_small<T, SIGN>(v, R, W, n);
        cpx<T>* dst = out + t * n + (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
    }
}

template <typename T, int SIGN>
__device__ inline void fft_stage_any(const cpx<T>* in, cpx<T>* out, const cpx<T>* W, int n,
                                     int nl, int Ns, int R, int tid, int bd)
{
    const int nbf = n / R;
    const int tws = n / (Ns * R);
    const int step = n / R;
    for (int w = tid; w < nl * nbf; w += bd)
    {
        const int t = w / nbf, j = w - t * nbf;
        const int k = j % Ns;
        const cpx<T>* src = in + t * n + j;
        cpx<T>* dst = out + t * n + (j / Ns) * Ns * R + k;
        for (int s = 0; s < R; ++s)
        {
            cpx<T> acc = {0, 0};
            for (int r = 0; r < R; ++r)
            {
                cpx<T> x = src[r * nbf];
                if (r) x = cmul(x, tw<T, SIGN>(W, r * k * tws));
                acc = cadd(acc, cmul(x, tw<T, SIGN>(W, ((r * s) % R) * step)));
            }
            dst[s * Ns] = acc;
        }
    }
}

// One axis pass over a batch of lines.  Line (o, i) element j lives at
//   data[o * ostride + i * lstride + j * estride],  i in [0, n_inner).
// A block takes tl consecutive i of one o.
template <typename T, int SIGN>
__global__ __launch_bounds__(256) void k_fft_lines(cpx<T>* __restrict__ data, int n,
                                                   int64_t n_inner, int64_t ostride,
                                                   int64_t lstride, int64_t estride,
                                                   int tl, FftStages st,
                                                   const cpx<T>* __restrict__ Wg)
{
    extern __shared__ char fft_smem[];
    cpx<T>* W = reinterpret_cast<cpx<T>*>(fft_smem);
    cpx<T>* b0 = W + n;
    cpx<T>* b1 = b0 + (int64_t)tl * n;
    const int tid = threadIdx.x, bd = blockDim.x;
    const int64_t tiles = (n_inner + tl - 1) / tl;
    const int64_t o = blockIdx.x / tiles;
    const int64_t i0 = (blockIdx.x % tiles) * tl;
    const int nl = (int)((n_inner - i0) < tl ? (n_inner - i0) : tl);
    cpx<T>* base = data + o * ostride + i0 * lstride;

    for (int t = tid; t < n; t += bd) W[t] = Wg[t];
    const int tot = nl * n;
    if (lstride == 1)
    {
        for (int f = tid; f < tot; f += bd)
        {
            int j = f / nl, t = f - j * nl;
            b0[t * n + j] = base[(int64_t)j * estride + t];
        }
    }
    else
    {
        for (int f = tid; f < tot; f += bd)
        {
            int t = f / n, j = f - t * n;
            b0[t * n + j] = base[(int64_t)t * lstride + (int64_t)j * estride];
        }
    }
    __syncthreads();

    cpx<T>* in = b0;
    cpx<T>* out = b1;
    int Ns = 1;
    for (int s = 0; s < st.nst; ++s)
    {
        const int R = st.radix[s];
        switch (R)
        {
            case 2: fft_stage<T, SIGN, 2>(in, out, W, n, nl, Ns, tid, bd); break;
            case 3: fft_stage<T, SIGN, 3>(in, out, W, n, nl, Ns, tid, bd); break;
            case 4: fft_stage<T, SIGN, 4>(in, out, W, n, nl, Ns, tid, bd); break;
            case 5: fft_stage<T, SIGN, 5>(in, out, W, n, nl, Ns, tid, bd); break;
            default: fft_stage_any<T, SIGN>(in, out, W, n, nl, Ns, R, tid, bd); break;
        }
        __syncthreads();
        cpx<T>* tmp = in;
        in = out;
        out = tmp;
        Ns *= R;
    }

    if (lstride == 1)
    {
        for (int f = tid; f < tot; f += bd)
        {
            int j = f / nl, t = f - j * nl;
            base[(int64_t)j * estride + t] = in[t * n + j];
        }
    }
    else
    {
        for (int f = tid; f < tot; f += bd)
        {
            int t = f / n, j = f - t * n;
            base[(int64_t)t * lstride + (int64_t)j * estride] = in[t * n + j];
        }
    }
}

// ---------------------------------------------------------------------------
// Compile-time-size variant: line length N, lines per block TL and every
// stage's radix are constants, so all index arithmetic reduces to shifts and
// multiplications.  Same factorisation rule as fft_factor (4s first, then
// primes in increasing order).
constexpr int fft_radix_c(int n, int s)
{
    int m = n, k = 0;
    while (m % 4 == 0)
    {
        if (k == s) return 4;
        m /= 4;
        ++k;
    }
    for (int p = 2; m > 1;)
    {
        if (m % p == 0)
        {
            if (k == s) return p;
            m /= p;
            ++k;
        }
        else
            ++p;
    }
    return 0;
}
constexpr int fft_ns_c(int n, int s) // product of the radices before stage s
{
    int r = 1;
    for (int i = 0; i < s; ++i) r *= fft_radix_c(n, i);
    return r;
}

// Lines sit LS elements apart in LDS (LS = N + 1 keeps the transposing
// loads and stores of k_fft_lines_c free of bank conflicts).  LB: the stage
// barriers order LDS only (global loads in flight stay outstanding).
template <typename T, int SIGN, int N, int TL, int S, int LS = N, bool LB = false>
struct FftStagesC
{
    static __device__ inline void run(cpx<T>*& in, cpx<T>*& out, const cpx<T>* W, int tid)
    {
        constexpr int R = fft_radix_c(N, S);
        if constexpr (R != 0)
        {
            constexpr int Ns = fft_ns_c(N, S);
            constexpr int NBF = N / R;
            constexpr int TWS = N / (Ns * R);
            for (int w = tid; w < TL * NBF; w += 256)
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
            if constexpr (LB)
                lds_barrier();
            else
                __syncthreads();
            cpx<T>* tmp = in;
            in = out;
            out = tmp;
            FftStagesC<T, SIGN, N, TL, S + 1, LS, LB>::run(in, out, W, tid);
        }
    }
};

// One axis pass, compile-time N and TL (see k_fft_lines for the addressing).
// Lines i >= n_inner of the last tile are computed on zeros and not stored.
// In LDS the lines are N + 1 elements apart: where adjacent lines are
// adjacent in memory (lstride 1), consecutive threads move one element of
// consecutive lines, which a stride of N (a multiple of the bank count for
// N = 96) would put on one bank.
// Pruning (both optional): blocks whose outer index o has ocnt[o % omod] == 0
// do nothing (their lines are all zero and are never read); elements j with
// emask[j] == 0 are taken as zero without being read.
template <typename T, int SIGN, int N, int TL>
__global__ __launch_bounds__(256) void k_fft_lines_c(cpx<T>* __restrict__ data, int64_t n_inner,
                                                     int64_t ostride, int64_t lstride,
                                                     int64_t estride,
                                                     const cpx<T>* __restrict__ Wg,
                                                     const int* __restrict__ ocnt, int omod,
                                                     const int* __restrict__ emask)
{
    constexpr int LS = N + 1;
    __shared__ cpx<T> W[N];
    __shared__ cpx<T> b0[TL * LS];
    __shared__ cpx<T> b1[TL * LS];
    const int tid = threadIdx.x;
    const int64_t tiles = (n_inner + TL - 1) / TL;
    const int64_t o = blockIdx.x / tiles;
    if (ocnt && ocnt[o % omod] == 0) return;
    const int64_t i0 = (blockIdx.x - o * tiles) * TL;
    const int nl = (int)((n_inner - i0) < TL ? (n_inner - i0) : TL);
    cpx<T>* base = data + o * ostride + i0 * lstride;

    for (int t = tid; t < N; t += 256) W[t] = Wg[t];
    if (lstride == 1)
    {
        for (int f = tid; f < TL * N; f += 256)
        {
            const int j = f / TL, t = f - j * TL;
            b0[t * LS + j] = (t < nl && (!emask || emask[j])) ? base[(int64_t)j * estride + t]
                                                              : cpx<T>{0, 0};
        }
    }
    else
    {
        for (int f = tid; f < TL * N; f += 256)
        {
            const int t = f / N, j = f - t * N;
            b0[t * LS + j] = (t < nl && (!emask || emask[j]))
                                 ? base[(int64_t)t * lstride + (int64_t)j * estride]
                                 : cpx<T>{0, 0};
        }
    }
    __syncthreads();
    cpx<T>* in = b0;
    cpx<T>* out = b1;
    FftStagesC<T, SIGN, N, TL, 0, LS>::run(in, out, W, tid);
    if (lstride == 1)
    {
        for (int f = tid; f < TL * N; f += 256)
        {
            const int j = f / TL, t = f - j * TL;
            if (t < nl) base[(int64_t)j * estride + t] = in[t * LS + j];
        }
    }
    else
    {
        for (int f = tid; f < TL * N; f += 256)
        {
            const int t = f / N, j = f - t * N;
            if (t < nl) base[(int64_t)t * lstride + (int64_t)j * estride] = in[t * LS + j];
        }
    }
}

// The same line transform where adjacent lines are adjacent in memory
// (lstride 1: the x and y passes over boxes), as a persistent kernel: a
// block walks tiles blockIdx.x, + gridDim.x, ... and keeps the elements of
// its next tile in flight (registers) while it transforms and stores the
// current one; its barriers order LDS only, so those loads stay outstanding
// across them.  Per line the arithmetic is that of k_fft_lines_c.
template <typename T, int SIGN, int N, int TL>
__global__ __launch_bounds__(256) void k_fft_lines_cp(cpx<T>* __restrict__ data, int64_t n_inner,
                                                      int64_t ostride, int64_t estride,
                                                      int64_t ntile,
                                                      const cpx<T>* __restrict__ Wg,
                                                      const int* __restrict__ ocnt, int omod,
                                                      const int* __restrict__ emask)
{
    constexpr int LS = N + 1;
    constexpr int PER = (TL * N + 255) / 256;
    __shared__ cpx<T> W[N];
    __shared__ cpx<T> b0[TL * LS];
    __shared__ cpx<T> b1[TL * LS];
    __shared__ int em[N];
    const int tid = threadIdx.x;
    const int64_t tiles = (n_inner + TL - 1) / TL;
    const int64_t step = gridDim.x;
    for (int t = tid; t < N; t += 256)
    {
        W[t] = Wg[t];
        em[t] = emask ? emask[t] : 1;
    }
    lds_barrier();
    // first tile at or after s whose outer index is not pruned
    auto next_tile = [&](int64_t s) {
        if (ocnt)
            while (s < ntile && ocnt[(s / tiles) % omod] == 0) s += step;
        return s;
    };
    cpx<T> nx[PER];
    auto fetch = [&](int64_t tile) {
        const int64_t o = tile / tiles;
        const int64_t i0 = (tile - o * tiles) * TL;
        const int nl = (int)((n_inner - i0) < TL ? (n_inner - i0) : TL);
        const cpx<T>* base = data + o * ostride + i0;
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            const int f = tid + k * 256, j = f / TL, t = f - j * TL;
            nx[k] = (f < TL * N && t < nl && em[j]) ? base[(int64_t)j * estride + t]
                                                    : cpx<T>{0, 0};
        }
    };
    int64_t tile = next_tile(blockIdx.x);
    if (tile < ntile) fetch(tile);
    while (tile < ntile)
    {
        const int64_t nxt = next_tile(tile + step);
        // the previous tile's stores have read their LDS buffer
        lds_barrier();
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            const int f = tid + k * 256, j = f / TL, t = f - j * TL;
            if (f < TL * N) b0[t * LS + j] = nx[k];
        }
        lds_barrier();
        if (nxt < ntile) fetch(nxt);
        cpx<T>* in = b0;
        cpx<T>* out = b1;
        FftStagesC<T, SIGN, N, TL, 0, LS, true>::run(in, out, W, tid);
        const int64_t o = tile / tiles;
        const int64_t i0 = (tile - o * tiles) * TL;
        const int nl = (int)((n_inner - i0) < TL ? (n_inner - i0) : TL);
        cpx<T>* base = data + o * ostride + i0;
        for (int f = tid; f < TL * N; f += 256)
        {
            const int j = f / TL, t = f - j * TL;
            if (t < nl) base[(int64_t)j * estride + t] = in[t * LS + j];
        }
        tile = nxt;
    }
}

// lines per block: two buffers of TL*(N+1) within an LDS budget of 56 KB
// (FP64) or 28 KB (FP32: more blocks per CU; at N = 96, 16 lines of 8-byte
// elements per block make the pass faster than 32)
template <typename T, int N>
constexpr int fft_tl_c()
{
    constexpr int budget = sizeof(T) == 4 ? 28 * 1024 : 56 * 1024;
    return ((N + 1) * (int)sizeof(cpx<T>) * 2 * 32 <= budget) ? 32
           : ((N + 1) * (int)sizeof(cpx<T>) * 2 * 16 <= budget) ? 16
           : 8;
}

// a line tile TL fits when the twiddles and both line buffers stay within
// 64 KB of static LDS
template <typename T, int N, int TL>
constexpr bool fft_tl_fits()
{
    return (N + 2 * TL * (N + 1)) * (int)sizeof(cpx<T>) <= 64 * 1024;
}

template <typename T, int N, int TL>
static bool fft_launch_tl(cpx<T>* data, int64_t n_outer, int64_t n_inner, int64_t ostride,
                          int64_t lstride, int64_t estride, int sign, const cpx<T>* W,
                          hipStream_t stream, const int* ocnt, int omod, const int* emask)
{
    if constexpr (!fft_tl_fits<T, N, TL>())
        return false;
    else
    {
        const int64_t nblk = cdiv(n_inner, TL) * n_outer;
        if (nblk > 0x7fffffffLL) return false;
        if (lstride == 1)
        {
            // persistent: as many blocks as stay resident
            auto fn = sign > 0 ? (const void*)k_fft_lines_cp<T, 1, N, TL>
                               : (const void*)k_fft_lines_cp<T, -1, N, TL>;
            int occ = 0, dev = 0, ncu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, 0) != hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                    hipSuccess)
                return false;
            const int64_t grid = std::min<int64_t>(nblk, (int64_t)std::max(1, occ) * ncu);
            if (sign > 0)
                hipLaunchKernelGGL((k_fft_lines_cp<T, 1, N, TL>), dim3((unsigned)grid), dim3(256),
                                   0, stream, data, n_inner, ostride, estride, nblk, W, ocnt,
                                   omod, emask);
            else
                hipLaunchKernelGGL((k_fft_lines_cp<T, -1, N, TL>), dim3((unsigned)grid),
                                   dim3(256), 0, stream, data, n_inner, ostride, estride, nblk, W,
                                   ocnt, omod, emask);
            return true;
        }
        if (sign > 0)
            hipLaunchKernelGGL((k_fft_lines_c<T, 1, N, TL>), dim3((unsigned)nblk), dim3(256), 0,
                               stream, data, n_inner, ostride, lstride, estride, W, ocnt, omod,
                               emask);
        else
            hipLaunchKernelGGL((k_fft_lines_c<T, -1, N, TL>), dim3((unsigned)nblk), dim3(256), 0,
                               stream, data, n_inner, ostride, lstride, estride, W, ocnt, omod,
                               emask);
        return true;
    }
}

// tl: lines per block (8, 16 or 32 where it fits; anything else takes the
// default of fft_tl_c).  The transform is the same for every tile size:
// the lines are independent, so the results are bit-identical.
template <typename T, int N>
static bool fft_launch_c(cpx<T>* data, int64_t n_outer, int64_t n_inner, int64_t ostride,
                         int64_t lstride, int64_t estride, int sign, const cpx<T>* W,
                         hipStream_t stream, const int* ocnt, int omod, const int* emask,
                         int tl = 0)
{
    if (tl == 8 && fft_launch_tl<T, N, 8>(data, n_outer, n_inner, ostride, lstride, estride, sign,
                                          W, stream, ocnt, omod, emask))
        return true;
    if (tl == 16 && fft_launch_tl<T, N, 16>(data, n_outer, n_inner, ostride, lstride, estride,
                                            sign, W, stream, ocnt, omod, emask))
        return true;
    if (tl == 32 && fft_launch_tl<T, N, 32>(data, n_outer, n_inner, ostride, lstride, estride,
                                            sign, W, stream, ocnt, omod, emask))
        return true;
    return fft_launch_tl<T, N, fft_tl_c<T, N>()>(data, n_outer, n_inner, ostride, lstride,
                                                 estride, sign, W, stream, ocnt, omod, emask);
}

// true if n has a compile-time line kernel
static inline bool fft_has_c(int64_t n)
{
    switch (n)
    {
        case 12: case 24: case 32: case 36: case 45: case 48:
        case 60: case 64: case 72: case 96: case 108: case 216: return true;
        default: return false;
    }
}

template <typename T>
static inline const cpx<T>* axis_tw(const FftAxis& a);

// compiled line lengths (the grids of the benchmark systems and their
// neighbours); other lengths use the generic kernel
template <typename T>
static bool fft_try_c(int n, cpx<T>* data, int64_t no, int64_t ni, int64_t os, int64_t ls,
                      int64_t es, int sign, const cpx<T>* W, hipStream_t st,
                      const int* ocnt = nullptr, int omod = 1, const int* emask = nullptr,
                      int tl = 0)
{
#define FFT_C_CASE(NV)                                                                        \
    case NV:                                                                                  \
        return fft_launch_c<T, NV>(data, no, ni, os, ls, es, sign, W, st, ocnt, omod, emask, tl);
    switch (n)
    {
        FFT_C_CASE(12)
        FFT_C_CASE(24)
        FFT_C_CASE(32)
        FFT_C_CASE(36)
        FFT_C_CASE(45)
        FFT_C_CASE(48)
        FFT_C_CASE(60)
        FFT_C_CASE(64)
        FFT_C_CASE(72)
        FFT_C_CASE(96)
        FFT_C_CASE(108)
        FFT_C_CASE(216)
        default: return false;
    }
#undef FFT_C_CASE
}

template <>
inline const z64* axis_tw<double>(const FftAxis& a)
{
    return a.w64;
}
template <>
inline const c32* axis_tw<float>(const FftAxis& a)
{
    return a.w32;
}

template <typename T>
static int fft_pass(const FftAxis& ax, cpx<T>* data, int64_t n_outer, int64_t n_inner,
                    int64_t ostride, int64_t lstride, int64_t estride, int sign,
                    hipStream_t stream, int lines = 0)
{
    const int n = ax.n;
    if (fft_try_c<T>(n, data, n_outer, n_inner, ostride, lstride, estride, sign, axis_tw<T>(ax),
                     stream, nullptr, 1, nullptr, lines))
    {
        HIP_TRY(hipGetLastError());
        return LELPH_OK;
    }
    const int64_t lds_budget = 60 * 1024;
    int64_t tl = (lds_budget / (int64_t)sizeof(cpx<T>) - n) / (2 * n);
    if (tl > 64) tl = 64;
    if (tl > n_inner) tl = n_inner;
    if (tl < 1) return LELPH_ERR_UNSUPPORTED;
    const int64_t tiles = cdiv(n_inner, tl);
    const int64_t nblk = tiles * n_outer;
    if (nblk > 0x7fffffffLL) return LELPH_ERR_UNSUPPORTED;
    const size_t lds = sizeof(cpx<T>) * (size_t)(n + 2 * tl * n);
    if (sign > 0)
        hipLaunchKernelGGL((k_fft_lines<T, 1>), dim3((unsigned)nblk), dim3(256), lds, stream,
                           data, n, n_inner, ostride, lstride, estride, (int)tl, ax.st,
                           axis_tw<T>(ax));
    else
        hipLaunchKernelGGL((k_fft_lines<T, -1>), dim3((unsigned)nblk), dim3(256), lds, stream,
                           data, n, n_inner, ostride, lstride, estride, (int)tl, ax.st,
                           axis_tw<T>(ax));
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

// 3-D transform of nsets boxes (Nx,Ny,Nz), in place.
template <typename T>
static int fft3d_batch(const FftAxis* ax, cpx<T>* data, int64_t nsets, int sign,
                       hipStream_t stream, int tl = 0)
{
    const int64_t Nx = ax[0].n, Ny = ax[1].n, Nz = ax[2].n;
    const int64_t nf = Nx * Ny * Nz;
    // z: contiguous lines
    LELPH_TRY(fft_pass<T>(ax[2], data, 1, nsets * Nx * Ny, 0, Nz, 1, sign, stream, tl));
    // y: outer (set, x), inner z
    LELPH_TRY(fft_pass<T>(ax[1], data, nsets * Nx, Nz, Ny * Nz, 1, Nz, sign, stream, tl));
    // x: outer set, inner (y, z)
    LELPH_TRY(fft_pass<T>(ax[0], data, nsets, Ny * Nz, nf, 1, Ny * Nz, sign, stream, tl));
    return LELPH_OK;
}
