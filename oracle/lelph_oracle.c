/*
 * lelph_oracle.c -- FP64 CPU restatement of LetzElPhC's e-ph hot path.
 *
 * TEST INFRASTRUCTURE (the checker).  Only tests/, the smoke check and the
 * bench's cpu_baseline leg load it; the product path never does.
 *
 * It follows the reference algorithm step by step (file:line cited on each
 * function), single MPI rank, COMPILE_ELPH_DOUBLE semantics, with the
 * reference's constants (ELPH_PI = 3.1415927, src/common/constants.h:9-12).
 * Differences, all without numerical effect beyond FP64 rounding:
 *   - FFTs use this file's own mixed-radix DFT instead of FFTW (same sign and
 *     normalisation conventions: BACKWARD = e^{+i}, unnormalised; FORWARD =
 *     e^{-i}, the caller multiplies by 1/N as src/fft/convolution.c:283-284).
 *   - ZGEMM is a plain triple loop with the row-major CBLAS semantics of
 *     matmul_cmplx (src/common/blas_lapack.c:109-127).
 *   - MPI reductions are absent (one rank owns everything).
 *
 * Parity pins (checked in tests/): FFT vs brute-force DFT and numpy; Y_lm
 * orthonormality and explicit l<=2 forms; Simpson/spline against closed forms
 * and the reference's own compiled numerical_func.c/spline.c/su2mat.c
 * (oracle/_ref); elphLocal against the real-space Parseval form and the
 * constant-dV closed form; non-local against the translation sum rule;
 * dVlocq against the Gaussian closed form and its sum rule.
 */
#include "lelph_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define RPI 3.1415927   /* ELPH_PI     src/common/constants.h:11 */
#define REPS 1e-5       /* ELPH_EPS    src/common/constants.h:9  */
#define REPS6 1e-6      /* ELPH_EPS6   src/common/constants.h:10 */
#define RE2 2.0         /* ELPH_e2     src/common/constants.h:13 */

static int g_threads = 1;       /* over k points (orc_compute_elphq)        */
static int g_inner_threads = 1; /* inside one elphLocal (bands, GEMM rows) */
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_threads(void) { return g_threads; }
void orc_set_inner_threads(int n) { g_inner_threads = n > 0 ? n : 1; }

/* ======================================================================
 * 3x3 helpers: MatVec3f / Gemm3x3f / det3x3 / reciprocal_vecs
 * (src/common/numerical_func.c:226-302, 367-423)
 * ==================================================================== */
static void matvec3(const double* M, const double* v, bool trans, double* out)
{
    if (trans)
    {
        out[0] = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
        out[1] = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
        out[2] = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
    }
    else
    {
        out[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
        out[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
        out[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
    }
}

static double det3(const double* m)
{
    double d = 0;
    d += m[0] * (m[4] * m[8] - m[5] * m[7]);
    d -= m[1] * (m[3] * m[8] - m[5] * m[6]);
    d += m[2] * (m[7] * m[3] - m[4] * m[6]);
    return d;
}

void orc_reciprocal_vecs(const double* a, double* b)
{
    double d = 2 * RPI / det3(a);
    b[0] = (a[4] * a[8] - a[7] * a[5]) * d;
    b[1] = (a[5] * a[6] - a[3] * a[8]) * d;
    b[2] = (a[3] * a[7] - a[6] * a[4]) * d;
    b[3] = (a[2] * a[7] - a[1] * a[8]) * d;
    b[4] = (a[0] * a[8] - a[2] * a[6]) * d;
    b[5] = (a[1] * a[6] - a[0] * a[7]) * d;
    b[6] = (a[1] * a[5] - a[2] * a[4]) * d;
    b[7] = (a[2] * a[3] - a[0] * a[5]) * d;
    b[8] = (a[0] * a[4] - a[1] * a[3]) * d;
}

/* ======================================================================
 * legendre / Ylm (src/common/numerical_func.c:29-152), real harmonics of
 * Blanco et al. without the Condon-Shortley phase.
 * ==================================================================== */
static double fact(int n) { return n < 0 ? 0.0 : tgamma(n + 1.0); }

double orc_legendre(int l, int m, double x)
{
    if (l == 0 && m == 0) return 1.0;
    if (l < 0) return orc_legendre(-l - 1, m, x);
    if (abs(m) > l) return 0.0;
    if (m < 0)
        return orc_legendre(l, -m, x) * pow(-1, -m) * fact(l + m) /
               fact(l - m);
    double p = 1.0;
    double s = sqrt(1 - x * x);
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

double orc_Ylm(int l, int m, const double* v)
{
    if (l < 0) return 0.0;
    double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nrm < REPS6) return l != 0 ? 0.0 : 1.0 / sqrt(4 * RPI);
    double cost = v[2] / nrm;
    double phi = atan2(v[1], v[0]);
    double sh;
    if (m > 0)
        sh = sqrt(2) * cos(m * phi);
    else if (m < 0)
    {
        m = -m;
        sh = sqrt(2) * sin(m * phi);
    }
    else
        sh = 1;
    sh *= orc_legendre(l, m, cost);
    sh *= sqrt(fact(l - m) * (2 * l + 1) / (4 * RPI * fact(l + m)));
    return sh;
}

/* simpson (src/common/numerical_func.c:155-192) */
double orc_simpson(const double* f, const double* dx, int64_t n)
{
    if (n < 3) return 0;
    double s = f[0] * dx[0];
    for (int64_t i = 1; i < n - 1; i += 2) s += 4.0 * f[i] * dx[i];
    for (int64_t i = 2; i < n - 1; i += 2) s += 2.0 * f[i] * dx[i];
    if (n % 2 != 0)
        s += f[n - 1] * dx[n - 1];
    else
    {
        s -= f[n - 3] * dx[n - 3] * 0.25;
        s += f[n - 2] * dx[n - 2];
        s += f[n - 1] * dx[n - 1] * 1.25;
    }
    return s / 3.0;
}

/* spline_interpolate / prepare_spline (src/common/spline.c:13-139):
 * not-a-knot cubic Hermite derivatives by one Thomas sweep. */
double orc_spline_interpolate(double x, int64_t i, const double* xi,
                              const double* yi, const double* dy)
{
    double h = xi[i + 1] - xi[i];
    double t = (x - xi[i]) / h;
    double a = dy[i] * h - (yi[i + 1] - yi[i]);
    double b = -dy[i + 1] * h + (yi[i + 1] - yi[i]);
    return (1 - t) * yi[i] + t * yi[i + 1] + t * (1 - t) * ((1 - t) * a + t * b);
}

void orc_prepare_spline(int64_t n, const double* x, const double* y,
                        double* dy)
{
    double* a = malloc(sizeof(double) * 4 * n);
    double* b = a + n;
    double* c = a + 2 * n;
    double* d = a + 3 * n;
    double h1 = x[1] - x[0], s1 = (y[1] - y[0]) / h1;
    double h2 = x[2] - x[1], s2 = (y[2] - y[1]) / h2;
    a[0] = h2 * h2;
    b[0] = h2 * h2 - h1 * h1;
    c[0] = -h1 * h1;
    d[0] = 2 * (s1 * h2 * h2 - s2 * h1 * h1);
    h1 = x[n - 2] - x[n - 3];
    s1 = (y[n - 2] - y[n - 3]) / h1;
    h2 = x[n - 1] - x[n - 2];
    s2 = (y[n - 1] - y[n - 2]) / h2;
    a[n - 1] = h2 * h2;
    b[n - 1] = h2 * h2 - h1 * h1;
    c[n - 1] = -h1 * h1;
    d[n - 1] = 2 * (s1 * h2 * h2 - s2 * h1 * h1);
    for (int64_t i = 1; i < n - 1; ++i)
    {
        double hi = x[i + 1] - x[i], hm = x[i] - x[i - 1];
        double si = (y[i + 1] - y[i]) / hi, sm = (y[i] - y[i - 1]) / hm;
        a[i] = hi;
        b[i] = 2.0 * (hi + hm);
        c[i] = hm;
        d[i] = 3 * (sm * hi + si * hm);
    }
    double al = -c[0] / c[1];
    a[0] += al * a[1];
    b[0] += al * b[1];
    c[0] += al * c[1];
    d[0] += al * d[1];
    c[0] = b[0];
    b[0] = a[0];
    a[0] = 0;
    al = -a[n - 1] / a[n - 2];
    a[n - 1] += al * a[n - 2];
    b[n - 1] += al * b[n - 2];
    c[n - 1] += al * c[n - 2];
    d[n - 1] += al * d[n - 2];
    a[n - 1] = b[n - 1];
    b[n - 1] = c[n - 1];
    c[n - 1] = 0;
    for (int64_t i = 1; i < n; ++i)
    {
        double w = a[i] / b[i - 1];
        b[i] -= w * c[i - 1];
        d[i] -= w * d[i - 1];
    }
    dy[n - 1] = d[n - 1] / b[n - 1];
    for (int64_t i = n - 2; i >= 0; --i) dy[i] = (d[i] - c[i] * dy[i + 1]) / b[i];
    free(a);
}

/* SU2mat (src/wfc/su2mat.c:16-117) */
int orc_SU2mat(const double* sym_in, int nspinor, int invert, int time_rev,
               double* out)
{
    cplx* su2 = (cplx*)out;
    if (nspinor == 1)
    {
        su2[0] = 1.0;
        return 0;
    }
    double det = det3(sym_in);
    if (fabs(fabs(det) - 1) > 1e-4) return LELPH_ERR_SYM;
    double s[9];
    for (int i = 0; i < 9; ++i) s[i] = sym_in[i] / det;
    const cplx p0[4] = {1, 0, 0, 1}, px[4] = {0, 1, 1, 0},
               py[4] = {0, -I, I, 0}, pz[4] = {1, 0, 0, -1};
    double al, be, de;
    if (fabs(s[6] + 1) < REPS)
    {
        al = 0;
        be = RPI / 2.0;
        de = atan2(s[1], s[2]);
    }
    else if (fabs(s[6] - 1) < REPS)
    {
        al = 0;
        be = -RPI / 2.0;
        de = atan2(-s[1], -s[2]);
    }
    else
    {
        be = -asin(s[6]);
        al = atan2(s[3] / cos(be), s[0] / cos(be));
        de = atan2(s[7] / cos(be), s[8] / cos(be));
    }
    al /= 2;
    be /= 2;
    de /= 2;
    cplx sd[4], sb[4], sa[4], t[4];
    for (int i = 0; i < 4; ++i)
    {
        sd[i] = cos(de) * p0[i] - I * sin(de) * px[i];
        sb[i] = cos(be) * p0[i] - I * sin(be) * py[i];
        sa[i] = cos(al) * p0[i] - I * sin(al) * pz[i];
    }
    t[0] = sb[0] * sd[0] + sb[1] * sd[2];
    t[1] = sb[0] * sd[1] + sb[1] * sd[3];
    t[2] = sb[2] * sd[0] + sb[3] * sd[2];
    t[3] = sb[2] * sd[1] + sb[3] * sd[3];
    su2[0] = sa[0] * t[0] + sa[1] * t[2];
    su2[1] = sa[0] * t[1] + sa[1] * t[3];
    su2[2] = sa[2] * t[0] + sa[3] * t[2];
    su2[3] = sa[2] * t[1] + sa[3] * t[3];
    if (time_rev)
    {
        cplx a1 = su2[0], b1 = su2[1], c1 = su2[2], d1 = su2[3];
        su2[0] = -c1;
        su2[1] = -d1;
        su2[2] = a1;
        su2[3] = b1;
    }
    if (invert)
    {
        cplx b1 = conj(su2[1]);
        su2[0] = conj(su2[0]);
        su2[3] = conj(su2[3]);
        su2[1] = conj(su2[2]);
        su2[2] = b1;
    }
    return 0;
}

/* su2rotate (src/wfc/su2rotate.c:13-41) */
static void su2rotate(int nspinor, int64_t npw, const cplx* su, cplx* w)
{
    if (nspinor != 2) return;
    cplx* up = w;
    cplx* dw = w + npw;
    for (int64_t i = 0; i < npw; ++i)
    {
        cplx t = up[i] * su[0] + dw[i] * su[1];
        dw[i] = up[i] * su[2] + dw[i] * su[3];
        up[i] = t;
    }
}

/* apply_trans_wfc (src/wfc/trans_wfc.c:15-52) */
static void apply_trans_wfc(const double* tr, const double* kvec,
                            int64_t nsets, int64_t npw, const double* gv,
                            cplx* w, bool cj)
{
    cplx kph = cexp(-I * 2 * RPI *
                    (kvec[0] * tr[0] + kvec[1] * tr[1] + kvec[2] * tr[2]));
    for (int64_t s = 0; s < nsets; ++s)
    {
        cplx* ws = w + s * npw;
        for (int64_t g = 0; g < npw; ++g)
        {
            const double* G = gv + 3 * g;
            cplx ph = kph * cexp(-I * 2 * RPI *
                                 (G[0] * tr[0] + G[1] * tr[1] + G[2] * tr[2]));
            ws[g] *= ph;
            if (cj) ws[g] = conj(ws[g]);
        }
    }
}

/* rotateGvecs (src/wfc/rotate_Gvecs.c:15-81), inverse = false */
static void rotateGvecs(const double* G, const double* S, int64_t n,
                        const double* lat, bool crystal, const double* G0in,
                        double* out)
{
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const double* L = crystal ? lat : I3;
    double M[9]; /* M = L^T S */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] =
                L[i] * S[j] + L[3 + i] * S[3 + j] + L[6 + i] * S[6 + j];
    for (int64_t g = 0; g < n; ++g)
        for (int i = 0; i < 3; ++i)
            out[3 * g + i] = G[3 * g] * M[3 * i] + G[3 * g + 1] * M[3 * i + 1] +
                             G[3 * g + 2] * M[3 * i + 2];
    if (G0in)
    {
        double G0[3];
        if (crystal)
            matvec3(lat, G0in, true, G0);
        else
            memcpy(G0, G0in, sizeof(G0));
        for (int64_t g = 0; g < n; ++g)
            for (int i = 0; i < 3; ++i) out[3 * g + i] += G0[i];
    }
}

/* ======================================================================
 * Sort_pw / Sorted_gvecs_idxs (src/wfc/sortpw.c:14-234,
 * src/wfc/gsort.c:16-47,95-128), one rank: sort G into z sticks (x, then y,
 * then z) and permute the coefficients.
 * ==================================================================== */
typedef struct
{
    long long x, y, z;
    int64_t i;
} gkey;

static int gkey_cmp(const void* a, const void* b)
{
    const gkey* p = a;
    const gkey* q = b;
    if (p->x != q->x) return (p->x > q->x) - (p->x < q->x);
    if (p->y != q->y) return (p->y > q->y) - (p->y < q->y);
    return (p->z > q->z) - (p->z < q->z);
}

static void sort_pw(int64_t npw, const double* gin, const cplx* win,
                    int64_t nsets, int* gout, cplx* wout)
{
    gkey* k = malloc(sizeof(gkey) * npw);
    for (int64_t g = 0; g < npw; ++g)
    {
        k[g].x = (long long)rint(gin[3 * g]);
        k[g].y = (long long)rint(gin[3 * g + 1]);
        k[g].z = (long long)rint(gin[3 * g + 2]);
        k[g].i = g;
    }
    qsort(k, npw, sizeof(gkey), gkey_cmp);
    for (int64_t g = 0; g < npw; ++g)
    {
        gout[3 * g] = (int)k[g].x;
        gout[3 * g + 1] = (int)k[g].y;
        gout[3 * g + 2] = (int)k[g].z;
    }
    for (int64_t s = 0; s < nsets; ++s)
        for (int64_t g = 0; g < npw; ++g)
            wout[s * npw + g] = win[s * npw + k[g].i];
    free(k);
}

/* ======================================================================
 * FFT: own mixed-radix decimation-in-time DFT (FFTW sign conventions).
 * ==================================================================== */
typedef struct
{
    int64_t n;
    int nf;
    int f[64];
    cplx* w; /* e^{+2 pi i t/n}, t in [0,n) */
} fplan;

static int fplan_make(fplan* p, int64_t n)
{
    p->n = n;
    p->nf = 0;
    int64_t m = n;
    for (int64_t d = 2; m > 1;)
    {
        if (m % d == 0)
        {
            if (p->nf >= 64) return -1;
            p->f[p->nf++] = (int)d;
            m /= d;
        }
        else
            ++d;
    }
    p->w = malloc(sizeof(cplx) * (n > 0 ? n : 1));
    for (int64_t t = 0; t < n; ++t)
    {
        double a = 2.0 * M_PI * (double)t / (double)n;
        p->w[t] = cos(a) + I * sin(a);
    }
    return 0;
}

static void fplan_free(fplan* p) { free(p->w); }

static void fft_rec(const cplx* in, int64_t is, cplx* out, int64_t n,
                    const int* f, const fplan* P, int sign)
{
    if (n == 1)
    {
        out[0] = in[0];
        return;
    }
    const int p = f[0];
    const int64_t m = n / p;
    for (int r = 0; r < p; ++r) fft_rec(in + r * is, is * p, out + r * m, m, f + 1, P, sign);
    const int64_t ws = P->n / n;
    cplx tmp[64];
    for (int64_t k = 0; k < m; ++k)
    {
        for (int r = 0; r < p; ++r) tmp[r] = out[r * m + k];
        for (int s = 0; s < p; ++s)
        {
            cplx acc = 0;
            int64_t kk = k + s * m;
            for (int r = 0; r < p; ++r)
            {
                int64_t t = ((int64_t)r * kk) % n;
                cplx w = P->w[t * ws];
                if (sign < 0) w = conj(w);
                acc += w * tmp[r];
            }
            out[kk] = acc;
        }
    }
}

int orc_fft1d(double* data, int64_t n, int sign)
{
    fplan P;
    if (fplan_make(&P, n)) return LELPH_ERR_ARG;
    for (int i = 0; i < P.nf; ++i)
        if (P.f[i] > 64)
        {
            fplan_free(&P);
            return LELPH_ERR_UNSUPPORTED;
        }
    cplx* out = malloc(sizeof(cplx) * n);
    fft_rec((cplx*)data, 1, out, n, P.f, &P, sign);
    memcpy(data, out, sizeof(cplx) * n);
    free(out);
    fplan_free(&P);
    return 0;
}

int orc_fft3d(double* boxd, const int64_t* dims, int sign)
{
    cplx* box = (cplx*)boxd;
    const int64_t n[3] = {dims[0], dims[1], dims[2]};
    const int64_t st[3] = {dims[1] * dims[2], dims[2], 1};
    for (int ax = 0; ax < 3; ++ax)
    {
        fplan P;
        if (fplan_make(&P, n[ax])) return LELPH_ERR_ARG;
        for (int i = 0; i < P.nf; ++i)
            if (P.f[i] > 64)
            {
                fplan_free(&P);
                return LELPH_ERR_UNSUPPORTED;
            }
        cplx* line = malloc(sizeof(cplx) * n[ax]);
        cplx* out = malloc(sizeof(cplx) * n[ax]);
        const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
        for (int64_t i1 = 0; i1 < n[a1]; ++i1)
            for (int64_t i2 = 0; i2 < n[a2]; ++i2)
            {
                cplx* base = box + i1 * st[a1] + i2 * st[a2];
                for (int64_t j = 0; j < n[ax]; ++j) line[j] = base[j * st[ax]];
                fft_rec(line, 1, out, n[ax], P.f, &P, sign);
                for (int64_t j = 0; j < n[ax]; ++j) base[j * st[ax]] = out[j];
            }
        free(line);
        free(out);
        fplan_free(&P);
    }
    return 0;
}

/* box index of an integer G (planner.c:119-137, bwdfft.c:158-162):
 * negative components wrap by +N, anything outside [0,N) is an error. */
static int box_index(const int* G, const int64_t* N, int64_t* idx)
{
    int64_t c[3];
    for (int i = 0; i < 3; ++i)
    {
        c[i] = G[i] < 0 ? G[i] + N[i] : G[i];
        if (c[i] < 0 || c[i] >= N[i]) return LELPH_ERR_RANGE;
    }
    *idx = (c[0] * N[1] + c[1]) * N[2] + c[2];
    return 0;
}

/* invfft3D (src/fft/bwdfft.c:13-120): sphere -> box (optional conj) then the
 * unnormalised backward transform; nsets sets of npw coefficients. */
static int invfft3D(const int* gv, int64_t npw, const int64_t* N,
                    int64_t nsets, const cplx* wG, cplx* wr, bool cj)
{
    const int64_t nf = N[0] * N[1] * N[2];
    for (int64_t s = 0; s < nsets; ++s)
    {
        cplx* box = wr + s * nf;
        memset(box, 0, sizeof(cplx) * nf);
        for (int64_t g = 0; g < npw; ++g)
        {
            int64_t id;
            if (box_index(gv + 3 * g, N, &id)) return LELPH_ERR_RANGE;
            box[id] = cj ? conj(wG[s * npw + g]) : wG[s * npw + g];
        }
        int e = orc_fft3d((double*)box, N, +1);
        if (e) return e;
    }
    return 0;
}

/* fft_convolution3D (src/fft/convolution.c:13-189): (V.psi)(r), forward
 * FFT, box -> sphere times 1/N (optional conj). */
static int fft_convolution3D(const int* gv, int64_t npw, const int64_t* N,
                             int nspinor, int nmag, const cplx* V,
                             const cplx* psir, cplx* wG, bool cj, cplx* work)
{
    if (nspinor == 1) nmag = 1;
    if (nmag != 1 && nmag != 4) return LELPH_ERR_ARG;
    const int64_t nf = N[0] * N[1] * N[2];
    const double norm = 1.0 / (double)(N[0] * N[1] * N[2]);
    for (int sp = 0; sp < nspinor; ++sp)
    {
        if (nmag == 1)
        {
            const cplx* ps = psir + sp * nf;
            for (int64_t i = 0; i < nf; ++i) work[i] = ps[i] * V[i];
        }
        else
        {
            const cplx* dV = V + (sp == 1 ? 2 * nf : 0);
            for (int64_t i = 0; i < nf; ++i)
                work[i] = psir[i] * dV[i] + psir[i + nf] * dV[i + nf];
        }
        int e = orc_fft3d((double*)work, N, -1);
        if (e) return e;
        cplx* out = wG + sp * npw;
        for (int64_t g = 0; g < npw; ++g)
        {
            int64_t id;
            if (box_index(gv + 3 * g, N, &id)) return LELPH_ERR_RANGE;
            out[g] = work[id] * norm;
            if (cj) out[g] = conj(out[g]);
        }
    }
    return 0;
}

/* matmul_cmplx (src/common/blas_lapack.c:109-127): row-major
 * C(m,n) = alpha op(A) op(B) + beta C, op in {'N','T','C'} */
static void zgemm_rm(char ta, char tb, const cplx* A, const cplx* B, cplx* Cm,
                     cplx alpha, cplx beta, int64_t lda, int64_t ldb,
                     int64_t ldc, int64_t m, int64_t n, int64_t k)
{
#pragma omp parallel for num_threads(g_inner_threads) schedule(static) if (g_inner_threads > 1)
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j)
        {
            cplx acc = 0;
            for (int64_t l = 0; l < k; ++l)
            {
                cplx a = (ta == 'N') ? A[i * lda + l] : A[l * lda + i];
                if (ta == 'C') a = conj(a);
                cplx b = (tb == 'N') ? B[l * ldb + j] : B[j * ldb + l];
                if (tb == 'C') b = conj(b);
                acc += a * b;
            }
            Cm[i * ldc + j] = alpha * acc + (beta == 0 ? 0 : beta * Cm[i * ldc + j]);
        }
}

/* ======================================================================
 * elphLocal (src/dvloc/compute_Vbare_loc.c:17-353)
 * ==================================================================== */
int orc_elphLocal(const double* qpt, const orc_wfc* wfcs,
                  const lelph_lattice* lat, int ikq, int ik, int kqsym,
                  int ksym, const double* dVlocrd, double* elph_kqd)
{
    const int64_t nspin = lat->nspin, nb = lat->nbnds, nspinor = lat->nspinor;
    const int64_t nmodes = lat->nmodes, nmag = lat->nmag;
    const int64_t* N = lat->fft_dims;
    const int64_t nf = N[0] * N[1] * N[2];
    const double* latv = lat->alat_vec;
    const double* blat = lat->blat_vec;
    const cplx* dVlocr = (const cplx*)dVlocrd;
    cplx* elph_kq = (cplx*)elph_kqd;
    int err = 0;

    const cplx* wfc_kq = (const cplx*)wfcs[ikq].wfc;
    const double* Gkq = wfcs[ikq].gvec;
    const int64_t npwkq = wfcs[ikq].npw;
    const cplx* wfc_k = (const cplx*)wfcs[ik].wfc;
    const double* Gk = wfcs[ik].gvec;
    const int64_t npwk = wfcs[ik].npw;

    const double* kqvec = lat->kpt_iredBZ + 3 * ikq;
    const double* kvec = lat->kpt_iredBZ + 3 * ik;
    const double* symkq = lat->syms[kqsym].Rmat;
    const double* symk = lat->syms[ksym].Rmat;
    const double* taukq = lat->syms[kqsym].tau;
    const double* tauk = lat->syms[ksym].tau;
    const bool trkq = lat->syms[kqsym].time_rev;
    const bool trk = lat->syms[ksym].time_rev;

    /* geometry, :93-144 */
    double tSkq[3], tSk[3], kvecSk[3], kvecSkq[3], qc[3], taukq_c[3], tauk_c[3],
        ulm[3];
    matvec3(symkq, kqvec, false, tSkq);
    matvec3(symk, kvec, false, tSk);
    matvec3(latv, tSk, true, kvecSk);
    for (int i = 0; i < 3; ++i) kvecSkq[i] = kvecSk[i] + qpt[i];
    for (int i = 0; i < 3; ++i) tSkq[i] -= tSk[i];
    matvec3(blat, qpt, false, qc);
    for (int i = 0; i < 3; ++i) qc[i] /= (2 * RPI);
    matvec3(blat, taukq, true, taukq_c);
    for (int i = 0; i < 3; ++i) taukq_c[i] /= (2 * RPI);
    matvec3(blat, tauk, true, tauk_c);
    for (int i = 0; i < 3; ++i) tauk_c[i] /= (2 * RPI);
    for (int i = 0; i < 3; ++i) ulm[i] = tSkq[i] - qc[i];

    /* rotate and shift G, :146-147 */
    double* gSkq = malloc(sizeof(double) * 3 * npwkq);
    double* gSk = malloc(sizeof(double) * 3 * npwk);
    rotateGvecs(Gkq, symkq, npwkq, latv, true, ulm, gSkq);
    rotateGvecs(Gk, symk, npwk, latv, true, NULL, gSk);

    cplx su2kq[4] = {1, 0, 0, 1}, su2k[4] = {1, 0, 0, 1};
    if ((err = orc_SU2mat(symkq, nspinor, 0, trkq, (double*)su2kq))) goto out0;
    if ((err = orc_SU2mat(symk, nspinor, 0, trk, (double*)su2k))) goto out0;

    /* Sort_pw, :170-176 */
    const int64_t nsets = nspin * nspinor * nb;
    int* gvkq = malloc(sizeof(int) * 3 * npwkq);
    int* gvk = malloc(sizeof(int) * 3 * npwk);
    cplx* wSkq = malloc(sizeof(cplx) * nsets * npwkq);
    cplx* wSk = malloc(sizeof(cplx) * nsets * npwk);
    sort_pw(npwkq, gSkq, wfc_kq, nsets, gvkq, wSkq);
    sort_pw(npwk, gSk, wfc_k, nsets, gvk, wSk);

    const int64_t maxpw = npwkq > npwk ? npwkq : npwk;
    double* gf = malloc(sizeof(double) * 3 * maxpw);
    for (int64_t i = 0; i < 3 * npwkq; ++i) gf[i] = gvkq[i];
    for (int64_t s = 0; s < nspin * nb; ++s)
    {
        cplx* w = wSkq + s * nspinor * npwkq;
        su2rotate(nspinor, npwkq, su2kq, w);
        apply_trans_wfc(taukq_c, kvecSkq, nspinor, npwkq, gf, w, false);
    }

    /* psi_k(r), :207-241 */
    cplx* wSk_r = malloc(sizeof(cplx) * nspin * nb * nspinor * nf);
    for (int64_t i = 0; i < 3 * npwk; ++i) gf[i] = gvk[i];
    int ierr = 0;
#pragma omp parallel for num_threads(g_inner_threads) schedule(dynamic, 1) reduction(| : ierr) if (g_inner_threads > 1)
    for (int64_t s = 0; s < nspin * nb; ++s)
    {
        cplx* w = wSk + s * nspinor * npwk;
        su2rotate(nspinor, npwk, su2k, w);
        apply_trans_wfc(tauk_c, kvecSk, nspinor, npwk, gf, w, false);
        ierr |= invfft3D(gvk, npwk, N, nspinor, w, wSk_r + s * nspinor * nf, trk);
    }
    if ((err = ierr)) goto out1;

    cplx* dVpsiG = malloc(sizeof(cplx) * nspin * nb * nspinor * npwkq);
    cplx* work = malloc(sizeof(cplx) * nf * g_inner_threads);
    cplx* mn = calloc(nb * nb, sizeof(cplx));
    for (int64_t i = 0; i < nmodes * nb * nb * nspin; ++i) elph_kq[i] = 0;
    if ((nmag == 2 && nspin != 2) || (nmag == 2 && nspinor != 1))
    {
        err = LELPH_ERR_ARG;
        goto out2;
    }
    /* :286-342 */
    for (int64_t iv = 0; iv < nmodes; ++iv)
    {
        const cplx* dv_nu = dVlocr + nmag * nf * iv;
        for (int64_t is = 0; is < nspin; ++is)
        {
            const cplx* psi = wSk_r + is * nb * nspinor * nf;
            const cplx* dV = dv_nu + nf * is;
            int cerr = 0;
#pragma omp parallel for num_threads(g_inner_threads) schedule(dynamic, 1) reduction(| : cerr) if (g_inner_threads > 1)
            for (int64_t ib = 0; ib < nb; ++ib)
            {
                int tn = 0;
#ifdef _OPENMP
                tn = omp_get_thread_num();
#endif
                cplx* o = dVpsiG + is * nb * nspinor * npwkq + ib * nspinor * npwkq;
                cerr |= fft_convolution3D(gvkq, npwkq, N, (int)nspinor, (int)nmag, dV,
                                          psi + ib * nspinor * nf, o, false,
                                          work + (int64_t)tn * nf);
            }
            if ((err = cerr)) goto out2;
            char bc = trkq ? 'T' : 'C';
            zgemm_rm('N', bc, dVpsiG + is * nb * nspinor * npwkq,
                     wSkq + is * nb * nspinor * npwkq, mn, 1.0, 0.0,
                     nspinor * npwkq, nspinor * npwkq, nb, nb, nb,
                     nspinor * npwkq);
            memcpy(elph_kq + (iv * nspin + is) * nb * nb, mn,
                   sizeof(cplx) * nb * nb);
        }
    }
out2:
    free(mn);
    free(work);
    free(dVpsiG);
out1:
    free(wSk_r);
    free(gf);
    free(wSkq);
    free(wSk);
    free(gvkq);
    free(gvk);
out0:
    free(gSkq);
    free(gSk);
    return err;
}

/* ======================================================================
 * add_elphNonLocal (src/nonloc/Vbare_nloc.c:35-577) and sum_VNL_KKp
 * (src/nonloc/sum_VNL_KKp.c:8-58)
 * ==================================================================== */
static void sum_VNL_KKp(const cplx* K, const cplx* Kp, const cplx* f,
                        int64_t nspin, int64_t nb, int64_t nspinor, cplx* out)
{
    const cplx* K3 = K + nspin * nb * nspinor;
    const cplx* Kp3 = Kp + nspin * nb * nspinor;
    for (int64_t ix = 0; ix < 3; ++ix)
        for (int64_t is = 0; is < nspin; ++is)
            for (int64_t b1 = 0; b1 < nb; ++b1)
            {
                int64_t s1 = nspinor * b1 + nb * nspinor * is;
                const cplx* k0 = Kp + s1;
                const cplx* k3 = Kp3 + s1 + nspin * nb * nspinor * ix;
                for (int64_t b2 = 0; b2 < nb; ++b2)
                {
                    int64_t s2 = nspinor * b2 + nb * nspinor * is;
                    const cplx* q0 = K + s2;
                    const cplx* q3 = K3 + s2 + nspin * nb * nspinor * ix;
                    cplx* o = out + b2 + b1 * nb + is * nb * nb + ix * nb * nb * nspin;
                    for (int64_t p1 = 0; p1 < nspinor; ++p1)
                        for (int64_t p2 = 0; p2 < nspinor; ++p2)
                            *o += f[nspinor * p1 + p2] * (q3[p1] * k0[p2] + q0[p1] * k3[p2]);
                }
            }
}

int orc_add_elphNonLocal(const orc_wfc* wfcs, const lelph_lattice* lat,
                         const lelph_pseudo* pp, int ikq, int ik, int kqsym,
                         int ksym, const double* eigVecd, double* elph_kqd)
{
    const int64_t nspin = lat->nspin, nb = lat->nbnds, nspinor = lat->nspinor;
    const int64_t ntype = pp->ntype, natom = lat->natom;
    const int64_t npwK = wfcs[ikq].npw, npwKp = wfcs[ik].npw;
    const cplx* eigVec = (const cplx*)eigVecd;
    cplx* elph_kq = (cplx*)elph_kqd;
    int err = 0;

    const int64_t szK = nspin * nb * nspinor * npwK;
    const int64_t szKp = nspin * nb * nspinor * npwKp;
    cplx* wK = malloc(sizeof(cplx) * szK);
    cplx* wKp = malloc(sizeof(cplx) * szKp);
    double* GK = malloc(sizeof(double) * 3 * npwK);
    double* GKp = malloc(sizeof(double) * 3 * npwKp);
    memcpy(wK, wfcs[ikq].wfc, sizeof(cplx) * szK);
    memcpy(wKp, wfcs[ik].wfc, sizeof(cplx) * szKp);
    const double* FK = wfcs[ikq].Fk;
    const double* FKp = wfcs[ik].Fk;
    const double* Kv = lat->kpt_iredBZ + 3 * ikq;
    const double* Kpv = lat->kpt_iredBZ + 3 * ik;
    const int lmax = pp->lmax;
    const int64_t nlmax = (lmax + 1) * (lmax + 1);
    const lelph_symmetry* sK = lat->syms + kqsym;
    const lelph_symmetry* sKp = lat->syms + ksym;

    /* K = S(k+q+G), K' = S(k+G), cart, :148-168 */
    for (int64_t g = 0; g < npwK; ++g)
    {
        const double* G = wfcs[ikq].gvec + 3 * g;
        double t[3] = {Kv[0] + G[0], Kv[1] + G[1], Kv[2] + G[2]};
        matvec3(sK->Rmat, t, false, GK + 3 * g);
    }
    for (int64_t g = 0; g < npwKp; ++g)
    {
        const double* G = wfcs[ik].gvec + 3 * g;
        double t[3] = {Kpv[0] + G[0], Kpv[1] + G[1], Kpv[2] + G[2]};
        matvec3(sKp->Rmat, t, false, GKp + 3 * g);
    }
    /* Ylm tables, :178-211 */
    double* YK = malloc(sizeof(double) * npwK * nlmax);
    double* YKp = malloc(sizeof(double) * npwKp * nlmax);
    for (int l = 0; l <= lmax; ++l)
        for (int im = 0; im <= 2 * l; ++im)
        {
            int m = im - l;
            int64_t id = l * l + im;
            for (int64_t g = 0; g < npwK; ++g) YK[id * npwK + g] = orc_Ylm(l, m, GK + 3 * g);
            for (int64_t g = 0; g < npwKp; ++g) YKp[id * npwKp + g] = orc_Ylm(l, m, GKp + 3 * g);
        }
    /* SU2 + translation, :216-245 */
    cplx su2K[4] = {1, 0, 0, 1}, su2Kp[4] = {1, 0, 0, 1};
    if ((err = orc_SU2mat(sK->Rmat, nspinor, 0, sK->time_rev, (double*)su2K))) goto o0;
    if ((err = orc_SU2mat(sKp->Rmat, nspinor, 0, sKp->time_rev, (double*)su2Kp))) goto o0;
    const double kz[3] = {0, 0, 0};
    for (int64_t s = 0; s < nspin * nb; ++s)
    {
        su2rotate(nspinor, npwK, su2K, wK + s * nspinor * npwK);
        apply_trans_wfc(sK->tau, kz, nspinor, npwK, GK, wK + s * nspinor * npwK, false);
    }
    for (int64_t s = 0; s < nspin * nb; ++s)
    {
        su2rotate(nspinor, npwKp, su2Kp, wKp + s * nspinor * npwKp);
        apply_trans_wfc(sKp->tau, kz, nspinor, npwKp, GKp, wKp + s * nspinor * npwKp, false);
    }

    const int64_t nlj = pp->nltimesj;
    const int64_t bstride = nspin * nb * nspinor * 4;
    const int64_t tl = nlj * (2 * lmax + 1);
    cplx* bbK = calloc(tl * bstride, sizeof(cplx));
    cplx* bbKp = calloc(tl * bstride, sizeof(cplx));
    cplx* bK = calloc(4 * tl * npwK, sizeof(cplx));
    cplx* bKp = calloc(4 * tl * npwKp, sizeof(cplx));
    const int64_t estride = 3 * nb * nb * nspin;
    cplx* ebuf = calloc(natom * estride, sizeof(cplx));

    for (int64_t ia = 0; ia < natom; ++ia)
    {
        const int64_t it = lat->atom_type[ia];
        const double* tau = lat->atomic_pos + 3 * ia;
        int64_t iK = 0, iKp = 0;
        for (int64_t li = 0; li < nlj; ++li)
        {
            const int l = (int)rint(pp->PP_table[ntype * 3 * li + it * 3] - 1);
            if (l < 0) continue;
            const double sg = pp->Fsign[li * ntype + it];
            const double* FKt = FK + (li * ntype + it) * npwK;
            const double* FKpt = FKp + (li * ntype + it) * npwKp;
            for (int m1 = 0; m1 <= 2 * l; ++m1)
            {
                const double* Yt = YK + (l * l + m1) * npwK;
                cplx* b = bK + iK * 4 * npwK;
                for (int64_t g = 0; g < npwK; ++g)
                {
                    const double* Gp = GK + 3 * g;
                    double td = Gp[0] * tau[0] + Gp[1] * tau[1] + Gp[2] * tau[2];
                    b[g] = FKt[g] * cexp(-I * 2 * RPI * td) * Yt[g];
                    b[npwK + g] = b[g] * Gp[0] * sg;
                    b[2 * npwK + g] = b[g] * Gp[1] * sg;
                    b[3 * npwK + g] = b[g] * Gp[2] * sg;
                }
                ++iK;
            }
            for (int m2 = 0; m2 <= 2 * l; ++m2)
            {
                const double* Yt = YKp + (l * l + m2) * npwKp;
                cplx* b = bKp + iKp * 4 * npwKp;
                for (int64_t g = 0; g < npwKp; ++g)
                {
                    const double* Gp = GKp + 3 * g;
                    double td = Gp[0] * tau[0] + Gp[1] * tau[1] + Gp[2] * tau[2];
                    b[g] = FKpt[g] * cexp(I * 2 * RPI * td) * conj(Yt[g]);
                    b[npwKp + g] = -b[g] * Gp[0] * sg;
                    b[2 * npwKp + g] = -b[g] * Gp[1] * sg;
                    b[3 * npwKp + g] = -b[g] * Gp[2] * sg;
                }
                ++iKp;
            }
        }
        char cK = sK->time_rev ? 'T' : 'C';
        char cKp = sKp->time_rev ? 'C' : 'T';
        /* :412-418 */
        zgemm_rm('N', cK, bK, wK, bbK, 1.0, 0.0, npwK, npwK, nspin * nspinor * nb,
                 4 * iK, nspin * nspinor * nb, npwK);
        zgemm_rm('N', cKp, bKp, wKp, bbKp, 1.0, 0.0, npwKp, npwKp,
                 nspin * nspinor * nb, 4 * iKp, nspin * nspinor * nb, npwKp);
        if (iK != iKp)
        {
            err = LELPH_ERR_ARG;
            goto o1;
        }
        /* :465-537 */
        cplx* et = ebuf + ia * estride;
        int64_t ilc = 0;
        for (int64_t li = 0; li < nlj; ++li)
        {
            int l = (int)rint(pp->PP_table[ntype * 3 * li + it * 3] - 1);
            int j = (int)rint(pp->PP_table[ntype * 3 * li + it * 3 + 1]);
            if (l < 0) continue;
            double soc = 1.0;
            if (j != 0 && !lat->is_soc_present) soc = (j + 1.0) / (4.0 * l + 2.0);
            const int tlp1 = 2 * l + 1;
            const cplx* ifc = NULL;
            if (pp->fCoeff) ifc = (const cplx*)pp->fCoeff[ntype * li + it];
            for (int m1 = 0; m1 <= 2 * l; ++m1)
            {
                const cplx* bpK = bbK + (ilc + m1) * bstride;
                for (int m2 = 0; m2 <= 2 * l; ++m2)
                {
                    const cplx* bpKp = bbKp + (ilc + m2) * bstride;
                    if (!ifc && m1 != m2) continue;
                    cplx tf[4] = {soc, 0.0, 0.0, soc};
                    const cplx* f = tf;
                    if (ifc) f = ifc + (tlp1 * m1 + m2) * nspinor * nspinor;
                    sum_VNL_KKp(bpK, bpKp, f, nspin, nb, nspinor, et);
                }
            }
            ilc += 2 * l + 1;
        }
    }
    {
        /* cart -> mode, :541-560 */
        const int64_t nmodes = lat->nmodes, es = nspin * nb * nb;
        const cplx pre = -2 * RPI * I * RE2;
        zgemm_rm('N', 'N', eigVec, ebuf, elph_kq, pre, 1.0, natom * 3, es, es,
                 nmodes, es, natom * 3);
    }
o1:
    free(ebuf);
    free(bbK);
    free(bbKp);
    free(bK);
    free(bKp);
o0:
    free(YK);
    free(YKp);
    free(wK);
    free(wKp);
    free(GK);
    free(GKp);
    return err;
}

/* ======================================================================
 * V_loc(G) table: Vloc_Gspace (src/dvloc/VlocG_atom.c:10-57),
 * create_vlocg_table (src/dvloc/vlocg_table.c:16-154)
 * ==================================================================== */
double orc_Vloc_Gspace(double* work, char cutoff, double Gn, int64_t ngrid,
                       const double* Vl, const double* r, const double* rab,
                       double Z, double eta, double vol)
{
    const double se = sqrt(eta);
    for (int64_t i = 0; i < ngrid; ++i)
    {
        double rad = r[i], fac;
        fac = (Gn < REPS) ? rad : sin(rad * Gn) / Gn;
        if (cutoff == '3' && Gn < REPS)
            fac *= (rad * Vl[i] + RE2 * Z);
        else
            fac *= (rad * Vl[i] + RE2 * Z * erf(rad * se));
        work[i] = fac;
    }
    return 4 * RPI * orc_simpson(work, rab, ngrid) / vol;
}

int orc_create_vlocg_table(const lelph_lattice* lat, const lelph_pseudo* pp,
                           orc_vloc_table* tab)
{
    const double* b = lat->blat_vec;
    double bi[3];
    bi[0] = sqrt(b[0] * b[0] + b[3] * b[3] + b[6] * b[6]);
    bi[1] = sqrt(b[1] * b[1] + b[4] * b[4] + b[7] * b[7]);
    bi[2] = sqrt(b[2] * b[2] + b[5] * b[5] + b[8] * b[8]);
    double bmin = bi[0], bmax = bi[0], Nmax = (double)lat->fft_dims[0];
    for (int i = 0; i < 3; ++i)
    {
        if (bi[i] < bmin) bmin = bi[i];
        if (bi[i] > bmax) bmax = bi[i];
        if (lat->fft_dims[i] > Nmax) Nmax = (double)lat->fft_dims[i];
    }
    if (bmin > 0.01) bmin = 0.01;
    double gmax = (1.7 * (Nmax + 4.0) + 3 * fabs(pp->qmax_abs)) * bmax;
    int64_t npts = (int64_t)ceil(gmax / bmin);
    double gmin = 0.0;
    const int64_t nt = pp->ntype;
    tab->npts = npts;
    tab->g_co = malloc(sizeof(double) * npts);
    tab->vlocg = malloc(sizeof(double) * npts * nt);
    tab->vploc = malloc(sizeof(double) * npts * nt);
    double diff = (gmax - gmin) / (npts - 1.0);
    for (int64_t i = 0; i < npts; ++i) tab->g_co[i] = gmin + i * diff;
    tab->dg = diff;
    int64_t ngmax = 0;
    for (int64_t t = 0; t < nt; ++t)
        if (pp->loc_pseudo[t].ngrid > ngmax) ngmax = pp->loc_pseudo[t].ngrid;
    for (int64_t t = 0; t < nt; ++t)
    {
        const lelph_local_pseudo* lp = pp->loc_pseudo + t;
        double* y = tab->vlocg + t * npts;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < npts; ++i)
        {
            double* w = malloc(sizeof(double) * (ngmax + 1));
            y[i] = orc_Vloc_Gspace(w, lat->dimension, tab->g_co[i], lp->ngrid,
                                   lp->Vloc_atomic, lp->r_grid, lp->rab_grid,
                                   lp->Zval, 1.0, lat->volume);
            free(w);
        }
        orc_prepare_spline(npts, tab->g_co, y, tab->vploc + t * npts);
    }
    return 0;
}

void orc_free_vloc_table(orc_vloc_table* t)
{
    free(t->g_co);
    free(t->vlocg);
    free(t->vploc);
    t->g_co = t->vlocg = t->vploc = NULL;
}

/* get_miller_idx (src/common/numerical_func.c:436-449) */
static int64_t miller(int64_t i, int64_t N)
{
    int64_t mid = (N - 1) / 2 + 1;
    return i < mid ? i : (i - mid) - N / 2;
}

/* dVlocq (src/dvloc/dVbare_loc.c:16-231) */
int orc_dVlocq(const double* qpt, const lelph_lattice* lat,
               const lelph_pseudo* pp, const orc_vloc_table* tab,
               const double* eigVecd, double* Vlocrd)
{
    const cplx* eig = (const cplx*)eigVecd;
    cplx* Vlocr = (cplx*)Vlocrd;
    const double* latv = lat->alat_vec;
    const int64_t ntype = pp->ntype, natom = lat->natom, nm = lat->nmodes;
    const int64_t Nx = lat->fft_dims[0], Ny = lat->fft_dims[1], Nz = lat->fft_dims[2];
    const int64_t nf = Nx * Ny * Nz;
    const double vol = lat->volume, eta = 1.0;
    const double* blat = lat->blat_vec;
    const int64_t npts = tab->npts;
    cplx* VG = calloc(nm * nf, sizeof(cplx));
    double* Vt = malloc(sizeof(double) * ntype);
    cplx* Vz = malloc(sizeof(cplx) * Nz * natom * 3);
    int err = 0;
    for (int64_t ig = 0; ig < Nx * Ny && !err; ++ig)
    {
        const int64_t ix = ig / Ny, jy = ig % Ny;
        for (int64_t kz = 0; kz < Nz; ++kz)
        {
            double qG[3] = {miller(ix, Nx) + qpt[0], miller(jy, Ny) + qpt[1],
                            miller(kz, Nz) + qpt[2]};
            double qc[3];
            matvec3(blat, qG, false, qc);
            double qn = sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2]);
            if (qn > tab->g_co[npts - 1])
            {
                err = LELPH_ERR_RANGE;
                break;
            }
            int64_t gi = (int64_t)floor(qn / tab->dg);
            cplx* tp = Vz + kz * natom * 3;
            double cf = 1;
            if (lat->dimension == '2')
            {
                double Gp = sqrt(qc[0] * qc[0] + qc[1] * qc[1]);
                double qGp = latv[8] * Gp;
                double cs = cos(qc[2] * latv[8] * 0.5);
                if (Gp > REPS) cs = cs - sin(qc[2] * latv[8] * 0.5) * qc[2] / Gp;
                cf = 1 - exp(-qGp * 0.5) * cs;
            }
            for (int64_t t = 0; t < ntype; ++t)
            {
                Vt[t] = orc_spline_interpolate(qn, gi, tab->g_co, tab->vlocg + t * npts,
                                               tab->vploc + t * npts);
                if (qn >= REPS)
                    Vt[t] -= (4 * RPI * RE2) * pp->loc_pseudo[t].Zval *
                             exp(-qn * qn * 0.25 / eta) * cf / (qn * qn * vol);
            }
            for (int64_t ia = 0; ia < natom; ++ia)
            {
                const int64_t t = lat->atom_type[ia];
                const double* p = lat->atomic_pos + 3 * ia;
                double qd = qc[0] * p[0] + qc[1] * p[1] + qc[2] * p[2];
                cplx fac = -I * cexp(-I * qd);
                fac *= Vt[t];
                tp[ia * 3] = fac * qc[0];
                tp[ia * 3 + 1] = fac * qc[1];
                tp[ia * 3 + 2] = fac * qc[2];
            }
        }
        if (err) break;
        zgemm_rm('N', 'T', eig, Vz, VG + Nz * ig, 1.0, 0.0, nm, nm, nf, nm, Nz, nm);
    }
    free(Vt);
    free(Vz);
    if (!err)
    {
        memcpy(Vlocr, VG, sizeof(cplx) * nm * nf);
        for (int64_t v = 0; v < nm && !err; ++v) err = orc_fft3d((double*)(Vlocr + v * nf), lat->fft_dims, +1);
    }
    free(VG);
    return err;
}

/* add_dvscf_qe (src/dvloc/add_dvscf.c:12-92) */
void orc_add_dvscf_qe(double* dVscfd, const double* dVlocd,
                      const lelph_lattice* lat)
{
    cplx* dVscf = (cplx*)dVscfd;
    const cplx* dVloc = (const cplx*)dVlocd;
    const int64_t nmag = lat->nmag, nm = lat->nmodes;
    const int64_t nf = lat->fft_dims[0] * lat->fft_dims[1] * lat->fft_dims[2];
    const int64_t mi = nmag == 2 ? 2 : 1;
    for (int64_t v = 0; v < nm; ++v)
    {
        cplx* m = dVscf + v * nf * nmag;
        const cplx* b = dVloc + v * nf;
        for (int64_t im = 0; im < mi; ++im)
            for (int64_t i = 0; i < nf; ++i) m[im * nf + i] += b[i];
        if (nmag == 4)
            for (int64_t i = 0; i < nf; ++i)
            {
                cplx V = m[i], Bx = m[nf + i], By = m[2 * nf + i], Bz = m[3 * nf + i];
                m[i] = V + Bz;
                m[nf + i] = Bx - I * By;
                m[2 * nf + i] = Bx + I * By;
                m[3 * nf + i] = V - Bz;
            }
    }
}

/* read_dvscf_qe's pattern -> mode rotation (src/io/qe/read_dvscf.c:119-158):
 * out = eig @ pats^H @ dvscf over the mode index. */
void orc_dvscf_pattern_to_mode(double* dvd, const double* eigd,
                               const double* patd, const lelph_lattice* lat)
{
    cplx* dv = (cplx*)dvd;
    const cplx* eig = (const cplx*)eigd;
    const cplx* pat = (const cplx*)patd;
    const int64_t nm = lat->nmodes;
    const int64_t nf = lat->fft_dims[0] * lat->fft_dims[1] * lat->fft_dims[2];
    const int64_t ld = lat->nmag * nf;
    cplx* ep = calloc(nm * nm, sizeof(cplx));
    zgemm_rm('N', 'C', eig, pat, ep, 1.0, 0.0, nm, nm, nm, nm, nm, nm);
    cplx* col = malloc(sizeof(cplx) * nm);
    for (int64_t j = 0; j < ld; ++j)
    {
        for (int64_t v = 0; v < nm; ++v)
        {
            cplx a = 0;
            for (int64_t u = 0; u < nm; ++u) a += ep[v * nm + u] * dv[u * ld + j];
            col[v] = a;
        }
        for (int64_t v = 0; v < nm; ++v) dv[v * ld + j] = col[v];
    }
    free(col);
    free(ep);
}

/* elph_q_rotate (src/symmetries/elphq_expand.c:16-73) */
void orc_elph_q_rotate(const double* Dld, const double* gd, const double* Drd,
                       int nmodes, int nspin, int nb, int tim_rev,
                       double* gSqd)
{
    const cplx* Dl = (const cplx*)Dld;
    const cplx* g = (const cplx*)gd;
    const cplx* Dr = (const cplx*)Drd;
    cplx* gSq = (cplx*)gSqd;
    cplx* t = calloc((int64_t)nb * nb, sizeof(cplx));
    char gc = tim_rev ? 'C' : 'T';
    for (int64_t s = 0; s < nspin; ++s)
    {
        const cplx* dr = Dr + s * nb * nb;
        const cplx* dl = Dl + s * nb * nb;
        for (int64_t m = 0; m < nmodes; ++m)
        {
            int64_t sh = (int64_t)nb * nb * (s + nspin * m);
            zgemm_rm(gc, 'C', g + sh, dr, t, 1.0, 0.0, nb, nb, nb, nb, nb, nb);
            zgemm_rm('T', 'T', t, dl, gSq + sh, 1.0, 0.0, nb, nb, nb, nb, nb, nb);
        }
    }
    free(t);
}

/* find_kidx_in_list / get_KplusQ_idxs (src/common/numerical_func.c:471-524) */
static int64_t find_kidx(int64_t nk, const double* kl, const double* k)
{
    for (int64_t i = 0; i < nk; ++i)
    {
        double s = 0;
        for (int j = 0; j < 3; ++j)
        {
            double d = kl[3 * i + j] - k[j];
            d = d - rint(d);
            s += d * d;
        }
        if (sqrt(s) < REPS) return i;
    }
    return -1;
}

int orc_get_KplusQ_idxs(int64_t nbz, const double* kp, const double* q, int* idx)
{
    for (int64_t i = 0; i < nbz; ++i)
    {
        double kq[3] = {kp[3 * i] + q[0], kp[3 * i + 1] + q[1], kp[3 * i + 2] + q[2]};
        idx[i] = (int)find_kidx(nbz, kp, kq);
        if (idx[i] < 0) return LELPH_ERR_NOKQ;
    }
    return 0;
}

/* the k loop of compute_and_write_elphq (src/elph/compute_elph_matq.c:
 * 20-234), without the netCDF writes and the star unfolding. */
int orc_compute_elphq(const double* qpt, const orc_wfc* wfcs,
                      const lelph_lattice* lat, const lelph_pseudo* pp,
                      const double* eigVec, const double* dVscf, int non_loc,
                      int kminusq, int64_t k_first, int64_t k_count,
                      double* g_out)
{
    const int64_t nbz = lat->nkpts_BZ;
    int* kq = malloc(sizeof(int) * nbz);
    double qt[3] = {qpt[0], qpt[1], qpt[2]};
    if (kminusq)
        for (int i = 0; i < 3; ++i) qt[i] = -qt[i];
    int err = orc_get_KplusQ_idxs(nbz, lat->kpt_fullBZ_crys, qt, kq);
    if (err)
    {
        free(kq);
        return err;
    }
    const int64_t gsz = (int64_t)lat->nmodes * lat->nspin * lat->nbnds * lat->nbnds * 2;
    if (k_count < 0) k_count = nbz - k_first;
    int errs = 0;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1) reduction(| : errs)
    for (int64_t ii = 0; ii < k_count; ++ii)
    {
        int64_t i = k_first + ii;
        int ik_ = (int)i, ikq_ = kq[i];
        if (kminusq)
        {
            int t = ik_;
            ik_ = ikq_;
            ikq_ = t;
        }
        int ik = lat->kmap[2 * ik_], ksym = lat->kmap[2 * ik_ + 1];
        int ikq = lat->kmap[2 * ikq_], kqsym = lat->kmap[2 * ikq_ + 1];
        double* g = g_out + ii * gsz;
        int e = orc_elphLocal(qpt, wfcs, lat, ikq, ik, kqsym, ksym, dVscf, g);
        if (!e && non_loc)
            e = orc_add_elphNonLocal(wfcs, lat, pp, ikq, ik, kqsym, ksym, eigVec, g);
        errs |= e;
    }
    free(kq);
    return errs;
}

/* ---- exported for the parity tests: the static helpers above ---- */
void orc_su2rotate(int nspinor, int64_t npw, const double* su2, double* wfc)
{
    su2rotate(nspinor, npw, (const cplx*)su2, (cplx*)wfc);
}

void orc_apply_trans_wfc(const double* tr, const double* kvec, int64_t nsets,
                         int64_t npw, const double* gvecs, double* wfc,
                         int conjugate)
{
    apply_trans_wfc(tr, kvec, nsets, npw, gvecs, (cplx*)wfc, conjugate != 0);
}

void orc_sum_VNL_KKp(const double* K, const double* Kp, const double* fcoeff,
                     int64_t nspin, int64_t nb, int64_t nspinor, double* out)
{
    sum_VNL_KKp((const cplx*)K, (const cplx*)Kp, (const cplx*)fcoeff, nspin,
                nb, nspinor, (cplx*)out);
}

/* ======================================================================
 * find_gvecs_idxs (src/wfc/gsort.c:49-95): index of every search vector in
 * the list (G compared after rint, ordered x, then y, then z), -1 if absent;
 * uses the z-stick ordering of sort_pw above.
 * ==================================================================== */
void orc_find_gvecs_idxs(int64_t nsearch, const double* search, int64_t npw,
                         const double* gvecs, int64_t* out)
{
    gkey* k = malloc(sizeof(gkey) * (npw > 0 ? npw : 1));
    for (int64_t i = 0; i < npw; ++i)
    {
        k[i].x = (long long)rint(gvecs[3 * i]);
        k[i].y = (long long)rint(gvecs[3 * i + 1]);
        k[i].z = (long long)rint(gvecs[3 * i + 2]);
        k[i].i = i;
    }
    qsort(k, npw, sizeof(gkey), gkey_cmp);
    for (int64_t s = 0; s < nsearch; ++s)
    {
        gkey q = {(long long)rint(search[3 * s]), (long long)rint(search[3 * s + 1]),
                  (long long)rint(search[3 * s + 2]), 0};
        gkey* hit = bsearch(&q, k, npw, sizeof(gkey), gkey_cmp);
        out[s] = hit ? hit->i : -1;
    }
    free(k);
}

/* ======================================================================
 * electronic_reps (src/symmetries/elec_reps.c:16-413), one rank:
 * Dkmn[s][b][a] = <S2 k2, b | U(R) U(S1) | k1, a> with k_BZ = S1 k1 and
 * R k_BZ = S2 k2 (up to a G vector).
 * ==================================================================== */
int orc_electronic_reps(const orc_wfc* wfcs, const lelph_lattice* lat,
                        const double* Rsym, const double* tauR, int tim_revR,
                        int64_t ikBZ, double* Dkmn_d)
{
    cplx* Dkmn = (cplx*)Dkmn_d;
    const double* alat = lat->alat_vec;
    const double* blat = lat->blat_vec;
    const int64_t nspin = lat->nspin, nb = lat->nbnds, nspinor = lat->nspinor;
    /* R k_BZ in crystal coordinates (:34-46); k_BZ cartesian from its
     * crystal coordinates, without 2 pi */
    const double* kbc = lat->kpt_fullBZ_crys + 3 * ikBZ;
    double kbz[3];
    for (int i = 0; i < 3; ++i)
        kbz[i] = (blat[3 * i] * kbc[0] + blat[3 * i + 1] * kbc[1] + blat[3 * i + 2] * kbc[2]) /
                 (2 * RPI);
    double Rk_tmp[3], Rk_vec[3];
    matvec3(Rsym, kbz, false, Rk_tmp);
    matvec3(alat, Rk_tmp, true, Rk_vec);
    const int64_t iRk = find_kidx(lat->nkpts_BZ, lat->kpt_fullBZ_crys, Rk_vec);
    if (iRk < 0) return LELPH_ERR_SYM;
    /* :69-80 */
    const int ik1 = lat->kmap[2 * ikBZ], iS1 = lat->kmap[2 * ikBZ + 1];
    const int ik2 = lat->kmap[2 * iRk], iS2 = lat->kmap[2 * iRk + 1];
    const double* S1 = lat->syms[iS1].Rmat;
    const double* tau1 = lat->syms[iS1].tau;
    const int tr1 = lat->syms[iS1].time_rev;
    const double* S2 = lat->syms[iS2].Rmat;
    const double* tau2 = lat->syms[iS2].tau;
    const int tr2 = lat->syms[iS2].time_rev;
    const double* k1 = lat->kpt_iredBZ + 3 * ik1;
    const double* k2 = lat->kpt_iredBZ + 3 * ik2;
    /* :83-103 */
    double kcrys[3], Rkcrys[3], t3[3];
    matvec3(S1, k1, false, t3);
    matvec3(alat, t3, true, kcrys);
    double tau1c[3], tauRc[3], tau2c[3];
    matvec3(blat, tau1, true, tau1c);
    matvec3(blat, tauR, true, tauRc);
    matvec3(blat, tau2, true, tau2c);
    for (int i = 0; i < 3; ++i)
    {
        tau1c[i] /= 2 * RPI;
        tauRc[i] /= 2 * RPI;
        tau2c[i] /= 2 * RPI;
    }
    const int64_t npw1 = wfcs[ik1].npw, npw2 = wfcs[ik2].npw;
    const cplx* w1 = (const cplx*)wfcs[ik1].wfc;
    const cplx* w2 = (const cplx*)wfcs[ik2].wfc;
    /* :121-127 */
    cplx su1[4] = {1, 0, 0, 1}, suR[4] = {1, 0, 0, 1}, su2[4] = {1, 0, 0, 1};
    int err;
    if ((err = orc_SU2mat(S1, nspinor, 0, tr1, (double*)su1))) return err;
    if ((err = orc_SU2mat(Rsym, nspinor, 0, tim_revR, (double*)suR))) return err;
    if ((err = orc_SU2mat(S2, nspinor, 0, tr2, (double*)su2))) return err;
    /* :129-176 */
    double* G_S1k1 = malloc(sizeof(double) * 3 * npw1);
    double* G_RS1k1 = malloc(sizeof(double) * 3 * npw1);
    double* G_S2k2 = malloc(sizeof(double) * 3 * npw2);
    double RS1[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            RS1[3 * i + j] = Rsym[3 * i] * S1[j] + Rsym[3 * i + 1] * S1[3 + j] +
                             Rsym[3 * i + 2] * S1[6 + j];
    double ulm[3], Rkc[3];
    matvec3(S2, k2, false, ulm);
    matvec3(RS1, k1, false, Rkc);
    matvec3(alat, Rkc, true, Rkcrys);
    for (int i = 0; i < 3; ++i) ulm[i] -= Rkc[i];
    rotateGvecs(wfcs[ik1].gvec, S1, npw1, alat, true, NULL, G_S1k1);
    rotateGvecs(wfcs[ik1].gvec, RS1, npw1, alat, true, NULL, G_RS1k1);
    rotateGvecs(wfcs[ik2].gvec, S2, npw2, alat, true, ulm, G_S2k2);
    /* :256-262 */
    int64_t* idx = malloc(sizeof(int64_t) * (npw2 > 0 ? npw2 : 1));
    orc_find_gvecs_idxs(npw2, G_S2k2, npw1, G_RS1k1, idx);
    /* :304-373 */
    const int64_t nsets = nspin * nb, L = nspinor * npw1;
    cplx* wRS = malloc(sizeof(cplx) * nsets * L);
    cplx* wS2 = calloc(nsets * L, sizeof(cplx));
    for (int64_t s = 0; s < nsets; ++s)
    {
        cplx* a = wRS + s * L;
        memcpy(a, w1 + s * L, sizeof(cplx) * L);
        su2rotate(nspinor, npw1, su1, a);
        apply_trans_wfc(tau1c, kcrys, nspinor, npw1, G_S1k1, a, tr1);
        su2rotate(nspinor, npw1, suR, a);
        apply_trans_wfc(tauRc, Rkcrys, nspinor, npw1, G_RS1k1, a, tim_revR);
        cplx* b = wS2 + s * L;
        for (int64_t sp = 0; sp < nspinor; ++sp)
            for (int64_t i = 0; i < npw2; ++i)
                if (idx[i] >= 0) b[sp * npw1 + idx[i]] = w2[s * nspinor * npw2 + sp * npw2 + i];
        su2rotate(nspinor, npw1, su2, b);
        apply_trans_wfc(tau2c, Rkcrys, nspinor, npw1, G_RS1k1, b, !tr2);
    }
    /* :377-394 */
    for (int64_t is = 0; is < nspin; ++is)
        zgemm_rm('N', 'T', wS2 + is * nb * L, wRS + is * nb * L, Dkmn + is * nb * nb, 1.0,
                 0.0, L, L, nb, nb, nb, L);
    free(G_S1k1);
    free(G_RS1k1);
    free(G_S2k2);
    free(idx);
    free(wRS);
    free(wS2);
    return 0;
}
