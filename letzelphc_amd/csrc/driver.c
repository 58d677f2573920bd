/*
 * driver.c -- host-side C mirror of the reference's k loop and helpers:
 *   compute_and_write_elphq k loop (src/elph/compute_elph_matq.c:20-234),
 *   get_KplusQ_idxs / find_kidx_in_list (src/common/numerical_func.c:471-524),
 *   elph_q_rotate (src/symmetries/elphq_expand.c:16-73).
 * Plain C99 over the C-ABI of lelph_hip.h.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lelph_hip.h"

#define DRV_EPS 1e-5 /* ELPH_EPS, src/common/constants.h:9 */

/* full-BZ grid of a context (lelph_hip.hip) */
int lelph_internal_grid(const lelph_ctx* ctx, int64_t* nbz, const double** kcrys,
                        const int** kmap);

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
        if (sqrt(s) < DRV_EPS) return i;
    }
    return -1;
}

int lelph_get_kplusq_idxs(int64_t nbz, const double* kp, const double q[3], int* idx)
{
    if (nbz < 1 || !kp || !q || !idx) return LELPH_ERR_ARG;
    for (int64_t i = 0; i < nbz; ++i)
    {
        double kq[3] = {kp[3 * i] + q[0], kp[3 * i + 1] + q[1], kp[3 * i + 2] + q[2]};
        int64_t j = find_kidx(nbz, kp, kq);
        if (j < 0) return LELPH_ERR_NOKQ;
        idx[i] = (int)j;
    }
    return LELPH_OK;
}

int lelph_compute_elphq(lelph_ctx* ctx, const double qpt[3], int non_loc, int kminusq,
                        double* g_out)
{
    if (!ctx || !qpt || !g_out) return LELPH_ERR_ARG;
    int64_t nbz = 0;
    const double* kcrys = NULL;
    const int* kmap = NULL;
    int s0 = lelph_internal_grid(ctx, &nbz, &kcrys, &kmap);
    if (s0 != LELPH_OK) return s0;
    int* kq = malloc(sizeof(int) * nbz);
    lelph_pair* pairs = malloc(sizeof(lelph_pair) * nbz);
    if (!kq || !pairs)
    {
        free(kq);
        free(pairs);
        return LELPH_ERR_ALLOC;
    }
    double qt[3] = {qpt[0], qpt[1], qpt[2]};
    if (kminusq)
        for (int i = 0; i < 3; ++i) qt[i] = -qt[i];
    int s = lelph_get_kplusq_idxs(nbz, kcrys, qt, kq);
    if (s == LELPH_OK)
    {
        for (int64_t i = 0; i < nbz; ++i)
        {
            int idx_k = (int)i, idx_kq = kq[i];
            if (kminusq)
            {
                int t = idx_k;
                idx_k = idx_kq;
                idx_kq = t;
            }
            pairs[i].ik = kmap[2 * idx_k];
            pairs[i].ksym = kmap[2 * idx_k + 1];
            pairs[i].ikq = kmap[2 * idx_kq];
            pairs[i].kqsym = kmap[2 * idx_kq + 1];
        }
        /* elphLocal uses the unflipped q even for kminusq
         * (src/elph/compute_elph_matq.c:135-136) */
        s = lelph_gkq_batch(ctx, qpt, nbz, pairs, non_loc, g_out);
    }
    free(kq);
    free(pairs);
    return s;
}

/* row-major C = op(A) op(B), n x n complex */
static void zmm(char ta, char tb, const double complex* A, const double complex* B,
                double complex* C, int n)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
        {
            double complex acc = 0;
            for (int l = 0; l < n; ++l)
            {
                double complex a = (ta == 'N') ? A[i * n + l] : A[l * n + i];
                if (ta == 'C') a = conj(a);
                double complex b = (tb == 'N') ? B[l * n + j] : B[j * n + l];
                if (tb == 'C') b = conj(b);
                acc += a * b;
            }
            C[i * n + j] = acc;
        }
}

int lelph_elph_q_rotate(const double* Dl, const double* gq, const double* Dr, int nmodes,
                        int nspin, int nb, int tim_rev, double* gSq)
{
    if (!Dl || !gq || !Dr || !gSq || nmodes < 1 || nspin < 1 || nb < 1) return LELPH_ERR_ARG;
    const double complex* DL = (const double complex*)Dl;
    const double complex* DR = (const double complex*)Dr;
    const double complex* G = (const double complex*)gq;
    double complex* O = (double complex*)gSq;
    double complex* t = malloc(sizeof(double complex) * (size_t)nb * nb);
    if (!t) return LELPH_ERR_ALLOC;
    const char gc = tim_rev ? 'C' : 'T';
    for (int s = 0; s < nspin; ++s)
        for (int m = 0; m < nmodes; ++m)
        {
            const size_t sh = (size_t)nb * nb * (s + (size_t)nspin * m);
            zmm(gc, 'C', G + sh, DR + (size_t)s * nb * nb, t, nb);
            zmm('T', 'T', t, DL + (size_t)s * nb * nb, O + sh, nb);
        }
    free(t);
    return LELPH_OK;
}

/* build_wigner_seitz_vectors (src/interpolation/wigner_seitz.c:27-202).
 * The reference queries a kd-tree of the superlattice points n_i * grid_i *
 * a_i, n_i in [-6, 6]; the same 13^3 candidates are scanned here: the
 * nearest distance d to r_m - r_n - R, then every candidate within
 * d + 1e-10 + eps |r_m - r_n - R| + 1e-8.  Vectors are stored in crystal
 * units of the unit cell (rounded), degenerate ones in scan order. */
int lelph_wigner_seitz(const int64_t grid[3], const double lat[9], double eps,
                       const double* rvec_m, int64_t nrvec_m, const double* rvec_n,
                       int64_t nrvec_n, int64_t cap, int64_t* ws_vecs, int64_t* ws_degen,
                       int64_t* nws)
{
    enum { S = 6, SIDE = 2 * S + 1, NP = SIDE * SIDE * SIDE };
    static const double zero[3] = {0, 0, 0};
    if (!grid || !lat || !ws_degen || !nws || cap < 0 || (cap && !ws_vecs)) return LELPH_ERR_ARG;
    for (int i = 0; i < 3; ++i)
        if (grid[i] < 1) return LELPH_ERR_ARG;
    if (!rvec_m) { rvec_m = zero; nrvec_m = 1; }
    if (!rvec_n) { rvec_n = zero; nrvec_n = 1; }
    if (nrvec_m < 1 || nrvec_n < 1) return LELPH_ERR_ARG;
    /* inverse of the lattice matrix: crystal = inv(A) cart */
    const double* a = lat;
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (fabs(det) < 1e-12) return LELPH_ERR_ARG;
    double inv[9];
    inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
    double* pts = malloc(sizeof(double) * 3 * NP);
    double* d2 = malloc(sizeof(double) * NP);
    int64_t (*cr)[3] = malloc(sizeof(int64_t) * 3 * NP);
    if (!pts || !d2 || !cr)
    {
        free(pts);
        free(d2);
        free(cr);
        return LELPH_ERR_ALLOC;
    }
    int np = 0;
    for (int i = -S; i <= S; ++i)
        for (int j = -S; j <= S; ++j)
            for (int k = -S; k <= S; ++k)
            {
                const double n[3] = {(double)(i * grid[0]), (double)(j * grid[1]),
                                     (double)(k * grid[2])};
                for (int x = 0; x < 3; ++x)
                    pts[3 * np + x] = a[3 * x] * n[0] + a[3 * x + 1] * n[1] + a[3 * x + 2] * n[2];
                for (int x = 0; x < 3; ++x)
                    cr[np][x] = (int64_t)rint(inv[3 * x] * pts[3 * np] +
                                              inv[3 * x + 1] * pts[3 * np + 1] +
                                              inv[3 * x + 2] * pts[3 * np + 2]);
                ++np;
            }
    const int64_t nR = grid[0] * grid[1] * grid[2];
    const int64_t yz = grid[1] * grid[2];
    int64_t nfound = 0;
    int st = LELPH_OK;
    for (int64_t r = 0; r < nR && st == LELPH_OK; ++r)
    {
        int64_t R[3] = {r / yz, (r % yz) / grid[2], r % grid[2]};
        for (int x = 0; x < 3; ++x)
        {
            const int64_t mid = (grid[x] - 1) / 2 + 1;
            R[x] = R[x] < mid ? R[x] : (R[x] - mid) - grid[x] / 2;
        }
        double Rv[3];
        for (int x = 0; x < 3; ++x) Rv[x] = a[3 * x] * R[0] + a[3 * x + 1] * R[1] + a[3 * x + 2] * R[2];
        for (int64_t im = 0; im < nrvec_m && st == LELPH_OK; ++im)
            for (int64_t in = 0; in < nrvec_n; ++in)
            {
                double qp[3];
                for (int x = 0; x < 3; ++x) qp[x] = rvec_m[3 * im + x] - rvec_n[3 * in + x] - Rv[x];
                const double qn = sqrt(qp[0] * qp[0] + qp[1] * qp[1] + qp[2] * qp[2]);
                double best = INFINITY;
                for (int p = 0; p < np; ++p)
                {
                    const double dx = pts[3 * p] - qp[0], dy = pts[3 * p + 1] - qp[1],
                                 dz = pts[3 * p + 2] - qp[2];
                    d2[p] = dx * dx + dy * dy + dz * dz;
                    if (d2[p] < best) best = d2[p];
                }
                const double rad = sqrt(best) + 1e-10 + fabs(eps * qn + 1e-8);
                int64_t cnt = 0;
                for (int p = 0; p < np; ++p)
                {
                    if (d2[p] > rad * rad) continue;
                    if (nfound < cap)
                        for (int x = 0; x < 3; ++x) ws_vecs[3 * nfound + x] = cr[p][x];
                    ++nfound;
                    ++cnt;
                }
                ws_degen[in + im * nrvec_n + r * nrvec_m * nrvec_n] = cnt;
            }
    }
    free(pts);
    free(d2);
    free(cr);
    *nws = nfound;
    return nfound > cap ? LELPH_ERR_RANGE : LELPH_OK;
}
