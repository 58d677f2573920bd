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
