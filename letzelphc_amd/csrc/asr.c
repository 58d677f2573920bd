/*
 * asr.c -- acoustic sum rules of the force constants and Born charges
 * (src/phonon/asr.c:155-512), host code.
 *
 * The "crystal", "all" and "all_huang" rules of the reference are the
 * orthogonal projection of the real force constants onto the null space of
 * a sparse constraint matrix A (translational 9 natom rows; rotational 9
 * natom more; Huang 15 more; C. Lin et al., npj Comput. Mater. 8, 236
 * (2022)).  The reference reaches it with LSMR on min |A^T l - b|; the
 * projection itself is unique, so here it is computed directly: A is
 * enumerated once as (row, column, value) entries, G = A A^T (nc x nc) and
 * c = A b are formed in long double, G l = c is solved by Cholesky with
 * symmetric pivoting (stopping at G's numerical rank, where the remaining
 * rows are combinations of the kept ones and c lies in G's range), and
 * b - A^T l is the result.
 *
 * Storage follows the reference: only R or -R of each pair is independent
 * (the -R block is the transpose), only the lower triangle at R = 0, the
 * real parts only; the full complex array is rebuilt from them.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lelph_hip.h"

static int64_t miller(int64_t i, int64_t n)
{
    const int64_t mid = (n - 1) / 2 + 1;
    return i < mid ? i : (i - mid) - n / 2;
}

/* the 15 independent (alpha, beta, gamma, delta) of the Huang conditions
 * (asr.c:808-849) */
static const int huang_idx[15][4] = {
    {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 1}, {0, 0, 1, 2}, {0, 0, 2, 2},
    {0, 1, 0, 2}, {0, 1, 1, 1}, {0, 1, 1, 2}, {0, 1, 2, 2}, {0, 2, 1, 1},
    {0, 2, 1, 2}, {0, 2, 2, 2}, {1, 1, 1, 2}, {1, 1, 2, 2}, {1, 2, 2, 2}};

struct entry
{
    int64_t col;
    int32_t row;
    double val;
};

struct elist
{
    struct entry* e;
    int64_t n, cap;
};

static int push(struct elist* L, int32_t row, int64_t col, double val)
{
    if (L->n == L->cap)
    {
        const int64_t nc = L->cap ? 2 * L->cap : 4096;
        struct entry* p = realloc(L->e, sizeof(struct entry) * (size_t)nc);
        if (!p) return LELPH_ERR_ALLOC;
        L->e = p;
        L->cap = nc;
    }
    L->e[L->n].row = row;
    L->e[L->n].col = col;
    L->e[L->n].val = val;
    ++L->n;
    return LELPH_OK;
}

static int by_col_row(const void* a, const void* b)
{
    const struct entry* x = a;
    const struct entry* y = b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return (x->row > y->row) - (x->row < y->row);
}

/* column of Phi(ia, alpha, ja, beta) at independent slot iig (asr.c:750-786):
 * -R slots hold the transpose, R = 0 only its lower triangle */
static int64_t fc_col(int64_t iig, int nm, int processR, int gamma, int ia, int al, int ja,
                      int be)
{
    int64_t i = processR ? 3 * ia + al : 3 * ja + be;
    int64_t j = processR ? 3 * ja + be : 3 * ia + al;
    if (gamma && j > i)
    {
        const int64_t t = i;
        i = j;
        j = t;
    }
    return iig * nm * nm + i * nm + j;
}

int lelph_asr_fc(int mode, const int64_t qgrid[3], int natom, double* frc,
                 const double* atomic_pos, const double* lat_vecs, const int64_t* ws_vecs,
                 int64_t n_ws_vecs, const int64_t* ws_degen)
{
    if (mode == LELPH_ASR_NONE) return LELPH_OK;
    if (!qgrid || natom < 1 || !frc || mode < 0 || mode > LELPH_ASR_ALL_HUANG)
        return LELPH_ERR_ARG;
    if (qgrid[0] < 1 || qgrid[1] < 1 || qgrid[2] < 1) return LELPH_ERR_ARG;
    const int nm = 3 * natom;
    const int64_t ng = qgrid[0] * qgrid[1] * qgrid[2], gyz = qgrid[1] * qgrid[2];
    const int64_t nn = (int64_t)nm * nm;
    if (ng < 1) return LELPH_ERR_ARG;
    if (mode == LELPH_ASR_SIMPLE)
    {
        /* asr.c:293-317: the on-site block takes minus the R-summed row */
        for (int ia = 0; ia < natom; ++ia)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                {
                    double sr = 0, si = 0;
                    const int64_t base = (int64_t)ia * 9 * natom + a * nm + b;
                    for (int64_t g = 0; g < ng; ++g)
                        for (int jb = 0; jb < natom; ++jb)
                        {
                            sr += frc[2 * (g * nn + base + 3 * jb)];
                            si += frc[2 * (g * nn + base + 3 * jb) + 1];
                        }
                    frc[2 * (base + 3 * ia)] -= sr;
                    frc[2 * (base + 3 * ia) + 1] -= si;
                }
        return LELPH_OK;
    }
    if (!atomic_pos || !ws_degen || (mode != LELPH_ASR_CRYSTAL && (!ws_vecs || !lat_vecs)))
        return LELPH_ERR_ARG;
    const int all = mode == LELPH_ASR_ALL || mode == LELPH_ASR_ALL_HUANG;
    const int huang = mode == LELPH_ASR_ALL_HUANG;
    int st = LELPH_ERR_ALLOC;
    int64_t* gmap = malloc(sizeof(int64_t) * (size_t)ng);
    char* doR = malloc((size_t)ng);
    double* b = NULL;
    long double *G = NULL, *cv = NULL, *lam = NULL;
    int64_t* perm = NULL;
    struct elist L = {0};
    if (!gmap || !doR) goto done;
    /* independent R: each pair {R, -R} once (asr.c:325-354) */
    memset(doR, 1, (size_t)ng);
    int64_t nind = 0;
    for (int64_t g = 0; g < ng; ++g)
    {
        if (!doR[g]) continue;
        const int64_t rx = g / gyz, ry = (g % gyz) / qgrid[2], rz = g % qgrid[2];
        const int64_t mg = ((qgrid[2] - rz) % qgrid[2]) + ((qgrid[1] - ry) % qgrid[1]) * qgrid[2] +
                           ((qgrid[0] - rx) % qgrid[0]) * gyz;
        gmap[g] = nind;
        if (mg != g)
        {
            doR[mg] = 0;
            gmap[mg] = nind;
        }
        ++nind;
    }
    const int64_t nx = nind * nn;
    b = malloc(sizeof(double) * (size_t)nx);
    if (!b) goto done;
    for (int64_t g = 0, k = 0; g < ng; ++g)
        if (doR[g])
        {
            for (int64_t i = 0; i < nn; ++i) b[k * nn + i] = frc[2 * (g * nn + i)];
            ++k;
        }
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < i; ++j) b[(int64_t)j * nm + i] = 0.0;
    const int nc = 9 * natom + (all ? 9 * natom : 0) + (huang ? 15 : 0);
    /* the constraint entries (asr_matvec, asr.c:569-748) */
    int64_t iws = 0;
    for (int64_t g = 0; g < ng; ++g)
    {
        const int64_t iig = gmap[g];
        const int pR = doR[g], gam = g == 0;
        const int64_t mx = miller(g / gyz, qgrid[0]), my = miller((g % gyz) / qgrid[2], qgrid[1]),
                      mz = miller(g % qgrid[2], qgrid[2]);
        for (int ia = 0; ia < natom; ++ia)
            for (int ja = 0; ja < natom; ++ja)
            {
                const int64_t dg = ws_degen[(g * natom + ia) * natom + ja];
                if (dg > 0)
                    for (int a = 0; a < 3; ++a)
                        for (int be = 0; be < 3; ++be)
                            if (push(&L, 9 * ia + 3 * a + be, fc_col(iig, nm, pR, gam, ia, a, ja, be),
                                     1.0))
                                goto done;
                if (!all) continue;
                const double inv = 1.0 / (double)dg;
                for (int64_t w = 0; w < dg; ++w, ++iws)
                {
                    if (iws >= n_ws_vecs)
                    {
                        st = LELPH_ERR_RANGE;
                        goto done;
                    }
                    const double R[3] = {(double)(mx + ws_vecs[3 * iws]),
                                         (double)(my + ws_vecs[3 * iws + 1]),
                                         (double)(mz + ws_vecs[3 * iws + 2])};
                    double tRj[3], tRij[3];
                    for (int x = 0; x < 3; ++x)
                    {
                        tRj[x] = lat_vecs[3 * x] * R[0] + lat_vecs[3 * x + 1] * R[1] +
                                 lat_vecs[3 * x + 2] * R[2] + atomic_pos[3 * ja + x];
                        tRij[x] = atomic_pos[3 * ia + x] - tRj[x];
                    }
                    for (int ic = 0; ic < 9; ++ic)
                    {
                        /* (beta, gamma) in {(1,0), (2,0), (2,1)} */
                        const int a = ic / 3, bg = ic % 3;
                        const int be = bg + 1 < 2 ? bg + 1 : 2, ga = bg / 2;
                        const int32_t row = 9 * natom + 9 * ia + ic;
                        if (push(&L, row, fc_col(iig, nm, pR, gam, ia, a, ja, be), inv * tRj[ga]) ||
                            push(&L, row, fc_col(iig, nm, pR, gam, ia, a, ja, ga), -inv * tRj[be]))
                            goto done;
                    }
                    if (huang)
                        for (int ic = 0; ic < 15; ++ic)
                        {
                            const int a = huang_idx[ic][0], be = huang_idx[ic][1],
                                      ga = huang_idx[ic][2], de = huang_idx[ic][3];
                            const int32_t row = 18 * natom + ic;
                            if (push(&L, row, fc_col(iig, nm, pR, gam, ia, a, ja, be),
                                     inv * tRij[ga] * tRij[de]) ||
                                push(&L, row, fc_col(iig, nm, pR, gam, ia, ga, ja, de),
                                     -inv * tRij[a] * tRij[be]))
                                goto done;
                        }
                }
            }
    }
    /* merge duplicates, then G = A A^T and c = A b */
    qsort(L.e, (size_t)L.n, sizeof(struct entry), by_col_row);
    int64_t m = 0;
    for (int64_t i = 0; i < L.n; ++i)
    {
        if (m > 0 && L.e[m - 1].col == L.e[i].col && L.e[m - 1].row == L.e[i].row)
            L.e[m - 1].val += L.e[i].val;
        else
            L.e[m++] = L.e[i];
    }
    L.n = m;
    G = calloc((size_t)nc * nc, sizeof(long double));
    cv = calloc((size_t)nc, sizeof(long double));
    lam = calloc((size_t)nc, sizeof(long double));
    perm = malloc(sizeof(int64_t) * (size_t)nc);
    if (!G || !cv || !lam || !perm) goto done;
    for (int64_t s = 0; s < L.n;)
    {
        int64_t t = s;
        while (t < L.n && L.e[t].col == L.e[s].col) ++t;
        const long double bx = b[L.e[s].col];
        for (int64_t u = s; u < t; ++u)
        {
            cv[L.e[u].row] += (long double)L.e[u].val * bx;
            for (int64_t v = s; v < t; ++v)
                G[(int64_t)L.e[u].row * nc + L.e[v].row] +=
                    (long double)L.e[u].val * L.e[v].val;
        }
        s = t;
    }
    /* Cholesky with symmetric pivoting, in place: the factor in the lower
     * triangle of the leading r columns, the trailing block kept whole */
    for (int i = 0; i < nc; ++i) perm[i] = i;
    long double dmax = 0;
    for (int i = 0; i < nc; ++i)
        if (G[(int64_t)i * nc + i] > dmax) dmax = G[(int64_t)i * nc + i];
    const long double tol = dmax * 1e-13L;
    int r = 0;
    for (; r < nc; ++r)
    {
        int p = r;
        for (int i = r + 1; i < nc; ++i)
            if (G[(int64_t)i * nc + i] > G[(int64_t)p * nc + p]) p = i;
        if (!(G[(int64_t)p * nc + p] > tol)) break;
        if (p != r)
        {
            for (int k = 0; k < nc; ++k)
            {
                const long double t = G[(int64_t)r * nc + k];
                G[(int64_t)r * nc + k] = G[(int64_t)p * nc + k];
                G[(int64_t)p * nc + k] = t;
            }
            for (int k = 0; k < nc; ++k)
            {
                const long double t = G[(int64_t)k * nc + r];
                G[(int64_t)k * nc + r] = G[(int64_t)k * nc + p];
                G[(int64_t)k * nc + p] = t;
            }
            const int64_t t = perm[r];
            perm[r] = perm[p];
            perm[p] = t;
        }
        const long double d = sqrtl(G[(int64_t)r * nc + r]);
        G[(int64_t)r * nc + r] = d;
        for (int i = r + 1; i < nc; ++i) G[(int64_t)i * nc + r] /= d;
        for (int i = r + 1; i < nc; ++i)
        {
            const long double li = G[(int64_t)i * nc + r];
            for (int j = r + 1; j < nc; ++j) G[(int64_t)i * nc + j] -= li * G[(int64_t)j * nc + r];
        }
    }
    /* solve L L^T y = P c on the leading r rows, zero the rest */
    long double* y = lam;
    for (int i = 0; i < r; ++i)
    {
        long double s = cv[perm[i]];
        for (int k = 0; k < i; ++k) s -= G[(int64_t)i * nc + k] * y[k];
        y[i] = s / G[(int64_t)i * nc + i];
    }
    for (int i = r - 1; i >= 0; --i)
    {
        long double s = y[i];
        for (int k = i + 1; k < r; ++k) s -= G[(int64_t)k * nc + i] * y[k];
        y[i] = s / G[(int64_t)i * nc + i];
    }
    long double* lfull = cv; /* reuse: lambda in constraint order */
    for (int i = 0; i < nc; ++i) lfull[i] = 0;
    for (int i = 0; i < r; ++i) lfull[perm[i]] = y[i];
    for (int64_t u = 0; u < L.n; ++u)
        b[L.e[u].col] -= (double)((long double)L.e[u].val * lfull[L.e[u].row]);
    /* rebuild the complex array (asr.c:464-499) */
    for (int64_t g = 0; g < ng; ++g)
    {
        if (!doR[g]) continue;
        const int64_t iig = gmap[g];
        const int64_t rx = g / gyz, ry = (g % gyz) / qgrid[2], rz = g % qgrid[2];
        const int64_t mg = ((qgrid[2] - rz) % qgrid[2]) + ((qgrid[1] - ry) % qgrid[1]) * qgrid[2] +
                           ((qgrid[0] - rx) % qgrid[0]) * gyz;
        for (int i = 0; i < nm; ++i)
            for (int j = 0; j < nm; ++j)
            {
                const double v = b[iig * nn + (int64_t)i * nm + j];
                frc[2 * (g * nn + (int64_t)i * nm + j)] = v;
                frc[2 * (g * nn + (int64_t)i * nm + j) + 1] = 0.0;
                if (mg != g)
                {
                    frc[2 * (mg * nn + (int64_t)j * nm + i)] = v;
                    frc[2 * (mg * nn + (int64_t)j * nm + i) + 1] = 0.0;
                }
            }
    }
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < i; ++j)
        {
            frc[2 * ((int64_t)j * nm + i)] = frc[2 * ((int64_t)i * nm + j)];
            frc[2 * ((int64_t)j * nm + i) + 1] = frc[2 * ((int64_t)i * nm + j) + 1];
        }
    st = LELPH_OK;
done:
    free(L.e);
    free(gmap);
    free(doR);
    free(b);
    free(G);
    free(cv);
    free(lam);
    free(perm);
    return st;
}

/* apply_acoustic_sum_rule_born_charges (asr.c:155-267): sum over atoms of
 * each Z*_{ab} made zero.  The reference's "simple" subtracts the mean; its
 * other modes project orthogonally onto the same 9 constraints
 * sum_a Z_a[ab] = 0, whose projection is that same mean subtraction. */
int lelph_asr_born(int mode, double* Zborn, int natom)
{
    if (mode == LELPH_ASR_NONE || !Zborn) return LELPH_OK;
    if (natom < 1 || mode < 0 || mode > LELPH_ASR_ALL_HUANG) return LELPH_ERR_ARG;
    for (int i = 0; i < 9; ++i)
    {
        double s = 0;
        for (int a = 0; a < natom; ++a) s += Zborn[9 * a + i];
        const double avg = s / natom;
        for (int a = 0; a < natom; ++a) Zborn[9 * a + i] -= avg;
    }
    return LELPH_OK;
}
