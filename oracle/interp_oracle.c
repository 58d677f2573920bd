/*
 * interp_oracle.c -- FP64 CPU restatement of LetzElPhC's Fourier
 * interpolation of dV_q (src/interpolation/, src/dvloc/dVpot_add_long_range.c,
 * src/dvloc/dVlong_range_kernel.c, src/dvloc/multiply_eikr.c,
 * src/dvloc/dvscf_basis_rotation.c, src/symmetries/rotate_dvscf.c).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, never by the
 * product path.  Every function names the reference file:line it restates.
 * The Wigner-Seitz search enumerates the same superlattice points the
 * reference puts into its kd-tree and applies the same distance test; the
 * order of the degenerate vectors may differ from the kd-tree's, which only
 * changes the summation order in the R -> q sum.
 *
 * Pin status of dyn_lr_internal (the long-range dynamical matrix): the 3-D
 * form, G != 0 terms and quadrupoles included, is pinned by Ewald
 * invariance; in the 2-D form the in-plane non-analytic term is pinned
 * against the published 2-D expression (2 pi e^2/A)(Z^T q)(Z^T q)^T/|q|,
 * while its 1 - tanh range separation, r_eff and the out-of-plane (Z_z,
 * quadrupole) terms are checked by invariants only: parity unpinned for
 * those terms.
 */
#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lelph_oracle.h"

typedef double complex cplx;

#define RPI 3.1415927 /* ELPH_PI   src/common/constants.h:11 */
#define REPS 1e-5     /* ELPH_EPS  src/common/constants.h:9  */
#define REPS6 1e-6    /* ELPH_EPS6 src/common/constants.h:10 */
#define RE2 2.0       /* ELPH_e2   src/common/constants.h:13 */
#define WS_SEARCH 6   /* WS_SUPERCELL_SEARCH_SIZE wigner_seitz.c:17 */

/* MatVec3f, src/common/numerical_func.c:226-241 */
static void mv3(const double* M, const double* v, bool trans, double* o)
{
    if (trans)
    {
        o[0] = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
        o[1] = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
        o[2] = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
    }
    else
    {
        o[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
        o[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
        o[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
    }
}

static double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* get_miller_idx, src/common/numerical_func.c:436-449 */
static int64_t miller(int64_t i, int64_t N)
{
    const int64_t mid = (N - 1) / 2 + 1;
    return i < mid ? i : (i - mid) - N / 2;
}

/* get_fft_box, src/wfc/gen_fftbox.c:8-90 (one rank) */
void orc_get_fft_box(double EcutRy, const double* blat, int64_t* box)
{
    const double b0 = sqrt(blat[0] * blat[0] + blat[3] * blat[3] + blat[6] * blat[6]);
    const double b1 = sqrt(blat[1] * blat[1] + blat[4] * blat[4] + blat[7] * blat[7]);
    const double b2 = sqrt(blat[2] * blat[2] + blat[5] * blat[5] + blat[8] * blat[8]);
    const double e = sqrt(EcutRy);
    box[0] = 2 * (int64_t)rint(e / b0 + 1) + 1;
    box[1] = 2 * (int64_t)rint(e / b1 + 1) + 1;
    box[2] = 2 * (int64_t)rint(e / b2 + 1) + 1;
}

/* ------------------------------------------------------------------------ */
/* Wigner-Seitz vectors, src/interpolation/wigner_seitz.c:27-313            */
/* ------------------------------------------------------------------------ */
int64_t orc_ws_vectors(const int64_t* grid, const double* lat, double eps,
                       const double* rvec_m, int64_t nrvec_m,
                       const double* rvec_n, int64_t nrvec_n, int64_t cap,
                       int64_t* ws_vecs, int64_t* ws_degen)
{
    static const double zero[3] = {0, 0, 0};
    if (!rvec_m)
    {
        rvec_m = zero;
        nrvec_m = 1;
    }
    if (!rvec_n)
    {
        rvec_n = zero;
        nrvec_n = 1;
    }
    /* superlattice points (setup_ws_tree, :212-252) */
    const int side = 2 * WS_SEARCH + 1;
    const int npts = side * side * side;
    double* pts = malloc(sizeof(double) * 3 * npts);
    double sl[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sl[j * 3 + i] = lat[j * 3 + i] * grid[i];
    int np = 0;
    for (int i = -WS_SEARCH; i <= WS_SEARCH; ++i)
        for (int j = -WS_SEARCH; j <= WS_SEARCH; ++j)
            for (int k = -WS_SEARCH; k <= WS_SEARCH; ++k)
            {
                pts[3 * np] = sl[0] * i + sl[1] * j + sl[2] * k;
                pts[3 * np + 1] = sl[3] * i + sl[4] * j + sl[5] * k;
                pts[3 * np + 2] = sl[6] * i + sl[7] * j + sl[8] * k;
                ++np;
            }
    double blat[9];
    orc_reciprocal_vecs(lat, blat);
    for (int x = 0; x < 9; ++x) blat[x] /= (2 * RPI);

    const int64_t nR = grid[0] * grid[1] * grid[2];
    const int64_t Gyz = grid[1] * grid[2];
    int64_t nfound = 0;
    for (int64_t i = 0; i < nR; ++i)
    {
        const int64_t Rx = miller(i / Gyz, grid[0]);
        const int64_t Ry = miller((i % Gyz) / grid[2], grid[1]);
        const int64_t Rz = miller((i % Gyz) % grid[2], grid[2]);
        double Rv[3];
        Rv[0] = lat[0] * Rx + lat[1] * Ry + lat[2] * Rz;
        Rv[1] = lat[3] * Rx + lat[4] * Ry + lat[5] * Rz;
        Rv[2] = lat[6] * Rx + lat[7] * Ry + lat[8] * Rz;
        for (int64_t im = 0; im < nrvec_m; ++im)
            for (int64_t in = 0; in < nrvec_n; ++in)
            {
                const int64_t ish = in + im * nrvec_n + i * nrvec_m * nrvec_n;
                double qp[3];
                for (int x = 0; x < 3; ++x)
                    qp[x] = rvec_m[3 * im + x] - rvec_n[3 * in + x] - Rv[x];
                const double qn = sqrt(dot3(qp, qp));
                const double e = eps * qn + 1e-8;
                /* get_ws_nearest_superlat (:253-313) */
                double best = INFINITY;
                for (int p = 0; p < np; ++p)
                {
                    double d[3] = {pts[3 * p] - qp[0], pts[3 * p + 1] - qp[1],
                                   pts[3 * p + 2] - qp[2]};
                    const double dd = dot3(d, d);
                    if (dd < best) best = dd;
                }
                const double range = sqrt(best) + 1e-10 + fabs(e);
                int64_t cnt = 0;
                for (int p = 0; p < np; ++p)
                {
                    double d[3] = {pts[3 * p] - qp[0], pts[3 * p + 1] - qp[1],
                                   pts[3 * p + 2] - qp[2]};
                    if (dot3(d, d) > range * range) continue;
                    if (nfound >= cap)
                    {
                        free(pts);
                        return -1;
                    }
                    double t[3];
                    t[0] = blat[0] * pts[3 * p] + blat[3] * pts[3 * p + 1] + blat[6] * pts[3 * p + 2];
                    t[1] = blat[1] * pts[3 * p] + blat[4] * pts[3 * p + 1] + blat[7] * pts[3 * p + 2];
                    t[2] = blat[2] * pts[3 * p] + blat[5] * pts[3 * p + 1] + blat[8] * pts[3 * p + 2];
                    for (int x = 0; x < 3; ++x) ws_vecs[3 * nfound + x] = (int64_t)rint(t[x]);
                    ++nfound;
                    ++cnt;
                }
                if (cnt == 0)
                {
                    free(pts);
                    return -2;
                }
                ws_degen[ish] = cnt;
            }
    }
    free(pts);
    return nfound;
}

/* ------------------------------------------------------------------------ */
/* long-range kernels, src/dvloc/dVlong_range_kernel.c                      */
/* ------------------------------------------------------------------------ */

/* long_range_3D_kernel, dVlong_range_kernel.c:139-230 */
static void lr3d(const double* qG, const double* Zval, const double* Zb,
                 const double* Qp, const double* eps, const double* tau,
                 double eta_bare, double eta_ind, cplx* out)
{
    if (!Zval && !Zb && !Qp) return;
    const double qq = dot3(qG, qG);
    if (sqrt(qq) < REPS) return;
    const cplx ph = cexp(-I * dot3(qG, tau));
    double qeq = qq;
    if (eps)
    {
        double t[3];
        mv3(eps, qG, false, t);
        qeq = dot3(t, qG);
    }
    double zv[3] = {0, 0, 0}, zb[3] = {0, 0, 0}, qpb[3] = {0, 0, 0};
    if (Zval)
        for (int i = 0; i < 3; ++i) zv[i] = -(*Zval) * qG[i] / qq;
    if (eps && Zb) mv3(Zb, qG, true, zb);
    if (eps && Qp)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) qpb[k] += qG[i] * qG[j] * Qp[k + 3 * j + 9 * i];
    const double dfb = exp(-qq * 0.25 / eta_bare);
    const double dfi = exp(-fabs(qeq) * 0.25 / eta_ind);
    for (int i = 0; i < 3; ++i)
        out[i] = ph * ((zb[i] - I * 0.5 * qpb[i]) * dfi / qeq + dfb * zv[i]);
}

/* long_range_2D_kernel, dVlong_range_kernel.c:232-458 */
static void lr2d(const double* qG, const double* Zval, const double* Zb,
                 const double* Qp, const double* eps, const double* tau,
                 double zlat, double eta_bare, double eta_ind, cplx* out)
{
    if (!Zval && !Zb && !Qp) return;
    const double qq = dot3(qG, qG);
    if (sqrt(qq) < REPS) return;
    const cplx ph = cexp(-I * dot3(qG, tau));
    const double gp = sqrt(qG[0] * qG[0] + qG[1] * qG[1]);
    const double qGp = zlat * gp;
    double csf = cos(qG[2] * zlat * 0.5);
    if (gp > REPS) csf = csf - sin(qG[2] * zlat * 0.5) * qG[2] / gp;
    const double cutf = 1.0 - exp(-qGp * 0.5) * csf;
    const double Leff = eps ? eta_ind * zlat * 1.001 * (1.0 - 1.0 / eps[8]) : zlat * 0.5;
    const double fq = 1.0 - tanh(gp * 0.5 * Leff);
    double etp = 1.0;
    if (eps)
    {
        double qaq = (eps[0] - 1.0) * qG[0] * qG[0] + (eps[4] - 1.0) * qG[1] * qG[1] +
                     (eps[1] + eps[3]) * qG[0] * qG[1];
        qaq *= 0.5 * zlat;
        if (gp > REPS) etp = 1.0 + fq * qaq / gp;
    }
    double zv[3] = {0, 0, 0};
    if (Zval)
        for (int i = 0; i < 3; ++i) zv[i] = -(*Zval) * qG[i];
    const double qpar[3] = {qG[0], qG[1], 0.0};
    double zb[3] = {0, 0, 0};
    if (eps && Zb) mv3(Zb, qpar, true, zb);
    double qin[3] = {0, 0, 0}, qzz[3] = {0, 0, 0}, qz[3] = {0, 0, 0};
    if (eps && Qp)
        for (int a = 0; a < 3; ++a)
        {
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) qin[a] += qpar[j] * qpar[k] * Qp[a + 3 * k + 9 * j];
            qzz[a] = Qp[a + 3 * 2 + 9 * 2];
            for (int k = 0; k < 2; ++k) qz[a] += qpar[k] * Qp[a + 3 * k + 9 * 2];
        }
    const double dfb = exp(-qq * 0.25 / eta_bare);
    const cplx pph = cexp(-I * (qG[0] * tau[0] + qG[1] * tau[1]));
    const double x = 0.95;
    const double Gz = qG[2];
    cplx fGz = 0.0;
    if (fabs(Gz) > REPS)
    {
        const double term =
            (sin(x * Gz * zlat * 0.5) - x * sin(Gz * zlat * 0.5)) / (Gz * Gz * (1.0 - x));
        fGz = -2.0 * I * term / zlat;
    }
    const double gGz = fabs(Gz) < REPS ? 1.0 : 0.0;
    for (int i = 0; i < 3; ++i)
    {
        const cplx vb = ph * cutf * zv[i] * dfb / qq;
        cplx vi = 0.0;
        if (gp > REPS)
        {
            const cplx tin = (2.0 * I * zb[i] + qin[i] - gp * gp * qzz[i]) * gGz / etp;
            double zz = 0.0;
            if (Zb) zz = Zb[6 + i] / eps[8];
            const cplx tout = 2.0 * gp * gp * (zz - I * qz[i]) * fGz;
            vi = (zlat / (4.0 * I)) * (fq / gp) * pph * (tin + tout);
        }
        out[i] = vb + vi;
    }
}

/* dVlong_range_kernel, dVlong_range_kernel.c:31-135; out (npw, natom, 3) */
void orc_dVlong_range_kernel(const double* qpt, const double* gvecs, int64_t npw,
                             const double* Zvals, const double* eps, const double* Zeu,
                             const double* Qpole, int64_t natom, const double* atom_pos,
                             char dimension, double volume, double zlat, double EcutRy,
                             double eta_bare, double eta_ind, double* outd)
{
    cplx* out = (cplx*)outd;
    for (int64_t i = 0; i < 3 * natom * npw; ++i) out[i] = 0.0;
    if (!eps && !Zvals) return;
    const cplx factor = 4.0 * RPI * I * RE2 / volume;
    for (int64_t ig = 0; ig < npw; ++ig)
    {
        double qG[3];
        for (int i = 0; i < 3; ++i) qG[i] = 2 * RPI * (qpt[i] + gvecs[3 * ig + i]);
        const double n2 = dot3(qG, qG);
        if (dimension == '3' && n2 > EcutRy) continue;
        if (dimension == '2')
        {
            const double gp = sqrt(qG[0] * qG[0] + qG[1] * qG[1]);
            const double Leff = eps ? eta_ind * zlat * 1.001 * (1.0 - 1.0 / eps[8]) : zlat * 0.5;
            const double fq = 1.0 - tanh(gp * 0.5 * Leff);
            if (fabs(fq) < REPS6 && (!Zvals || n2 > EcutRy)) continue;
        }
        for (int64_t ia = 0; ia < natom; ++ia)
        {
            const double* zv = Zvals ? Zvals + ia : NULL;
            const double* zb = Zeu ? Zeu + 9 * ia : NULL;
            const double* qp = Qpole ? Qpole + 27 * ia : NULL;
            cplx* o = out + ia * 3 + ig * natom * 3;
            if (dimension == '3')
                lr3d(qG, zv, zb, qp, eps, atom_pos + 3 * ia, eta_bare, eta_ind, o);
            else if (dimension == '2')
                lr2d(qG, zv, zb, qp, eps, atom_pos + 3 * ia, zlat, eta_bare, eta_ind, o);
            for (int i = 0; i < 3; ++i) o[i] *= factor;
        }
    }
}

/* dV_add_longrange, src/dvloc/dVpot_add_long_range.c:18-241 */
int orc_dV_add_longrange(const double* qpt, const lelph_lattice* lat,
                         const lelph_longrange* lr, const double* eigd, int sign,
                         double* dvd)
{
    if (lat->dimension == '2' && fabs(qpt[2]) > REPS) return LELPH_ERR_ARG;
    const int64_t natom = lat->natom, nm = lat->nmodes, nmag = lat->nmag;
    const int64_t* N = lat->fft_dims;
    const int64_t nf = N[0] * N[1] * N[2];
    int64_t Gb[3];
    orc_get_fft_box(lr->EcutRy, lat->blat_vec, Gb);
    for (int i = 0; i < 3; ++i) Gb[i] = Gb[i] < N[i] ? Gb[i] : N[i];
    if (lat->dimension == '2') Gb[2] = N[2];
    const int64_t nG = Gb[0] * Gb[1] * Gb[2];
    cplx* VG = calloc(nm * nG, sizeof(cplx));
    int64_t* gv = malloc(sizeof(int64_t) * 3 * nG);
    double* Gz = malloc(sizeof(double) * 3 * Gb[2]);
    cplx* Vz = calloc(Gb[2] * 3 * natom, sizeof(cplx));
    double qc[3];
    mv3(lat->blat_vec, qpt, false, qc);
    for (int i = 0; i < 3; ++i) qc[i] /= 2.0 * RPI;
    const double* Zv = lr->only_induced ? NULL : lr->Zvals;
    const cplx* eig = (const cplx*)eigd;
    for (int64_t ig = 0; ig < Gb[0] * Gb[1]; ++ig)
    {
        const int64_t ix = miller(ig / Gb[1], Gb[0]);
        const int64_t jy = miller(ig % Gb[1], Gb[1]);
        for (int64_t kz = 0; kz < Gb[2]; ++kz)
        {
            const int64_t gz = miller(kz, Gb[2]);
            int64_t* g = gv + 3 * (Gb[2] * ig + kz);
            g[0] = ix;
            g[1] = jy;
            g[2] = gz;
            const double m[3] = {(double)ix, (double)jy, (double)gz};
            mv3(lat->blat_vec, m, false, Gz + 3 * kz);
            for (int i = 0; i < 3; ++i) Gz[3 * kz + i] /= 2.0 * RPI;
        }
        orc_dVlong_range_kernel(qc, Gz, Gb[2], Zv, lr->epsilon, lr->Zborn, lr->Qpole, natom,
                                lat->atomic_pos, lat->dimension, lat->volume,
                                lat->alat_vec[8], lr->EcutRy, lr->eta_bare, lr->eta_induced,
                                (double*)Vz);
        for (int64_t kz = 0; kz < Gb[2]; ++kz)
            for (int64_t v = 0; v < nm; ++v)
            {
                cplx acc = 0.0;
                if (eig)
                    for (int64_t j = 0; j < nm; ++j) acc += eig[v * nm + j] * Vz[kz * nm + j];
                else
                    acc = Vz[kz * nm + v];
                VG[v * nG + Gb[2] * ig + kz] = acc;
            }
    }
    cplx* box = malloc(sizeof(cplx) * nf);
    cplx* dv = (cplx*)dvd;
    const double f = sign < 0 ? -1.0 : 1.0;
    for (int64_t v = 0; v < nm; ++v)
    {
        memset(box, 0, sizeof(cplx) * nf);
        for (int64_t g = 0; g < nG; ++g)
        {
            int64_t b[3];
            for (int i = 0; i < 3; ++i) b[i] = ((gv[3 * g + i] % N[i]) + N[i]) % N[i];
            box[(b[0] * N[1] + b[1]) * N[2] + b[2]] = VG[v * nG + g];
        }
        orc_fft3d((double*)box, N, +1);
        for (int64_t im = 0; im < nmag; ++im)
        {
            if (!lr->nmags_add[im]) continue;
            cplx* d = dv + (v * nmag + im) * nf;
            for (int64_t i = 0; i < nf; ++i) d[i] = d[i] + f * box[i];
        }
    }
    free(box);
    free(Vz);
    free(Gz);
    free(gv);
    free(VG);
    return 0;
}

/* multiply_eikr, src/dvloc/multiply_eikr.c:14-64 */
void orc_multiply_eikr(double* potd, const double* q, const int64_t* N, int64_t nsets, int sign)
{
    cplx* pot = (cplx*)potd;
    const int64_t nf = N[0] * N[1] * N[2];
    cplx fac = 2 * I * RPI;
    if (sign < 0) fac = -fac;
    cplx* e = malloc(sizeof(cplx) * nf);
    const int64_t yz = N[1] * N[2];
    for (int64_t ig = 0; ig < nf; ++ig)
    {
        const int64_t ix = ig / yz, iy = (ig % yz) / N[2], iz = (ig % yz) % N[2];
        const double r[3] = {(double)ix / N[0], (double)iy / N[1], (double)iz / N[2]};
        e[ig] = cexp(fac * (q[0] * r[0] + q[1] * r[1] + q[2] * r[2]));
    }
    for (int64_t s = 0; s < nsets; ++s)
        for (int64_t ig = 0; ig < nf; ++ig) pot[s * nf + ig] *= e[ig];
    free(e);
}

/* dVscf_change_basis, src/dvloc/dvscf_basis_rotation.c:12-51;
 * dvscf (nsets, nmodes, nmag, nf), rot (nmodes, nmodes) */
void orc_dvscf_change_basis(double* dvd, const double* rotd, int64_t nsets, int64_t nm,
                            int64_t nmag, int64_t nf, char op)
{
    cplx* dv = (cplx*)dvd;
    const cplx* rot = (const cplx*)rotd;
    cplx* tmp = malloc(sizeof(cplx) * nm);
    const int64_t st = nmag * nf;
    for (int64_t s = 0; s < nsets; ++s)
        for (int64_t p = 0; p < st; ++p)
        {
            cplx* d = dv + s * nm * st + p;
            for (int64_t a = 0; a < nm; ++a)
            {
                cplx acc = 0.0;
                for (int64_t b = 0; b < nm; ++b)
                {
                    cplx r;
                    if (op == 'N')
                        r = rot[a * nm + b];
                    else if (op == 'T')
                        r = rot[b * nm + a];
                    else
                        r = conj(rot[b * nm + a]);
                    acc += r * d[b * st];
                }
                tmp[a] = acc;
            }
            for (int64_t a = 0; a < nm; ++a) d[a * st] = tmp[a];
        }
    free(tmp);
}

/* rotate_dvscf, src/symmetries/rotate_dvscf.c:22-368 (one rank) */
int orc_rotate_dvscf(const double* qpt, const double* dvin, const lelph_symmetry* sym,
                     const lelph_lattice* lat, int composite, double* dvout)
{
    const int64_t* N = lat->fft_dims;
    const int64_t nf = N[0] * N[1] * N[2];
    const int64_t nsets = (int64_t)lat->nmodes * lat->nmag;
    const double* blat = lat->blat_vec;
    double tr[3], sr[9], tmp[9];
    mv3(blat, sym->tau, true, tr);
    for (int x = 0; x < 3; ++x)
    {
        tr[x] /= 2 * RPI;
        if (sym->time_rev) tr[x] = -tr[x];
    }
    /* Gemm3x3f(R, 'T', alat, 'N') then Gemm3x3f(blat, 'T', ., 'N') */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += sym->Rmat[k * 3 + i] * lat->alat_vec[k * 3 + j];
            tmp[i * 3 + j] = a / (2 * RPI) * (sym->time_rev ? -1.0 : 1.0);
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += blat[k * 3 + i] * tmp[k * 3 + j];
            sr[i * 3 + j] = a;
        }
    int64_t* rot = malloc(sizeof(int64_t) * nf);
    const int64_t yz = N[1] * N[2];
    for (int64_t i = 0; i < nf; ++i)
    {
        double t[3] = {(double)(i / yz), (double)((i % yz) / N[2]), (double)((i % yz) % N[2])};
        for (int x = 0; x < 3; ++x) t[x] = t[x] / N[x] - tr[x];
        double r[3];
        mv3(sr, t, false, r);
        int64_t id[3];
        for (int x = 0; x < 3; ++x)
        {
            r[x] -= floor(r[x]);
            r[x] -= floor(r[x] + REPS);
            r[x] -= REPS;
            r[x] *= N[x];
            if (fabs(r[x] - rint(r[x])) > 1e-3 + 5e-5 * fabs(r[x]))
            {
                free(rot);
                return LELPH_ERR_SYM;
            }
            id[x] = (int64_t)rint(r[x]);
        }
        rot[i] = id[0] * yz + id[1] * N[2] + id[2];
        if (rot[i] >= nf || rot[i] < 0)
        {
            free(rot);
            return LELPH_ERR_RANGE;
        }
    }
    double qc[3], Rq[3];
    mv3(blat, qpt, false, qc);
    mv3(sym->Rmat, qc, false, Rq);
    const cplx ph = cexp(-I * dot3(Rq, sym->tau));
    const cplx* in = (const cplx*)dvin;
    cplx* out = (cplx*)dvout;
    for (int64_t s = 0; s < nsets; ++s)
        for (int64_t i = 0; i < nf; ++i) out[s * nf + i] = ph * in[s * nf + rot[i]];
    free(rot);
    if (lat->nmag == 4)
    {
        cplx su[4];
        orc_SU2mat(sym->Rmat, lat->nspinor, 0, sym->time_rev, (double*)su);
        for (int64_t m = 0; m < lat->nmodes; ++m)
        {
            cplx* d0 = out + (m * 4 + 0) * nf;
            cplx* d1 = out + (m * 4 + 1) * nf;
            cplx* d2 = out + (m * 4 + 2) * nf;
            cplx* d3 = out + (m * 4 + 3) * nf;
            for (int64_t i = 0; i < nf; ++i)
            {
                if (!composite)
                {
                    const cplx v = d0[i], bx = d1[i], by = d2[i], bz = d3[i];
                    d0[i] = v + bz;
                    d1[i] = bx - I * by;
                    d2[i] = bx + I * by;
                    d3[i] = v - bz;
                }
                const cplx M0 = su[0] * (d0[i] * conj(su[0]) + d1[i] * conj(su[1])) +
                                su[1] * (d2[i] * conj(su[0]) + d3[i] * conj(su[1]));
                const cplx M1 = su[0] * (d0[i] * conj(su[2]) + d1[i] * conj(su[3])) +
                                su[1] * (d2[i] * conj(su[2]) + d3[i] * conj(su[3]));
                const cplx M2 = su[2] * (d0[i] * conj(su[0]) + d1[i] * conj(su[1])) +
                                su[3] * (d2[i] * conj(su[0]) + d3[i] * conj(su[1]));
                const cplx M3 = su[2] * (d0[i] * conj(su[2]) + d1[i] * conj(su[3])) +
                                su[3] * (d2[i] * conj(su[2]) + d3[i] * conj(su[3]));
                if (composite)
                {
                    d0[i] = M0;
                    d1[i] = M1;
                    d2[i] = M2;
                    d3[i] = M3;
                }
                else
                {
                    d0[i] = (M0 + M3) / 2.0;
                    d1[i] = (M1 + M2) / 2.0;
                    d2[i] = -I * (M2 - M1) / 2.0;
                    d3[i] = (M0 - M3) / 2.0;
                }
            }
        }
    }
    if (sym->time_rev)
        for (int64_t i = 0; i < nsets * nf; ++i) out[i] = conj(out[i]);
    return 0;
}

/* fft_q2R, src/interpolation/interpolation_utilities.c:25-76:
 * data (qx, qy, qz, nsets), forward transform over the q axes, / nq */
void orc_fft_q2R(double* datad, const int64_t* qgrid, int64_t nsets)
{
    cplx* data = (cplx*)datad;
    const int64_t nq = qgrid[0] * qgrid[1] * qgrid[2];
    cplx* box = malloc(sizeof(cplx) * nq);
    for (int64_t s = 0; s < nsets; ++s)
    {
        for (int64_t i = 0; i < nq; ++i) box[i] = data[i * nsets + s];
        orc_fft3d((double*)box, qgrid, -1);
        for (int64_t i = 0; i < nq; ++i) data[i * nsets + s] = box[i] / (double)nq;
    }
    free(box);
}

/* fft_R2q_dvscf, interpolation_utilities.c:165-238: dataR (R, nmodes, nsets)
 * -> dataq (nmodes, nsets) */
int orc_fft_R2q_dvscf(const double* dataRd, const double* q, const int64_t* qgrid,
                      int64_t natom, int64_t nsets, const int64_t* ws, int64_t nws,
                      const int64_t* degen, double* dataqd)
{
    const cplx* dR = (const cplx*)dataRd;
    cplx* dq = (cplx*)dataqd;
    const int64_t nR = qgrid[0] * qgrid[1] * qgrid[2];
    const int64_t nm = 3 * natom;
    for (int64_t i = 0; i < nm * nsets; ++i) dq[i] = 0.0;
    const int64_t yz = qgrid[1] * qgrid[2];
    int64_t iw = 0;
    for (int64_t i = 0; i < nR; ++i)
    {
        const double R[3] = {(double)miller(i / yz, qgrid[0]),
                             (double)miller((i % yz) / qgrid[2], qgrid[1]),
                             (double)miller((i % yz) % qgrid[2], qgrid[2])};
        const cplx eR = cexp(I * 2 * RPI * dot3(q, R));
        for (int64_t ia = 0; ia < natom; ++ia)
        {
            const int64_t dg = degen[i * natom + ia];
            cplx eT = 0.0;
            for (int64_t t = 0; t < dg; ++t)
            {
                if (iw >= nws) return LELPH_ERR_RANGE;
                const double d = q[0] * ws[3 * iw] + q[1] * ws[3 * iw + 1] + q[2] * ws[3 * iw + 2];
                eT += cexp(I * 2 * RPI * d);
                ++iw;
            }
            eT *= eR;
            eT *= 1.0 / dg;
            for (int64_t s = 0; s < 3 * nsets; ++s)
                dq[ia * 3 * nsets + s] += eT * dR[(i * nm + ia * 3) * nsets + s];
        }
    }
    return 0;
}

/* find_qpt_grid, interpolation_utilities.c:280-331 */
int orc_find_qpt_grid(int64_t nq, const double* q, int64_t* grid)
{
    bool gamma = false;
    for (int i = 0; i < 3; ++i) grid[i] = 1;
    for (int64_t i = 0; i < nq; ++i)
    {
        bool isg = true;
        for (int x = 0; x < 3; ++x)
        {
            double v = q[3 * i + x] - floor(q[3 * i + x]);
            if (fabs(v) > REPS && fabs(v - 1) > REPS)
            {
                isg = false;
                v = 1.0 / v;
                const int64_t vi = (int64_t)rint(v);
                v -= vi;
                if (fabs(v) < REPS) grid[x] = vi > grid[x] ? vi : grid[x];
            }
        }
        if (isg) gamma = true;
    }
    if (!gamma) return LELPH_ERR_ARG;
    if (nq != grid[0] * grid[1] * grid[2]) return LELPH_ERR_ARG;
    return 0;
}

/* qpt_sort_cmp, interpolation_utilities.c:451-492 */
static int qcmp(const void* a, const void* b)
{
    const double* v1 = *(const double* const*)a;
    const double* v2 = *(const double* const*)b;
    double d[3];
    for (int i = 0; i < 3; ++i)
    {
        double k1 = v1[i];
        k1 -= floor(k1);
        k1 += REPS;
        k1 -= floor(k1);
        double k2 = v2[i];
        k2 -= floor(k2);
        k2 += REPS;
        k2 -= floor(k2);
        d[i] = k1 - k2;
    }
    for (int i = 0; i < 3; ++i)
        if (fabs(d[i]) >= REPS) return d[i] > 0.0 ? 1 : -1;
    return 0;
}

/* Sorted_qpts_idxs, interpolation_utilities.c:240-271 */
void orc_sorted_qpts_idxs(int64_t nq, const double* q, int64_t* idx)
{
    const double** p = malloc(sizeof(double*) * nq);
    for (int64_t i = 0; i < nq; ++i) p[i] = q + 3 * i;
    qsort(p, nq, sizeof(double*), qcmp);
    for (int64_t i = 0; i < nq; ++i) idx[i] = (p[i] - q) / 3;
    free(p);
}

/* read_dvscf_qe, src/io/qe/read_dvscf.c:17-172: file sets (Nz, Ny, Nx) ->
 * (Nx, Ny, Nz), then eig @ pats^H @ dV; out (nmodes, nmag, Nx, Ny, Nz) */
int orc_read_dvscf_qe(const char* path, const lelph_lattice* lat, const double* eigd,
                      const double* patd, double* outd)
{
    const int64_t* N = lat->fft_dims;
    const int64_t nf = N[0] * N[1] * N[2], nm = lat->nmodes, nmag = lat->nmag;
    FILE* fp = fopen(path, "rb");
    if (!fp) return LELPH_ERR_ARG;
    cplx* buf = malloc(sizeof(cplx) * nf);
    cplx* out = (cplx*)outd;
    for (int64_t s = 0; s < nm * nmag; ++s)
    {
        if (fread(buf, sizeof(cplx), nf, fp) != (size_t)nf)
        {
            free(buf);
            fclose(fp);
            return LELPH_ERR_ARG;
        }
        cplx* d = out + s * nf;
        for (int64_t iy = 0; iy < N[1]; ++iy)
            for (int64_t ix = 0; ix < N[0]; ++ix)
                for (int64_t iz = 0; iz < N[2]; ++iz)
                    d[(ix * N[1] + iy) * N[2] + iz] = buf[(iz * N[1] + iy) * N[0] + ix];
    }
    free(buf);
    fclose(fp);
    const cplx* eig = (const cplx*)eigd;
    const cplx* pat = (const cplx*)patd;
    cplx* ep = malloc(sizeof(cplx) * nm * nm);
    for (int64_t i = 0; i < nm; ++i)
        for (int64_t j = 0; j < nm; ++j)
        {
            cplx a = 0.0;
            for (int64_t k = 0; k < nm; ++k) a += eig[i * nm + k] * conj(pat[j * nm + k]);
            ep[i * nm + j] = a;
        }
    orc_dvscf_change_basis(outd, (const double*)ep, 1, nm, nmag, nf, 'N');
    free(ep);
    return 0;
}

/* write_dvscf_qe, src/io/qe/write_dyn_dvscf_qe.c:99-185 */
int orc_write_dvscf_qe(const char* path, const lelph_lattice* lat, const double* dvd)
{
    const int64_t* N = lat->fft_dims;
    const int64_t nf = N[0] * N[1] * N[2], ns = (int64_t)lat->nmodes * lat->nmag;
    const cplx* dv = (const cplx*)dvd;
    FILE* fp = fopen(path, "wb");
    if (!fp) return LELPH_ERR_ARG;
    cplx* buf = malloc(sizeof(cplx) * nf);
    int st = 0;
    for (int64_t s = 0; s < ns && !st; ++s)
    {
        const cplx* d = dv + s * nf;
        for (int64_t iy = 0; iy < N[1]; ++iy)
            for (int64_t ix = 0; ix < N[0]; ++ix)
                for (int64_t iz = 0; iz < N[2]; ++iz)
                    buf[(iz * N[1] + iy) * N[0] + ix] = d[(ix * N[1] + iy) * N[2] + iz];
        if (fwrite(buf, sizeof(cplx), nf, fp) != (size_t)nf) st = LELPH_ERR_ARG;
    }
    free(buf);
    if (fclose(fp)) st = LELPH_ERR_ARG;
    return st;
}

/* ------------------------------------------------------------------------ */
/* dynamical matrices (src/phonon/, interpolation_utilities.c:78-163)      */
/* ------------------------------------------------------------------------ */

/* pol_vecs_to_dyn, src/phonon/pol_vecs_to_dyn.c:8-40 (with
 * mass_normalize_pol_vecs power 0.5, mass_normalize_phonons.c:11-34) */
void orc_pol_vecs_to_dyn(const double* omega, int64_t natom, const double* masses,
                         double* pold)
{
    const int64_t nm = 3 * natom;
    cplx* pol = (cplx*)pold;
    if (masses)
        for (int64_t v = 0; v < nm; ++v)
            for (int64_t j = 0; j < nm; ++j) pol[v * nm + j] *= pow(masses[j / 3], 0.5);
    cplx* e = malloc(sizeof(cplx) * nm * nm);
    memcpy(e, pol, sizeof(cplx) * nm * nm);
    for (int64_t a = 0; a < nm; ++a)
        for (int64_t b = 0; b < nm; ++b)
        {
            cplx acc = 0.0;
            for (int64_t i = 0; i < nm; ++i)
            {
                const double f = (omega[i] > 0 ? 1.0 : -1.0) * omega[i] * omega[i];
                acc += e[i * nm + a] * (f * conj(e[i * nm + b]));
            }
            pol[a * nm + b] = acc;
        }
    free(e);
}

/* add_ph_dyn_long_range_internal, src/phonon/dyn_long_range.c:125-404 */
static void dyn_lr_internal(const double* qpt, const lelph_lattice* lat, const double* eps,
                            const double* Zb, const double* Qp, const int64_t* Gg, int sign,
                            const double* masses, double eta, cplx* dyn)
{
    if (!eps || (!Zb && !Qp)) return;
    const int64_t na = lat->natom, nm = 3 * na;
    double factor = 4.0 * RPI * RE2 / lat->volume;
    if (sign < 0) factor = -factor;
    double ea[9];
    memcpy(ea, eps, sizeof(ea));
    const double zlat = lat->alat_vec[8];
    const bool d2 = lat->dimension == '2';
    if (d2)
    {
        factor = factor * zlat / 2.0;
        for (int i = 0; i < 9; ++i) ea[i] = 0.0;
        ea[0] = 0.5 * zlat * (eps[0] - 1);
        ea[1] = 0.5 * zlat * eps[1];
        ea[3] = 0.5 * zlat * eps[3];
        ea[4] = 0.5 * zlat * (eps[4] - 1);
    }
    const int64_t GZ = d2 ? 1 : Gg[2];
    const int64_t nG = Gg[0] * Gg[1] * GZ;
    cplx* X1 = calloc(nG * nm, sizeof(cplx));
    cplx* X2 = calloc(nG * nm, sizeof(cplx));
    cplx* P1 = d2 ? calloc(nG * nm, sizeof(cplx)) : NULL;
    cplx* P2 = d2 ? calloc(nG * nm, sizeof(cplx)) : NULL;
    for (int64_t ig = 0; ig < nG; ++ig)
    {
        const int64_t gx = ig / (Gg[1] * GZ), gy = (ig % (Gg[1] * GZ)) / GZ,
                      gz = (ig % (Gg[1] * GZ)) % GZ;
        double t[3] = {qpt[0] + miller(gx, Gg[0]), qpt[1] + miller(gy, Gg[1]),
                       qpt[2] + miller(gz, GZ)};
        double qG[3];
        mv3(lat->blat_vec, t, false, qG);
        const double qn = sqrt(dot3(qG, qG));
        if (qn < REPS) continue;
        double tb[3];
        mv3(ea, qG, false, tb);
        double qeq = dot3(tb, qG);
        double qeqp = 1.0;
        double decay = exp(-fabs(qeq) * 0.125 / eta);
        if (d2)
        {
            const double Leff = eta * zlat * 1.001 * (1.0 - 1.0 / eps[8]);
            decay = 1 - tanh(qn * 0.5 * Leff);
            if (fabs(decay) < REPS6) continue;
            qeq *= decay;
            qeq += qn;
            qeqp = 1.0;
            qeqp = fabs(qeqp) < REPS ? 0.0 : -qn / qeqp;
            decay = sqrt(decay);
        }
        for (int64_t ia = 0; ia < na; ++ia)
        {
            const double* Z = Zb ? Zb + 9 * ia : NULL;
            const double* Q = Qp ? Qp + 27 * ia : NULL;
            double zq[3] = {0, 0, 0}, qb[3] = {0, 0, 0};
            if (Z) mv3(Z, qG, true, zq);
            if (Q)
            {
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        for (int k = 0; k < 3; ++k) qb[k] += qG[i] * qG[j] * Q[k + 3 * j + 9 * i];
                if (d2)
                    for (int i = 0; i < 3; ++i) qb[i] -= qn * qn * Q[i + 24];
            }
            cplx ph = cexp(-I * dot3(qG, lat->atomic_pos + 3 * ia));
            ph /= qeq;
            ph *= decay;
            if (masses) ph /= sqrt(masses[ia]);
            for (int i = 0; i < 3; ++i)
            {
                X1[ig * nm + 3 * ia + i] = (zq[i] - I * 0.5 * qb[i]) * ph;
                X2[ig * nm + 3 * ia + i] = X1[ig * nm + 3 * ia + i] * qeq;
            }
            if (d2)
            {
                ph *= qeq;
                for (int i = 0; i < 3; ++i)
                {
                    if (Z)
                    {
                        P1[ig * nm + 3 * ia + i] += Z[i + 6] * ph * qeqp / eps[8];
                        P2[ig * nm + 3 * ia + i] += Z[i + 6] * ph / eps[8];
                    }
                    if (Q)
                        for (int x = 0; x < 3; ++x)
                        {
                            P1[ig * nm + 3 * ia + i] -= I * qG[x] * Q[i + 18 + 3 * x] * ph * qeqp;
                            P2[ig * nm + 3 * ia + i] -= qG[x] * I * Q[i + 18 + 3 * x] * ph;
                        }
                }
            }
        }
    }
    /* dyn += factor X2^H X1 (and the out-of-plane pair in 2-D) */
    for (int64_t a = 0; a < nm; ++a)
        for (int64_t b = 0; b < nm; ++b)
        {
            cplx acc = 0.0, accp = 0.0;
            for (int64_t ig = 0; ig < nG; ++ig)
            {
                acc += conj(X2[ig * nm + a]) * X1[ig * nm + b];
                if (d2) accp += conj(P2[ig * nm + a]) * P1[ig * nm + b];
            }
            dyn[a * nm + b] += factor * acc;
            if (d2) dyn[a * nm + b] += factor * accp;
        }
    free(X1);
    free(X2);
    free(P1);
    free(P2);
}

/* compute_dyn_lr_asr_correction, dyn_long_range.c:70-123: out (natom,3,3) */
void orc_dyn_lr_asr(const lelph_lattice* lat, const double* eps, const double* Zb,
                    const double* Qp, const int64_t* Gg, double eta, double* outd)
{
    const int64_t na = lat->natom, nm = 3 * na;
    cplx* out = (cplx*)outd;
    for (int64_t i = 0; i < 9 * na; ++i) out[i] = 0.0;
    if (!eps || (!Zb && !Qp)) return;
    cplx* t = calloc(nm * nm, sizeof(cplx));
    const double q0[3] = {0, 0, 0};
    dyn_lr_internal(q0, lat, eps, Zb, Qp, Gg, 1, NULL, eta, t);
    for (int64_t ia = 0; ia < na; ++ia)
        for (int64_t ja = 0; ja < na; ++ja)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) out[ia * 9 + 3 * i + j] += t[(ia * 3 + i) * nm + ja * 3 + j];
    free(t);
}

/* add_ph_dyn_long_range, dyn_long_range.c:27-68 */
int orc_add_ph_dyn_long_range(const double* qpt, const lelph_lattice* lat, const double* eps,
                              const double* Zb, const double* Qp, const int64_t* Gg, int sign,
                              const double* masses, const double* asrd, double eta,
                              double* dynd)
{
    if (lat->dimension == '2' && fabs(qpt[2]) > REPS && eps && (Zb || Qp)) return LELPH_ERR_ARG;
    cplx* dyn = (cplx*)dynd;
    dyn_lr_internal(qpt, lat, eps, Zb, Qp, Gg, sign, masses, eta, dyn);
    if (!asrd) return 0;
    const cplx* asr = (const cplx*)asrd;
    const double f = sign < 0 ? 1.0 : -1.0;
    const int64_t na = lat->natom, nm = 3 * na;
    for (int64_t ia = 0; ia < na; ++ia)
    {
        const double m = masses ? 1.0 / masses[ia] : 1.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                dyn[(ia * 3 + i) * nm + ia * 3 + j] += f * asr[ia * 9 + 3 * i + j] * m;
    }
    return 0;
}

/* fft_R2q_dyn, interpolation_utilities.c:78-163: dataR (R, nm, nm) */
int orc_fft_R2q_dyn(const double* dataRd, const double* q, const int64_t* qgrid, int64_t natom,
                    const int64_t* ws, int64_t nws, const int64_t* degen, double* dataqd)
{
    const cplx* dR = (const cplx*)dataRd;
    cplx* dq = (cplx*)dataqd;
    const int64_t nR = qgrid[0] * qgrid[1] * qgrid[2], nm = 3 * natom;
    for (int64_t i = 0; i < nm * nm; ++i) dq[i] = 0.0;
    const int64_t yz = qgrid[1] * qgrid[2];
    int64_t iw = 0;
    for (int64_t i = 0; i < nR; ++i)
    {
        const double R[3] = {(double)miller(i / yz, qgrid[0]),
                             (double)miller((i % yz) / qgrid[2], qgrid[1]),
                             (double)miller((i % yz) % qgrid[2], qgrid[2])};
        const cplx eR = cexp(I * 2 * RPI * dot3(q, R));
        for (int64_t ia = 0; ia < natom; ++ia)
            for (int64_t ja = 0; ja < natom; ++ja)
            {
                const int64_t dg = degen[i * natom * natom + ia * natom + ja];
                cplx eT = 0.0;
                for (int64_t t = 0; t < dg; ++t)
                {
                    if (iw >= nws) return LELPH_ERR_RANGE;
                    eT += cexp(I * 2 * RPI * (q[0] * ws[3 * iw] + q[1] * ws[3 * iw + 1] +
                                              q[2] * ws[3 * iw + 2]));
                    ++iw;
                }
                eT *= eR;
                eT *= 1.0 / dg;
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y)
                    {
                        const int64_t im = x + 3 * ia, jm = y + 3 * ja;
                        dq[im * nm + jm] += eT * dR[i * nm * nm + im * nm + jm];
                    }
            }
    }
    return 0;
}

/* mass_normalize_force_constants, mass_normalize_phonons.c:36-71 */
void orc_mass_normalize_fc(const double* masses, int64_t nsets, int64_t natom, double power,
                           double* frcd)
{
    if (!masses) return;
    cplx* frc = (cplx*)frcd;
    const int64_t nm = 3 * natom;
    for (int64_t ia = 0; ia < natom; ++ia)
    {
        const double Ma = pow(masses[ia], power);
        for (int64_t ib = 0; ib < natom; ++ib)
        {
            const double Mb = pow(masses[ib], power);
            for (int64_t s = 0; s < nsets; ++s)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        frc[s * nm * nm + ia * 9 * natom + 3 * ib + j + i * nm] *= Ma * Mb;
        }
    }
}

/* apply_acoustic_sum_rule_fc, ASR_SIMPLE branch, src/phonon/asr.c:294-318 */
void orc_asr_simple_fc(const int64_t* qgrid, int64_t natom, double* frcd)
{
    cplx* frc = (cplx*)frcd;
    const int64_t ng = qgrid[0] * qgrid[1] * qgrid[2], nm = 3 * natom;
    for (int64_t ia = 0; ia < natom; ++ia)
        for (int64_t al = 0; al < 3; ++al)
            for (int64_t be = 0; be < 3; ++be)
            {
                cplx sum = 0.0;
                const int64_t base = ia * 9 * natom + al * nm + be;
                for (int64_t g = 0; g < ng; ++g)
                    for (int64_t ib = 0; ib < natom; ++ib) sum += frc[g * nm * nm + base + ib * 3];
                frc[base + ia * 3] -= sum;
            }
}
