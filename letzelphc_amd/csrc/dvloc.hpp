// dvloc.hpp -- bare local dV: V_loc(|G|) table (src/dvloc/vlocg_table.c:16-154,
// src/dvloc/VlocG_atom.c:10-57), dVlocq (src/dvloc/dVbare_loc.c:16-231),
// add_dvscf_qe (src/dvloc/add_dvscf.c:12-92) and the pattern -> mode
// rotation of the screened dV (src/io/qe/read_dvscf.c:107-158).  All FP64.
#pragma once

#include "common.hpp"

// V_loc(G_i) for every table point: one wave per (type, point); the radial
// Simpson weights w (ngrid) carry the 1/3 and the rule's end corrections.
__global__ void k_vloc_table(const double* __restrict__ gco, int64_t npts,
                             const double* __restrict__ V, const double* __restrict__ r,
                             const double* __restrict__ rab, const double* __restrict__ w,
                             int64_t ngrid, double Z, char cutoff, double volume,
                             double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= npts) return;
    const double G = gco[i];
    double acc = 0;
    for (int64_t j = lane; j < ngrid; j += 64)
    {
        const double rad = r[j];
        double fac = (G < LELPH_EPS) ? rad : sin(rad * G) / G;
        if (cutoff == '3' && G < LELPH_EPS)
            fac *= (rad * V[j] + LELPH_E2 * Z);
        else
            fac *= (rad * V[j] + LELPH_E2 * Z * erf(rad));
        acc += fac * rab[j] * w[j];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) out[i] = 4 * LELPH_RPI * acc / volume;
}

__device__ inline int64_t d_miller(int64_t i, int64_t N)
{
    const int64_t mid = (N - 1) / 2 + 1;
    return i < mid ? i : (i - mid) - N / 2;
}

struct DvlocArgs
{
    double q[3];
    double blat[9];
    double latz; // alat_vec[8]
    double volume;
    int Nx, Ny, Nz;
    int ntype, natom, nm;
    char cutoff;
    int64_t npts;
    double dg;
};

#define DVLOC_MAXATOM 64
#define DVLOC_MAXTYPE 16

// VlocG[nu][box] of dVbare_loc.c:128-214 for every box point.
__global__ void k_dvloc_G(DvlocArgs a, const double* __restrict__ gco,
                          const double* __restrict__ vlocg, const double* __restrict__ vploc,
                          const double* __restrict__ Zval, const int* __restrict__ atype,
                          const double* __restrict__ apos, const z64* __restrict__ eig,
                          z64* __restrict__ VG, int* __restrict__ err)
{
    const int64_t nf = (int64_t)a.Nx * a.Ny * a.Nz;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nf) return;
    const int64_t kz = id % a.Nz;
    const int64_t ixy = id / a.Nz;
    const int64_t jy = ixy % a.Ny, ix = ixy / a.Ny;
    const double qG[3] = {d_miller(ix, a.Nx) + a.q[0], d_miller(jy, a.Ny) + a.q[1],
                          d_miller(kz, a.Nz) + a.q[2]};
    double qc[3];
    for (int i = 0; i < 3; ++i)
        qc[i] = a.blat[3 * i] * qG[0] + a.blat[3 * i + 1] * qG[1] + a.blat[3 * i + 2] * qG[2];
    const double qn = sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2]);
    if (qn > gco[a.npts - 1])
    {
        atomicOr(err, 2);
        for (int v = 0; v < a.nm; ++v) VG[(int64_t)v * nf + id] = {0, 0};
        return;
    }
    int64_t gi = (int64_t)floor(qn / a.dg);
    if (gi > a.npts - 2) gi = a.npts - 2; // |q+G| == table end
    double cf = 1;
    if (a.cutoff == '2')
    {
        const double Gp = sqrt(qc[0] * qc[0] + qc[1] * qc[1]);
        const double qGp = a.latz * Gp;
        double cs = cos(qc[2] * a.latz * 0.5);
        if (Gp > LELPH_EPS) cs = cs - sin(qc[2] * a.latz * 0.5) * qc[2] / Gp;
        cf = 1 - exp(-qGp * 0.5) * cs;
    }
    double Vt[DVLOC_MAXTYPE];
    for (int t = 0; t < a.ntype; ++t)
    {
        // spline_interpolate (src/common/spline.c:13-24)
        const double* yi = vlocg + (int64_t)t * a.npts;
        const double* dy = vploc + (int64_t)t * a.npts;
        const double h = gco[gi + 1] - gco[gi];
        const double tx = (qn - gco[gi]) / h;
        const double ax = dy[gi] * h - (yi[gi + 1] - yi[gi]);
        const double bx = -dy[gi + 1] * h + (yi[gi + 1] - yi[gi]);
        double v = (1 - tx) * yi[gi] + tx * yi[gi + 1] + tx * (1 - tx) * ((1 - tx) * ax + tx * bx);
        if (qn >= LELPH_EPS)
            v -= (4 * LELPH_RPI * LELPH_E2) * Zval[t] * exp(-qn * qn * 0.25) * cf /
                 (qn * qn * a.volume);
        Vt[t] = v;
    }
    z64 fac[DVLOC_MAXATOM];
    for (int ia = 0; ia < a.natom; ++ia)
    {
        const double* p = apos + 3 * ia;
        const double qd = qc[0] * p[0] + qc[1] * p[1] + qc[2] * p[2];
        double s, c;
        sincos(-qd, &s, &c);
        // -i e^{-i q.tau} V
        const double V = Vt[atype[ia]];
        fac[ia] = {s * V, -c * V};
    }
    const int n3 = 3 * a.natom;
    for (int v = 0; v < a.nm; ++v)
    {
        z64 acc = {0, 0};
        const z64* e = eig + (int64_t)v * n3;
        for (int ia = 0; ia < a.natom; ++ia)
        {
            z64 d = {0, 0};
            for (int x = 0; x < 3; ++x) d = cadd(d, cscale(e[3 * ia + x], qc[x]));
            acc = cadd(acc, cmul(d, fac[ia]));
        }
        VG[(int64_t)v * nf + id] = acc;
    }
}

// dVscf[nu][im] += dVloc[nu] (im < mag_iter), then the 2x2 form for nmag = 4
__global__ void k_add_dvscf(z64* __restrict__ dv, const z64* __restrict__ vl, int nm, int nmag,
                            int64_t nf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nm * nf) return;
    const int v = (int)(t / nf);
    const int64_t i = t - (int64_t)v * nf;
    z64* m = dv + (int64_t)v * nmag * nf;
    const z64 b = vl[t];
    const int mi = nmag == 2 ? 2 : 1;
    for (int im = 0; im < mi; ++im) m[im * nf + i] = cadd(m[im * nf + i], b);
    if (nmag == 4)
    {
        const z64 V = m[i], Bx = m[nf + i], By = m[2 * nf + i], Bz = m[3 * nf + i];
        const z64 iBy = {-By.y, By.x};
        m[i] = cadd(V, Bz);
        m[nf + i] = csub(Bx, iBy);
        m[2 * nf + i] = cadd(Bx, iBy);
        m[3 * nf + i] = csub(V, Bz);
    }
}

// out[nu][j] = sum_u EP[nu][u] in[u][j], EP = eig pats^H (nm x nm)
__global__ void k_mode_rotate(const z64* __restrict__ in, const z64* __restrict__ EP, int nm,
                              int64_t ld, z64* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nm * ld) return;
    const int v = (int)(t / ld);
    const int64_t j = t - (int64_t)v * ld;
    z64 acc = {0, 0};
    for (int u = 0; u < nm; ++u) acc = cadd(acc, cmul(EP[v * nm + u], in[(int64_t)u * ld + j]));
    out[t] = acc;
}
