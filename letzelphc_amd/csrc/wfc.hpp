// wfc.hpp -- G-vector rotation, k+q umklapp shift and sphere -> box scatter.
//
// Replaces rotateGvecs + Sort_pw + su2rotate + apply_trans_wfc + the
// sphere->box step of invfft3D (src/wfc/rotate_Gvecs.c:15-81,
// src/wfc/sortpw.c:14-234, src/wfc/su2rotate.c:13-41,
// src/wfc/trans_wfc.c:15-52, src/fft/bwdfft.c:145-170).  On one GPU no sort
// into z sticks is needed: each coefficient goes straight to
// box[(G_x mod Nx, G_y mod Ny, G_z mod Nz)]; the order of the sphere has no
// numerical effect on the result (the real-space contraction sums over the
// box).
#pragma once

#include "common.hpp"

// per-side geometry, computed on the host in FP64 (compute_Vbare_loc.c:93-147)
struct SideGeom
{
    double M[9];    // lat^T S : cartesian G (no 2pi) -> crystal, rotated
    double G0c[3];  // crystal G shift (k+q side: lat^T (S_kq k_kq - S_k k - q))
    double kc[3];   // crystal k of the rotated point (for the phase)
    double tauc[3]; // fractional translation, crystal
    int conj;       // time reversal: conjugate the coefficients
    z64 su2[4];     // spinor rotation
};

// integer rotated G -> box index and translation phase.  Out-of-box G set
// bit 1 of *err (planner.c:128-137: "gvec > Nfft dim").
__global__ void k_gindex(const double* __restrict__ gvec, int64_t npw, SideGeom geo,
                         int Nx, int Ny, int Nz, int* __restrict__ idx,
                         z64* __restrict__ ph, int* __restrict__ err)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npw) return;
    const double G[3] = {gvec[3 * g], gvec[3 * g + 1], gvec[3 * g + 2]};
    int gi[3];
    const int N[3] = {Nx, Ny, Nz};
    int bad = 0;
    for (int i = 0; i < 3; ++i)
    {
        double c = G[0] * geo.M[3 * i] + G[1] * geo.M[3 * i + 1] + G[2] * geo.M[3 * i + 2] +
                   geo.G0c[i];
        gi[i] = (int)rint(c);
        int b = gi[i] < 0 ? gi[i] + N[i] : gi[i];
        if (b < 0 || b >= N[i]) bad = 1;
        gi[i] = b;
    }
    if (bad)
    {
        atomicOr(err, 1);
        idx[g] = 0;
        ph[g] = {0.0, 0.0};
        return;
    }
    idx[g] = (gi[0] * Ny + gi[1]) * Nz + gi[2];
    // e^{-i 2 pi (k + G).tau} with the integer rotated G (trans_wfc.c:146-158)
    double Gr[3];
    for (int i = 0; i < 3; ++i)
    {
        double c = G[0] * geo.M[3 * i] + G[1] * geo.M[3 * i + 1] + G[2] * geo.M[3 * i + 2] +
                   geo.G0c[i];
        Gr[i] = rint(c);
    }
    double kd = geo.kc[0] * geo.tauc[0] + geo.kc[1] * geo.tauc[1] + geo.kc[2] * geo.tauc[2];
    double gd = Gr[0] * geo.tauc[0] + Gr[1] * geo.tauc[1] + Gr[2] * geo.tauc[2];
    double s1, c1, s2, c2;
    sincos(-2.0 * LELPH_RPI * kd, &s1, &c1);
    sincos(-2.0 * LELPH_RPI * gd, &s2, &c2);
    ph[g] = cmul(z64{c1, s1}, z64{c2, s2});
}

// scatter (nspin*nb) sets of nspinor*npw coefficients into boxes
// out[(set*nspinor + sigma)][Nf]: su2 rotation, phase, optional conjugate.
template <typename T>
__global__ void k_scatter(const z64* __restrict__ wfc, int64_t nsets, int nspinor,
                          int64_t npw, const int* __restrict__ idx,
                          const z64* __restrict__ ph, SideGeom geo, int64_t nf,
                          cpx<T>* __restrict__ out)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsets * npw) return;
    const int64_t s = t / npw, g = t - s * npw;
    const z64* w = wfc + s * nspinor * npw;
    z64 p = ph[g];
    const int64_t id = idx[g];
    if (nspinor == 1)
    {
        z64 c = cmul(w[g], p);
        if (geo.conj) c.y = -c.y;
        out[s * nf + id] = {(T)c.x, (T)c.y};
    }
    else
    {
        z64 u = w[g], d = w[npw + g];
        z64 u2 = cadd(cmul(u, geo.su2[0]), cmul(d, geo.su2[1]));
        z64 d2 = cadd(cmul(u, geo.su2[2]), cmul(d, geo.su2[3]));
        u2 = cmul(u2, p);
        d2 = cmul(d2, p);
        if (geo.conj)
        {
            u2.y = -u2.y;
            d2.y = -d2.y;
        }
        out[(2 * s) * nf + id] = {(T)u2.x, (T)u2.y};
        out[(2 * s + 1) * nf + id] = {(T)d2.x, (T)d2.y};
    }
}

// dst[b][p] = src[b][p] e^{sgn 2 pi i G0 . r_p} for every box b (src == dst
// in place): the backward transform of coefficients shifted by the integer
// G0 (shift theorem), so a psi(r) box built for one umklapp vector serves
// another.  Exact integer reduction of G0 . r before the FP64 sincospi.
template <typename T>
__global__ __launch_bounds__(256) void k_box_shift(const cpx<T>* src, cpx<T>* dst, int64_t nbox,
                                                   int Nx, int Ny, int Nz, int g0, int g1, int g2,
                                                   int sgn)
{
    const int64_t nf = (int64_t)Nx * Ny * Nz;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nf) return;
    const cpx<T> e = box_phase<T>(p, Nx, Ny, Nz, g0, g1, g2, sgn);
    for (int64_t b = 0; b < nbox; ++b) dst[b * nf + p] = cmul(src[b * nf + p], e);
}
