// interp.hpp -- Fourier interpolation of dV_q on the device
// (src/interpolation/interpolation_driver.c:190-420 and :790-835).
//
//   coarse q:  dV_cart(r) = e^{iqr} * rot^H dV_mode(r) after the long-range
//              part is removed                     (k_coarse_cart)
//   q -> R:    dV_R = 1/Nq sum_q e^{-2 pi i q.R} dV_q (k_phase_gemm)
//   R -> q:    dV_q = e^{-iqr} sum_R W_a(q,R) dV_R    (k_phase_gemm, phase
//              epilogue), W_a(q,R) = e^{iqR} (1/deg) sum_T e^{iqT} over the
//              Wigner-Seitz vectors T of atom a      (k_ws_weights)
//   long range: G-box coefficients (k_lr_box, the 3-D and 2-D kernels of
//              src/dvloc/dVlong_range_kernel.c), batched backward FFTs, and
//              a masked add into the nmag components (k_lr_add).
//
// The q -> R and R -> q sums are small-K complex GEMMs (K = coarse grid
// size) with a long N (modes x nmag x N points); k_phase_gemm keeps MB rows
// of accumulators per thread, stages the phase tile through LDS (a wave-wide
// broadcast read) and streams the N axis with coalesced 16 B loads, so the
// long axis is read Nrows/MB times.
#pragma once

#include "common.hpp"

// ---- geometry of the long-range G box (dVpot_add_long_range.c:95-176) ----
struct LrGeom
{
    double blat[9]; // reciprocal vectors (2 pi included), b_i = blat[:, i]
    double eps[9];
    double latz, volume, EcutRy, eta_bare, eta_ind;
    int has_eps, has_zv, has_zb, has_qp;
    int natom, N[3], lo[3], hi[3];
    int dim; // '3' or '2'
};

__device__ inline z64 lr_cexp(double a)
{
    double s, c;
    sincos(a, &s, &c);
    return {c, s};
}

__device__ inline double lr_dot(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// long_range_3D_kernel (dVlong_range_kernel.c:139-230)
__device__ inline void lr_3d(const double* qG, const double* zv, const double* zb,
                             const double* qp, const double* eps, const double* tau,
                             double eta_bare, double eta_ind, z64* out)
{
    if (!zv && !zb && !qp) return;
    const double qq = lr_dot(qG, qG);
    if (sqrt(qq) < LELPH_EPS) return;
    const z64 ph = lr_cexp(-lr_dot(qG, tau));
    double qeq = qq;
    if (eps)
    {
        double t[3];
        for (int i = 0; i < 3; ++i) t[i] = eps[3 * i] * qG[0] + eps[3 * i + 1] * qG[1] + eps[3 * i + 2] * qG[2];
        qeq = lr_dot(t, qG);
    }
    double v[3] = {0, 0, 0}, b[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
    if (zv)
        for (int i = 0; i < 3; ++i) v[i] = -(*zv) * qG[i] / qq;
    if (eps && zb)
        for (int i = 0; i < 3; ++i) b[i] = zb[i] * qG[0] + zb[3 + i] * qG[1] + zb[6 + i] * qG[2];
    if (eps && qp)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) Q[k] += qG[i] * qG[j] * qp[k + 3 * j + 9 * i];
    const double dfb = exp(-qq * 0.25 / eta_bare);
    const double dfi = exp(-fabs(qeq) * 0.25 / eta_ind);
    for (int i = 0; i < 3; ++i)
    {
        // ph * ((b - i Q/2) dfi / qeq + dfb v)
        const z64 t = {b[i] * dfi / qeq + dfb * v[i], -0.5 * Q[i] * dfi / qeq};
        out[i] = cmul(ph, t);
    }
}

// long_range_2D_kernel (dVlong_range_kernel.c:232-458)
__device__ inline void lr_2d(const double* qG, const double* zv, const double* zb,
                             const double* qp, const double* eps, const double* tau, double zlat,
                             double eta_bare, double eta_ind, z64* out)
{
    if (!zv && !zb && !qp) return;
    const double qq = lr_dot(qG, qG);
    if (sqrt(qq) < LELPH_EPS) return;
    const z64 ph = lr_cexp(-lr_dot(qG, tau));
    const double gp = sqrt(qG[0] * qG[0] + qG[1] * qG[1]);
    double csf = cos(qG[2] * zlat * 0.5);
    if (gp > LELPH_EPS) csf = csf - sin(qG[2] * zlat * 0.5) * qG[2] / gp;
    const double cutf = 1.0 - exp(-zlat * gp * 0.5) * csf;
    const double Leff = eps ? eta_ind * zlat * 1.001 * (1.0 - 1.0 / eps[8]) : zlat * 0.5;
    const double fq = 1.0 - tanh(gp * 0.5 * Leff);
    double etp = 1.0;
    if (eps)
    {
        double qaq = (eps[0] - 1.0) * qG[0] * qG[0] + (eps[4] - 1.0) * qG[1] * qG[1] +
                     (eps[1] + eps[3]) * qG[0] * qG[1];
        qaq *= 0.5 * zlat;
        if (gp > LELPH_EPS) etp = 1.0 + fq * qaq / gp;
    }
    double v[3] = {0, 0, 0};
    if (zv)
        for (int i = 0; i < 3; ++i) v[i] = -(*zv) * qG[i];
    double b[3] = {0, 0, 0};
    if (eps && zb)
        for (int i = 0; i < 3; ++i) b[i] = zb[i] * qG[0] + zb[3 + i] * qG[1];
    double qin[3] = {0, 0, 0}, qzz[3] = {0, 0, 0}, qz[3] = {0, 0, 0};
    if (eps && qp)
        for (int a = 0; a < 3; ++a)
        {
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) qin[a] += qG[j] * qG[k] * qp[a + 3 * k + 9 * j];
            qzz[a] = qp[a + 3 * 2 + 9 * 2];
            for (int k = 0; k < 2; ++k) qz[a] += qG[k] * qp[a + 3 * k + 9 * 2];
        }
    const double dfb = exp(-qq * 0.25 / eta_bare);
    const z64 pph = lr_cexp(-(qG[0] * tau[0] + qG[1] * tau[1]));
    const double x = 0.95, Gz = qG[2];
    double fGz_im = 0.0; // f(G_z) is purely imaginary
    if (fabs(Gz) > LELPH_EPS)
    {
        const double term =
            (sin(x * Gz * zlat * 0.5) - x * sin(Gz * zlat * 0.5)) / (Gz * Gz * (1.0 - x));
        fGz_im = -2.0 * term / zlat;
    }
    const double gGz = fabs(Gz) < LELPH_EPS ? 1.0 : 0.0;
    for (int i = 0; i < 3; ++i)
    {
        const z64 vb = cscale(ph, cutf * v[i] * dfb / qq);
        z64 vi = {0, 0};
        if (gp > LELPH_EPS)
        {
            // (2 i b + qin - gp^2 qzz) g / etp
            const z64 tin = {(qin[i] - gp * gp * qzz[i]) * gGz / etp, 2.0 * b[i] * gGz / etp};
            const double zz = zb ? zb[6 + i] / eps[8] : 0.0;
            // 2 gp^2 (zz - i qz) * (i fGz_im)
            const z64 tout = cmul(z64{2.0 * gp * gp * zz, -2.0 * gp * gp * qz[i]}, z64{0.0, fGz_im});
            // zlat / (4 i) = -i zlat / 4
            const z64 pre = cscale(pph, fq / gp);
            vi = cmul(cmul(z64{0.0, -zlat / 4.0}, pre), cadd(tin, tout));
        }
        out[i] = cadd(vb, vi);
    }
}

// One thread per box point and q: the long-range coefficient of every
// (atom, x) at the G of this point (zero outside the G box or past the
// cut-off), x (4 pi i e^2 / Omega), written (nq, 3 natom, N).
__global__ __launch_bounds__(256) void k_lr_box(LrGeom g, const double* __restrict__ qs,
                                                const double* __restrict__ apos,
                                                const double* __restrict__ zvals,
                                                const double* __restrict__ zborn,
                                                const double* __restrict__ qpole,
                                                z64* __restrict__ box)
{
    const int64_t nf = (int64_t)g.N[0] * g.N[1] * g.N[2];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nf) return;
    const int iq = blockIdx.y;
    const int nm = 3 * g.natom;
    z64* out = box + (int64_t)iq * nm * nf + p;
    const int b[3] = {(int)(p / ((int64_t)g.N[1] * g.N[2])), (int)((p / g.N[2]) % g.N[1]),
                      (int)(p % g.N[2])};
    int m[3];
    bool in = true;
    for (int d = 0; d < 3; ++d)
    {
        if (b[d] <= g.hi[d])
            m[d] = b[d];
        else if (b[d] - g.N[d] >= g.lo[d])
            m[d] = b[d] - g.N[d];
        else
            in = false;
    }
    bool skip = !in || (!g.has_eps && !g.has_zv);
    double qG[3], n2 = 0;
    if (!skip)
    {
        const double* q = qs + 3 * iq;
        const double tp = 2.0 * LELPH_RPI;
        for (int i = 0; i < 3; ++i)
        {
            const double qc = (g.blat[3 * i] * q[0] + g.blat[3 * i + 1] * q[1] + g.blat[3 * i + 2] * q[2]) / tp;
            const double gc = (g.blat[3 * i] * m[0] + g.blat[3 * i + 1] * m[1] + g.blat[3 * i + 2] * m[2]) / tp;
            qG[i] = tp * (qc + gc);
        }
        n2 = lr_dot(qG, qG);
        if (g.dim == '3' && n2 > g.EcutRy) skip = true;
        if (g.dim == '2')
        {
            const double gp = sqrt(qG[0] * qG[0] + qG[1] * qG[1]);
            const double Leff = g.has_eps ? g.eta_ind * g.latz * 1.001 * (1.0 - 1.0 / g.eps[8])
                                          : g.latz * 0.5;
            const double fq = 1.0 - tanh(gp * 0.5 * Leff);
            if (fabs(fq) < LELPH_EPS6 && (!g.has_zv || n2 > g.EcutRy)) skip = true;
        }
    }
    const double fac = 4.0 * LELPH_RPI * LELPH_E2 / g.volume; // times i
    for (int a = 0; a < g.natom; ++a)
    {
        z64 o[3] = {{0, 0}, {0, 0}, {0, 0}};
        if (!skip)
        {
            const double* zv = g.has_zv ? zvals + a : nullptr;
            const double* zb = g.has_zb ? zborn + 9 * a : nullptr;
            const double* qp = g.has_qp ? qpole + 27 * a : nullptr;
            const double* eps = g.has_eps ? g.eps : nullptr;
            if (g.dim == '3')
                lr_3d(qG, zv, zb, qp, eps, apos + 3 * a, g.eta_bare, g.eta_ind, o);
            else
                lr_2d(qG, zv, zb, qp, eps, apos + 3 * a, g.latz, g.eta_bare, g.eta_ind, o);
        }
        for (int x = 0; x < 3; ++x)
            out[(int64_t)(3 * a + x) * nf] = cmul(o[x], z64{0.0, fac});
    }
}

// dv[q][v][mag][p] += sgn * box[q][v][p] for the magnetic components in mask
__global__ __launch_bounds__(256) void k_lr_add(const z64* __restrict__ box, z64* __restrict__ dv,
                                                int nm, int nmag, int64_t nf, int mask, double sgn)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nf) return;
    const int64_t qv = blockIdx.y; // q * nm + v
    const z64 b = box[qv * nf + p];
    z64* d = dv + qv * nmag * nf + p;
    for (int m = 0; m < nmag; ++m)
        if (mask >> m & 1) d[m * nf] = cadd(d[m * nf], cscale(b, sgn));
}

// e^{+-2 pi i q.r} at grid point p with the reference's pi (multiply_eikr.c:14-64)
__device__ inline z64 eikr(const double* q, int64_t p, int Nx, int Ny, int Nz, double sgn)
{
    const int ix = (int)(p / ((int64_t)Ny * Nz)), iy = (int)((p / Nz) % Ny), iz = (int)(p % Nz);
    const double r0 = (double)ix / Nx, r1 = (double)iy / Ny, r2 = (double)iz / Nz;
    const double d = q[0] * r0 + q[1] * r1 + q[2] * r2;
    return lr_cexp(sgn * 2.0 * LELPH_RPI * d);
}

// coarse q: out[j][mag][p] = e^{iqr} sum_v conj(rot[v][j]) in[v][mag][p]
// (dVscf_change_basis 'C', dvscf_basis_rotation.c:12-51, then multiply_eikr)
__global__ __launch_bounds__(256) void k_coarse_cart(const z64* __restrict__ in,
                                                     const z64* __restrict__ rot, int nm,
                                                     int nmag, int Nx, int Ny, int Nz, double q0,
                                                     double q1, double q2, z64* __restrict__ out)
{
    const int64_t nf = (int64_t)Nx * Ny * Nz;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nmag * nf) return;
    const int64_t p = t % nf;
    const double q[3] = {q0, q1, q2};
    const z64 e = eikr(q, p, Nx, Ny, Nz, 1.0);
    const int64_t ld = (int64_t)nmag * nf;
    for (int j = 0; j < nm; ++j)
    {
        z64 acc = {0, 0};
        for (int v = 0; v < nm; ++v) acc = cadd(acc, cmul(cconj(rot[v * nm + j]), in[v * ld + t]));
        out[j * ld + t] = cmul(acc, e);
    }
}

// q -> R phases: P[R][q] = e^{-2 pi i (q_idx / Nq) . R} / nq with exact
// integer reduction (fft_q2R's FFTW_FORWARD over the sorted grid,
// interpolation_utilities.c:25-76)
__global__ void k_q2r_phases(int Q0, int Q1, int Q2, z64* __restrict__ P)
{
    const int nq = Q0 * Q1 * Q2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * nq) return;
    const int R = t / nq, q = t % nq;
    const int R0 = R / (Q1 * Q2), R1 = (R / Q2) % Q1, R2 = R % Q2;
    const int a0 = q / (Q1 * Q2), a1 = (q / Q2) % Q1, a2 = q % Q2;
    const double f = (double)((a0 * R0) % Q0) / Q0 + (double)((a1 * R1) % Q1) / Q1 +
                     (double)((a2 * R2) % Q2) / Q2;
    double s, c;
    sincospi(-2.0 * f, &s, &c);
    P[t] = {c / nq, s / nq};
}

// Wigner-Seitz weights W[q][R][a] = e^{iqR} (1/deg) sum_T e^{iqT}
// (fft_R2q_dvscf, interpolation_utilities.c:165-238; reference pi)
__global__ void k_ws_weights(const double* __restrict__ qs, int nq, int Q0, int Q1, int Q2,
                             int natom, const int64_t* __restrict__ off,
                             const int* __restrict__ ws, z64* __restrict__ W)
{
    const int nR = Q0 * Q1 * Q2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nq * nR * natom) return;
    const int a = (int)(t % natom);
    const int R = (int)((t / natom) % nR);
    const int iq = (int)(t / ((int64_t)natom * nR));
    const double* q = qs + 3 * iq;
    auto mil = [](int i, int n) { const int mid = (n - 1) / 2 + 1; return i < mid ? i : (i - mid) - n / 2; };
    const double Rm[3] = {(double)mil(R / (Q1 * Q2), Q0), (double)mil((R / Q2) % Q1, Q1),
                          (double)mil(R % Q2, Q2)};
    const z64 eR = lr_cexp(2.0 * LELPH_RPI * lr_dot(q, Rm));
    const int64_t b = off[(int64_t)R * natom + a], e = off[(int64_t)R * natom + a + 1];
    z64 eT = {0, 0};
    for (int64_t i = b; i < e; ++i)
    {
        const double d = q[0] * ws[3 * i] + q[1] * ws[3 * i + 1] + q[2] * ws[3 * i + 2];
        eT = cadd(eT, lr_cexp(2.0 * LELPH_RPI * d));
    }
    eT = cmul(eT, eR);
    W[t] = cscale(eT, 1.0 / (double)(e - b));
}

// Y[m][a*cpa + c] = sum_k P[(m*K + k)*natP + a'] X[k][a*cpa + c]
// (a' = a when natP > 1, else 0), optional epilogue e^{-2 pi i q_m . r},
// r the grid point of column c (c % nf).  Block: 256 columns of one atom
// slice, MB rows; phase tiles of MB x 64 through LDS.
template <int MB>
__global__ __launch_bounds__(256) void k_phase_gemm(const z64* __restrict__ P, int M, int K,
                                                    int natP, const z64* __restrict__ X,
                                                    int natom, int64_t cpa,
                                                    z64* __restrict__ Y,
                                                    const double* __restrict__ qe, int Nx,
                                                    int Ny, int Nz)
{
    __shared__ z64 ps[MB][64];
    const int64_t chunks = (cpa + 255) / 256;
    const int a = (int)(blockIdx.x / chunks);
    const int64_t c = (blockIdx.x % chunks) * 256 + threadIdx.x;
    const int m0 = blockIdx.y * MB;
    const int ap = natP > 1 ? a : 0;
    const int64_t ld = (int64_t)natom * cpa;
    const bool ok = c < cpa;
    const z64* x = X + (int64_t)a * cpa + (ok ? c : 0);
    z64 acc[MB];
#pragma unroll
    for (int i = 0; i < MB; ++i) acc[i] = {0, 0};
    for (int k0 = 0; k0 < K; k0 += 64)
    {
        __syncthreads();
        for (int i = threadIdx.x; i < MB * 64; i += 256)
        {
            const int r = i / 64, kk = i % 64;
            const int m = m0 + r, k = k0 + kk;
            ps[r][kk] = (m < M && k < K) ? P[((int64_t)m * K + k) * natP + ap] : z64{0, 0};
        }
        __syncthreads();
        const int kn = K - k0 < 64 ? K - k0 : 64;
        for (int kk = 0; kk < kn; ++kk)
        {
            const z64 v = x[(int64_t)(k0 + kk) * ld];
#pragma unroll
            for (int i = 0; i < MB; ++i) acc[i] = cadd(acc[i], cmul(ps[i][kk], v));
        }
    }
    if (!ok) return;
    const int64_t nf = (int64_t)Nx * Ny * Nz;
#pragma unroll
    for (int i = 0; i < MB; ++i)
    {
        const int m = m0 + i;
        if (m >= M) break;
        z64 y = acc[i];
        if (qe) y = cmul(y, eikr(qe + 3 * m, c % nf, Nx, Ny, Nz, -1.0));
        Y[(int64_t)m * ld + (int64_t)a * cpa + c] = y;
    }
}

// rotate_dvscf (rotate_dvscf.c:120-220): grid map r -> S^-1 (r - tau) in
// reduced units with the reference's wrap and integer test
__global__ void k_rot_index(const double* __restrict__ sr, double t0, double t1, double t2, int Nx,
                            int Ny, int Nz, int64_t* __restrict__ idx, int* __restrict__ err)
{
    const int64_t nf = (int64_t)Nx * Ny * Nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int N[3] = {Nx, Ny, Nz};
    const double tr[3] = {t0, t1, t2};
    double t[3] = {(double)(i / ((int64_t)Ny * Nz)), (double)((i / Nz) % Ny), (double)(i % Nz)};
    for (int x = 0; x < 3; ++x) t[x] = t[x] / N[x] - tr[x];
    int64_t id[3];
    for (int x = 0; x < 3; ++x)
    {
        double r = sr[3 * x] * t[0] + sr[3 * x + 1] * t[1] + sr[3 * x + 2] * t[2];
        r -= floor(r);
        r -= floor(r + LELPH_EPS);
        r -= LELPH_EPS;
        r *= N[x];
        if (fabs(r - rint(r)) > 1e-3 + 5e-5 * fabs(r))
        {
            atomicOr(err, 2);
            idx[i] = 0;
            return;
        }
        id[x] = (int64_t)rint(r);
    }
    const int64_t o = (id[0] * Ny + id[1]) * Nz + id[2];
    if (o < 0 || o >= nf)
    {
        atomicOr(err, 4);
        idx[i] = 0;
        return;
    }
    idx[i] = o;
}

// gather with the translation phase, SU(2) sandwich for nmag 4, conjugate
// for time reversal (rotate_dvscf.c:222-345); one thread per (mode, point)
__global__ __launch_bounds__(256) void k_rot_gather(const z64* __restrict__ in,
                                                    const int64_t* __restrict__ idx, int nm,
                                                    int nmag, int64_t nf, z64 ph, z64 s0, z64 s1,
                                                    z64 s2, z64 s3, int composite, int tr,
                                                    z64* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nm * nf) return;
    const int64_t v = t / nf, p = t % nf;
    const int64_t src = idx[p];
    const z64* a = in + v * nmag * nf;
    z64* o = out + v * nmag * nf;
    z64 d[4];
    for (int m = 0; m < nmag; ++m) d[m] = cmul(ph, a[m * nf + src]);
    if (nmag == 4)
    {
        if (!composite)
        {
            const z64 V = d[0], bx = d[1], by = d[2], bz = d[3];
            d[0] = cadd(V, bz);
            d[1] = cadd(bx, z64{by.y, -by.x});  // bx - i by
            d[2] = cadd(bx, z64{-by.y, by.x});  // bx + i by
            d[3] = csub(V, bz);
        }
        const z64 c0 = cconj(s0), c1 = cconj(s1), c2 = cconj(s2), c3 = cconj(s3);
        const z64 u0 = cadd(cmul(d[0], c0), cmul(d[1], c1));
        const z64 w0 = cadd(cmul(d[2], c0), cmul(d[3], c1));
        const z64 u2 = cadd(cmul(d[0], c2), cmul(d[1], c3));
        const z64 w2 = cadd(cmul(d[2], c2), cmul(d[3], c3));
        const z64 M0 = cadd(cmul(s0, u0), cmul(s1, w0));
        const z64 M1 = cadd(cmul(s0, u2), cmul(s1, w2));
        const z64 M2 = cadd(cmul(s2, u0), cmul(s3, w0));
        const z64 M3 = cadd(cmul(s2, u2), cmul(s3, w2));
        if (composite)
        {
            d[0] = M0;
            d[1] = M1;
            d[2] = M2;
            d[3] = M3;
        }
        else
        {
            d[0] = cscale(cadd(M0, M3), 0.5);
            d[1] = cscale(cadd(M1, M2), 0.5);
            const z64 df = csub(M2, M1); // -i (M2 - M1) / 2
            d[2] = cscale(z64{df.y, -df.x}, 0.5);
            d[3] = cscale(csub(M0, M3), 0.5);
        }
    }
    for (int m = 0; m < nmag; ++m) o[m * nf + p] = tr ? cconj(d[m]) : d[m];
}

// QE dvscf file layout (nsets, Nz, Ny, Nx) <-> (nsets, Nx, Ny, Nz)
// (read_dvscf.c:83-101, write_dyn_dvscf_qe.c:145-161): per (set, y) a 2-D
// transpose of Nz x Nx through a 32 x 32 LDS tile, coalesced on both sides.
// TO_XYZ: in (z, y, x) -> out (x, y, z); otherwise the inverse.
template <bool TO_XYZ>
__global__ __launch_bounds__(256) void k_qe_transpose(const z64* __restrict__ in, int Nx, int Ny,
                                                      int Nz, z64* __restrict__ out)
{
    __shared__ z64 t[32][33];
    const int tx = (Nx + 31) / 32;
    const int bx = blockIdx.x % tx, bz = blockIdx.x / tx;
    const int y = blockIdx.y;
    const int64_t nf = (int64_t)Nx * Ny * Nz;
    const z64* src = in + (int64_t)blockIdx.z * nf;
    z64* dst = out + (int64_t)blockIdx.z * nf;
    const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5; // 32 x 8
    // the file side is x-fastest, the grid side z-fastest
    for (int r = ly; r < 32; r += 8)
    {
        if (TO_XYZ)
        {
            const int z = bz * 32 + r, x = bx * 32 + lx;
            if (z < Nz && x < Nx) t[r][lx] = src[((int64_t)z * Ny + y) * Nx + x];
        }
        else
        {
            const int x = bx * 32 + r, z = bz * 32 + lx;
            if (z < Nz && x < Nx) t[r][lx] = src[((int64_t)x * Ny + y) * Nz + z];
        }
    }
    __syncthreads();
    for (int r = ly; r < 32; r += 8)
    {
        if (TO_XYZ)
        {
            const int x = bx * 32 + r, z = bz * 32 + lx;
            if (z < Nz && x < Nx) dst[((int64_t)x * Ny + y) * Nz + z] = t[lx][r];
        }
        else
        {
            const int z = bz * 32 + r, x = bx * 32 + lx;
            if (z < Nz && x < Nx) dst[((int64_t)z * Ny + y) * Nx + x] = t[lx][r];
        }
    }
}
