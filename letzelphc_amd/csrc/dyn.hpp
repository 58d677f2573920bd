// dyn.hpp -- Fourier interpolation of dynamical matrices on the device
// (src/interpolation/interpolation_driver.c:420-530, 855-1000;
// src/phonon/dyn_long_range.c, asr.c, mass_normalize_phonons.c;
// interpolation_utilities.c:78-163).
//
//   long range   k_dyn_lr_vec: per (q, G) the dipole (+ quadrupole) vectors
//                X1 = (Z^T(q+G) - i/2 Q(q+G,q+G)) e^{-i(q+G).tau}/((q+G).eps.(q+G))
//                x decay (/sqrt(M)), X2 = X1 (q+G).eps.(q+G) (and the 2-D
//                out-of-plane pair); k_dyn_lr_acc: one block per q,
//                dyn += factor X2^H X1 over G with G tiles in LDS, and the
//                ASR correction of the long range on the diagonal blocks.
//   q -> R / R -> q: the phase GEMM of interp.hpp (q -> R) and
//                k_dyn_r2q (Wigner-Seitz weights per atom pair).
//   force constants: k_fc_asr (mass factors, simple acoustic sum rule).
//   k_herm_jacobi: one block per q, the Hermitian matrix in LDS, cyclic
//                complex Jacobi rotations, eigenvalues ascending, phonon
//                frequencies sign(w) sqrt|w| and eigenvectors / sqrt(M).
#pragma once

#include "interp.hpp"

#define DYN_MAX_NM 64

struct DynLrGeom
{
    double blat[9];
    double ea[9];  // eps (3-D) or alpha = c/2 (eps - 1) in-plane (2-D)
    double eps22;  // eps_zz
    double zlat, eta, factor;
    int Gg[3], GZ;
    int natom, has_zb, has_qp, d2, use_mass;
};

// k_dyn_lr_vec: X1, X2 (and P1, P2 for 2-D) (nq, nG, nm); one thread per (q, G)
// (add_ph_dyn_long_range_internal, dyn_long_range.c:210-370)
__global__ __launch_bounds__(256) void k_dyn_lr_vec(DynLrGeom g, const double* __restrict__ qs,
                                                    const double* __restrict__ apos,
                                                    const double* __restrict__ zborn,
                                                    const double* __restrict__ qpole,
                                                    const double* __restrict__ mass,
                                                    z64* __restrict__ X1, z64* __restrict__ X2,
                                                    z64* __restrict__ P1, z64* __restrict__ P2)
{
    const int64_t nG = (int64_t)g.Gg[0] * g.Gg[1] * g.GZ;
    const int64_t ig = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ig >= nG) return;
    const int iq = blockIdx.y;
    const int nm = 3 * g.natom;
    const int64_t o = ((int64_t)iq * nG + ig) * nm;
    auto mil = [](int i, int n) { const int mid = (n - 1) / 2 + 1; return i < mid ? i : (i - mid) - n / 2; };
    const int gx = (int)(ig / ((int64_t)g.Gg[1] * g.GZ));
    const int gy = (int)((ig / g.GZ) % g.Gg[1]);
    const int gz = (int)(ig % g.GZ);
    const double* q = qs + 3 * iq;
    const double t[3] = {q[0] + mil(gx, g.Gg[0]), q[1] + mil(gy, g.Gg[1]), q[2] + mil(gz, g.GZ)};
    double qG[3];
    for (int i = 0; i < 3; ++i) qG[i] = g.blat[3 * i] * t[0] + g.blat[3 * i + 1] * t[1] + g.blat[3 * i + 2] * t[2];
    const double qn = sqrt(lr_dot(qG, qG));
    bool skip = qn < LELPH_EPS;
    double qeq = 0, qeqp = 1.0, decay = 0;
    if (!skip)
    {
        double tb[3];
        for (int i = 0; i < 3; ++i) tb[i] = g.ea[3 * i] * qG[0] + g.ea[3 * i + 1] * qG[1] + g.ea[3 * i + 2] * qG[2];
        qeq = lr_dot(tb, qG);
        decay = exp(-fabs(qeq) * 0.125 / g.eta);
        if (g.d2)
        {
            const double Leff = g.eta * g.zlat * 1.001 * (1.0 - 1.0 / g.eps22);
            decay = 1 - tanh(qn * 0.5 * Leff);
            if (fabs(decay) < LELPH_EPS6)
                skip = true;
            else
            {
                qeq *= decay;
                qeq += qn;
                qeqp = -qn; // -|q+G| / 1 (the out-of-plane screening kept at order 0)
                decay = sqrt(decay);
            }
        }
    }
    for (int a = 0; a < g.natom; ++a)
    {
        z64 x1[3] = {{0, 0}, {0, 0}, {0, 0}}, p1[3] = {{0, 0}, {0, 0}, {0, 0}},
            p2[3] = {{0, 0}, {0, 0}, {0, 0}};
        double x2s = 0;
        if (!skip)
        {
            const double* Z = g.has_zb ? zborn + 9 * a : nullptr;
            const double* Q = g.has_qp ? qpole + 27 * a : nullptr;
            double zq[3] = {0, 0, 0}, qb[3] = {0, 0, 0};
            if (Z)
                for (int i = 0; i < 3; ++i) zq[i] = Z[i] * qG[0] + Z[3 + i] * qG[1] + Z[6 + i] * qG[2];
            if (Q)
            {
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        for (int k = 0; k < 3; ++k) qb[k] += qG[i] * qG[j] * Q[k + 3 * j + 9 * i];
                if (g.d2)
                    for (int i = 0; i < 3; ++i) qb[i] -= qn * qn * Q[i + 24];
            }
            z64 ph = lr_cexp(-lr_dot(qG, apos + 3 * a));
            double sc = decay / qeq;
            if (g.use_mass) sc /= sqrt(mass[a]);
            ph = cscale(ph, sc);
            for (int i = 0; i < 3; ++i) x1[i] = cmul(z64{zq[i], -0.5 * qb[i]}, ph);
            x2s = qeq;
            if (g.d2)
            {
                const z64 ph2 = cscale(ph, qeq);
                for (int i = 0; i < 3; ++i)
                {
                    if (Z)
                    {
                        p1[i] = cadd(p1[i], cscale(ph2, Z[i + 6] * qeqp / g.eps22));
                        p2[i] = cadd(p2[i], cscale(ph2, Z[i + 6] / g.eps22));
                    }
                    if (Q)
                        for (int x = 0; x < 3; ++x)
                        {
                            // -= i qG_x Q ph2 (qeqp) : i * s * ph2 = (-s ph2.y, s ph2.x)
                            const double s1 = qG[x] * Q[i + 18 + 3 * x] * qeqp;
                            const double s2 = qG[x] * Q[i + 18 + 3 * x];
                            p1[i] = csub(p1[i], z64{-s1 * ph2.y, s1 * ph2.x});
                            p2[i] = csub(p2[i], z64{-s2 * ph2.y, s2 * ph2.x});
                        }
                }
            }
        }
        for (int i = 0; i < 3; ++i)
        {
            X1[o + 3 * a + i] = x1[i];
            X2[o + 3 * a + i] = cscale(x1[i], x2s);
            if (g.d2)
            {
                P1[o + 3 * a + i] = p1[i];
                P2[o + 3 * a + i] = p2[i];
            }
        }
    }
}

// k_dyn_lr_acc: one block per q; dyn[q] += factor (X2^H X1 + P2^H P1) over
// G (tiles of 32 G in LDS), then the long-range ASR correction
// asr (natom, 3, 3) x asr_f / M_a on the diagonal atom blocks
// (add_ph_dyn_long_range, dyn_long_range.c:27-68)
__global__ __launch_bounds__(256) void k_dyn_lr_acc(const z64* __restrict__ X1,
                                                    const z64* __restrict__ X2,
                                                    const z64* __restrict__ P1,
                                                    const z64* __restrict__ P2, int64_t nG, int nm,
                                                    double factor, const z64* __restrict__ asr,
                                                    double asr_f, const double* __restrict__ mass,
                                                    int use_mass, z64* __restrict__ dyn)
{
    __shared__ z64 s1[32][DYN_MAX_NM], s2[32][DYN_MAX_NM];
    const int iq = blockIdx.x;
    const int nn = nm * nm;
    const int per = (nn + 255) / 256; // entries per thread (<= 16 for nm <= 64)
    z64 acc[16];
    for (int e = 0; e < 16; ++e) acc[e] = {0, 0};
    const int npass = P1 ? 2 : 1;
    for (int pass = 0; pass < npass; ++pass)
    {
        const z64* A = (pass == 0 ? X2 : P2) + (int64_t)iq * nG * nm;
        const z64* B = (pass == 0 ? X1 : P1) + (int64_t)iq * nG * nm;
        for (int64_t g0 = 0; g0 < nG; g0 += 32)
        {
            __syncthreads();
            for (int i = threadIdx.x; i < 32 * nm; i += 256)
            {
                const int r = i / nm, c = i % nm;
                const int64_t gg = g0 + r;
                s1[r][c] = gg < nG ? A[gg * nm + c] : z64{0, 0};
                s2[r][c] = gg < nG ? B[gg * nm + c] : z64{0, 0};
            }
            __syncthreads();
            const int gn = nG - g0 < 32 ? (int)(nG - g0) : 32;
            for (int e = 0; e < per; ++e)
            {
                const int idx = threadIdx.x + 256 * e;
                if (idx >= nn) break;
                const int a = idx / nm, b = idx % nm;
                z64 s = acc[e];
                for (int r = 0; r < gn; ++r) s = cadd(s, cmul(cconj(s1[r][a]), s2[r][b]));
                acc[e] = s;
            }
        }
    }
    z64* d = dyn + (int64_t)iq * nn;
    for (int e = 0; e < per; ++e)
    {
        const int idx = threadIdx.x + 256 * e;
        if (idx >= nn) break;
        const int a = idx / nm, b = idx % nm;
        z64 v = cadd(d[idx], cscale(acc[e], factor));
        if (asr && a / 3 == b / 3)
        {
            const int ia = a / 3;
            const double m = use_mass ? 1.0 / mass[ia] : 1.0;
            v = cadd(v, cscale(asr[ia * 9 + 3 * (a % 3) + (b % 3)], asr_f * m));
        }
        d[idx] = v;
    }
}

// Wigner-Seitz weights of atom pairs W[q][R][a][b] = e^{iqR} (1/deg) sum_T e^{iqT}
// and R -> q: dyn[q][i][j] = sum_R W[q][R][i/3][j/3] fc[R][i][j]
// (fft_R2q_dyn, interpolation_utilities.c:78-163); one thread per (q, i, j)
__global__ __launch_bounds__(256) void k_dyn_r2q(const double* __restrict__ qs, int nq, int Q0,
                                                 int Q1, int Q2, int natom,
                                                 const int64_t* __restrict__ off,
                                                 const int* __restrict__ ws,
                                                 const z64* __restrict__ fc, z64* __restrict__ dyn)
{
    const int nm = 3 * natom, nn = nm * nm;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nq * nn) return;
    const int iq = (int)(t / nn), e = (int)(t % nn);
    const int i = e / nm, j = e % nm, a = i / 3, b = j / 3;
    const int nR = Q0 * Q1 * Q2;
    const double* q = qs + 3 * iq;
    auto mil = [](int x, int n) { const int mid = (n - 1) / 2 + 1; return x < mid ? x : (x - mid) - n / 2; };
    z64 acc = {0, 0};
    for (int R = 0; R < nR; ++R)
    {
        const double Rm[3] = {(double)mil(R / (Q1 * Q2), Q0), (double)mil((R / Q2) % Q1, Q1),
                              (double)mil(R % Q2, Q2)};
        const z64 eR = lr_cexp(2.0 * LELPH_RPI * lr_dot(q, Rm));
        const int64_t p = ((int64_t)R * natom + a) * natom + b;
        const int64_t s = off[p], en = off[p + 1];
        z64 eT = {0, 0};
        for (int64_t w = s; w < en; ++w)
            eT = cadd(eT, lr_cexp(2.0 * LELPH_RPI *
                                  (q[0] * ws[3 * w] + q[1] * ws[3 * w + 1] + q[2] * ws[3 * w + 2])));
        eT = cscale(cmul(eT, eR), 1.0 / (double)(en - s));
        acc = cadd(acc, cmul(eT, fc[(int64_t)R * nn + e]));
    }
    dyn[t] = acc;
}

// force constants (nR, nm, nm): x sqrt(Ma Mb), the simple acoustic sum rule
// on the R = 0 block (asr.c:294-318), / sqrt(Ma Mb)
// (interpolation_driver.c:510-532); one thread per (a, alpha, beta)
__global__ void k_fc_asr(z64* __restrict__ fc, int nR, int natom, const double* __restrict__ mass,
                         int use_mass, int asr)
{
    const int nm = 3 * natom;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * natom) return;
    const int a = t / 9, al = (t / 3) % 3, be = t % 3;
    if (!asr) return;
    // the sum over R and b of the mass-unnormalised constants
    z64 sum = {0, 0};
    for (int R = 0; R < nR; ++R)
        for (int b = 0; b < natom; ++b)
        {
            const double m = use_mass ? sqrt(mass[a] * mass[b]) : 1.0;
            sum = cadd(sum, cscale(fc[(int64_t)R * nm * nm + (3 * a + al) * nm + 3 * b + be], m));
        }
    // R = 0 self term: (C - sum) / M_a
    const double ma = use_mass ? mass[a] : 1.0;
    z64& self = fc[(3 * a + al) * nm + 3 * a + be];
    self = cscale(csub(cscale(self, ma), sum), 1.0 / ma);
}

// one block of 64 threads per q: Hermitian part from the lower triangle
// (interpolation_driver.c:945-956 with heev 'L'), cyclic complex Jacobi,
// eigenvalues ascending; w -> sign(w) sqrt|w| and eigenvector rows / sqrt(M)
// (:975-992)
__global__ __launch_bounds__(64) void k_herm_jacobi(const z64* __restrict__ dyn, int nm,
                                                    const double* __restrict__ mass, int use_mass,
                                                    double* __restrict__ freq,
                                                    z64* __restrict__ eig, int max_sweeps)
{
    __shared__ z64 A[DYN_MAX_NM][DYN_MAX_NM + 1];
    __shared__ z64 V[DYN_MAX_NM][DYN_MAX_NM + 1];
    __shared__ double wv[DYN_MAX_NM];
    __shared__ int ord[DYN_MAX_NM];
    __shared__ double red[64];
    __shared__ int go;
    const int iq = blockIdx.x, k = threadIdx.x;
    const z64* D = dyn + (int64_t)iq * nm * nm;
    for (int i = 0; i < nm; ++i)
        if (k < nm)
        {
            // lower triangle (i >= k): 0.5 (D[i][k] + conj(D[k][i])); upper from Hermiticity
            const int r = i >= k ? i : k, c = i >= k ? k : i;
            const z64 lo = cscale(cadd(D[r * nm + c], cconj(D[c * nm + r])), 0.5);
            A[i][k] = i >= k ? lo : cconj(lo);
            if (i == k) A[i][k].y = 0.0;
            V[i][k] = (i == k) ? z64{1, 0} : z64{0, 0};
        }
    __syncthreads();
    for (int sweep = 0; sweep < max_sweeps; ++sweep)
    {
        // off-diagonal norm against the total norm
        double off = 0, tot = 0;
        if (k < nm)
            for (int i = 0; i < nm; ++i)
            {
                const double m2 = A[i][k].x * A[i][k].x + A[i][k].y * A[i][k].y;
                tot += m2;
                if (i != k) off += m2;
            }
        red[k] = off;
        __syncthreads();
        if (k == 0)
        {
            double o = 0;
            for (int i = 0; i < 64; ++i) o += red[i];
            red[0] = o;
        }
        __syncthreads();
        const double offs = red[0];
        __syncthreads();
        red[k] = tot;
        __syncthreads();
        if (k == 0)
        {
            double o = 0;
            for (int i = 0; i < 64; ++i) o += red[i];
            go = offs > 1e-30 * o;
        }
        __syncthreads();
        if (!go) break;
        for (int p = 0; p < nm - 1; ++p)
            for (int q = p + 1; q < nm; ++q)
            {
                const z64 apq = A[p][q];
                const double gm = sqrt(apq.x * apq.x + apq.y * apq.y);
                if (gm < 1e-300) continue; // uniform: every thread reads the same LDS word
                const double app = A[p][p].x, aqq = A[q][q].x;
                const double tau = (aqq - app) / (2.0 * gm);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                const z64 e = {apq.x / gm, -apq.y / gm}; // e^{-i phi}
                __syncthreads();
                // columns p, q of A J and V J (J = [[c, s], [-s e, c e]])
                if (k < nm)
                {
                    const z64 akp = A[k][p], akq = A[k][q];
                    const z64 vkp = V[k][p], vkq = V[k][q];
                    const z64 se = cscale(e, s), ce = cscale(e, c);
                    if (k != p && k != q)
                    {
                        const z64 np = csub(cscale(akp, c), cmul(akq, se));
                        const z64 nq = cadd(cscale(akp, s), cmul(akq, ce));
                        A[k][p] = np;
                        A[k][q] = nq;
                        A[p][k] = cconj(np);
                        A[q][k] = cconj(nq);
                    }
                    V[k][p] = csub(cscale(vkp, c), cmul(vkq, se));
                    V[k][q] = cadd(cscale(vkp, s), cmul(vkq, ce));
                }
                __syncthreads();
                if (k == 0)
                {
                    A[p][p] = {app - t * gm, 0.0};
                    A[q][q] = {aqq + t * gm, 0.0};
                    A[p][q] = {0, 0};
                    A[q][p] = {0, 0};
                }
                __syncthreads();
            }
    }
    // ascending order (a stable rank sort)
    if (k < nm) wv[k] = A[k][k].x;
    __syncthreads();
    if (k < nm)
    {
        int r = 0;
        for (int i = 0; i < nm; ++i)
            if (wv[i] < wv[k] || (wv[i] == wv[k] && i < k)) ++r;
        ord[r] = k;
    }
    __syncthreads();
    if (k < nm)
    {
        const double w = wv[ord[k]];
        const double f = sqrt(fabs(w));
        freq[(int64_t)iq * nm + k] = w < 0 ? -f : f;
        // mode k: column ord[k] of V as a row, / sqrt(M)
        for (int j = 0; j < nm; ++j)
        {
            const double m = use_mass ? 1.0 / sqrt(mass[j / 3]) : 1.0;
            eig[((int64_t)iq * nm + k) * nm + j] = cscale(V[j][ord[k]], m);
        }
    }
}
