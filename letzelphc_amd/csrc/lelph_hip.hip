// lelph_hip.hip -- C-ABI implementation of the MI355X e-ph matrix-element path.
//
// Host orchestration of the gfx950 kernels in fft.hpp, wfc.hpp,
// contract.hpp, nonloc.hpp and dvloc.hpp.  One context owns one HIP stream
// and all device buffers; the iBZ wavefunctions and the current dV_q stay
// resident in HBM across (k, q) pairs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "contract.hpp"
#include "dvloc.hpp"
#include "fft.hpp"
#include "nonloc.hpp"
#include "split.hpp"
#include "splitpc.hpp"
#include "boxes.hpp"
#include "dmats.hpp"
#include "interp.hpp"
#include "dyn.hpp"

#include <type_traits>
#include "wfc.hpp"

namespace
{

struct DevWfc
{
    int64_t npw = 0;
    z64* wfc = nullptr;
    double* gvec = nullptr;
    double* Fk = nullptr;
};

// a growable device buffer
#define LELPH_OUT_SLOTS 3

struct DBuf
{
    void* p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= n) return LELPH_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess)
        {
            p = nullptr;
            return LELPH_ERR_ALLOC;
        }
        n = bytes;
        return LELPH_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename X>
    X* as()
    {
        return reinterpret_cast<X*>(p);
    }
};

} // namespace

struct lelph_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = LELPH_FP64;
    lelph_tuning tune{0, 0, 0, 0, 0};

    // lattice (host copy)
    lelph_lattice lat{};
    std::vector<double> atomic_pos, kpt_ibz, kpt_bz;
    std::vector<int> atom_type, kmap;
    std::vector<lelph_symmetry> syms;

    // pseudo (host copy)
    int ntype = 0, lmax = 0;
    int64_t nlj = 0;
    std::vector<double> PP, Fsign, Zval, socfac;
    std::vector<std::vector<double>> locV, locR, locRab;
    double qmax_abs = 0;
    std::vector<z64> fco;
    std::vector<int64_t> fco_off;
    bool has_fco = false;

    // projectors: plist (natom, pmax, 3) = (lidx, l, l*l + m index) or -1
    int pmax = 0;
    std::vector<int> plist;

    // device constant tables
    int* d_atype = nullptr;
    double* d_apos = nullptr;
    double* d_Fsign = nullptr;
    double* d_socfac = nullptr;
    double* d_Zval = nullptr;
    int* d_plist = nullptr;
    z64* d_fco = nullptr;
    int64_t* d_fco_off = nullptr;
    int* d_err = nullptr;

    FftAxis ax[3];
    int64_t nf = 0;

    std::vector<DevWfc> wf;
    int64_t npw_max = 0;

    // V_loc table
    int64_t npts = 0;
    double dg = 0;
    std::vector<double> h_gco, h_vlocg, h_vploc;
    double* d_gco = nullptr;
    double* d_vlocg = nullptr;
    double* d_vploc = nullptr;

    // current potential
    bool dv_set = false;
    double q[3] = {0, 0, 0};
    std::vector<z64> eig;
    DBuf dv64, dv32, eig_d, vloc_tmp, rot_tmp;
    bool dv32_valid = false;

    // workspaces
    DBuf idx_k, idx_kq, ph_k, ph_kq, box_k, box_kq, part, g, gbatch, presplit;
    // pruned transforms: per side x-plane counts, offsets, cursors, G order
    DBuf pl_cnt[2], pl_off[2], pl_cur[2], pl_perm[2];
    // electronic representation matrices
    DBuf dm_map, dm_ph, dm_pos, dm_wRS, dm_wS2, dm_part, dm_out;
    // q-star rotation
    DBuf qr_dl, qr_dr, qr_g, qr_t, qr_out;
    // Fourier interpolation of dV_q
    bool ip_on = false, ip_ready = false, ip_last_ok = false, ip_has_lr = false;
    int64_t ip_grid[3] = {0, 0, 0};
    int64_t ip_nq = 0;
    std::vector<double> ip_mass;
    std::vector<char> ip_have;
    double ip_last_q[3] = {0, 0, 0};
    lelph_longrange ip_lr{};
    std::vector<double> ip_eps, ip_zb, ip_qp, ip_zv;
    DBuf ip_dvq, ip_dvR, ip_last, ip_tmp, ip_out, ip_ws, ip_off, ip_W, ip_P, ip_qs, ip_rot,
        ip_eig, lr_box, lr_box2, lr_q, lr_zv, lr_zb, lr_qp, rot_idx, rot_sr;
    // dynamical-matrix interpolation
    bool dy_on = false, dy_ready = false;
    int64_t dy_grid[3] = {0, 0, 0}, dy_nq = 0, dy_Gg[3] = {0, 0, 0};
    double dy_eta = 1.0;
    int dy_asr_kind = 0;
    std::vector<double> dy_mass, dy_eps, dy_zb, dy_qp;
    std::vector<char> dy_have;
    DBuf dy_dyn, dy_fc, dy_X1, dy_X2, dy_P1, dy_P2, dy_asr, dy_ws, dy_off, dy_qs, dy_qs2, dy_out,
        dy_freq, dy_eig, dy_zbd, dy_qpd, dy_massd, dy_P, dy_rmass, dy_lr, dy_fclr;
    // host copies for the projected sum rules: Wigner-Seitz vectors / degeneracies
    std::vector<int64_t> dy_ws_h, dy_degen_h;
    bool dy_keep_lr = false; // long-range part kept per coarse q (dy_lr)
    // device time of the last lelph_interp_dvscf: weights, R -> q, long range, total, nq
    hipEvent_t ip_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double ip_ms[5] = {0, 0, 0, 0, 0};
    DBuf nlK, nlKp, Y_K, Y_Kp, wr_K, wr_Kp, beta_K, beta_Kp, bb_K, bb_Kp, bbpart, ebuf;
    // non-local part along a batch's chains: <beta_K|psi> of the previous
    // pair (key: its iBZ k+q and symmetry) serves as this pair's K' side
    DBuf bb_Kprev;
    bool nl_chain = false; // set while a batch issues its pairs along chains
    bool nl_prev_valid = false;
    int nl_prev_ik = -1, nl_prev_sym = -1;
    // <beta_K'|psi> of the chain's head (key: its iBZ k and symmetry): the
    // K side of a closed cycle's last pair
    DBuf bb_Khead;
    bool nl_head_valid = false;
    int nl_head_ik = -1, nl_head_sym = -1;
    // reduced-precision projector operands: FP32 copies, or bf16 piece planes
    DBuf beta32_K, beta32_Kp, wr32_K, wr32_Kp, nlsplitA, nlsplitB;

    // psi box reuse along a batch: the k+q box of the previous pair is the k
    // box of the next when their (iBZ k, symmetry) agree, up to the umklapp
    // shift of the previous k+q side
    bool box_reuse = false, box_prev_valid = false;
    int box_prev_ik = -1, box_prev_sym = -1, box_prev_prec = -1;
    int box_prev_g0[3] = {0, 0, 0};
    // the psi_k box of the chain's head stays put (box_head points into
    // box_k, box_kq or the spare box_x) until the next head: the last pair
    // of a closed k -> k+q cycle, whose k+q side is the head's k side, takes
    // it over with the shift theorem
    DBuf box_x;
    const void* box_head = nullptr;
    int box_head_ik = -1, box_head_sym = -1, box_head_prec = -1;
    int64_t box_sides = 0; // psi sides transformed since the call began
    // k+q box in k-side form, its umklapp phase left to the presplit (set
    // around build_boxes by the bf16 split paths); where build_boxes left
    // the k+q box and the umklapp vector its consumer applies
    bool box_defer = false;
    const void* kq_view = nullptr;
    int kq_g0[3] = {0, 0, 0};

    // per-pair events: [0] start, [1] boxes built, [2] contraction done,
    // [3] local done, [4] non-local done
    std::vector<hipEvent_t> evs;
    int64_t ev_pairs = 0;
    double timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};

    // copy-out of a batch: each pair's g leaves on its own stream into a
    // pinned ring while later pairs compute
    hipStream_t out_stream = nullptr;
    hipStream_t nl_stream = nullptr; // batch: non-local parts beside the local ones
    void* out_ring = nullptr;
    size_t out_ring_bytes = 0;
    hipEvent_t out_ev[LELPH_OUT_SLOTS] = {};
};

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */

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
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[7] * m[3] - m[4] * m[6]);
}

static z64 zc(double re, double im) { return z64{re, im}; }

// SU2mat (src/wfc/su2mat.c:16-117), invert_sym = false
static int su2mat(const double* sym_in, int nspinor, int time_rev, z64* su)
{
    su[0] = zc(1, 0);
    su[1] = zc(0, 0);
    su[2] = zc(0, 0);
    su[3] = zc(1, 0);
    if (nspinor == 1) return LELPH_OK;
    const double det = det3(sym_in);
    if (fabs(fabs(det) - 1) > 1e-4) return LELPH_ERR_SYM;
    double s[9];
    for (int i = 0; i < 9; ++i) s[i] = sym_in[i] / det;
    double al, be, de;
    if (fabs(s[6] + 1) < LELPH_EPS)
    {
        al = 0;
        be = LELPH_RPI / 2.0;
        de = atan2(s[1], s[2]);
    }
    else if (fabs(s[6] - 1) < LELPH_EPS)
    {
        al = 0;
        be = -LELPH_RPI / 2.0;
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
    // exp(-i a sz) exp(-i b sy) exp(-i d sx)
    const z64 sd[4] = {zc(cos(de), 0), zc(0, -sin(de)), zc(0, -sin(de)), zc(cos(de), 0)};
    const z64 sb[4] = {zc(cos(be), 0), zc(-sin(be), 0), zc(sin(be), 0), zc(cos(be), 0)};
    const z64 sa[4] = {zc(cos(al), -sin(al)), zc(0, 0), zc(0, 0), zc(cos(al), sin(al))};
    z64 t[4];
    t[0] = cadd(cmul(sb[0], sd[0]), cmul(sb[1], sd[2]));
    t[1] = cadd(cmul(sb[0], sd[1]), cmul(sb[1], sd[3]));
    t[2] = cadd(cmul(sb[2], sd[0]), cmul(sb[3], sd[2]));
    t[3] = cadd(cmul(sb[2], sd[1]), cmul(sb[3], sd[3]));
    su[0] = cadd(cmul(sa[0], t[0]), cmul(sa[1], t[2]));
    su[1] = cadd(cmul(sa[0], t[1]), cmul(sa[1], t[3]));
    su[2] = cadd(cmul(sa[2], t[0]), cmul(sa[3], t[2]));
    su[3] = cadd(cmul(sa[2], t[1]), cmul(sa[3], t[3]));
    if (time_rev)
    {
        const z64 a1 = su[0], b1 = su[1], c1 = su[2], d1 = su[3];
        su[0] = zc(-c1.x, -c1.y);
        su[1] = zc(-d1.x, -d1.y);
        su[2] = a1;
        su[3] = b1;
    }
    return LELPH_OK;
}

// geometry of elphLocal (src/dvloc/compute_Vbare_loc.c:93-154)
static int local_geometry(const lelph_ctx* c, const double* qpt, int ikq, int ik, int kqsym,
                          int ksym, SideGeom* gkq, SideGeom* gk)
{
    const double* lat = c->lat.alat_vec;
    const double* blat = c->lat.blat_vec;
    const double* kqv = c->kpt_ibz.data() + 3 * ikq;
    const double* kv = c->kpt_ibz.data() + 3 * ik;
    const lelph_symmetry& Skq = c->syms[kqsym];
    const lelph_symmetry& Sk = c->syms[ksym];
    double tSkq[3], tSk[3], kSk[3], kSkq[3], qc[3], ulm[3];
    matvec3(Skq.Rmat, kqv, false, tSkq);
    matvec3(Sk.Rmat, kv, false, tSk);
    matvec3(lat, tSk, true, kSk);
    for (int i = 0; i < 3; ++i) kSkq[i] = kSk[i] + qpt[i];
    for (int i = 0; i < 3; ++i) tSkq[i] -= tSk[i];
    matvec3(blat, qpt, false, qc);
    for (int i = 0; i < 3; ++i) qc[i] /= (2 * LELPH_RPI);
    for (int i = 0; i < 3; ++i) ulm[i] = tSkq[i] - qc[i];
    matvec3(blat, Skq.tau, true, gkq->tauc);
    matvec3(blat, Sk.tau, true, gk->tauc);
    for (int i = 0; i < 3; ++i)
    {
        gkq->tauc[i] /= (2 * LELPH_RPI);
        gk->tauc[i] /= (2 * LELPH_RPI);
    }
    // rotateGvecs: M = lat^T S, G0 -> lat^T G0 (src/wfc/rotate_Gvecs.c:94-115)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            gkq->M[3 * i + j] = lat[i] * Skq.Rmat[j] + lat[3 + i] * Skq.Rmat[3 + j] +
                                lat[6 + i] * Skq.Rmat[6 + j];
            gk->M[3 * i + j] =
                lat[i] * Sk.Rmat[j] + lat[3 + i] * Sk.Rmat[3 + j] + lat[6 + i] * Sk.Rmat[6 + j];
        }
    matvec3(lat, ulm, true, gkq->G0c);
    gk->G0c[0] = gk->G0c[1] = gk->G0c[2] = 0;
    for (int i = 0; i < 3; ++i)
    {
        gkq->kc[i] = kSkq[i];
        gk->kc[i] = kSk[i];
    }
    gkq->conj = Skq.time_rev ? 1 : 0;
    gk->conj = Sk.time_rev ? 1 : 0;
    LELPH_TRY(su2mat(Skq.Rmat, c->lat.nspinor, Skq.time_rev, gkq->su2));
    LELPH_TRY(su2mat(Sk.Rmat, c->lat.nspinor, Sk.time_rev, gk->su2));
    return LELPH_OK;
}

template <typename T>
static cpx<T>* dv_ptr(lelph_ctx* c);
template <>
z64* dv_ptr<double>(lelph_ctx* c)
{
    return c->dv64.as<z64>();
}
template <>
c32* dv_ptr<float>(lelph_ctx* c)
{
    return c->dv32.as<c32>();
}

static int ensure_dv32(lelph_ctx* c)
{
    if (c->dv32_valid) return LELPH_OK;
    const int64_t n = (int64_t)c->lat.nmodes * c->lat.nmag * c->nf;
    LELPH_TRY(c->dv32.ensure(sizeof(c32) * n));
    hipLaunchKernelGGL(k_z64_to_c32, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, c->stream,
                       c->dv64.as<z64>(), c->dv32.as<c32>(), n);
    HIP_TRY(hipGetLastError());
    c->dv32_valid = true;
    return LELPH_OK;
}

// psi_k(r), psi_kq(r) boxes: scatter + backward FFT
// one side of the pruned FP32 transform (boxes.hpp): group the sphere by x
// plane, gather + z transform per occupied plane (and its y transform in
// LDS when the y-z plane is square), else a y pass on occupied planes; the
// x pass reads occupied planes only
static int zero_async(void* p, size_t bytes, hipStream_t st);

static int pruned_side(lelph_ctx* c, const DevWfc& w, const SideGeom& geo, const int* idx,
                       const z64* ph, int side, c32* box)
{
    const lelph_lattice& L = c->lat;
    const int Nx = (int)L.fft_dims[0], Ny = (int)L.fft_dims[1], Nz = (int)L.fft_dims[2];
    const int nyz = Ny * Nz;
    const int nbox = L.nspin * L.nbnds * L.nspinor;
    LELPH_TRY(c->pl_cnt[side].ensure(sizeof(int) * Nx));
    LELPH_TRY(c->pl_off[side].ensure(sizeof(int) * Nx));
    LELPH_TRY(c->pl_cur[side].ensure(sizeof(int) * Nx));
    LELPH_TRY(c->pl_perm[side].ensure(sizeof(int) * std::max<int64_t>(1, w.npw)));
    int* cnt = c->pl_cnt[side].as<int>();
    LELPH_TRY(zero_async(cnt, sizeof(int) * Nx, c->stream));
    const unsigned gb = (unsigned)cdiv(w.npw, 4096);
    hipLaunchKernelGGL(k_plane_count, dim3(gb), dim3(256), 0, c->stream, idx, w.npw, nyz, Nx, cnt);
    hipLaunchKernelGGL(k_plane_scan, dim3(1), dim3(1024), 0, c->stream, cnt, Nx,
                       c->pl_off[side].as<int>(), c->pl_cur[side].as<int>());
    hipLaunchKernelGGL(k_plane_fill, dim3(gb), dim3(256), 0, c->stream, idx, w.npw, nyz, Nx,
                       c->pl_cur[side].as<int>(), c->pl_perm[side].as<int>());
    HIP_TRY(hipGetLastError());
    // cubic y-z planes: the y transform joins the z one in LDS
    const bool fused = plane_z_launch(Nz, Nx, nbox, plane_lds_bytes(Ny, Nz, true), w.wfc,
                                      L.nspinor, w.npw, idx, ph, geo, c->pl_perm[side].as<int>(),
                                      c->pl_off[side].as<int>(), cnt, Ny, c->nf, c->ax[2].w32,
                                      c->ax[1].w32, box, c->stream);
    if (!fused)
    {
        if (!plane_z_launch(Nz, Nx, nbox, plane_lds_bytes(Ny, Nz), w.wfc, L.nspinor, w.npw, idx,
                            ph, geo, c->pl_perm[side].as<int>(), c->pl_off[side].as<int>(), cnt,
                            Ny, c->nf, c->ax[2].w32, nullptr, box, c->stream))
            return LELPH_ERR_UNSUPPORTED;
        HIP_TRY(hipGetLastError());
        if (!fft_try_c<float>(Ny, box, (int64_t)nbox * Nx, Nz, nyz, 1, Nz, +1, c->ax[1].w32,
                              c->stream, cnt, Nx, nullptr, c->tune.fft_lines))
            return LELPH_ERR_UNSUPPORTED;
    }
    HIP_TRY(hipGetLastError());
    if (!fft_try_c<float>(Nx, box, nbox, nyz, c->nf, 1, nyz, +1, c->ax[0].w32, c->stream, nullptr,
                          1, cnt, c->tune.fft_lines))
        return LELPH_ERR_UNSUPPORTED;
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

// orders pairs along k -> k+q chains: a pair follows the one whose k+q side
// (iBZ k, symmetry) is its own k side; chain heads first, then what is left
// of cycles
static void chain_order(const lelph_ctx* c, const lelph_pair* pairs, int64_t npairs,
                        std::vector<int64_t>& order)
{
    order.clear();
    order.reserve((size_t)npairs);
    std::vector<int64_t> next((size_t)npairs, -1);
    std::vector<char> has_prev((size_t)npairs, 0), used((size_t)npairs, 0);
    std::vector<std::pair<int64_t, int64_t>> keys((size_t)npairs);
    for (int64_t i = 0; i < npairs; ++i)
        keys[i] = {(int64_t)pairs[i].ik * c->lat.nsym + pairs[i].ksym, i};
    std::sort(keys.begin(), keys.end());
    for (int64_t i = 0; i < npairs; ++i)
    {
        const int64_t want = (int64_t)pairs[i].ikq * c->lat.nsym + pairs[i].kqsym;
        auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(want, (int64_t)-1));
        if (it != keys.end() && it->first == want && it->second != i && !has_prev[it->second])
        {
            next[i] = it->second;
            has_prev[it->second] = 1;
        }
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t i = 0; i < npairs; ++i)
        {
            if (used[i] || (pass == 0 && has_prev[i])) continue;
            for (int64_t j = i; j >= 0 && !used[j]; j = next[j])
            {
                used[j] = 1;
                order.push_back(j);
            }
        }
}

// psi_k and psi_k+q boxes of a pair into c->box_k, c->box_kq.  k_from: the
// k side is the k+q box of the pair before it in a k -> k+q chain (its
// umklapp vector k_from_g0): taken over with the shift theorem (as is, when
// that vector is zero and k_from is c->box_k itself) instead of
// transformed.  kq_from: the k+q side is the (unshifted) psi_k box of the
// chain's head, the pair whose k side it is; shifted by this pair's umklapp
// vector instead of transformed.  g0_out: this pair's k+q umklapp vector.
// With c->box_defer (FP32 boxes whose k+q side is consumed only through the
// bf16 presplit) the k+q box is built in k-side form, the box of (iBZ k+q,
// symmetry) with no umklapp vector, and the consumer applies e^{+2 pi i
// G0.r} (c->kq_g0) itself: every box is then a plain transform of its key,
// taken over along chains and at a cycle's close without any shift.
// c->kq_view: where the k+q box is (box_kq, or a kept head's box).
template <typename T>
static int build_boxes(lelph_ctx* c, const double* qpt, const lelph_pair& pr,
                       const void* k_from = nullptr, const int* k_from_g0 = nullptr,
                       int* g0_out = nullptr, const void* kq_from = nullptr)
{
    const lelph_lattice& L = c->lat;
    if (pr.ik < 0 || pr.ikq < 0 || pr.ik >= (int)c->wf.size() || pr.ikq >= (int)c->wf.size())
        return LELPH_ERR_ARG;
    if (pr.ksym < 0 || pr.kqsym < 0 || pr.ksym >= L.nsym || pr.kqsym >= L.nsym) return LELPH_ERR_ARG;
    const DevWfc& wk = c->wf[pr.ik];
    const DevWfc& wkq = c->wf[pr.ikq];
    if (!wk.wfc || !wkq.wfc) return LELPH_ERR_STATE;
    SideGeom gkq, gk;
    LELPH_TRY(local_geometry(c, qpt, pr.ikq, pr.ik, pr.kqsym, pr.ksym, &gkq, &gk));
    const bool defer = c->box_defer && std::is_same<T, float>::value;
    // k-side geometry of (ikq, kqsym): the same box the next pair of a chain
    // builds for its k side (the k+q side's translation phase uses
    // S_kq k_kq + R G and does not depend on the umklapp vector)
    SideGeom gkq_k;
    if (defer)
    {
        SideGeom unused;
        LELPH_TRY(local_geometry(c, qpt, pr.ikq, pr.ikq, pr.kqsym, pr.kqsym, &unused, &gkq_k));
    }
    const int64_t nsb = (int64_t)L.nspin * L.nbnds;
    const int64_t nbox = nsb * L.nspinor;
    const size_t bb = sizeof(cpx<T>) * (size_t)(nbox * c->nf);
    LELPH_TRY(c->box_k.ensure(bb));
    LELPH_TRY(c->box_kq.ensure(bb));
    LELPH_TRY(c->idx_k.ensure(sizeof(int) * c->npw_max));
    LELPH_TRY(c->idx_kq.ensure(sizeof(int) * c->npw_max));
    LELPH_TRY(c->ph_k.ensure(sizeof(z64) * c->npw_max));
    LELPH_TRY(c->ph_kq.ensure(sizeof(z64) * c->npw_max));
    const int Nx = (int)L.fft_dims[0], Ny = (int)L.fft_dims[1], Nz = (int)L.fft_dims[2];
    // FP32 boxes: pruned transforms (boxes.hpp) when every axis has a
    // compile-time line kernel and a y-z plane fits in LDS
    const bool pruned = std::is_same<T, float>::value && fft_has_c(Nx) && fft_has_c(Ny) &&
                        fft_has_c(Nz) && Nx <= 1024 && plane_lds_bytes(Ny, Nz) <= 160 * 1024 &&
                        nbox <= 65535;
    // the k side of this pair is the k+q side of the previous one: swap the
    // boxes and undo that side's umklapp shift (none in k-side form) instead
    // of transforming again; boxes of the two forms are never mixed
    const int prec_tag = std::is_same<T, float>::value ? (defer ? 2 : 1) : 0;
    const bool reuse = c->box_reuse && c->box_prev_valid && c->box_prev_ik == pr.ik &&
                       c->box_prev_sym == pr.ksym && c->box_prev_prec == prec_tag;
    int g0[3];
    for (int i = 0; i < 3; ++i) g0[i] = (int)std::lrint(gkq.G0c[i]);
    if (g0_out) std::memcpy(g0_out, g0, sizeof(g0));
    c->box_prev_valid = false;
    const bool linked = k_from != nullptr && !reuse;
    // the head's box is never handed out as a k+q box while it is kept
    auto swap_bufs = [](DBuf& a, DBuf& b) {
        std::swap(a.p, b.p);
        std::swap(a.n, b.n);
    };
    if (reuse)
    {
        swap_bufs(c->box_k, c->box_kq);
        if (c->box_head && c->box_head == c->box_kq.p)
        {
            LELPH_TRY(c->box_x.ensure(bb));
            swap_bufs(c->box_kq, c->box_x);
        }
    }
    else if (c->box_reuse && !linked)
    {
        c->box_head = c->box_k.p;
        c->box_head_ik = pr.ik;
        c->box_head_sym = pr.ksym;
        c->box_head_prec = prec_tag;
    }
    // the k+q side closes a cycle: it is the head's k side
    const void* kq_src = kq_from;
    if (!kq_src && c->box_reuse && c->box_head && c->box_head_ik == pr.ikq &&
        c->box_head_sym == pr.kqsym && c->box_head_prec == prec_tag)
        kq_src = c->box_head;
    const bool build_kq = kq_src == nullptr;
    if (reuse)
    {
        const int* pg = c->box_prev_g0;
        if (pg[0] || pg[1] || pg[2])
        {
            hipLaunchKernelGGL(k_box_shift<T>, dim3((unsigned)cdiv(c->nf, 256)), dim3(256), 0,
                               c->stream, c->box_k.as<cpx<T>>(), c->box_k.as<cpx<T>>(), nbox, Nx,
                               Ny, Nz, pg[0], pg[1], pg[2], -1);
            HIP_TRY(hipGetLastError());
        }
        if (!pruned && build_kq) LELPH_TRY(zero_async(c->box_kq.p, bb, c->stream));
    }
    else if (linked)
    {
        const int* pg = k_from_g0;
        if (pg[0] || pg[1] || pg[2])
        {
            hipLaunchKernelGGL(k_box_shift<T>, dim3((unsigned)cdiv(c->nf, 256)), dim3(256), 0,
                               c->stream, static_cast<const cpx<T>*>(k_from),
                               c->box_k.as<cpx<T>>(), nbox, Nx, Ny, Nz, pg[0], pg[1], pg[2], -1);
            HIP_TRY(hipGetLastError());
        }
        else if (k_from != c->box_k.p)
            HIP_TRY(hipMemcpyAsync(c->box_k.p, k_from, bb, hipMemcpyDeviceToDevice, c->stream));
        if (!pruned && build_kq) LELPH_TRY(zero_async(c->box_kq.p, bb, c->stream));
    }
    else
    {
        if (!pruned)
        {
            LELPH_TRY(zero_async(c->box_k.p, bb, c->stream));
            if (build_kq) LELPH_TRY(zero_async(c->box_kq.p, bb, c->stream));
        }
        hipLaunchKernelGGL(k_gindex, dim3((unsigned)cdiv(wk.npw, 256)), dim3(256), 0, c->stream,
                           wk.gvec, wk.npw, gk, Nx, Ny, Nz, c->idx_k.as<int>(),
                           c->ph_k.as<z64>(), c->d_err);
    }
    // the k+q side's G + G0 must fit the box (checked also where the side
    // is not transformed or is built in k-side form)
    hipLaunchKernelGGL(k_gindex, dim3((unsigned)cdiv(wkq.npw, 256)), dim3(256), 0, c->stream,
                       wkq.gvec, wkq.npw, gkq, Nx, Ny, Nz, c->idx_kq.as<int>(),
                       c->ph_kq.as<z64>(), c->d_err);
    const SideGeom& gkq_b = defer ? gkq_k : gkq;
    if (defer && build_kq)
        hipLaunchKernelGGL(k_gindex, dim3((unsigned)cdiv(wkq.npw, 256)), dim3(256), 0, c->stream,
                           wkq.gvec, wkq.npw, gkq_b, Nx, Ny, Nz, c->idx_kq.as<int>(),
                           c->ph_kq.as<z64>(), c->d_err);
    const bool build_k = !reuse && !linked;
    c->box_sides += (build_k ? 1 : 0) + (build_kq ? 1 : 0);
    bool done = false;
    if constexpr (std::is_same<T, float>::value)
        if (pruned)
        {
            if (build_k)
                LELPH_TRY(pruned_side(c, wk, gk, c->idx_k.as<int>(), c->ph_k.as<z64>(), 0,
                                      c->box_k.as<c32>()));
            if (build_kq)
                LELPH_TRY(pruned_side(c, wkq, gkq_b, c->idx_kq.as<int>(), c->ph_kq.as<z64>(), 1,
                                      c->box_kq.as<c32>()));
            done = true;
        }
    if (!done)
    {
        if (build_k)
            hipLaunchKernelGGL(k_scatter<T>, dim3((unsigned)cdiv(nsb * wk.npw, 256)), dim3(256),
                               0, c->stream, wk.wfc, nsb, L.nspinor, wk.npw, c->idx_k.as<int>(),
                               c->ph_k.as<z64>(), gk, c->nf, c->box_k.as<cpx<T>>());
        if (build_kq)
            hipLaunchKernelGGL(k_scatter<T>, dim3((unsigned)cdiv(nsb * wkq.npw, 256)), dim3(256),
                               0, c->stream, wkq.wfc, nsb, L.nspinor, wkq.npw,
                               c->idx_kq.as<int>(), c->ph_kq.as<z64>(), gkq_b, c->nf,
                               c->box_kq.as<cpx<T>>());
        HIP_TRY(hipGetLastError());
        if (build_k)
            LELPH_TRY(fft3d_batch<T>(c->ax, c->box_k.as<cpx<T>>(), nbox, +1, c->stream,
                                     c->tune.fft_lines));
        if (build_kq)
            LELPH_TRY(fft3d_batch<T>(c->ax, c->box_kq.as<cpx<T>>(), nbox, +1, c->stream,
                                     c->tune.fft_lines));
    }
    const void* kq_view = c->box_kq.p;
    if (!build_kq && defer)
    {
        // k-side form: the head's box is this pair's k+q box as it is; kept
        // in the spare, it becomes box_kq so that a chain may go on from it
        if (kq_src == c->box_x.p)
            swap_bufs(c->box_kq, c->box_x);
        kq_view = kq_src;
    }
    else if (!build_kq)
    {
        if (g0[0] || g0[1] || g0[2])
        {
            hipLaunchKernelGGL(k_box_shift<T>, dim3((unsigned)cdiv(c->nf, 256)), dim3(256), 0,
                               c->stream, static_cast<const cpx<T>*>(kq_src),
                               c->box_kq.as<cpx<T>>(), nbox, Nx, Ny, Nz, g0[0], g0[1], g0[2], +1);
            HIP_TRY(hipGetLastError());
        }
        else if (kq_src != c->box_kq.p)
            HIP_TRY(hipMemcpyAsync(c->box_kq.p, kq_src, bb, hipMemcpyDeviceToDevice, c->stream));
    }
    c->kq_view = kq_view;
    for (int i = 0; i < 3; ++i) c->kq_g0[i] = defer ? g0[i] : 0;
    if (c->box_reuse && kq_view == c->box_kq.p)
    {
        c->box_prev_valid = true;
        c->box_prev_ik = pr.ikq;
        c->box_prev_sym = pr.kqsym;
        c->box_prev_prec = prec_tag;
        for (int i = 0; i < 3; ++i) c->box_prev_g0[i] = defer ? 0 : g0[i];
    }
    return LELPH_OK;
}

// zeroes device memory with a kernel (graph-capturable, unlike a memset
// node); bytes a multiple of 4, 16-byte units when aligned
static int zero_async(void* p, size_t bytes, hipStream_t st)
{
    if (!bytes) return LELPH_OK;
    if (((uintptr_t)p % 16) == 0 && bytes % 16 == 0)
    {
        const int64_t n = (int64_t)(bytes / 16);
        const unsigned nb = (unsigned)std::min<int64_t>(cdiv(n, 256), 65536);
        hipLaunchKernelGGL(k_zero_u4, dim3(nb), dim3(256), 0, st, static_cast<uint4*>(p), n);
    }
    else
    {
        const int64_t n = (int64_t)(bytes / 4);
        const unsigned nb = (unsigned)std::min<int64_t>(cdiv(n, 256), 65536);
        hipLaunchKernelGGL(k_zero_words, dim3(nb), dim3(256), 0, st, static_cast<unsigned*>(p), n);
    }
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

static int pick_splits(int64_t base_blocks, int64_t nchunks, int64_t out_elems, int64_t* cps,
                       int* nsplit, int64_t target = 2048)
{
    int64_t ns = cdiv(target > 0 ? target : 2048, base_blocks > 0 ? base_blocks : 1);
    if (ns < 1) ns = 1;
    // keep the partial buffer below ~512 MB
    const int64_t cap = std::max<int64_t>(1, (512ll << 20) / (16 * std::max<int64_t>(out_elems, 1)));
    ns = std::min(ns, cap);
    ns = std::min(ns, nchunks);
    if (ns < 1) ns = 1;
    *cps = cdiv(nchunks, ns);
    *nsplit = (int)cdiv(nchunks, *cps);
    return LELPH_OK;
}

// true when the contraction of this context runs k_contract_sep
template <typename T>
static bool contract_sep_form(const lelph_ctx* c)
{
    const int nmag_eff = (c->lat.nspinor == 1 && c->lat.nmag == 4) ? 1 : c->lat.nmag;
    return std::is_same<T, float>::value && !(nmag_eff == 4 && c->lat.nspinor == 2) &&
           c->tune.fp32_form == 0;
}

// block tile of the FP64 / FP32 contraction (contract.hpp): 1 CtWide (2
// modes x 32 x 64), 2 Ct32 (4 x 32 x 32), 3 Ct16x48 (4 x 16 x 48), 4 Ct16 (4 x
// 16 x 16).  All give bit-identical g; they differ in padding.
static const int CT_SHAPES[4][3] = {{2, 32, 64}, {4, 32, 32}, {4, 16, 48}, {4, 16, 16}};

static int64_t ct_tiles(int shape, int nm, int nb)
{
    const int* s = CT_SHAPES[shape - 1];
    return (int64_t)cdiv(nm, s[0]) * cdiv(nb, s[1]) * cdiv(nb, s[2]);
}

// FP32: the 16 x 16 tile (its blocks are small enough for many waves per
// SIMD: 57.0 against 76.5 ms per c2 pair with the wide tile, and the
// autotuner's pick on the 8-cell workload's FP32 grids).  FP64: the shape
// with the least padding for (nm, nb), weighed by how well its wave tiles
// reuse their fragments (a 16 x 16 wave tile reads twice the LDS per MFMA
// of a 32 x 32 one); ties go to the larger tiles (the wide tile at c2:
// 129.5 ms, the others 137..141 ms)
static int ct_auto(int nm, int nb, bool fp32)
{
    if (fp32) return 4;
    static const double reuse[4] = {1.0, 0.97, 0.85, 0.7};
    static const int order[4] = {1, 3, 2, 4};
    int best = 1;
    double bs = -1.0;
    for (int k : order)
    {
        const int* s = CT_SHAPES[k - 1];
        const double util = (double)nm / (cdiv(nm, s[0]) * s[0]) *
                            ((double)nb / (cdiv(nb, s[1]) * s[1])) *
                            ((double)nb / (cdiv(nb, s[2]) * s[2]));
        const double sc = util * reuse[k - 1];
        if (sc > bs * 1.01)
        {
            bs = sc;
            best = k;
        }
    }
    return best;
}

static int ct_shape(const lelph_ctx* c, bool fp32)
{
    return c->tune.ctile ? c->tune.ctile : ct_auto(c->lat.nmodes, c->lat.nbnds, fp32);
}

template <typename T, typename S>
static void launch_contract_s(bool sep, dim3 grid, hipStream_t st, const cpx<T>* psik,
                              const cpx<T>* psikq, const cpx<T>* dv, z64* part, int nm, int nb,
                              int nspin, int nspinor, int nmag, int64_t N, int64_t cps)
{
    if constexpr (std::is_same<T, float>::value)
        if (sep)
        {
            hipLaunchKernelGGL((k_contract_sep<T, S>), grid, dim3(256), 0, st, psik, psikq, dv,
                               part, nm, nb, nspin, nspinor, nmag, N, cps);
            return;
        }
    hipLaunchKernelGGL((k_contract<T, S>), grid, dim3(256), 0, st, psik, psikq, dv, part, nm, nb,
                       nspin, nspinor, nmag, N, cps);
}

template <typename T>
static void launch_contract(int shape, bool sep, dim3 grid, hipStream_t st, const cpx<T>* psik,
                            const cpx<T>* psikq, const cpx<T>* dv, z64* part, int nm, int nb,
                            int nspin, int nspinor, int nmag, int64_t N, int64_t cps)
{
    switch (shape)
    {
    case 2:
        launch_contract_s<T, Ct32>(sep, grid, st, psik, psikq, dv, part, nm, nb, nspin, nspinor,
                                   nmag, N, cps);
        break;
    case 3:
        launch_contract_s<T, Ct16x48>(sep, grid, st, psik, psikq, dv, part, nm, nb, nspin, nspinor,
                                   nmag, N, cps);
        break;
    case 4:
        launch_contract_s<T, Ct16>(sep, grid, st, psik, psikq, dv, part, nm, nb, nspin, nspinor,
                                   nmag, N, cps);
        break;
    default:
        launch_contract_s<T, CtWide>(sep, grid, st, psik, psikq, dv, part, nm, nb, nspin,
                                     nspinor, nmag, N, cps);
    }
}

// local part into g (device): elphLocal
template <typename T>
static int run_local(lelph_ctx* c, const double* qpt, const lelph_pair& pr, z64* g,
                     hipEvent_t* ev)
{
    const lelph_lattice& L = c->lat;
    if (!c->dv_set) return LELPH_ERR_STATE;
    if (L.nmag == 2 && (L.nspin != 2 || L.nspinor != 1)) return LELPH_ERR_ARG;
    LELPH_TRY(build_boxes<T>(c, qpt, pr));
    HIP_TRY(hipEventRecord(ev[1], c->stream));
    const int nm = L.nmodes, nb = L.nbnds, nspin = L.nspin;
    const int nmag_eff = (L.nspinor == 1 && L.nmag == 4) ? 1 : L.nmag;
    const int64_t nchunks = cdiv(c->nf, CT_KC);
    // the K splits follow the wide tile's count whatever the shape, so every
    // shape reduces the same K ranges (bit-identical g)
    const int64_t tiles = ct_tiles(1, nm, nb);
    const int shape = ct_shape(c, std::is_same<T, float>::value);
    const int64_t gsz = (int64_t)nm * nspin * nb * nb;
    int64_t cps;
    int nsplit;
    pick_splits(tiles * nspin, nchunks, gsz, &cps, &nsplit, c->tune.split_target);
    LELPH_TRY(c->part.ensure(sizeof(z64) * (size_t)(nsplit * gsz)));
    // FP32: dV and psi_k staged separately unless the potential is a 2x2
    // spinor one (87.0 -> 75.5 ms per c2 pair) or the tuning asks for the
    // element-wise form; FP64 forms A element by element (its FP64 products
    // beside the FP64 MFMAs made the separate form 21 % slower)
    launch_contract<T>(shape, contract_sep_form<T>(c),
                       dim3((unsigned)ct_tiles(shape, nm, nb), (unsigned)nsplit, (unsigned)nspin),
                       c->stream, c->box_k.as<cpx<T>>(), c->box_kq.as<cpx<T>>(), dv_ptr<T>(c),
                       c->part.as<z64>(), nm, nb, nspin, L.nspinor, nmag_eff, c->nf, cps);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[2], c->stream));
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv(gsz, 256)), dim3(256), 0, c->stream,
                       c->part.as<z64>(), nsplit, gsz, 1.0 / (double)c->nf, g);
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

// non-local part added onto g (device): add_elphNonLocal
// non-local part of one pair on stream st; its last kernel adds into g and
// waits for mode_after first when that is given (the local part of the same
// pair, which writes g, runs on the context stream)
static int effective_precision(const lelph_ctx* c);

// one side's <beta|psi> projector GEMM into out (rows, ncol) from the beta /
// psi pieces of k_nl_presplit on BF16 MFMA (NP pieces, NPROD products),
// FP64 split-K partials and a fixed-order FP64 reduction
template <int NP, int NPROD>
static int nl_gemm_split(lelph_ctx* c, const z64* A, const z64* B, int rows, int64_t ncol,
                         int64_t npw, int conjB, z64* out, hipStream_t st)
{
    const int Mp = (int)(cdiv(rows, 64) * 64), Np = (int)(cdiv(ncol, 64) * 64);
    const int64_t Kp = cdiv(npw, 32) * 32;
    LELPH_TRY(c->nlsplitA.ensure(sizeof(__bf16) * (size_t)NP * 2 * Mp * Kp));
    LELPH_TRY(c->nlsplitB.ensure(sizeof(__bf16) * (size_t)NP * 2 * Np * Kp));
    hipLaunchKernelGGL((k_nl_presplit<NP>), dim3((unsigned)cdiv(Kp, 256), (unsigned)Mp), dim3(256),
                       0, st, A, rows, npw, Mp, Kp, 0, c->nlsplitA.as<__bf16>());
    hipLaunchKernelGGL((k_nl_presplit<NP>), dim3((unsigned)cdiv(Kp, 256), (unsigned)Np), dim3(256),
                       0, st, B, (int)ncol, npw, Np, Kp, conjB, c->nlsplitB.as<__bf16>());
    HIP_TRY(hipGetLastError());
    const int64_t nchunks = Kp / 32;
    const int64_t tiles = (int64_t)(Mp / 64) * (Np / 64);
    int64_t cps;
    int nsplit;
    pick_splits(tiles, nchunks, (int64_t)rows * ncol, &cps, &nsplit);
    const size_t need = sizeof(z64) * (size_t)nsplit * rows * ncol;
    if (need > c->bbpart.n && st != c->stream) HIP_TRY(hipStreamSynchronize(st));
    LELPH_TRY(c->bbpart.ensure(need));
    hipLaunchKernelGGL((k_zgemm_nt_bf<NP, NPROD>), dim3((unsigned)tiles, (unsigned)nsplit),
                       dim3(256), 0, st, c->nlsplitA.as<__bf16>(), c->nlsplitB.as<__bf16>(), rows,
                       (int)ncol, Mp, Np, Kp, cps, c->bbpart.as<z64>());
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv((int64_t)rows * ncol, 256)), dim3(256),
                       0, st, c->bbpart.as<z64>(), nsplit, (int64_t)rows * ncol, 1.0, out);
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

// the cart -> mode rotation of the non-local part, accumulated into g
// (src/nonloc/Vbare_nloc.c:541-558): g += -2 pi i e^2 eig ebuf
static int nl_mode(lelph_ctx* c, const z64* ebuf, z64* g, hipStream_t st)
{
    const lelph_lattice& L = c->lat;
    const int64_t es = (int64_t)L.nspin * L.nbnds * L.nbnds;
    const z64 pre = {0.0, -2 * LELPH_RPI * LELPH_E2};
    hipLaunchKernelGGL(k_nl_mode, dim3((unsigned)cdiv((int64_t)L.nmodes * es, 256)), dim3(256), 0,
                       st, ebuf, c->eig_d.as<z64>(), L.nmodes, 3 * L.natom, es, pre, g);
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

// add_elphNonLocal into g on stream st; the final mode rotation waits for
// mode_after when given.  With ebuf_out the per-atom K/K' sums go there and
// the rotation is left to the caller (nl_mode), so g is not touched.
static int run_nonlocal(lelph_ctx* c, const lelph_pair& pr, z64* g, hipStream_t st,
                        hipEvent_t mode_after, z64* ebuf_out = nullptr)
{
    const lelph_lattice& L = c->lat;
    if (!c->dv_set) return LELPH_ERR_STATE;
    const DevWfc& wK = c->wf[pr.ikq];
    const DevWfc& wKp = c->wf[pr.ik];
    if (!wK.Fk || !wKp.Fk) return LELPH_ERR_STATE;
    const int nlm = (c->lmax + 1) * (c->lmax + 1);
    const int64_t nsb = (int64_t)L.nspin * L.nbnds;
    const int64_t ncol = nsb * L.nspinor;
    const int natom = L.natom;
    const int rows = natom * c->pmax * 4;
    const int64_t npm = c->npw_max;
    LELPH_TRY(c->nlK.ensure(sizeof(double) * 3 * npm));
    LELPH_TRY(c->nlKp.ensure(sizeof(double) * 3 * npm));
    LELPH_TRY(c->Y_K.ensure(sizeof(double) * nlm * npm));
    LELPH_TRY(c->Y_Kp.ensure(sizeof(double) * nlm * npm));
    LELPH_TRY(c->wr_K.ensure(sizeof(z64) * ncol * npm));
    LELPH_TRY(c->wr_Kp.ensure(sizeof(z64) * ncol * npm));
    LELPH_TRY(c->beta_K.ensure(sizeof(z64) * (size_t)rows * npm));
    LELPH_TRY(c->beta_Kp.ensure(sizeof(z64) * (size_t)rows * npm));
    LELPH_TRY(c->bb_K.ensure(sizeof(z64) * (size_t)rows * ncol));
    LELPH_TRY(c->bb_Kp.ensure(sizeof(z64) * (size_t)rows * ncol));
    const int64_t es = (int64_t)L.nspin * L.nbnds * L.nbnds;
    LELPH_TRY(c->ebuf.ensure(sizeof(z64) * (size_t)(3 * natom * es)));
    // in a batch, the K' side of a pair whose k side is the previous pair's
    // k+q side comes from that pair's K-side projections (k_nl_mirror)
    const bool mirror = c->nl_chain && c->nl_prev_valid && c->nl_prev_ik == pr.ik &&
                        c->nl_prev_sym == pr.ksym;
    // and the K side of a pair whose k+q side is the k side of the chain's
    // head (a closed cycle) from the head's K' projections
    const bool close_k = c->nl_chain && c->nl_head_valid && c->nl_head_ik == pr.ikq &&
                         c->nl_head_sym == pr.kqsym;
    c->nl_prev_valid = false;
    auto skip_side = [&](int side) { return side == 0 ? close_k : mirror; };

    const lelph_symmetry& SK = c->syms[pr.kqsym];
    const lelph_symmetry& SKp = c->syms[pr.ksym];
    NlSide sK{}, sKp{};
    std::memcpy(sK.S, SK.Rmat, sizeof(sK.S));
    std::memcpy(sKp.S, SKp.Rmat, sizeof(sKp.S));
    std::memcpy(sK.k, c->kpt_ibz.data() + 3 * pr.ikq, sizeof(sK.k));
    std::memcpy(sKp.k, c->kpt_ibz.data() + 3 * pr.ik, sizeof(sKp.k));
    std::memcpy(sK.tau, SK.tau, sizeof(sK.tau));
    std::memcpy(sKp.tau, SKp.tau, sizeof(sKp.tau));
    sK.conj_w = SK.time_rev ? 0 : 1;   // 'C' unless time reversed (Vbare_nloc.c:397-407)
    sKp.conj_w = SKp.time_rev ? 1 : 0; // 'T' unless time reversed
    LELPH_TRY(su2mat(SK.Rmat, L.nspinor, SK.time_rev, sK.su2));
    LELPH_TRY(su2mat(SKp.Rmat, L.nspinor, SKp.time_rev, sKp.su2));

    // projector GEMMs in FP32 (FP32 contexts; bf16 ones unless nl_form 1):
    // the rotated psi is written in FP32 and beta built in FP32 only
    const int prec = effective_precision(c);
    const bool nl32 = prec != LELPH_FP64 &&
                      !((prec == LELPH_BF16X3 || prec == LELPH_BF16X6) && c->tune.nl_form == 1);
    // K, Y_lm and the rotated psi of both sides; the rotation in groups of
    // NL_SPY sets per block row
    constexpr int NL_SPY = 16;
    const unsigned gy = (unsigned)std::max<int64_t>(1, cdiv(nsb, NL_SPY));
    for (int side = 0; side < 2; ++side)
    {
        if (skip_side(side)) continue;
        const DevWfc& w = side == 0 ? wK : wKp;
        const NlSide& sd = side == 0 ? sK : sKp;
        double* kc = side == 0 ? c->nlK.as<double>() : c->nlKp.as<double>();
        double* y = side == 0 ? c->Y_K.as<double>() : c->Y_Kp.as<double>();
        if (nl32)
        {
            DBuf& w32 = side == 0 ? c->wr32_K : c->wr32_Kp;
            LELPH_TRY(w32.ensure(sizeof(c32) * (size_t)ncol * npm));
            hipLaunchKernelGGL(k_nl_geom<float>, dim3((unsigned)cdiv(w.npw, 128), gy), dim3(128),
                               0, st, w.gvec, w.npw, sd, c->lmax, w.wfc, nsb, L.nspinor, kc, y,
                               w32.as<c32>(), NL_SPY);
        }
        else
            hipLaunchKernelGGL(k_nl_geom<double>, dim3((unsigned)cdiv(w.npw, 128), gy),
                               dim3(128), 0, st, w.gvec, w.npw, sd, c->lmax, w.wfc, nsb,
                               L.nspinor, kc, y,
                               side == 0 ? c->wr_K.as<z64>() : c->wr_Kp.as<z64>(), NL_SPY);
    }
    NlProj pj{natom, c->pmax, c->ntype};
    if (!nl32)
    {
        if (!close_k)
            hipLaunchKernelGGL(k_nl_beta<double>,
                               dim3((unsigned)cdiv(wK.npw, 256), natom * c->pmax), dim3(256), 0,
                               st, c->nlK.as<double>(), c->Y_K.as<double>(), wK.Fk, wK.npw, pj,
                               c->d_plist, c->d_atype, c->d_apos, c->d_Fsign, -1,
                               c->beta_K.as<z64>());
        if (!mirror)
            hipLaunchKernelGGL(k_nl_beta<double>,
                               dim3((unsigned)cdiv(wKp.npw, 256), natom * c->pmax), dim3(256), 0,
                               st, c->nlKp.as<double>(), c->Y_Kp.as<double>(), wKp.Fk, wKp.npw,
                               pj, c->d_plist, c->d_atype, c->d_apos, c->d_Fsign, +1,
                               c->beta_Kp.as<z64>());
    }
    HIP_TRY(hipGetLastError());
    for (int side = 0; side < 2 && prec != LELPH_FP64; ++side)
    {
        if (skip_side(side)) continue;
        const DevWfc& w = side == 0 ? wK : wKp;
        const z64* A = side == 0 ? c->beta_K.as<z64>() : c->beta_Kp.as<z64>();
        const z64* B = side == 0 ? c->wr_K.as<z64>() : c->wr_Kp.as<z64>();
        z64* out = side == 0 ? c->bb_K.as<z64>() : c->bb_Kp.as<z64>();
        const int conjB = side == 0 ? sK.conj_w : sKp.conj_w;
        // bf16 contexts: FP32 operands on FP32 MFMA by default (c2: 0.16 ms
        // per GEMM against 0.36 + 0.10 ms for the bf16 pieces, and closer:
        // 1.3-1.9e-7 against 2.1-2.6e-7 of the non-local part alone); the
        // non-local kernels run in the gaps between contractions, so their
        // time counts.  nl_form 1: bf16 pieces, 3 of them and 6 products for
        // both split precisions (the 2-piece form left 8.5e-6)
        if (!nl32)
            LELPH_TRY((nl_gemm_split<3, 6>(c, A, B, rows, ncol, w.npw, conjB, out, st)));
        else
        {
            // FP32: beta from the FP64 tables, stored in FP32; psi rotated
            // into FP32 by k_nl_geom
            DBuf& b32 = side == 0 ? c->beta32_K : c->beta32_Kp;
            DBuf& w32 = side == 0 ? c->wr32_K : c->wr32_Kp;
            LELPH_TRY(b32.ensure(sizeof(c32) * (size_t)rows * npm));
            hipLaunchKernelGGL(k_nl_beta<float>,
                               dim3((unsigned)cdiv(w.npw, 256), natom * c->pmax), dim3(256), 0, st,
                               side == 0 ? c->nlK.as<double>() : c->nlKp.as<double>(),
                               side == 0 ? c->Y_K.as<double>() : c->Y_Kp.as<double>(), w.Fk,
                               w.npw, pj, c->d_plist, c->d_atype, c->d_apos, c->d_Fsign,
                               side == 0 ? -1 : +1, b32.as<c32>());
            HIP_TRY(hipGetLastError());
            const int64_t nchunks = cdiv(w.npw, CT_KC);
            const int64_t tiles = cdiv(rows, GZ_MB) * cdiv(ncol, GZ_NB);
            int64_t cps;
            int nsplit;
            pick_splits(tiles, nchunks, (int64_t)rows * ncol, &cps, &nsplit);
            const size_t need = sizeof(z64) * (size_t)nsplit * rows * ncol;
            if (need > c->bbpart.n && st != c->stream) HIP_TRY(hipStreamSynchronize(st));
            LELPH_TRY(c->bbpart.ensure(need));
            hipLaunchKernelGGL(k_zgemm_nt<float>, dim3((unsigned)tiles, (unsigned)nsplit),
                               dim3(256), 0, st, b32.as<c32>(), w32.as<c32>(), rows, (int)ncol,
                               w.npw, conjB, cps, c->bbpart.as<z64>());
            hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv((int64_t)rows * ncol, 256)),
                               dim3(256), 0, st, c->bbpart.as<z64>(), nsplit,
                               (int64_t)rows * ncol, 1.0, out);
            HIP_TRY(hipGetLastError());
        }
    }
    for (int side = 0; side < 2 && prec == LELPH_FP64; ++side)
    {
        if (skip_side(side)) continue;
        const DevWfc& w = side == 0 ? wK : wKp;
        const z64* A = side == 0 ? c->beta_K.as<z64>() : c->beta_Kp.as<z64>();
        const z64* B = side == 0 ? c->wr_K.as<z64>() : c->wr_Kp.as<z64>();
        z64* out = side == 0 ? c->bb_K.as<z64>() : c->bb_Kp.as<z64>();
        const int conjB = side == 0 ? sK.conj_w : sKp.conj_w;
        const int64_t nchunks = cdiv(w.npw, CT_KC);
        const int64_t tiles = cdiv(rows, GZ_MB) * cdiv(ncol, GZ_NB);
        int64_t cps;
        int nsplit;
        pick_splits(tiles, nchunks, (int64_t)rows * ncol, &cps, &nsplit);
        // the split count follows the side's npw: a larger partial buffer is
        // allocated only once the previous pair's non-local part is done
        const size_t need = sizeof(z64) * (size_t)nsplit * rows * ncol;
        if (need > c->bbpart.n && st != c->stream) HIP_TRY(hipStreamSynchronize(st));
        LELPH_TRY(c->bbpart.ensure(need));
        hipLaunchKernelGGL(k_zgemm_nt<double>, dim3((unsigned)tiles, (unsigned)nsplit), dim3(256),
                           0, st, A, B, rows, (int)ncol, w.npw, conjB, cps,
                           c->bbpart.as<z64>());
        hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv((int64_t)rows * ncol, 256)),
                           dim3(256), 0, st, c->bbpart.as<z64>(), nsplit,
                           (int64_t)rows * ncol, 1.0, out);
        HIP_TRY(hipGetLastError());
    }
    if (mirror)
        hipLaunchKernelGGL(k_nl_mirror, dim3((unsigned)cdiv((int64_t)rows * ncol, 256)),
                           dim3(256), 0, st, c->bb_Kprev.as<z64>(), (int64_t)rows, ncol,
                           c->bb_Kp.as<z64>());
    if (close_k)
        hipLaunchKernelGGL(k_nl_mirror, dim3((unsigned)cdiv((int64_t)rows * ncol, 256)),
                           dim3(256), 0, st, c->bb_Khead.as<z64>(), (int64_t)rows, ncol,
                           c->bb_K.as<z64>());
    HIP_TRY(hipGetLastError());
    const int64_t ne = (int64_t)natom * 3 * es;
    z64* eb = ebuf_out ? ebuf_out : c->ebuf.as<z64>();
    hipLaunchKernelGGL(k_nl_kkp, dim3((unsigned)cdiv(ne, 256)), dim3(256), 0, st,
                       c->bb_K.as<z64>(), c->bb_Kp.as<z64>(), pj, c->d_plist, c->d_atype,
                       c->d_socfac, c->d_fco, c->has_fco ? c->d_fco_off : nullptr, L.nspin,
                       L.nbnds, L.nspinor, eb);
    HIP_TRY(hipGetLastError());
    if (c->nl_chain)
    {
        // this pair's K-side projections stay for the next pair of the chain
        // (the kernels above are ordered before any later write on st)
        LELPH_TRY(c->bb_Kprev.ensure(c->bb_K.n));
        std::swap(c->bb_K.p, c->bb_Kprev.p);
        std::swap(c->bb_K.n, c->bb_Kprev.n);
        c->nl_prev_valid = true;
        c->nl_prev_ik = pr.ikq;
        c->nl_prev_sym = pr.kqsym;
        // a chain's head keeps its K' projections for the cycle's close
        if (!mirror)
        {
            LELPH_TRY(c->bb_Khead.ensure(c->bb_Kp.n));
            std::swap(c->bb_Kp.p, c->bb_Khead.p);
            std::swap(c->bb_Kp.n, c->bb_Khead.n);
            c->nl_head_valid = true;
            c->nl_head_ik = pr.ik;
            c->nl_head_sym = pr.ksym;
        }
    }
    if (ebuf_out) return LELPH_OK;
    if (mode_after) HIP_TRY(hipStreamWaitEvent(st, mode_after, 0));
    return nl_mode(c, eb, g, st);
}

// local part with split-product bf16 MFMA (split.hpp)
template <int NP, int NPROD>
static int run_local_split_t(lelph_ctx* c, const double* qpt, const lelph_pair& pr, z64* g,
                             hipEvent_t* ev)
{
    const lelph_lattice& L = c->lat;
    if (!c->dv_set) return LELPH_ERR_STATE;
    if (L.nmag == 2 && (L.nspin != 2 || L.nspinor != 1)) return LELPH_ERR_ARG;
    // psi_k+q in k-side form; its umklapp phase goes into the presplit
    c->box_defer = true;
    const int bst = build_boxes<float>(c, qpt, pr);
    c->box_defer = false;
    LELPH_TRY(bst);
    const int nm = L.nmodes, nb = L.nbnds, nspin = L.nspin;
    const int nmag_eff = (L.nspinor == 1 && L.nmag == 4) ? 1 : L.nmag;
    const int64_t nchunks = cdiv(c->nf, SP_KC);
    const int nbp = (int)(cdiv(nb, SP_PAD) * SP_PAD);
    // psi_k+q -> bf16 pieces, chunk-major
    const int64_t nBp = presplit_elems(nspin, L.nspinor, nchunks, NP, nbp);
    LELPH_TRY(c->presplit.ensure(sizeof(__bf16) * (size_t)nBp));
    hipLaunchKernelGGL((k_presplit_B<NP>),
                       dim3((unsigned)nchunks, (unsigned)(nbp / PS_ROWS),
                            (unsigned)(nspin * L.nspinor)),
                       dim3(256), 0, c->stream, static_cast<const c32*>(c->kq_view), nb, nbp,
                       L.nspinor, c->nf, nchunks, c->presplit.as<__bf16>(),
                       (int)L.fft_dims[0], (int)L.fft_dims[1], (int)L.fft_dims[2], c->kq_g0[0],
                       c->kq_g0[1], c->kq_g0[2]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[1], c->stream));
    SplitArgs sa{};
    sa.nm = nm;
    sa.nb = nb;
    sa.nbp = nbp;
    sa.nspin = nspin;
    sa.nspinor = L.nspinor;
    sa.nmag = nmag_eff;
    const bool fast = L.nspinor == 1 && nmag_eff != 4 && c->nf % SP_KC == 0;
    // producer/consumer kernel: bf16x3 on the aligned path with enough bands
    // to fill its 128-column tiles
    const bool pc = fast && NP == 2 && nb >= 64;
    sa.nmt = pc ? (int)cdiv(nm, PC_MG) : nm;
    sa.ntn = pc ? (int)cdiv(nb, PC_TR / PC_MG) : nbp / SP_TR;
    sa.ntm = pc ? nbp / PC_TC : (int)cdiv(nb, SP_TC);
    sa.N = c->nf;
    sa.nchunks = nchunks;
    // nspin * nsplit must be a multiple of 8 (XCD grouping); ~768+ blocks
    const int per_group = sa.nmt * sa.ntn * sa.ntm;
    int64_t ns = cdiv(pc ? PC_BLOCKS : 768, (int64_t)per_group * nspin);
    ns = std::max<int64_t>(8, cdiv(ns, 8) * 8);
    ns = std::min<int64_t>(ns, std::max<int64_t>(8, (nchunks / 8) * 8));
    sa.nsplit = (int)ns;
    sa.chunks_per_split = cdiv(nchunks, ns);
    const int64_t gsz = (int64_t)nm * nspin * nb * nb;
    LELPH_TRY(c->part.ensure(sizeof(z64) * (size_t)(ns * gsz)));
    const size_t lds = pc     ? pc_lds_bytes<NP>()
                       : fast ? split_lds_bytes<NP, true>()
                              : split_lds_bytes<NP, false>();
    const int64_t nblk = (int64_t)nspin * ns * per_group;
    if (pc)
    {
        HIP_TRY(hipFuncSetAttribute((const void*)k_contract_split_pc<NP>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_contract_split_pc<NP>), dim3((unsigned)nblk), dim3(SP_THREADS), lds,
                           c->stream, c->box_k.as<c32>(), c->presplit.as<__bf16>(),
                           c->dv32.as<c32>(), c->part.as<z64>(), sa);
    }
    else if (fast)
    {
        HIP_TRY(hipFuncSetAttribute((const void*)k_contract_split<NP, NPROD, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_contract_split<NP, NPROD, true>), dim3((unsigned)nblk),
                           dim3(SP_THREADS), lds, c->stream, c->box_k.as<c32>(),
                           c->presplit.as<__bf16>(), c->dv32.as<c32>(), c->part.as<z64>(), sa);
    }
    else
    {
        HIP_TRY(hipFuncSetAttribute((const void*)k_contract_split<NP, NPROD, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((k_contract_split<NP, NPROD, false>), dim3((unsigned)nblk),
                           dim3(SP_THREADS), lds, c->stream, c->box_k.as<c32>(),
                           c->presplit.as<__bf16>(), c->dv32.as<c32>(), c->part.as<z64>(), sa);
    }
    HIP_TRY(hipGetLastError());
    const int nsplit = (int)ns;
    HIP_TRY(hipEventRecord(ev[2], c->stream));
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv(gsz, 256)), dim3(256), 0, c->stream,
                       c->part.as<z64>(), nsplit, gsz, 1.0 / (double)c->nf, g);
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

// The precision a context's contractions actually run in.  bf16x3 (two
// bf16 pieces, 16 mantissa bits per operand) holds |dg|/|g|max <= 1e-6 only
// with the producer/consumer kernel's 512-point FP32 chains on shapes where
// the error is measured (one spinor, N a multiple of 32, nb >= 64); every
// other shape runs the 3-piece bf16x6 form instead (2^-25 per product).
static int effective_precision(const lelph_ctx* c)
{
    if (c->precision != LELPH_BF16X3) return c->precision;
    const bool pc = c->lat.nspinor == 1 && c->nf % SP_KC == 0 && c->lat.nbnds >= 64;
    return pc ? LELPH_BF16X3 : LELPH_BF16X6;
}

static int run_local_split(lelph_ctx* c, const double* qpt, const lelph_pair& pr, z64* g,
                           hipEvent_t* ev)
{
    if (effective_precision(c) == LELPH_BF16X3) return run_local_split_t<2, 3>(c, qpt, pr, g, ev);
    return run_local_split_t<3, 6>(c, qpt, pr, g, ev);
}

static int run_local_any(lelph_ctx* c, const double* qpt, const lelph_pair& pr, z64* g,
                         hipEvent_t* ev)
{
    switch (c->precision)
    {
        case LELPH_FP64: return run_local<double>(c, qpt, pr, g, ev);
        case LELPH_FP32:
            LELPH_TRY(ensure_dv32(c));
            return run_local<float>(c, qpt, pr, g, ev);
        case LELPH_BF16X3:
        case LELPH_BF16X6:
            LELPH_TRY(ensure_dv32(c));
            return run_local_split(c, qpt, pr, g, ev);
        default: return LELPH_ERR_ARG;
    }
}

// reads and clears the device error word (bit 0: G outside the box or table
// range, bit 1: symmetry incompatible with the FFT grid, bit 2: rotated
// index out of range)
static int read_err(lelph_ctx* c, int* e)
{
    *e = 0;
    HIP_TRY(hipMemcpyAsync(e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (*e) HIP_TRY(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    return LELPH_OK;
}

static int check_err(lelph_ctx* c)
{
    int e = 0;
    LELPH_TRY(read_err(c, &e));
    if (e & 2) return LELPH_ERR_SYM;
    if (e) return LELPH_ERR_RANGE;
    return LELPH_OK;
}

/* ------------------------------------------------------------------------ */
/* C-ABI                                                                     */
/* ------------------------------------------------------------------------ */

extern "C" {

const char* lelph_strerror(int s)
{
    switch (s)
    {
        case LELPH_OK: return "ok";
        case LELPH_ERR_ARG: return "bad argument";
        case LELPH_ERR_ALLOC: return "allocation failed";
        case LELPH_ERR_HIP: return "HIP runtime error";
        case LELPH_ERR_RANGE: return "G vector outside the FFT box or table range";
        case LELPH_ERR_UNSUPPORTED: return "unsupported configuration";
        case LELPH_ERR_STATE: return "call order violated (missing wfc or dV)";
        case LELPH_ERR_NOKQ:
            return "K+Q point cannot be found in kgrid. The k-grid and q-grid are not "
                   "commensurate";
        case LELPH_ERR_SYM: return "Wrong symmetric matrices";
        case LELPH_ERR_SCRATCH:
            return "size exceeds the static scratch of the no-allocation entry point "
                   "(nmodes * nspin * nbnds^2 <= LELPH_NOALLOC_MAX_ELEMS)";
        default: return "unknown status";
    }
}

int lelph_get_device_count(int* count)
{
    if (!count) return LELPH_ERR_ARG;
    HIP_TRY(hipGetDeviceCount(count));
    return LELPH_OK;
}

static void free_ctx(lelph_ctx* c)
{
    if (!c) return;
    for (auto& w : c->wf)
    {
        if (w.wfc) (void)hipFree(w.wfc);
        if (w.gvec) (void)hipFree(w.gvec);
        if (w.Fk) (void)hipFree(w.Fk);
    }
    for (int i = 0; i < 3; ++i) fft_axis_free(&c->ax[i]);
    void* ptrs[] = {c->d_atype, c->d_apos, c->d_Fsign, c->d_socfac, c->d_Zval, c->d_plist,
                    c->d_fco,   c->d_fco_off, c->d_err, c->d_gco,  c->d_vlocg, c->d_vploc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    DBuf* bufs[] = {&c->dv64, &c->dv32, &c->eig_d, &c->vloc_tmp, &c->rot_tmp, &c->idx_k,
                    &c->idx_kq, &c->ph_k, &c->ph_kq, &c->box_k, &c->box_kq, &c->box_x, &c->part, &c->g,
                    &c->gbatch, &c->presplit, &c->nlK, &c->nlKp, &c->Y_K, &c->Y_Kp, &c->wr_K, &c->wr_Kp,
                    &c->beta_K, &c->beta_Kp, &c->bb_K, &c->bb_Kp, &c->bb_Kprev, &c->bb_Khead, &c->bbpart, &c->ebuf,
                    &c->pl_cnt[0], &c->pl_cnt[1], &c->pl_off[0], &c->pl_off[1], &c->pl_cur[0],
                    &c->pl_cur[1], &c->pl_perm[0], &c->pl_perm[1], &c->dm_map, &c->dm_ph,
                    &c->dm_pos, &c->dm_wRS, &c->dm_wS2, &c->dm_part, &c->dm_out, &c->qr_dl,
                    &c->qr_dr, &c->qr_g, &c->qr_t, &c->qr_out, &c->ip_dvq, &c->ip_dvR,
                    &c->ip_last, &c->ip_tmp, &c->ip_out, &c->ip_ws, &c->ip_off, &c->ip_W,
                    &c->ip_P, &c->ip_qs, &c->ip_rot, &c->ip_eig, &c->lr_box, &c->lr_box2,
                    &c->lr_q, &c->lr_zv, &c->lr_zb, &c->lr_qp, &c->rot_idx, &c->rot_sr,
                    &c->dy_dyn, &c->dy_fc, &c->dy_X1, &c->dy_X2, &c->dy_P1, &c->dy_P2,
                    &c->dy_asr, &c->dy_ws, &c->dy_off, &c->dy_qs, &c->dy_qs2, &c->dy_out,
                    &c->dy_freq, &c->dy_eig, &c->dy_zbd, &c->dy_qpd, &c->dy_massd, &c->dy_P,
                    &c->dy_rmass, &c->dy_lr, &c->dy_fclr};
    for (DBuf* b : bufs) b->release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->evs)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->ip_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->out_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->out_stream) (void)hipStreamDestroy(c->out_stream);
    if (c->nl_stream) (void)hipStreamDestroy(c->nl_stream);
    if (c->out_ring) (void)hipHostFree(c->out_ring);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void lelph_destroy(lelph_ctx* c) { free_ctx(c); }

extern "C++" {
template <typename X>
static int upload(X** d, const X* h, size_t n)
{
    HIP_TRY(hipMalloc(d, sizeof(X) * (n > 0 ? n : 1)));
    if (n) HIP_TRY(hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice));
    return LELPH_OK;
}
}

static int scratch_ptr(double** out);

static int create_impl(const lelph_lattice* L, const lelph_pseudo* P, int precision,
                       lelph_ctx* c)
{
    if (L->natom < 1 || L->nbnds < 1 || L->nmodes < 1 || L->nsym < 1) return LELPH_ERR_ARG;
    if (L->nspin < 1 || L->nspin > 2 || L->nspinor < 1 || L->nspinor > 2) return LELPH_ERR_ARG;
    if (L->nmag != 1 && L->nmag != 2 && L->nmag != 4) return LELPH_ERR_ARG;
    if (L->natom > DVLOC_MAXATOM || P->ntype > DVLOC_MAXTYPE) return LELPH_ERR_UNSUPPORTED;
    if (precision < LELPH_FP64 || precision > LELPH_BF16X6) return LELPH_ERR_ARG;
    c->precision = precision;
    HIP_TRY(hipGetDevice(&c->device));
    {
        double* sp = nullptr;
        LELPH_TRY(scratch_ptr(&sp));
    }
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_TRY(hipEventCreate(&e));
    c->lat = *L;
    c->atomic_pos.assign(L->atomic_pos, L->atomic_pos + 3 * L->natom);
    c->atom_type.assign(L->atom_type, L->atom_type + L->natom);
    c->kpt_ibz.assign(L->kpt_iredBZ, L->kpt_iredBZ + 3 * L->nkpts_iBZ);
    if (L->nkpts_BZ > 0 && L->kpt_fullBZ_crys && L->kmap)
    {
        c->kpt_bz.assign(L->kpt_fullBZ_crys, L->kpt_fullBZ_crys + 3 * L->nkpts_BZ);
        c->kmap.assign(L->kmap, L->kmap + 2 * L->nkpts_BZ);
    }
    c->syms.assign(L->syms, L->syms + L->nsym);
    c->lat.atomic_pos = c->atomic_pos.data();
    c->lat.atom_type = c->atom_type.data();
    c->lat.kpt_iredBZ = c->kpt_ibz.data();
    c->lat.kpt_fullBZ_crys = c->kpt_bz.empty() ? nullptr : c->kpt_bz.data();
    c->lat.kmap = c->kmap.empty() ? nullptr : c->kmap.data();
    c->lat.syms = c->syms.data();
    for (int i = 0; i < L->natom; ++i)
        if (L->atom_type[i] < 0 || L->atom_type[i] >= P->ntype) return LELPH_ERR_ARG;

    c->ntype = P->ntype;
    c->lmax = P->lmax;
    c->nlj = P->nltimesj;
    c->PP.assign(P->PP_table, P->PP_table + 3 * P->nltimesj * P->ntype);
    c->Fsign.assign(P->Fsign, P->Fsign + P->nltimesj * P->ntype);
    c->qmax_abs = P->qmax_abs;
    for (int t = 0; t < P->ntype; ++t)
    {
        const lelph_local_pseudo& lp = P->loc_pseudo[t];
        c->locV.emplace_back(lp.Vloc_atomic, lp.Vloc_atomic + lp.ngrid);
        c->locR.emplace_back(lp.r_grid, lp.r_grid + lp.ngrid);
        c->locRab.emplace_back(lp.rab_grid, lp.rab_grid + lp.ngrid);
        c->Zval.push_back(lp.Zval);
    }
    // projector lists and j-average factors (Vbare_nloc.c:318-395, 472-491)
    c->socfac.assign(P->nltimesj * P->ntype, 1.0);
    c->fco_off.assign(P->nltimesj * P->ntype, -1);
    int pmax = 0;
    for (int a = 0; a < L->natom; ++a)
    {
        const int t = L->atom_type[a];
        int np = 0;
        for (int64_t li = 0; li < P->nltimesj; ++li)
        {
            const int l = (int)rint(P->PP_table[P->ntype * 3 * li + t * 3] - 1);
            if (l < 0) continue;
            if (l > P->lmax) return LELPH_ERR_ARG;
            np += 2 * l + 1;
        }
        pmax = std::max(pmax, np);
    }
    c->pmax = std::max(pmax, 1);
    c->plist.assign((size_t)L->natom * c->pmax * 3, -1);
    for (int a = 0; a < L->natom; ++a)
    {
        const int t = L->atom_type[a];
        int p = 0;
        for (int64_t li = 0; li < P->nltimesj; ++li)
        {
            const int l = (int)rint(P->PP_table[P->ntype * 3 * li + t * 3] - 1);
            if (l < 0) continue;
            for (int m = 0; m <= 2 * l; ++m, ++p)
            {
                int* e = &c->plist[3 * ((size_t)a * c->pmax + p)];
                e[0] = (int)li;
                e[1] = l;
                e[2] = l * l + m;
            }
        }
    }
    for (int64_t li = 0; li < P->nltimesj; ++li)
        for (int t = 0; t < P->ntype; ++t)
        {
            const int l = (int)rint(P->PP_table[P->ntype * 3 * li + t * 3] - 1);
            const int j = (int)rint(P->PP_table[P->ntype * 3 * li + t * 3 + 1]);
            if (l >= 0 && j != 0 && !L->is_soc_present)
                c->socfac[li * P->ntype + t] = (j + 1.0) / (4.0 * l + 2.0);
            if (P->fCoeff && l >= 0 && P->fCoeff[P->ntype * li + t])
            {
                const int n = (2 * l + 1) * (2 * l + 1) * L->nspinor * L->nspinor;
                c->fco_off[li * P->ntype + t] = (int64_t)c->fco.size();
                const double* f = P->fCoeff[P->ntype * li + t];
                for (int i = 0; i < n; ++i) c->fco.push_back(z64{f[2 * i], f[2 * i + 1]});
                c->has_fco = true;
            }
        }
    LELPH_TRY(upload(&c->d_atype, c->atom_type.data(), c->atom_type.size()));
    LELPH_TRY(upload(&c->d_apos, c->atomic_pos.data(), c->atomic_pos.size()));
    LELPH_TRY(upload(&c->d_Fsign, c->Fsign.data(), c->Fsign.size()));
    LELPH_TRY(upload(&c->d_socfac, c->socfac.data(), c->socfac.size()));
    LELPH_TRY(upload(&c->d_Zval, c->Zval.data(), c->Zval.size()));
    LELPH_TRY(upload(&c->d_plist, c->plist.data(), c->plist.size()));
    LELPH_TRY(upload(&c->d_fco, c->fco.data(), c->fco.size()));
    LELPH_TRY(upload(&c->d_fco_off, c->fco_off.data(), c->fco_off.size()));
    HIP_TRY(hipMalloc(&c->d_err, sizeof(int)));
    HIP_TRY(hipMemset(c->d_err, 0, sizeof(int)));

    for (int i = 0; i < 3; ++i)
    {
        if (L->fft_dims[i] < 1 || L->fft_dims[i] > 4096) return LELPH_ERR_ARG;
        LELPH_TRY(fft_axis_init(&c->ax[i], (int)L->fft_dims[i]));
    }
    c->nf = L->fft_dims[0] * L->fft_dims[1] * L->fft_dims[2];
    c->wf.resize(L->nkpts_iBZ);
    return LELPH_OK;
}

int lelph_create(const lelph_lattice* L, const lelph_pseudo* P, int precision, lelph_ctx** out)
{
    if (!L || !P || !out) return LELPH_ERR_ARG;
    *out = nullptr;
    lelph_ctx* c = new (std::nothrow) lelph_ctx();
    if (!c) return LELPH_ERR_ALLOC;
    int s = create_impl(L, P, precision, c);
    if (s != LELPH_OK)
    {
        free_ctx(c);
        return s;
    }
    *out = c;
    return LELPH_OK;
}

int lelph_set_precision(lelph_ctx* c, int p)
{
    if (!c || p < LELPH_FP64 || p > LELPH_BF16X6) return LELPH_ERR_ARG;
    c->precision = p;
    return LELPH_OK;
}

int lelph_set_tuning(lelph_ctx* c, const lelph_tuning* t)
{
    if (!c || !t) return LELPH_ERR_ARG;
    if (t->split_target < 0 || t->split_target > (1 << 20)) return LELPH_ERR_ARG;
    if (t->fft_lines != 0 && t->fft_lines != 8 && t->fft_lines != 16 && t->fft_lines != 32)
        return LELPH_ERR_ARG;
    if (t->fp32_form != 0 && t->fp32_form != 1) return LELPH_ERR_ARG;
    if (t->ctile < 0 || t->ctile > 4) return LELPH_ERR_ARG;
    if (t->nl_form != 0 && t->nl_form != 1) return LELPH_ERR_ARG;
    c->tune = *t;
    return LELPH_OK;
}

int lelph_get_tuning(lelph_ctx* c, lelph_tuning* t)
{
    if (!c || !t) return LELPH_ERR_ARG;
    *t = c->tune;
    return LELPH_OK;
}

int lelph_get_precision(lelph_ctx* c, int* requested, int* effective)
{
    if (!c) return LELPH_ERR_ARG;
    if (requested) *requested = c->precision;
    if (effective) *effective = effective_precision(c);
    return LELPH_OK;
}

int lelph_upload_wfc(lelph_ctx* c, int64_t ik, const double* wfc, const double* gvec,
                     const double* Fk, int64_t npw)
{
    if (!c || ik < 0 || ik >= (int64_t)c->wf.size() || !wfc || !gvec || npw < 1)
        return LELPH_ERR_ARG;
    DevWfc& w = c->wf[ik];
    if (w.wfc) (void)hipFree(w.wfc);
    if (w.gvec) (void)hipFree(w.gvec);
    if (w.Fk) (void)hipFree(w.Fk);
    w = DevWfc{};
    const int64_t nset = (int64_t)c->lat.nspin * c->lat.nbnds * c->lat.nspinor;
    w.npw = npw;
    HIP_TRY(hipMalloc(&w.wfc, sizeof(z64) * nset * npw));
    HIP_TRY(hipMemcpy(w.wfc, wfc, sizeof(z64) * nset * npw, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&w.gvec, sizeof(double) * 3 * npw));
    HIP_TRY(hipMemcpy(w.gvec, gvec, sizeof(double) * 3 * npw, hipMemcpyHostToDevice));
    if (Fk)
    {
        const size_t n = (size_t)c->nlj * c->ntype * npw;
        HIP_TRY(hipMalloc(&w.Fk, sizeof(double) * n));
        HIP_TRY(hipMemcpy(w.Fk, Fk, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    c->npw_max = std::max(c->npw_max, npw);
    return LELPH_OK;
}

// prepare_spline (src/common/spline.c:26-139): not-a-knot derivatives
static void prepare_spline(int64_t n, const double* x, const double* y, double* dy)
{
    std::vector<double> A(n), B(n), Cc(n), D(n);
    double h1 = x[1] - x[0], s1 = (y[1] - y[0]) / h1;
    double h2 = x[2] - x[1], s2 = (y[2] - y[1]) / h2;
    A[0] = h2 * h2;
    B[0] = h2 * h2 - h1 * h1;
    Cc[0] = -h1 * h1;
    D[0] = 2 * (s1 * h2 * h2 - s2 * h1 * h1);
    h1 = x[n - 2] - x[n - 3];
    s1 = (y[n - 2] - y[n - 3]) / h1;
    h2 = x[n - 1] - x[n - 2];
    s2 = (y[n - 1] - y[n - 2]) / h2;
    A[n - 1] = h2 * h2;
    B[n - 1] = h2 * h2 - h1 * h1;
    Cc[n - 1] = -h1 * h1;
    D[n - 1] = 2 * (s1 * h2 * h2 - s2 * h1 * h1);
    for (int64_t i = 1; i < n - 1; ++i)
    {
        const double hi = x[i + 1] - x[i], hm = x[i] - x[i - 1];
        const double si = (y[i + 1] - y[i]) / hi, sm = (y[i] - y[i - 1]) / hm;
        A[i] = hi;
        B[i] = 2.0 * (hi + hm);
        Cc[i] = hm;
        D[i] = 3 * (sm * hi + si * hm);
    }
    double al = -Cc[0] / Cc[1];
    A[0] += al * A[1];
    B[0] += al * B[1];
    Cc[0] += al * Cc[1];
    D[0] += al * D[1];
    Cc[0] = B[0];
    B[0] = A[0];
    A[0] = 0;
    al = -A[n - 1] / A[n - 2];
    A[n - 1] += al * A[n - 2];
    B[n - 1] += al * B[n - 2];
    Cc[n - 1] += al * Cc[n - 2];
    D[n - 1] += al * D[n - 2];
    A[n - 1] = B[n - 1];
    B[n - 1] = Cc[n - 1];
    Cc[n - 1] = 0;
    for (int64_t i = 1; i < n; ++i)
    {
        const double w = A[i] / B[i - 1];
        B[i] -= w * Cc[i - 1];
        D[i] -= w * D[i - 1];
    }
    dy[n - 1] = D[n - 1] / B[n - 1];
    for (int64_t i = n - 2; i >= 0; --i) dy[i] = (D[i] - Cc[i] * dy[i + 1]) / B[i];
}

// Simpson weights (src/common/numerical_func.c:155-192), 1/3 included
static std::vector<double> simpson_weights(int64_t n)
{
    std::vector<double> w(n, 0.0);
    if (n < 3) return w;
    w[0] = 1;
    for (int64_t i = 1; i < n - 1; i += 2) w[i] += 4.0;
    for (int64_t i = 2; i < n - 1; i += 2) w[i] += 2.0;
    if (n % 2 != 0)
        w[n - 1] += 1.0;
    else
    {
        w[n - 3] -= 0.25;
        w[n - 2] += 1.0;
        w[n - 1] += 1.25;
    }
    for (auto& x : w) x /= 3.0;
    return w;
}

int lelph_create_vlocg_table(lelph_ctx* c, int64_t* npts_out, double* g_co, double* vlocg,
                             double* vploc)
{
    if (!c) return LELPH_ERR_ARG;
    const double* b = c->lat.blat_vec;
    double bi[3];
    bi[0] = sqrt(b[0] * b[0] + b[3] * b[3] + b[6] * b[6]);
    bi[1] = sqrt(b[1] * b[1] + b[4] * b[4] + b[7] * b[7]);
    bi[2] = sqrt(b[2] * b[2] + b[5] * b[5] + b[8] * b[8]);
    double bmin = bi[0], bmax = bi[0], Nmax = (double)c->lat.fft_dims[0];
    for (int i = 0; i < 3; ++i)
    {
        bmin = std::min(bmin, bi[i]);
        bmax = std::max(bmax, bi[i]);
        Nmax = std::max(Nmax, (double)c->lat.fft_dims[i]);
    }
    if (bmin > 0.01) bmin = 0.01;
    const double gmax = (1.7 * (Nmax + 4.0) + 3 * fabs(c->qmax_abs)) * bmax;
    const int64_t npts = (int64_t)ceil(gmax / bmin);
    if (npts < 4) return LELPH_ERR_ARG;
    const double diff = gmax / (npts - 1.0);
    c->h_gco.resize(npts);
    for (int64_t i = 0; i < npts; ++i) c->h_gco[i] = 0.0 + i * diff;
    c->npts = npts;
    c->dg = diff;
    c->h_vlocg.assign((size_t)npts * c->ntype, 0.0);
    c->h_vploc.assign((size_t)npts * c->ntype, 0.0);
    if (c->d_gco) (void)hipFree(c->d_gco);
    if (c->d_vlocg) (void)hipFree(c->d_vlocg);
    if (c->d_vploc) (void)hipFree(c->d_vploc);
    c->d_gco = c->d_vlocg = c->d_vploc = nullptr;
    LELPH_TRY(upload(&c->d_gco, c->h_gco.data(), (size_t)npts));
    HIP_TRY(hipMalloc(&c->d_vlocg, sizeof(double) * npts * c->ntype));
    for (int t = 0; t < c->ntype; ++t)
    {
        const int64_t ng = (int64_t)c->locR[t].size();
        std::vector<double> w = simpson_weights(ng);
        double *dV, *dr, *drab, *dw;
        LELPH_TRY(upload(&dV, c->locV[t].data(), ng));
        LELPH_TRY(upload(&dr, c->locR[t].data(), ng));
        LELPH_TRY(upload(&drab, c->locRab[t].data(), ng));
        LELPH_TRY(upload(&dw, w.data(), ng));
        hipLaunchKernelGGL(k_vloc_table, dim3((unsigned)cdiv(npts, 4)), dim3(256), 0, c->stream,
                           c->d_gco, npts, dV, dr, drab, dw, ng, c->Zval[t], c->lat.dimension,
                           c->lat.volume, c->d_vlocg + (int64_t)t * npts);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
        (void)hipFree(dV);
        (void)hipFree(dr);
        (void)hipFree(drab);
        (void)hipFree(dw);
    }
    HIP_TRY(hipMemcpy(c->h_vlocg.data(), c->d_vlocg, sizeof(double) * npts * c->ntype,
                      hipMemcpyDeviceToHost));
    for (int t = 0; t < c->ntype; ++t)
        prepare_spline(npts, c->h_gco.data(), c->h_vlocg.data() + (size_t)t * npts,
                       c->h_vploc.data() + (size_t)t * npts);
    LELPH_TRY(upload(&c->d_vploc, c->h_vploc.data(), (size_t)npts * c->ntype));
    if (npts_out) *npts_out = npts;
    if (g_co) std::memcpy(g_co, c->h_gco.data(), sizeof(double) * npts);
    if (vlocg) std::memcpy(vlocg, c->h_vlocg.data(), sizeof(double) * npts * c->ntype);
    if (vploc) std::memcpy(vploc, c->h_vploc.data(), sizeof(double) * npts * c->ntype);
    return LELPH_OK;
}

static int set_eig(lelph_ctx* c, const double* eigVec)
{
    const int nm = c->lat.nmodes;
    c->eig.resize((size_t)nm * nm);
    std::memcpy(c->eig.data(), eigVec, sizeof(z64) * nm * nm);
    LELPH_TRY(c->eig_d.ensure(sizeof(z64) * nm * nm));
    HIP_TRY(hipMemcpyAsync(c->eig_d.p, c->eig.data(), sizeof(z64) * nm * nm,
                           hipMemcpyHostToDevice, c->stream));
    return LELPH_OK;
}

// dVlocq on the device; result in c->vloc_tmp (nm, nf)
static int dvlocq_dev(lelph_ctx* c, const double* qpt)
{
    if (c->npts == 0) LELPH_TRY(lelph_create_vlocg_table(c, nullptr, nullptr, nullptr, nullptr));
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    LELPH_TRY(c->vloc_tmp.ensure(sizeof(z64) * nm * c->nf));
    DvlocArgs a{};
    std::memcpy(a.q, qpt, sizeof(a.q));
    std::memcpy(a.blat, L.blat_vec, sizeof(a.blat));
    a.latz = L.alat_vec[8];
    a.volume = L.volume;
    a.Nx = (int)L.fft_dims[0];
    a.Ny = (int)L.fft_dims[1];
    a.Nz = (int)L.fft_dims[2];
    a.ntype = c->ntype;
    a.natom = L.natom;
    a.nm = nm;
    a.cutoff = L.dimension;
    a.npts = c->npts;
    a.dg = c->dg;
    hipLaunchKernelGGL(k_dvloc_G, dim3((unsigned)cdiv(c->nf, 128)), dim3(128), 0, c->stream, a,
                       c->d_gco, c->d_vlocg, c->d_vploc, c->d_Zval, c->d_atype, c->d_apos,
                       c->eig_d.as<z64>(), c->vloc_tmp.as<z64>(), c->d_err);
    HIP_TRY(hipGetLastError());
    LELPH_TRY(fft3d_batch<double>(c->ax, c->vloc_tmp.as<z64>(), nm, +1, c->stream));
    return LELPH_OK;
}

int lelph_dvlocq(lelph_ctx* c, const double qpt[3], const double* eigVec, double* Vlocr)
{
    if (!c || !qpt || !eigVec) return LELPH_ERR_ARG;
    LELPH_TRY(set_eig(c, eigVec));
    LELPH_TRY(dvlocq_dev(c, qpt));
    LELPH_TRY(check_err(c));
    if (Vlocr)
        HIP_TRY(hipMemcpy(Vlocr, c->vloc_tmp.p, sizeof(z64) * c->lat.nmodes * c->nf,
                          hipMemcpyDeviceToHost));
    return LELPH_OK;
}

static int finish_dv(lelph_ctx* c, const double* qpt);

int lelph_set_dvscf(lelph_ctx* c, const double qpt[3], const double* eigVec, const double* dvscf,
                    const double* patterns, int kernel)
{
    if (!c || !qpt || !eigVec) return LELPH_ERR_ARG;
    if (kernel < LELPH_KERNEL_DFPT || kernel > LELPH_KERNEL_BARE_LOCAL) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    const int64_t ld = (int64_t)L.nmag * c->nf;
    const size_t bytes = sizeof(z64) * (size_t)nm * ld;
    c->dv_set = false;
    c->dv32_valid = false;
    LELPH_TRY(set_eig(c, eigVec));
    LELPH_TRY(c->dv64.ensure(bytes));
    const bool screened = (kernel == LELPH_KERNEL_DFPT || kernel == LELPH_KERNEL_DFPT_LOCAL);
    if (screened && dvscf)
    {
        HIP_TRY(hipMemcpyAsync(c->dv64.p, dvscf, bytes, hipMemcpyHostToDevice, c->stream));
        if (patterns)
        {
            // eig @ pats^H (read_dvscf.c:121-127), FP64 on the host (nm x nm)
            std::vector<z64> ep((size_t)nm * nm);
            const z64* pat = reinterpret_cast<const z64*>(patterns);
            for (int i = 0; i < nm; ++i)
                for (int j = 0; j < nm; ++j)
                {
                    z64 acc = {0, 0};
                    for (int k = 0; k < nm; ++k)
                        acc = cadd(acc, cmul(c->eig[(size_t)i * nm + k],
                                             cconj(pat[(size_t)j * nm + k])));
                    ep[(size_t)i * nm + j] = acc;
                }
            z64* dep;
            LELPH_TRY(upload(&dep, ep.data(), ep.size()));
            LELPH_TRY(c->rot_tmp.ensure(bytes));
            hipLaunchKernelGGL(k_mode_rotate, dim3((unsigned)cdiv((int64_t)nm * ld, 256)),
                               dim3(256), 0, c->stream, c->dv64.as<z64>(), dep, nm, ld,
                               c->rot_tmp.as<z64>());
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(c->dv64.p, c->rot_tmp.p, bytes, hipMemcpyDeviceToDevice,
                                   c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            (void)hipFree(dep);
        }
    }
    else
        HIP_TRY(hipMemsetAsync(c->dv64.p, 0, bytes, c->stream));
    return finish_dv(c, qpt);
}

// every kernel adds the bare local part (set_kernel.c: bare_loc = true)
static int finish_dv(lelph_ctx* c, const double* qpt)
{
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    LELPH_TRY(dvlocq_dev(c, qpt));
    hipLaunchKernelGGL(k_add_dvscf, dim3((unsigned)cdiv((int64_t)nm * c->nf, 256)), dim3(256), 0,
                       c->stream, c->dv64.as<z64>(), c->vloc_tmp.as<z64>(), nm, L.nmag, c->nf);
    HIP_TRY(hipGetLastError());
    LELPH_TRY(check_err(c));
    std::memcpy(c->q, qpt, sizeof(c->q));
    c->dv_set = true;
    return LELPH_OK;
}

int lelph_get_dvscf(lelph_ctx* c, double* dvscf)
{
    if (!c || !dvscf) return LELPH_ERR_ARG;
    if (!c->dv_set) return LELPH_ERR_STATE;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(dvscf, c->dv64.p, sizeof(z64) * c->lat.nmodes * c->lat.nmag * c->nf,
                      hipMemcpyDeviceToHost));
    return LELPH_OK;
}

static int ensure_events(lelph_ctx* c, int64_t npairs)
{
    const size_t need = (size_t)std::max<int64_t>(npairs, 1) * 5;
    while (c->evs.size() < need)
    {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        c->evs.push_back(e);
    }
    return LELPH_OK;
}

// local (+ non-local) of one pair into device g; events ev[0..4], ev[4]
// marking the end of both parts.  overlap: the non-local part runs on
// nl_stream, beside this pair's local part and the next pair's boxes.
static int gkq_device(lelph_ctx* c, const double* qpt, const lelph_pair& pr, int local,
                      int nonloc, z64* g, hipEvent_t* ev, bool overlap = false)
{
    HIP_TRY(hipEventRecord(ev[0], c->stream));
    if (local)
        LELPH_TRY(run_local_any(c, qpt, pr, g, ev));
    else
    {
        HIP_TRY(hipEventRecord(ev[1], c->stream));
        HIP_TRY(hipEventRecord(ev[2], c->stream));
    }
    HIP_TRY(hipEventRecord(ev[3], c->stream));
    if (nonloc && overlap)
    {
        LELPH_TRY(run_nonlocal(c, pr, g, c->nl_stream, ev[3]));
        HIP_TRY(hipEventRecord(ev[4], c->nl_stream));
        return LELPH_OK;
    }
    if (nonloc) LELPH_TRY(run_nonlocal(c, pr, g, c->stream, nullptr));
    HIP_TRY(hipEventRecord(ev[4], c->stream));
    return LELPH_OK;
}

// sums the per-pair phase times of the last call (after the stream drained)
// the batch copy-out's stream, events and pinned ring of LELPH_OUT_SLOTS slices
static int ensure_out_ring(lelph_ctx* c, size_t slice)
{
    if (!c->out_stream) HIP_TRY(hipStreamCreateWithFlags(&c->out_stream, hipStreamNonBlocking));
    for (auto& e : c->out_ev)
        if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const size_t need = slice * LELPH_OUT_SLOTS;
    if (c->out_ring_bytes < need)
    {
        if (c->out_ring) (void)hipHostFree(c->out_ring);
        c->out_ring = nullptr;
        c->out_ring_bytes = 0;
        if (hipHostMalloc(&c->out_ring, need) != hipSuccess)
        {
            c->out_ring = nullptr;
            return LELPH_ERR_ALLOC;
        }
        c->out_ring_bytes = need;
    }
    return LELPH_OK;
}

static int finish_timings(lelph_ctx* c, int64_t npairs)
{
    double t[5] = {0, 0, 0, 0, 0};
    for (int64_t i = 0; i < npairs; ++i)
    {
        hipEvent_t* e = c->evs.data() + 5 * i;
        float a = 0;
        for (int k = 0; k < 4; ++k)
        {
            HIP_TRY(hipEventElapsedTime(&a, e[k], e[k + 1]));
            t[k] += a;
        }
        HIP_TRY(hipEventElapsedTime(&a, e[0], e[4]));
        t[4] += a;
    }
    for (int k = 0; k < 5; ++k) c->timings[k] = t[k];
    c->timings[5] = (double)npairs;
    c->timings[6] = (double)c->box_sides;
    return LELPH_OK;
}

static int one_pair(lelph_ctx* c, const double* qpt, int ikq, int ik, int kqsym, int ksym,
                    int local, int nonloc, double* host_g, int accumulate)
{
    if (!c || !qpt || !host_g) return LELPH_ERR_ARG;
    const lelph_pair pr{ikq, ik, kqsym, ksym};
    if (ik < 0 || ikq < 0 || ik >= (int)c->wf.size() || ikq >= (int)c->wf.size())
        return LELPH_ERR_ARG;
    if (ksym < 0 || kqsym < 0 || ksym >= c->lat.nsym || kqsym >= c->lat.nsym) return LELPH_ERR_ARG;
    const int64_t gsz = (int64_t)c->lat.nmodes * c->lat.nspin * c->lat.nbnds * c->lat.nbnds;
    LELPH_TRY(c->g.ensure(sizeof(z64) * gsz));
    LELPH_TRY(ensure_events(c, 1));
    c->box_sides = 0;
    if (accumulate)
        HIP_TRY(hipMemcpyAsync(c->g.p, host_g, sizeof(z64) * gsz, hipMemcpyHostToDevice,
                               c->stream));
    LELPH_TRY(gkq_device(c, qpt, pr, local, nonloc, c->g.as<z64>(), c->evs.data()));
    HIP_TRY(hipMemcpyAsync(host_g, c->g.p, sizeof(z64) * gsz, hipMemcpyDeviceToHost, c->stream));
    LELPH_TRY(check_err(c));
    return finish_timings(c, 1);
}

int lelph_elph_local(lelph_ctx* c, const double qpt[3], int ikq, int ik, int kqsym, int ksym,
                     double* elph_kq)
{
    return one_pair(c, qpt, ikq, ik, kqsym, ksym, 1, 0, elph_kq, 0);
}

int lelph_add_elph_nonlocal(lelph_ctx* c, int ikq, int ik, int kqsym, int ksym,
                            const double* eigVec, double* elph_kq)
{
    if (!c || !eigVec) return LELPH_ERR_ARG;
    if (!c->dv_set) return LELPH_ERR_STATE;
    LELPH_TRY(set_eig(c, eigVec));
    return one_pair(c, c->q, ikq, ik, kqsym, ksym, 0, 1, elph_kq, 1);
}

int lelph_gkq(lelph_ctx* c, const double qpt[3], int ikq, int ik, int kqsym, int ksym, int non_loc,
              double* elph_kq)
{
    return one_pair(c, qpt, ikq, ik, kqsym, ksym, 1, non_loc ? 1 : 0, elph_kq, 0);
}

int lelph_gkq_batch(lelph_ctx* c, const double qpt[3], int64_t npairs, const lelph_pair* pairs,
                    int non_loc, double* g_out)
{
    if (!c || !qpt || npairs < 0 || (npairs > 0 && (!pairs || !g_out))) return LELPH_ERR_ARG;
    const int64_t gsz = (int64_t)c->lat.nmodes * c->lat.nspin * c->lat.nbnds * c->lat.nbnds;
    LELPH_TRY(c->gbatch.ensure(sizeof(z64) * std::max<int64_t>(1, npairs) * gsz));
    LELPH_TRY(ensure_events(c, npairs));
    for (int64_t i = 0; i < npairs; ++i)
    {
        const lelph_pair& p = pairs[i];
        if (p.ik < 0 || p.ikq < 0 || p.ik >= (int)c->wf.size() || p.ikq >= (int)c->wf.size())
            return LELPH_ERR_ARG;
        if (p.ksym < 0 || p.kqsym < 0 || p.ksym >= c->lat.nsym || p.kqsym >= c->lat.nsym)
            return LELPH_ERR_ARG;
    }
    // order the pairs along k -> k+q chains, so every second psi box is
    // reused (build_boxes); results keep the caller's order
    std::vector<int64_t> order;
    chain_order(c, pairs, npairs, order);
    // copy-out: pair n's g goes device -> pinned slot n % LELPH_OUT_SLOTS on
    // the copy stream once its last event has fired, and the host moves it
    // into g_out while the GPU works on later pairs
    const size_t slice = sizeof(z64) * (size_t)gsz;
    LELPH_TRY(ensure_out_ring(c, slice));
    if (!c->nl_stream) HIP_TRY(hipStreamCreateWithFlags(&c->nl_stream, hipStreamNonBlocking));
    z64* ring = static_cast<z64*>(c->out_ring);
    auto land = [&](int64_t m) -> int {
        const int slot = (int)(m % LELPH_OUT_SLOTS);
        HIP_TRY(hipEventSynchronize(c->out_ev[slot]));
        std::memcpy(g_out + 2 * order[m] * gsz, ring + slot * gsz, slice);
        return LELPH_OK;
    };
    c->box_reuse = true;
    c->box_prev_valid = false;
    c->box_head = nullptr;
    c->box_sides = 0;
    c->nl_chain = true;
    c->nl_prev_valid = false;
    c->nl_head_valid = false;
    int st = LELPH_OK;
    int64_t landed = 0;
    for (int64_t n = 0; n < npairs && st == LELPH_OK; ++n)
    {
        const int64_t i = order[n];
        st = gkq_device(c, qpt, pairs[i], 1, non_loc, c->gbatch.as<z64>() + i * gsz,
                        c->evs.data() + 5 * i, true);
        if (st != LELPH_OK) break;
        if (n >= LELPH_OUT_SLOTS)
        {
            st = land(landed++);
            if (st != LELPH_OK) break;
        }
        const int slot = (int)(n % LELPH_OUT_SLOTS);
        if (hipStreamWaitEvent(c->out_stream, c->evs[5 * i + 4], 0) != hipSuccess ||
            hipMemcpyAsync(ring + slot * gsz, c->gbatch.as<z64>() + i * gsz, slice,
                           hipMemcpyDeviceToHost, c->out_stream) != hipSuccess ||
            hipEventRecord(c->out_ev[slot], c->out_stream) != hipSuccess)
            st = LELPH_ERR_HIP;
    }
    c->box_reuse = false;
    c->box_prev_valid = false;
    c->box_head = nullptr;
    c->nl_chain = false;
    c->nl_prev_valid = false;
    c->nl_head_valid = false;
    if (st != LELPH_OK)
    {
        (void)hipStreamSynchronize(c->nl_stream);
        (void)hipStreamSynchronize(c->out_stream);
        (void)hipStreamSynchronize(c->stream);
        return st;
    }
    while (landed < npairs) LELPH_TRY(land(landed++));
    LELPH_TRY(check_err(c));
    return finish_timings(c, npairs);
}

/* ------------------------------------------------------------------------ */
/* Uneven batches over several contexts (cells): every pair's boxes, one     */
/* work-queue contraction launch per precision over the tiles of all pairs, */
/* the reductions and non-local parts, the copy-out, captured in one HIP     */
/* graph and replayed                                                        */
/* ------------------------------------------------------------------------ */
} // extern "C"

namespace {
struct MultiPair
{
    int cell = 0;
    int64_t index = 0; // position in its cell's output
    lelph_pair pr{};
    int prec = LELPH_FP64;
    int sep = 0;
    int nsplit = 1;
    int64_t cps = 1, gsz = 0, tiles = 0;
    size_t bb = 0;     // bytes of one psi box
    int shape = 1;     // contraction tile shape (ct_shape of its context)
    int64_t link = -1;  // the pair before it in its k -> k+q chain (index in pairs)
    int64_t close = -1; // the chain head whose psi_k box is its psi_k+q box
    int g0[3] = {0, 0, 0}; // umklapp vector of its k+q side
    void* kbox = nullptr;  // its psi_k box: box_k, or the linked pair's box_kq
    DBuf box_k, box_kq, part, ebuf;
};
} // namespace

struct lelph_multi
{
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<lelph_ctx*> ctx;
    std::vector<std::array<double, 3>> qpt;
    std::vector<int64_t> npairs;
    std::vector<MultiPair> pairs;
    std::vector<int64_t> border; // box build order: each cell's pairs along chains
    int non_loc = 1;
    std::vector<DBuf> gdev;    // per cell: (npairs, nmodes, nspin, nb, nb) complex128
    std::vector<void*> ghost;  // per cell: pinned copy of gdev
    std::vector<size_t> gbytes;
    // one work queue per (precision, tile shape): pair table, items, size,
    // grid; counter holds one int per queue
    struct Queue
    {
        int prec = LELPH_FP64, shape = 1;
        bool sep = false;
        DBuf qp, qi;
        int n = 0, grid = 0;
    };
    std::vector<Queue> queues;
    DBuf counter;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // the non-local parts run on a second stream: forked at the start,
    // their mode rotations wait for the reductions, joined before the copy
    hipStream_t nl = nullptr;
    hipEvent_t fork = nullptr, reduced = nullptr, joined = nullptr;
    double ms[4] = {0, 0, 0, 0};
};

namespace {
template <typename T, bool SEP, typename S>
static const void* queue_kernel()
{
    return reinterpret_cast<const void*>(k_contract_queue<T, SEP, S>);
}

// the work-queue kernel of (precision, separate staging, tile shape)
static const void* queue_fn(int prec, bool sep, int shape)
{
    if (prec == LELPH_FP64)
        switch (shape)
        {
        case 2: return queue_kernel<double, false, Ct32>();
        case 3: return queue_kernel<double, false, Ct16x48>();
        case 4: return queue_kernel<double, false, Ct16>();
        default: return queue_kernel<double, false, CtWide>();
        }
    if (sep)
        switch (shape)
        {
        case 2: return queue_kernel<float, true, Ct32>();
        case 3: return queue_kernel<float, true, Ct16x48>();
        case 4: return queue_kernel<float, true, Ct16>();
        default: return queue_kernel<float, true, CtWide>();
        }
    switch (shape)
    {
    case 2: return queue_kernel<float, false, Ct32>();
    case 3: return queue_kernel<float, false, Ct16x48>();
    case 4: return queue_kernel<float, false, Ct16>();
    default: return queue_kernel<float, false, CtWide>();
    }
}

static hipError_t launch_queue(int prec, bool sep, int shape, int grid, hipStream_t st,
                               void* pairs, void* items, int n, int* counter)
{
    void* args[] = {&pairs, &items, &n, &counter};
    return hipLaunchKernel(queue_fn(prec, sep, shape), dim3((unsigned)grid), dim3(256), args, 0,
                           st);
}

// issues the whole batch on m->stream (eagerly with events, or under capture)
static int multi_enqueue(lelph_multi* m, bool timed)
{
    struct Saved
    {
        hipStream_t st;
        DBuf bk, bkq;
        bool reuse, prev;
    };
    std::vector<Saved> saved(m->ctx.size());
    for (size_t i = 0; i < m->ctx.size(); ++i)
    {
        lelph_ctx* c = m->ctx[i];
        saved[i] = {c->stream, c->box_k, c->box_kq, c->box_reuse, c->box_prev_valid};
        c->stream = m->stream;
        c->box_reuse = false;
        c->box_prev_valid = false;
    }
    int st = LELPH_OK;
    auto body = [&]() -> int {
        if (timed) HIP_TRY(hipEventRecord(m->ev[0], m->stream));
        // the non-local G-space work of every pair (its own buffers and the
        // device wavefunctions) on the second stream, beside the boxes and
        // contractions
        if (m->non_loc)
        {
            HIP_TRY(hipEventRecord(m->fork, m->stream));
            HIP_TRY(hipStreamWaitEvent(m->nl, m->fork, 0));
            // along each cell's chains, as lelph_gkq_batch: a pair's K'
            // projections are the previous pair's K ones, mirrored
            for (lelph_ctx* c : m->ctx)
            {
                c->nl_chain = true;
                c->nl_prev_valid = false;
                c->nl_head_valid = false;
            }
            int nst = LELPH_OK;
            for (int64_t bi : m->border)
            {
                MultiPair& p = m->pairs[(size_t)bi];
                nst = run_nonlocal(m->ctx[p.cell], p.pr, nullptr, m->nl, nullptr,
                                   p.ebuf.as<z64>());
                if (nst != LELPH_OK) break;
            }
            for (lelph_ctx* c : m->ctx)
            {
                c->nl_chain = false;
                c->nl_prev_valid = false;
                c->nl_head_valid = false;
            }
            LELPH_TRY(nst);
        }
        // every pair's psi_k+q box into its own slot, along each cell's
        // k -> k+q chains: a linked pair's psi_k box is the k+q box of the
        // pair before it (the same box, or a shifted copy where that pair
        // has an umklapp vector), the last pair of a closed cycle takes its
        // psi_k+q box from the head's psi_k box, as in lelph_gkq_batch; a
        // chain head transforms both sides
        for (int64_t bi : m->border)
        {
            MultiPair& p = m->pairs[(size_t)bi];
            lelph_ctx* c = m->ctx[p.cell];
            c->box_k.p = p.kbox;
            c->box_k.n = p.bb;
            c->box_kq = p.box_kq;
            const double* q = m->qpt[p.cell].data();
            const MultiPair* lp = p.link >= 0 ? &m->pairs[(size_t)p.link] : nullptr;
            const void* kf = lp ? lp->box_kq.p : nullptr;
            const int* kg = lp ? lp->g0 : nullptr;
            const void* kqf = p.close >= 0 ? m->pairs[(size_t)p.close].kbox : nullptr;
            if (p.prec == LELPH_FP64)
                LELPH_TRY(build_boxes<double>(c, q, p.pr, kf, kg, nullptr, kqf));
            else
                LELPH_TRY(build_boxes<float>(c, q, p.pr, kf, kg, nullptr, kqf));
        }
        if (timed) HIP_TRY(hipEventRecord(m->ev[1], m->stream));
        // the work-queue contractions
        LELPH_TRY(zero_async(m->counter.p, sizeof(int) * m->queues.size(), m->stream));
        for (size_t qi = 0; qi < m->queues.size(); ++qi)
        {
            auto& qu = m->queues[qi];
            if (qu.n)
                HIP_TRY(launch_queue(qu.prec, qu.sep, qu.shape, qu.grid, m->stream, qu.qp.p,
                                     qu.qi.p, qu.n, m->counter.as<int>() + qi));
        }
        HIP_TRY(hipGetLastError());
        if (timed) HIP_TRY(hipEventRecord(m->ev[2], m->stream));
        // fixed-order reductions, then the non-local mode rotations (second
        // stream, after the reductions), joined, then the copy-out
        for (auto& p : m->pairs)
        {
            lelph_ctx* c = m->ctx[p.cell];
            z64* g = m->gdev[p.cell].as<z64>() + p.index * p.gsz;
            hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv(p.gsz, 256)), dim3(256), 0,
                               m->stream, p.part.as<z64>(), p.nsplit, p.gsz, 1.0 / (double)c->nf,
                               g);
            HIP_TRY(hipGetLastError());
        }
        if (m->non_loc)
        {
            HIP_TRY(hipEventRecord(m->reduced, m->stream));
            HIP_TRY(hipStreamWaitEvent(m->nl, m->reduced, 0));
            for (auto& p : m->pairs)
                LELPH_TRY(nl_mode(m->ctx[p.cell], p.ebuf.as<z64>(),
                                  m->gdev[p.cell].as<z64>() + p.index * p.gsz, m->nl));
            HIP_TRY(hipEventRecord(m->joined, m->nl));
            HIP_TRY(hipStreamWaitEvent(m->stream, m->joined, 0));
        }
        for (size_t i = 0; i < m->ctx.size(); ++i)
            if (m->gbytes[i])
                HIP_TRY(hipMemcpyAsync(m->ghost[i], m->gdev[i].p, m->gbytes[i],
                                       hipMemcpyDeviceToHost, m->stream));
        if (timed) HIP_TRY(hipEventRecord(m->ev[3], m->stream));
        return LELPH_OK;
    };
    st = body();
    for (size_t i = 0; i < m->ctx.size(); ++i)
    {
        lelph_ctx* c = m->ctx[i];
        c->stream = saved[i].st;
        c->box_k = saved[i].bk;
        c->box_kq = saved[i].bkq;
        c->box_reuse = saved[i].reuse;
        c->box_prev_valid = false;
    }
    return st;
}

static int queue_grid(int prec, bool sep, int shape, int nitems, int dev, int* grid)
{
    int occ = 0, ncu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, queue_fn(prec, sep, shape), 256,
                                                         0));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    *grid = std::max(1, std::min(nitems, std::max(1, occ) * std::max(1, ncu)));
    return LELPH_OK;
}

// the pair and item lists of one precision's queue, items largest first
template <typename T>
static int build_queue(lelph_multi* m, int prec, int shape, DBuf& qp, DBuf& qi, int* nitems,
                       bool* any_sep)
{
    std::vector<QueuePair<T>> hp;
    std::vector<std::pair<int64_t, QueueItem>> items;
    *any_sep = false;
    for (auto& p : m->pairs)
    {
        if (p.prec != prec || p.shape != shape) continue;
        lelph_ctx* c = m->ctx[p.cell];
        const lelph_lattice& L = c->lat;
        const int nmag_eff = (L.nspinor == 1 && L.nmag == 4) ? 1 : L.nmag;
        QueuePair<T> d{};
        d.psik = static_cast<cpx<T>*>(p.kbox);
        d.psikq = p.box_kq.as<cpx<T>>();
        d.dv = dv_ptr<T>(c);
        d.part = p.part.as<z64>();
        d.nm = L.nmodes;
        d.nb = L.nbnds;
        d.nspin = L.nspin;
        d.nspinor = L.nspinor;
        d.nmag = nmag_eff;
        d.sep = p.sep;
        d.N = c->nf;
        d.chunks_per_split = p.cps;
        *any_sep = *any_sep || p.sep;
        const int pi = (int)hp.size();
        hp.push_back(d);
        const int64_t nchunks = cdiv(c->nf, CT_KC);
        for (int64_t t = 0; t < p.tiles; ++t)
            for (int sp = 0; sp < p.nsplit; ++sp)
                for (int is = 0; is < L.nspin; ++is)
                {
                    const int64_t c0 = (int64_t)sp * p.cps;
                    const int64_t cost = std::max<int64_t>(0, std::min(p.cps, nchunks - c0)) *
                                         L.nspinor;
                    items.push_back({cost, QueueItem{pi, (int)t, sp, is}});
                }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<int64_t, QueueItem>& a,
                        const std::pair<int64_t, QueueItem>& b) { return a.first > b.first; });
    if (items.size() > 0x7fffffff) return LELPH_ERR_UNSUPPORTED;
    *nitems = (int)items.size();
    if (!*nitems) return LELPH_OK;
    std::vector<QueueItem> hi(items.size());
    for (size_t i = 0; i < items.size(); ++i) hi[i] = items[i].second;
    LELPH_TRY(qp.ensure(sizeof(QueuePair<T>) * hp.size()));
    LELPH_TRY(qi.ensure(sizeof(QueueItem) * hi.size()));
    HIP_TRY(hipMemcpy(qp.p, hp.data(), sizeof(QueuePair<T>) * hp.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(qi.p, hi.data(), sizeof(QueueItem) * hi.size(), hipMemcpyHostToDevice));
    return LELPH_OK;
}

static void multi_free(lelph_multi* m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    if (m->nl) (void)hipStreamSynchronize(m->nl);
    for (auto& p : m->pairs)
    {
        p.box_k.release();
        p.box_kq.release();
        p.part.release();
        p.ebuf.release();
    }
    for (auto& g : m->gdev) g.release();
    for (void* h : m->ghost)
        if (h) (void)hipHostFree(h);
    for (auto& qu : m->queues)
    {
        qu.qp.release();
        qu.qi.release();
    }
    m->counter.release();
    for (auto& e : m->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {m->fork, m->reduced, m->joined})
        if (e) (void)hipEventDestroy(e);
    if (m->nl) (void)hipStreamDestroy(m->nl);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}
} // namespace

extern "C" {

int lelph_multi_create(int ncells, lelph_ctx* const* ctxs, const double* qpts,
                       const int64_t* npairs, const lelph_pair* const* pairs, int non_loc,
                       lelph_multi** out)
{
    if (!out || ncells < 1 || !ctxs || !qpts || !npairs || !pairs) return LELPH_ERR_ARG;
    *out = nullptr;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    for (int i = 0; i < ncells; ++i)
    {
        lelph_ctx* c = ctxs[i];
        if (!c || npairs[i] < 0 || (npairs[i] > 0 && !pairs[i])) return LELPH_ERR_ARG;
        if (c->device != dev) return LELPH_ERR_ARG;
        if (!c->dv_set) return LELPH_ERR_STATE;
        const int pe = effective_precision(c);
        if (pe != LELPH_FP64 && pe != LELPH_FP32) return LELPH_ERR_UNSUPPORTED;
        if (c->lat.nmag == 2 && (c->lat.nspin != 2 || c->lat.nspinor != 1)) return LELPH_ERR_ARG;
        for (int64_t j = 0; j < npairs[i]; ++j)
        {
            const lelph_pair& p = pairs[i][j];
            if (p.ik < 0 || p.ikq < 0 || p.ik >= (int)c->wf.size() || p.ikq >= (int)c->wf.size())
                return LELPH_ERR_ARG;
            if (p.ksym < 0 || p.kqsym < 0 || p.ksym >= c->lat.nsym || p.kqsym >= c->lat.nsym)
                return LELPH_ERR_ARG;
            if (!c->wf[p.ik].wfc || !c->wf[p.ikq].wfc) return LELPH_ERR_STATE;
        }
    }
    lelph_multi* m = new (std::nothrow) lelph_multi();
    if (!m) return LELPH_ERR_ALLOC;
    auto fail = [&](int st) {
        multi_free(m);
        return st;
    };
    m->non_loc = non_loc ? 1 : 0;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(LELPH_ERR_HIP);
    for (auto& e : m->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(LELPH_ERR_HIP);
    if (hipStreamCreateWithFlags(&m->nl, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->reduced, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->joined, hipEventDisableTiming) != hipSuccess)
        return fail(LELPH_ERR_HIP);
    for (int i = 0; i < ncells; ++i)
    {
        lelph_ctx* c = ctxs[i];
        m->ctx.push_back(c);
        m->qpt.push_back({qpts[3 * i], qpts[3 * i + 1], qpts[3 * i + 2]});
        m->npairs.push_back(npairs[i]);
        const int prec = effective_precision(c);
        if (prec == LELPH_FP32)
        {
            int st = ensure_dv32(c);
            if (st != LELPH_OK) return fail(st);
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(LELPH_ERR_HIP);
        const lelph_lattice& L = c->lat;
        const int64_t gsz = (int64_t)L.nmodes * L.nspin * L.nbnds * L.nbnds;
        const int64_t nbox = (int64_t)L.nspin * L.nbnds * L.nspinor;
        const size_t bb = (prec == LELPH_FP64 ? sizeof(z64) : sizeof(c32)) * (size_t)(nbox * c->nf);
        const int64_t nchunks = cdiv(c->nf, CT_KC);
        const int64_t tiles = ct_tiles(1, L.nmodes, L.nbnds); // sets the K splits
        const int shape = ct_shape(c, prec == LELPH_FP32);
        for (int64_t j = 0; j < npairs[i]; ++j)
        {
            MultiPair p;
            p.cell = i;
            p.index = j;
            p.pr = pairs[i][j];
            p.prec = prec;
            p.sep = prec == LELPH_FP32 ? (contract_sep_form<float>(c) ? 1 : 0) : 0;
            p.gsz = gsz;
            p.shape = shape;
            p.tiles = ct_tiles(shape, L.nmodes, L.nbnds);
            pick_splits(tiles * L.nspin, nchunks, gsz, &p.cps, &p.nsplit, c->tune.split_target);
            p.bb = bb;
            const int64_t es = (int64_t)L.nspin * L.nbnds * L.nbnds;
            // the k+q side's umklapp vector (it shifts the box a linked
            // pair takes over)
            SideGeom gkq, gk;
            int gs = local_geometry(c, m->qpt.back().data(), p.pr.ikq, p.pr.ik, p.pr.kqsym,
                                    p.pr.ksym, &gkq, &gk);
            if (gs != LELPH_OK)
            {
                m->pairs.push_back(p);
                return fail(gs);
            }
            for (int d = 0; d < 3; ++d) p.g0[d] = (int)std::lrint(gkq.G0c[d]);
            if (p.box_kq.ensure(bb) || p.part.ensure(sizeof(z64) * (size_t)(p.nsplit * gsz)) ||
                (non_loc && p.ebuf.ensure(sizeof(z64) * (size_t)(3 * L.natom * es))))
            {
                m->pairs.push_back(p);
                return fail(LELPH_ERR_ALLOC);
            }
            m->pairs.push_back(p);
        }
        // chains: a pair whose k side is the k+q side of the pair before it
        // takes that box over (aliased when the umklapp vector is zero);
        // every other pair gets a psi_k box of its own
        {
            const int64_t first = (int64_t)m->pairs.size() - npairs[i];
            std::vector<int64_t> order;
            chain_order(c, pairs[i], npairs[i], order);
            int64_t head = first;
            for (size_t n = 0; n < order.size(); ++n)
            {
                MultiPair& p = m->pairs[(size_t)(first + order[n])];
                m->border.push_back(first + order[n]);
                if (n > 0)
                {
                    const MultiPair& q = m->pairs[(size_t)(first + order[n - 1])];
                    if (p.pr.ik == q.pr.ikq && p.pr.ksym == q.pr.kqsym)
                        p.link = first + order[n - 1];
                }
                const MultiPair* lp = p.link >= 0 ? &m->pairs[(size_t)p.link] : nullptr;
                if (lp && !lp->g0[0] && !lp->g0[1] && !lp->g0[2])
                    p.kbox = lp->box_kq.p;
                else
                {
                    if (p.box_k.ensure(bb)) return fail(LELPH_ERR_ALLOC);
                    p.kbox = p.box_k.p;
                }
                // a pair whose k+q side is the k side of its chain's head
                // closes a cycle: that box, shifted, is its k+q box
                if (!lp) head = first + order[n];
                const MultiPair& h = m->pairs[(size_t)head];
                if (p.pr.ikq == h.pr.ik && p.pr.kqsym == h.pr.ksym) p.close = head;
            }
        }
        m->gdev.emplace_back();
        m->gbytes.push_back(sizeof(z64) * (size_t)(npairs[i] * gsz));
        m->ghost.push_back(nullptr);
        if (m->gbytes[i])
        {
            if (m->gdev[i].ensure(m->gbytes[i])) return fail(LELPH_ERR_ALLOC);
            if (hipHostMalloc(&m->ghost[i], m->gbytes[i]) != hipSuccess)
            {
                m->ghost[i] = nullptr;
                return fail(LELPH_ERR_ALLOC);
            }
        }
    }
    int st = LELPH_OK;
    for (int prec : {LELPH_FP64, LELPH_FP32})
        for (int shape = 1; shape <= 4 && st == LELPH_OK; ++shape)
        {
            m->queues.emplace_back();
            auto& qu = m->queues.back();
            qu.prec = prec;
            qu.shape = shape;
            st = prec == LELPH_FP64
                     ? build_queue<double>(m, prec, shape, qu.qp, qu.qi, &qu.n, &qu.sep)
                     : build_queue<float>(m, prec, shape, qu.qp, qu.qi, &qu.n, &qu.sep);
            if (st == LELPH_OK && qu.n) st = queue_grid(prec, qu.sep, shape, qu.n, dev, &qu.grid);
        }
    if (st == LELPH_OK && m->counter.ensure(sizeof(int) * m->queues.size())) st = LELPH_ERR_ALLOC;
    if (st != LELPH_OK) return fail(st);
    // eager pass: every workspace of the contexts reaches its size here,
    // so the captured pass allocates nothing
    st = multi_enqueue(m, true);
    if (st == LELPH_OK && hipStreamSynchronize(m->stream) != hipSuccess) st = LELPH_ERR_HIP;
    if (st != LELPH_OK) return fail(st);
    float a = 0;
    for (int k = 0; k < 3; ++k)
    {
        if (hipEventElapsedTime(&a, m->ev[k], m->ev[k + 1]) != hipSuccess) return fail(LELPH_ERR_HIP);
        m->ms[k] = a;
    }
    for (lelph_ctx* c : m->ctx)
    {
        st = check_err(c);
        if (st != LELPH_OK) return fail(st);
    }
    if (hipStreamBeginCapture(m->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
        return fail(LELPH_ERR_HIP);
    st = multi_enqueue(m, false);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(m->stream, &gr);
    m->graph = gr;
    if (st != LELPH_OK) return fail(st);
    if (ec != hipSuccess || !gr) return fail(LELPH_ERR_HIP);
    if (hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0) != hipSuccess)
    {
        m->exec = nullptr;
        return fail(LELPH_ERR_HIP);
    }
    *out = m;
    return LELPH_OK;
}

int lelph_multi_run(lelph_multi* m, double* const* g_out)
{
    if (!m || !g_out) return LELPH_ERR_ARG;
    for (size_t i = 0; i < m->ctx.size(); ++i)
        if (m->gbytes[i] && !g_out[i]) return LELPH_ERR_ARG;
    HIP_TRY(hipEventRecord(m->ev[0], m->stream));
    HIP_TRY(hipGraphLaunch(m->exec, m->stream));
    HIP_TRY(hipEventRecord(m->ev[3], m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    float a = 0;
    HIP_TRY(hipEventElapsedTime(&a, m->ev[0], m->ev[3]));
    m->ms[3] = a;
    for (size_t i = 0; i < m->ctx.size(); ++i)
    {
        if (m->gbytes[i]) std::memcpy(g_out[i], m->ghost[i], m->gbytes[i]);
        LELPH_TRY(check_err(m->ctx[i]));
    }
    return LELPH_OK;
}

int lelph_multi_timings(lelph_multi* m, double* ms)
{
    if (!m || !ms) return LELPH_ERR_ARG;
    for (int k = 0; k < 4; ++k) ms[k] = m->ms[k];
    for (int k = 4; k < 8; ++k) ms[k] = 0;
    for (const auto& qu : m->queues)
    {
        const int f = qu.prec == LELPH_FP64 ? 0 : 1;
        ms[4 + f] += qu.n;
        ms[6 + f] = std::max(ms[6 + f], (double)qu.grid);
    }
    return LELPH_OK;
}

void lelph_multi_destroy(lelph_multi* m) { multi_free(m); }

// electronic_reps (src/symmetries/elec_reps.c:16-413)
int lelph_electronic_reps(lelph_ctx* c, const double Rsym[9], const double tauR[3], int tim_revR,
                          int64_t ikBZ, double* Dkmn)
{
    if (!c || !Rsym || !tauR || !Dkmn) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    if (c->kpt_bz.empty()) return LELPH_ERR_STATE;
    if (ikBZ < 0 || ikBZ >= L.nkpts_BZ) return LELPH_ERR_ARG;
    const double* alat = L.alat_vec;
    const double* blat = L.blat_vec;
    // R k_BZ (crystal) and its index in the full-BZ list (:34-53)
    const double* kbc = c->kpt_bz.data() + 3 * ikBZ;
    double kbz[3], t[3], Rk[3];
    for (int i = 0; i < 3; ++i)
        kbz[i] = (blat[3 * i] * kbc[0] + blat[3 * i + 1] * kbc[1] + blat[3 * i + 2] * kbc[2]) /
                 (2 * LELPH_RPI);
    matvec3(Rsym, kbz, false, t);
    matvec3(alat, t, true, Rk);
    int64_t iRk = -1;
    for (int64_t i = 0; i < L.nkpts_BZ && iRk < 0; ++i)
    {
        double s2 = 0;
        for (int j = 0; j < 3; ++j)
        {
            double d = c->kpt_bz[3 * i + j] - Rk[j];
            d -= rint(d);
            s2 += d * d;
        }
        if (sqrt(s2) < LELPH_EPS) iRk = i;
    }
    if (iRk < 0) return LELPH_ERR_SYM;
    const int ik1 = c->kmap[2 * ikBZ], iS1 = c->kmap[2 * ikBZ + 1];
    const int ik2 = c->kmap[2 * iRk], iS2 = c->kmap[2 * iRk + 1];
    if (ik1 < 0 || ik2 < 0 || ik1 >= (int)c->wf.size() || ik2 >= (int)c->wf.size())
        return LELPH_ERR_ARG;
    const DevWfc& w1 = c->wf[ik1];
    const DevWfc& w2 = c->wf[ik2];
    if (!w1.wfc || !w2.wfc) return LELPH_ERR_STATE;
    const lelph_symmetry& sy1 = c->syms[iS1];
    const lelph_symmetry& sy2 = c->syms[iS2];
    const double* k1 = c->kpt_ibz.data() + 3 * ik1;
    const double* k2 = c->kpt_ibz.data() + 3 * ik2;
    // geometry (:83-176)
    DmGeom geo{};
    matvec3(sy1.Rmat, k1, false, t);
    matvec3(alat, t, true, geo.kcrys);
    matvec3(blat, sy1.tau, true, geo.tau1c);
    matvec3(blat, tauR, true, geo.tauRc);
    matvec3(blat, sy2.tau, true, geo.tau2c);
    for (int i = 0; i < 3; ++i)
    {
        geo.tau1c[i] /= 2 * LELPH_RPI;
        geo.tauRc[i] /= 2 * LELPH_RPI;
        geo.tau2c[i] /= 2 * LELPH_RPI;
    }
    double RS1[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            RS1[3 * i + j] = Rsym[3 * i] * sy1.Rmat[j] + Rsym[3 * i + 1] * sy1.Rmat[3 + j] +
                             Rsym[3 * i + 2] * sy1.Rmat[6 + j];
    double ulm[3], Rkc[3];
    matvec3(sy2.Rmat, k2, false, ulm);
    matvec3(RS1, k1, false, Rkc);
    matvec3(alat, Rkc, true, geo.Rkcrys);
    for (int i = 0; i < 3; ++i) ulm[i] -= Rkc[i];
    matvec3(alat, ulm, true, geo.ulmc);
    // cartesian G -> crystal: M = alat^T S (rotate_Gvecs.c:15-81)
    const double* Ss[3] = {sy1.Rmat, RS1, sy2.Rmat};
    double* Ms[3] = {geo.M1, geo.M2, geo.M3};
    for (int m = 0; m < 3; ++m)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Ms[m][3 * i + j] = alat[i] * Ss[m][j] + alat[3 + i] * Ss[m][3 + j] +
                                   alat[6 + i] * Ss[m][6 + j];
    geo.tr1 = sy1.time_rev;
    geo.trR = tim_revR ? 1 : 0;
    geo.tr2 = sy2.time_rev;
    LELPH_TRY(su2mat(sy1.Rmat, L.nspinor, sy1.time_rev, geo.su1));
    LELPH_TRY(su2mat(Rsym, L.nspinor, tim_revR, geo.suR));
    LELPH_TRY(su2mat(sy2.Rmat, L.nspinor, sy2.time_rev, geo.su2));
    // device work
    const int Nx = (int)L.fft_dims[0], Ny = (int)L.fft_dims[1], Nz = (int)L.fft_dims[2];
    const int64_t nsets = (int64_t)L.nspin * L.nbnds, Lk = (int64_t)L.nspinor * w1.npw;
    const int nb = L.nbnds;
    LELPH_TRY(c->dm_map.ensure(sizeof(int) * c->nf));
    LELPH_TRY(c->dm_ph.ensure(sizeof(z64) * 3 * std::max<int64_t>(1, w1.npw)));
    LELPH_TRY(c->dm_pos.ensure(sizeof(int) * std::max<int64_t>(1, w2.npw)));
    LELPH_TRY(c->dm_wRS.ensure(sizeof(z64) * nsets * Lk));
    LELPH_TRY(c->dm_wS2.ensure(sizeof(z64) * nsets * Lk));
    HIP_TRY(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    HIP_TRY(hipMemsetAsync(c->dm_map.p, 0xFF, sizeof(int) * c->nf, c->stream));
    HIP_TRY(hipMemsetAsync(c->dm_wS2.p, 0, sizeof(z64) * nsets * Lk, c->stream));
    hipLaunchKernelGGL(k_dm_side1, dim3((unsigned)cdiv(w1.npw, 256)), dim3(256), 0, c->stream,
                       w1.gvec, w1.npw, geo, Nx, Ny, Nz, c->dm_map.as<int>(), c->dm_ph.as<z64>(),
                       c->d_err);
    hipLaunchKernelGGL(k_dm_side2, dim3((unsigned)cdiv(w2.npw, 256)), dim3(256), 0, c->stream,
                       w2.gvec, w2.npw, geo, Nx, Ny, Nz, c->dm_map.as<int>(), c->dm_pos.as<int>());
    hipLaunchKernelGGL(k_dm_rot1, dim3((unsigned)cdiv(nsets * w1.npw, 256)), dim3(256), 0,
                       c->stream, w1.wfc, nsets, L.nspinor, w1.npw, c->dm_ph.as<z64>(), geo,
                       c->dm_wRS.as<z64>());
    hipLaunchKernelGGL(k_dm_rot2, dim3((unsigned)cdiv(nsets * w2.npw, 256)), dim3(256), 0,
                       c->stream, w2.wfc, nsets, L.nspinor, w2.npw, w1.npw, c->dm_pos.as<int>(),
                       c->dm_ph.as<z64>(), geo, c->dm_wS2.as<z64>());
    HIP_TRY(hipGetLastError());
    // D[s] = wS2[s] (nb x Lk) . wRS[s]^T (:377-394), split over the sphere
    const int64_t nchunks = cdiv(Lk, CT_KC);
    const int64_t tiles = cdiv(nb, GZ_MB) * cdiv(nb, GZ_NB);
    int64_t cps;
    int nsplit;
    pick_splits(tiles, nchunks, (int64_t)nb * nb, &cps, &nsplit);
    LELPH_TRY(c->dm_part.ensure(sizeof(z64) * (size_t)nsplit * nb * nb));
    LELPH_TRY(c->dm_out.ensure(sizeof(z64) * (size_t)L.nspin * nb * nb));
    for (int is = 0; is < L.nspin; ++is)
    {
        hipLaunchKernelGGL(k_zgemm_nt<double>, dim3((unsigned)tiles, (unsigned)nsplit), dim3(256),
                           0, c->stream, c->dm_wS2.as<z64>() + is * nb * Lk,
                           c->dm_wRS.as<z64>() + is * nb * Lk, nb, nb, Lk, 0, cps,
                           c->dm_part.as<z64>());
        hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)cdiv((int64_t)nb * nb, 256)), dim3(256),
                           0, c->stream, c->dm_part.as<z64>(), nsplit, (int64_t)nb * nb, 1.0,
                           c->dm_out.as<z64>() + (int64_t)is * nb * nb);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(Dkmn, c->dm_out.p, sizeof(z64) * L.nspin * nb * nb,
                           hipMemcpyDeviceToHost, c->stream));
    return check_err(c);
}

// elph_q_rotate (src/symmetries/elphq_expand.c:16-73) on the device:
// gSq[m][s] = (D_l[s] op(g[m][s]) D_r[s]^H)^T = conj(D_r[s]) op'(g) D_l[s]^T,
// op = transpose (conjugate transpose under time reversal), so op' = 1 (conj)
int lelph_elph_q_rotate_gpu(lelph_ctx* c, const double* Dmats_l, const double* elph_mat_q,
                            const double* Dmats_r, int nmodes, int nspin, int nb, int tim_rev,
                            double* elph_mat_Sq)
{
    if (!c || !Dmats_l || !elph_mat_q || !Dmats_r || !elph_mat_Sq) return LELPH_ERR_ARG;
    if (nmodes <= 0 || nspin <= 0 || nb <= 0 || nspin > 2) return LELPH_ERR_ARG;
    const int64_t mat = (int64_t)nb * nb;
    const int64_t gsz = (int64_t)nmodes * nspin * mat;
    LELPH_TRY(c->qr_dl.ensure(sizeof(z64) * nspin * mat));
    LELPH_TRY(c->qr_dr.ensure(sizeof(z64) * nspin * mat));
    LELPH_TRY(c->qr_g.ensure(sizeof(z64) * gsz));
    LELPH_TRY(c->qr_t.ensure(sizeof(z64) * gsz));
    LELPH_TRY(c->qr_out.ensure(sizeof(z64) * gsz));
    HIP_TRY(hipMemcpyAsync(c->qr_dl.p, Dmats_l, sizeof(z64) * nspin * mat, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->qr_dr.p, Dmats_r, sizeof(z64) * nspin * mat, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->qr_g.p, elph_mat_q, sizeof(z64) * gsz, hipMemcpyHostToDevice,
                           c->stream));
    const int tiles = (nb + 31) / 32;
    for (int is = 0; is < nspin; ++is)
    {
        // batch over modes: matrix (m, is) sits at (m * nspin + is) * mat
        const z64* dr = c->qr_dr.as<z64>() + is * mat;
        const z64* dl = c->qr_dl.as<z64>() + is * mat;
        const z64* g = c->qr_g.as<z64>() + is * mat;
        z64* t = c->qr_t.as<z64>() + is * mat;
        z64* o = c->qr_out.as<z64>() + is * mat;
        const dim3 grid((unsigned)(tiles * tiles), (unsigned)nmodes);
        if (tim_rev)
            hipLaunchKernelGGL((k_zgemm_bat<true, true, false>), grid, dim3(256), 0, c->stream, dr,
                               (int64_t)0, g, nspin * mat, t, nspin * mat, nb);
        else
            hipLaunchKernelGGL((k_zgemm_bat<true, false, false>), grid, dim3(256), 0, c->stream,
                               dr, (int64_t)0, g, nspin * mat, t, nspin * mat, nb);
        hipLaunchKernelGGL((k_zgemm_bat<false, false, true>), grid, dim3(256), 0, c->stream, t,
                           nspin * mat, dl, (int64_t)0, o, nspin * mat, nb);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(elph_mat_Sq, c->qr_out.p, sizeof(z64) * gsz, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LELPH_OK;
}

/* ------------------------------------------------------------------------ */
/* q-star rotation without allocation                                       */
/* ------------------------------------------------------------------------ */

// The intermediate conj(D_r) op(g) of the no-allocation q-star rotation:
// one static array per device, held until the process ends.  Host code never
// uses the symbol itself (on the host it is only a stand-in); it uses the
// device address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[2 * (size_t)LELPH_NOALLOC_MAX_ELEMS];

#define NOALLOC_MAX_DEVICES 64
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of `scratch` on the current device, looked up once per
// device (the first call on a device, or lelph_create on it)
static int scratch_ptr(double** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return LELPH_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev])
    {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return LELPH_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<double*>(g_scratch_ptr[dev]);
    return LELPH_OK;
}

int lelph_elph_q_rotate_noalloc(const double* d_Dmats_l, const double* d_elph_mat_q,
                                const double* d_Dmats_r, int nmodes, int nspin, int nb,
                                int tim_rev, double* d_elph_mat_Sq, void* stream)
{
    if (!d_Dmats_l || !d_elph_mat_q || !d_Dmats_r || !d_elph_mat_Sq) return LELPH_ERR_ARG;
    if (nmodes <= 0 || nspin <= 0 || nb <= 0 || nspin > 2 || nmodes > 65535)
        return LELPH_ERR_ARG;
    const int64_t mat = (int64_t)nb * nb;
    const int64_t gsz = (int64_t)nmodes * nspin * mat;
    if (gsz > (int64_t)LELPH_NOALLOC_MAX_ELEMS) return LELPH_ERR_SCRATCH;
    double* sp = nullptr;
    LELPH_TRY(scratch_ptr(&sp));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const z64* Dl = reinterpret_cast<const z64*>(d_Dmats_l);
    const z64* Dr = reinterpret_cast<const z64*>(d_Dmats_r);
    const z64* G = reinterpret_cast<const z64*>(d_elph_mat_q);
    z64* T = reinterpret_cast<z64*>(sp);
    z64* O = reinterpret_cast<z64*>(d_elph_mat_Sq);
    const int tiles = (nb + 31) / 32;
    const dim3 grid((unsigned)(tiles * tiles), (unsigned)nmodes);
    // the same two batched GEMMs as lelph_elph_q_rotate_gpu; every element
    // of T is written by the first before the second reads it
    for (int is = 0; is < nspin; ++is)
    {
        if (tim_rev)
            hipLaunchKernelGGL((k_zgemm_bat<true, true, false>), grid, dim3(256), 0, st,
                               Dr + is * mat, (int64_t)0, G + is * mat, nspin * mat, T + is * mat,
                               nspin * mat, nb);
        else
            hipLaunchKernelGGL((k_zgemm_bat<true, false, false>), grid, dim3(256), 0, st,
                               Dr + is * mat, (int64_t)0, G + is * mat, nspin * mat, T + is * mat,
                               nspin * mat, nb);
        hipLaunchKernelGGL((k_zgemm_bat<false, false, true>), grid, dim3(256), 0, st,
                           T + is * mat, nspin * mat, Dl + is * mat, (int64_t)0, O + is * mat,
                           nspin * mat, nb);
        HIP_TRY(hipGetLastError());
    }
    return LELPH_OK;
}

/* ------------------------------------------------------------------------ */
/* Fourier interpolation of dV_q                                             */
/* ------------------------------------------------------------------------ */

// get_fft_box (src/wfc/gen_fftbox.c:8-90) for one rank
static void lr_fft_box(double EcutRy, const double* blat, int64_t* box)
{
    for (int i = 0; i < 3; ++i)
    {
        const double bn =
            std::sqrt(blat[i] * blat[i] + blat[3 + i] * blat[3 + i] + blat[6 + i] * blat[6 + i]);
        box[i] = 2 * (int64_t)std::rint(std::sqrt(EcutRy) / bn + 1) + 1;
    }
}

// copies the long-range parameters (arrays owned by the context)
static void lr_copy(lelph_ctx* c, const lelph_longrange* lr)
{
    const int na = c->lat.natom;
    c->ip_lr = *lr;
    auto cp = [](std::vector<double>& v, const double* p, size_t n) {
        v.assign(p ? p : nullptr, p ? p + n : nullptr);
    };
    cp(c->ip_eps, lr->epsilon, 9);
    cp(c->ip_zb, lr->Zborn, (size_t)9 * na);
    cp(c->ip_qp, lr->Qpole, (size_t)27 * na);
    cp(c->ip_zv, lr->Zvals, (size_t)na);
    c->ip_lr.epsilon = lr->epsilon ? c->ip_eps.data() : nullptr;
    c->ip_lr.Zborn = lr->Zborn ? c->ip_zb.data() : nullptr;
    c->ip_lr.Qpole = lr->Qpole ? c->ip_qp.data() : nullptr;
    c->ip_lr.Zvals = lr->Zvals ? c->ip_zv.data() : nullptr;
}

// dV_add_longrange (dVpot_add_long_range.c:18-241) for nq points at once:
// dv (nq, nm, nmag, nf) on the device += sgn * long range; eig (device,
// nm x nm) rotates to the mode basis and needs nq == 1.
static int lr_apply(lelph_ctx* c, const lelph_longrange& lr, const double* qs, int64_t nq,
                    const z64* eig, z64* dv, double sgn)
{
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes, na = L.natom;
    if (eig && nq != 1) return LELPH_ERR_ARG;
    if (L.dimension != '3' && L.dimension != '2') return LELPH_OK; // no kernel in the reference
    if (L.dimension == '2')
        for (int64_t i = 0; i < nq; ++i)
            if (std::fabs(qs[3 * i + 2]) > LELPH_EPS) return LELPH_ERR_ARG;
    if (lr.Zborn && !lr.epsilon) return LELPH_ERR_ARG;
    const double* zv = lr.only_induced ? nullptr : lr.Zvals;
    if (!lr.epsilon && !zv) return LELPH_OK; // the kernel returns zeros
    int mask = 0;
    for (int m = 0; m < L.nmag && m < 4; ++m)
        if (lr.nmags_add[m]) mask |= 1 << m;
    if (!mask) return LELPH_OK;
    LrGeom g{};
    std::memcpy(g.blat, L.blat_vec, sizeof(g.blat));
    if (lr.epsilon) std::memcpy(g.eps, lr.epsilon, sizeof(g.eps));
    g.latz = L.alat_vec[8];
    g.volume = L.volume;
    g.EcutRy = lr.EcutRy;
    g.eta_bare = lr.eta_bare;
    g.eta_ind = lr.eta_induced;
    g.has_eps = lr.epsilon != nullptr;
    g.has_zv = zv != nullptr;
    g.has_zb = lr.Zborn != nullptr;
    g.has_qp = lr.Qpole != nullptr;
    g.natom = na;
    g.dim = L.dimension;
    int64_t gb[3];
    lr_fft_box(lr.EcutRy, L.blat_vec, gb);
    for (int i = 0; i < 3; ++i)
    {
        g.N[i] = (int)L.fft_dims[i];
        if (gb[i] > L.fft_dims[i]) gb[i] = L.fft_dims[i];
    }
    if (L.dimension == '2') gb[2] = L.fft_dims[2];
    for (int i = 0; i < 3; ++i)
    {
        g.lo[i] = -(int)(gb[i] / 2);
        g.hi[i] = (int)((gb[i] - 1) / 2);
    }
    if (zv) LELPH_TRY(c->lr_zv.ensure(sizeof(double) * na));
    if (lr.Zborn) LELPH_TRY(c->lr_zb.ensure(sizeof(double) * 9 * na));
    if (lr.Qpole) LELPH_TRY(c->lr_qp.ensure(sizeof(double) * 27 * na));
    if (zv)
        HIP_TRY(hipMemcpyAsync(c->lr_zv.p, zv, sizeof(double) * na, hipMemcpyHostToDevice,
                               c->stream));
    if (lr.Zborn)
        HIP_TRY(hipMemcpyAsync(c->lr_zb.p, lr.Zborn, sizeof(double) * 9 * na,
                               hipMemcpyHostToDevice, c->stream));
    if (lr.Qpole)
        HIP_TRY(hipMemcpyAsync(c->lr_qp.p, lr.Qpole, sizeof(double) * 27 * na,
                               hipMemcpyHostToDevice, c->stream));
    const int64_t nf = c->nf;
    // q chunks keep the coefficient boxes within 2 GiB
    const int64_t per_q = (int64_t)nm * nf * (int64_t)sizeof(z64);
    int64_t chunk = std::max<int64_t>(1, std::min<int64_t>((int64_t)2 << 30, per_q * nq) / per_q);
    chunk = std::min<int64_t>(chunk, std::max<int64_t>(1, 65535 / nm));
    LELPH_TRY(c->lr_q.ensure(sizeof(double) * 3 * nq));
    HIP_TRY(hipMemcpyAsync(c->lr_q.p, qs, sizeof(double) * 3 * nq, hipMemcpyHostToDevice,
                           c->stream));
    LELPH_TRY(c->lr_box.ensure((size_t)per_q * chunk));
    if (eig) LELPH_TRY(c->lr_box2.ensure((size_t)per_q));
    for (int64_t q0 = 0; q0 < nq; q0 += chunk)
    {
        const int64_t nqc = std::min(chunk, nq - q0);
        hipLaunchKernelGGL(k_lr_box, dim3((unsigned)cdiv(nf, 256), (unsigned)nqc), dim3(256), 0,
                           c->stream, g, c->lr_q.as<double>() + 3 * q0, c->d_apos,
                           c->lr_zv.as<double>(), c->lr_zb.as<double>(), c->lr_qp.as<double>(),
                           c->lr_box.as<z64>());
        HIP_TRY(hipGetLastError());
        z64* box = c->lr_box.as<z64>();
        if (eig)
        {
            hipLaunchKernelGGL(k_mode_rotate, dim3((unsigned)cdiv((int64_t)nm * nf, 256)),
                               dim3(256), 0, c->stream, box, eig, nm, nf, c->lr_box2.as<z64>());
            HIP_TRY(hipGetLastError());
            box = c->lr_box2.as<z64>();
        }
        LELPH_TRY(fft3d_batch<double>(c->ax, box, nqc * nm, +1, c->stream));
        hipLaunchKernelGGL(k_lr_add, dim3((unsigned)cdiv(nf, 256), (unsigned)(nqc * nm)),
                           dim3(256), 0, c->stream, box, dv + q0 * nm * L.nmag * nf, nm, L.nmag,
                           nf, mask, sgn);
        HIP_TRY(hipGetLastError());
    }
    return LELPH_OK;
}

int lelph_dv_longrange(lelph_ctx* c, const double qpt[3], const lelph_longrange* lr,
                       const double* eigVec, int sign, double* dvscf)
{
    if (!c || !qpt || !lr || !dvscf) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    const size_t bytes = sizeof(z64) * (size_t)nm * L.nmag * c->nf;
    LELPH_TRY(c->ip_tmp.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(c->ip_tmp.p, dvscf, bytes, hipMemcpyHostToDevice, c->stream));
    const z64* eig = nullptr;
    if (eigVec)
    {
        LELPH_TRY(c->ip_eig.ensure(sizeof(z64) * nm * nm));
        HIP_TRY(hipMemcpyAsync(c->ip_eig.p, eigVec, sizeof(z64) * nm * nm, hipMemcpyHostToDevice,
                               c->stream));
        eig = c->ip_eig.as<z64>();
    }
    LELPH_TRY(lr_apply(c, *lr, qpt, 1, eig, c->ip_tmp.as<z64>(), sign < 0 ? -1.0 : 1.0));
    HIP_TRY(hipMemcpyAsync(dvscf, c->ip_tmp.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return check_err(c);
}

// rotate_dvscf on the device: out (nm, nmag, nf) from in
static int rotate_dev(lelph_ctx* c, const double* qpt, const lelph_symmetry* sym, int composite,
                      const z64* in, z64* out)
{
    const lelph_lattice& L = c->lat;
    const double* blat = L.blat_vec;
    const double tp = 2.0 * LELPH_RPI;
    double tr[3], tmp[9], sr[9];
    for (int i = 0; i < 3; ++i)
    {
        tr[i] = (blat[i] * sym->tau[0] + blat[3 + i] * sym->tau[1] + blat[6 + i] * sym->tau[2]) / tp;
        if (sym->time_rev) tr[i] = -tr[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += sym->Rmat[k * 3 + i] * L.alat_vec[k * 3 + j];
            tmp[i * 3 + j] = a / tp * (sym->time_rev ? -1.0 : 1.0);
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += blat[k * 3 + i] * tmp[k * 3 + j];
            sr[i * 3 + j] = a;
        }
    const int64_t nf = c->nf;
    LELPH_TRY(c->rot_sr.ensure(sizeof(sr)));
    LELPH_TRY(c->rot_idx.ensure(sizeof(int64_t) * nf));
    HIP_TRY(hipMemcpyAsync(c->rot_sr.p, sr, sizeof(sr), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_rot_index, dim3((unsigned)cdiv(nf, 256)), dim3(256), 0, c->stream,
                       c->rot_sr.as<double>(), tr[0], tr[1], tr[2], (int)L.fft_dims[0],
                       (int)L.fft_dims[1], (int)L.fft_dims[2], c->rot_idx.as<int64_t>(),
                       c->d_err);
    HIP_TRY(hipGetLastError());
    int e = 0;
    LELPH_TRY(read_err(c, &e));
    if (e & 2) return LELPH_ERR_SYM;
    if (e) return LELPH_ERR_RANGE;
    double qc[3], Rq[3];
    matvec3(blat, qpt, false, qc);
    matvec3(sym->Rmat, qc, false, Rq);
    const double a = -(Rq[0] * sym->tau[0] + Rq[1] * sym->tau[1] + Rq[2] * sym->tau[2]);
    const z64 ph = {std::cos(a), std::sin(a)};
    z64 su[4] = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
    if (L.nmag == 4) LELPH_TRY(su2mat(sym->Rmat, L.nspinor, sym->time_rev, su));
    hipLaunchKernelGGL(k_rot_gather, dim3((unsigned)cdiv((int64_t)L.nmodes * nf, 256)), dim3(256),
                       0, c->stream, in, c->rot_idx.as<int64_t>(), L.nmodes, L.nmag, nf, ph,
                       su[0], su[1], su[2], su[3], composite, sym->time_rev ? 1 : 0, out);
    HIP_TRY(hipGetLastError());
    return LELPH_OK;
}

int lelph_rotate_dvscf(lelph_ctx* c, const double qpt[3], const double* dvscf_in,
                       const lelph_symmetry* sym, int composite_form, double* dvscf_out)
{
    if (!c || !qpt || !dvscf_in || !sym || !dvscf_out) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const size_t bytes = sizeof(z64) * (size_t)L.nmodes * L.nmag * c->nf;
    LELPH_TRY(c->ip_tmp.ensure(bytes));
    LELPH_TRY(c->rot_tmp.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(c->ip_tmp.p, dvscf_in, bytes, hipMemcpyHostToDevice, c->stream));
    LELPH_TRY(rotate_dev(c, qpt, sym, composite_form, c->ip_tmp.as<z64>(), c->rot_tmp.as<z64>()));
    HIP_TRY(hipMemcpyAsync(dvscf_out, c->rot_tmp.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LELPH_OK;
}

int lelph_interp_begin(lelph_ctx* c, const int64_t qgrid[3], const double* masses,
                       const lelph_longrange* lr)
{
    if (!c || !qgrid) return LELPH_ERR_ARG;
    for (int i = 0; i < 3; ++i)
        if (qgrid[i] < 1 || qgrid[i] > 1024) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int na = L.natom;
    c->ip_on = c->ip_ready = c->ip_last_ok = false;
    std::memcpy(c->ip_grid, qgrid, sizeof(c->ip_grid));
    c->ip_nq = qgrid[0] * qgrid[1] * qgrid[2];
    c->ip_mass.assign(masses ? masses : nullptr, masses ? masses + na : nullptr);
    c->ip_has_lr = lr != nullptr;
    if (lr) lr_copy(c, lr);
    // Wigner-Seitz vectors of every (R, atom) (interpolation_driver.c:212-216)
    const int64_t nR = c->ip_nq;
    std::vector<int64_t> degen((size_t)nR * na);
    int64_t cap = 16 * nR * na, nws = 0;
    std::vector<int64_t> ws;
    for (;;)
    {
        ws.resize((size_t)3 * cap);
        const int st = lelph_wigner_seitz(qgrid, L.alat_vec, LELPH_EPS6, c->atomic_pos.data(), na,
                                          nullptr, 0, cap, ws.data(), degen.data(), &nws);
        if (st == LELPH_OK) break;
        if (st != LELPH_ERR_RANGE) return st;
        cap *= 4;
    }
    std::vector<int> wsi((size_t)3 * std::max<int64_t>(nws, 1));
    for (int64_t i = 0; i < 3 * nws; ++i) wsi[i] = (int)ws[i];
    c->dy_ws_h.assign(ws.begin(), ws.begin() + 3 * nws);
    c->dy_degen_h = degen;
    std::vector<int64_t> off((size_t)nR * na + 1, 0);
    for (int64_t i = 0; i < nR * na; ++i) off[i + 1] = off[i] + degen[i];
    LELPH_TRY(c->ip_ws.ensure(sizeof(int) * wsi.size()));
    LELPH_TRY(c->ip_off.ensure(sizeof(int64_t) * off.size()));
    HIP_TRY(hipMemcpy(c->ip_ws.p, wsi.data(), sizeof(int) * wsi.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->ip_off.p, off.data(), sizeof(int64_t) * off.size(),
                      hipMemcpyHostToDevice));
    const size_t per = sizeof(z64) * (size_t)L.nmodes * L.nmag * c->nf;
    c->ip_dvR.release();
    LELPH_TRY(c->ip_dvq.ensure(per * c->ip_nq));
    c->ip_have.assign((size_t)c->ip_nq, 0);
    c->ip_on = true;
    return LELPH_OK;
}

// slot of a coarse q in the grid (the order of Sorted_qpts_idxs)
static int ip_slot(const lelph_ctx* c, const double* q, int64_t* slot)
{
    int64_t id[3];
    for (int i = 0; i < 3; ++i)
    {
        const double v = (q[i] - std::floor(q[i])) * c->ip_grid[i];
        const double r = std::rint(v);
        if (std::fabs(v - r) > 1e-4) return LELPH_ERR_ARG;
        id[i] = ((int64_t)r) % c->ip_grid[i];
    }
    *slot = (id[0] * c->ip_grid[1] + id[1]) * c->ip_grid[2] + id[2];
    return LELPH_OK;
}

// basis change with eig * M (mass_normalize_pol_vecs power +1), e^{iqr},
// into the slot of q
static int ip_store(lelph_ctx* c, const double* q, const double* eig, const z64* in)
{
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    int64_t slot;
    LELPH_TRY(ip_slot(c, q, &slot));
    if (c->ip_have[slot]) return LELPH_ERR_ARG;
    std::vector<z64> rot((size_t)nm * nm);
    const z64* e = reinterpret_cast<const z64*>(eig);
    for (int v = 0; v < nm; ++v)
        for (int j = 0; j < nm; ++j)
        {
            const double m = c->ip_mass.empty() ? 1.0 : c->ip_mass[j / 3];
            rot[(size_t)v * nm + j] = cscale(e[(size_t)v * nm + j], m);
        }
    LELPH_TRY(c->ip_rot.ensure(sizeof(z64) * nm * nm));
    HIP_TRY(hipMemcpyAsync(c->ip_rot.p, rot.data(), sizeof(z64) * nm * nm, hipMemcpyHostToDevice,
                           c->stream));
    const int64_t ld = (int64_t)L.nmag * c->nf;
    hipLaunchKernelGGL(k_coarse_cart, dim3((unsigned)cdiv(ld, 256)), dim3(256), 0, c->stream, in,
                       c->ip_rot.as<z64>(), nm, L.nmag, (int)L.fft_dims[0], (int)L.fft_dims[1],
                       (int)L.fft_dims[2], q[0], q[1], q[2],
                       c->ip_dvq.as<z64>() + slot * nm * ld);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ip_have[slot] = 1;
    return LELPH_OK;
}

int lelph_interp_add_coarse(lelph_ctx* c, const double qpt[3], const double* eigVec,
                            const double* dvscf)
{
    if (!c || !qpt || !eigVec || !dvscf) return LELPH_ERR_ARG;
    if (!c->ip_on || c->ip_ready) return LELPH_ERR_STATE;
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    int64_t slot;
    LELPH_TRY(ip_slot(c, qpt, &slot));
    if (c->ip_have[slot]) return LELPH_ERR_ARG;
    const size_t bytes = sizeof(z64) * (size_t)nm * L.nmag * c->nf;
    LELPH_TRY(c->ip_last.ensure(bytes));
    LELPH_TRY(c->ip_eig.ensure(sizeof(z64) * nm * nm));
    c->ip_last_ok = false;
    HIP_TRY(hipMemcpyAsync(c->ip_last.p, dvscf, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->ip_eig.p, eigVec, sizeof(z64) * nm * nm, hipMemcpyHostToDevice,
                           c->stream));
    if (c->ip_has_lr)
        LELPH_TRY(lr_apply(c, c->ip_lr, qpt, 1, c->ip_eig.as<z64>(), c->ip_last.as<z64>(), -1.0));
    LELPH_TRY(ip_store(c, qpt, eigVec, c->ip_last.as<z64>()));
    LELPH_TRY(check_err(c));
    std::memcpy(c->ip_last_q, qpt, sizeof(c->ip_last_q));
    c->ip_last_ok = true;
    return LELPH_OK;
}

int lelph_interp_add_star(lelph_ctx* c, const lelph_symmetry* sym, const double qpt_Sq[3],
                          const double* eig_Sq)
{
    if (!c || !sym || !qpt_Sq || !eig_Sq) return LELPH_ERR_ARG;
    if (!c->ip_on || c->ip_ready || !c->ip_last_ok) return LELPH_ERR_STATE;
    const lelph_lattice& L = c->lat;
    const size_t bytes = sizeof(z64) * (size_t)L.nmodes * L.nmag * c->nf;
    LELPH_TRY(c->ip_tmp.ensure(bytes));
    // QE stores [V, Bx, By, Bz] (interpolation_driver.c:264-271)
    LELPH_TRY(rotate_dev(c, c->ip_last_q, sym, 0, c->ip_last.as<z64>(), c->ip_tmp.as<z64>()));
    return ip_store(c, qpt_Sq, eig_Sq, c->ip_tmp.as<z64>());
}

int lelph_interp_finish(lelph_ctx* c)
{
    if (!c) return LELPH_ERR_ARG;
    if (!c->ip_on || c->ip_ready) return LELPH_ERR_STATE;
    for (char h : c->ip_have)
        if (!h) return LELPH_ERR_STATE;
    const lelph_lattice& L = c->lat;
    const int64_t nq = c->ip_nq;
    const int64_t cols = (int64_t)L.nmodes * L.nmag * c->nf;
    LELPH_TRY(c->ip_P.ensure(sizeof(z64) * nq * nq));
    LELPH_TRY(c->ip_dvR.ensure(sizeof(z64) * nq * cols));
    hipLaunchKernelGGL(k_q2r_phases, dim3((unsigned)cdiv(nq * nq, 256)), dim3(256), 0, c->stream,
                       (int)c->ip_grid[0], (int)c->ip_grid[1], (int)c->ip_grid[2],
                       c->ip_P.as<z64>());
    HIP_TRY(hipGetLastError());
    constexpr int MB = 16;
    hipLaunchKernelGGL((k_phase_gemm<MB>), dim3((unsigned)cdiv(cols, 256), (unsigned)cdiv(nq, MB)),
                       dim3(256), 0, c->stream, c->ip_P.as<z64>(), (int)nq, (int)nq, 1,
                       c->ip_dvq.as<z64>(), 1, cols, c->ip_dvR.as<z64>(), nullptr, 0, 0, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ip_dvq.release();
    c->ip_ready = true;
    return LELPH_OK;
}

// R -> q for nq points into out (device, (nq, nm, nmag, nf))
static int ip_r2q(lelph_ctx* c, int64_t nq, const double* qs, z64* out)
{
    const lelph_lattice& L = c->lat;
    const int na = L.natom;
    const int64_t nR = c->ip_nq;
    const int64_t cpa = (int64_t)3 * L.nmag * c->nf;
    LELPH_TRY(c->ip_qs.ensure(sizeof(double) * 3 * nq));
    LELPH_TRY(c->ip_W.ensure(sizeof(z64) * nq * nR * na));
    HIP_TRY(hipMemcpyAsync(c->ip_qs.p, qs, sizeof(double) * 3 * nq, hipMemcpyHostToDevice,
                           c->stream));
    for (auto& e : c->ip_ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(c->ip_ev[0], c->stream));
    hipLaunchKernelGGL(k_ws_weights, dim3((unsigned)cdiv(nq * nR * na, 256)), dim3(256), 0,
                       c->stream, c->ip_qs.as<double>(), (int)nq, (int)c->ip_grid[0],
                       (int)c->ip_grid[1], (int)c->ip_grid[2], na, c->ip_off.as<int64_t>(),
                       c->ip_ws.as<int>(), c->ip_W.as<z64>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ip_ev[1], c->stream));
    constexpr int MB = 16;
    hipLaunchKernelGGL((k_phase_gemm<MB>),
                       dim3((unsigned)(na * cdiv(cpa, 256)), (unsigned)cdiv(nq, MB)), dim3(256), 0,
                       c->stream, c->ip_W.as<z64>(), (int)nq, (int)nR, na, c->ip_dvR.as<z64>(), na,
                       cpa, out, c->ip_qs.as<double>(), (int)L.fft_dims[0], (int)L.fft_dims[1],
                       (int)L.fft_dims[2]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ip_ev[2], c->stream));
    if (c->ip_has_lr) LELPH_TRY(lr_apply(c, c->ip_lr, qs, nq, nullptr, out, 1.0));
    HIP_TRY(hipEventRecord(c->ip_ev[3], c->stream));
    HIP_TRY(hipEventSynchronize(c->ip_ev[3]));
    float t[3];
    for (int i = 0; i < 3; ++i) HIP_TRY(hipEventElapsedTime(&t[i], c->ip_ev[i], c->ip_ev[i + 1]));
    c->ip_ms[0] += t[0];
    c->ip_ms[1] += t[1];
    c->ip_ms[2] += t[2];
    c->ip_ms[3] += t[0] + t[1] + t[2];
    c->ip_ms[4] += (double)nq;
    return LELPH_OK;
}

int lelph_interp_dvscf(lelph_ctx* c, int64_t nq, const double* qpts, double* dvscf_out)
{
    if (!c || nq < 0 || (nq && (!qpts || !dvscf_out))) return LELPH_ERR_ARG;
    if (!c->ip_ready) return LELPH_ERR_STATE;
    const lelph_lattice& L = c->lat;
    const size_t per = sizeof(z64) * (size_t)L.nmodes * L.nmag * c->nf;
    for (double& t : c->ip_ms) t = 0.0;
    // chunks of at most 2 GiB of output and 65535 grid rows
    int64_t chunk = std::max<int64_t>(1, ((int64_t)2 << 30) / (int64_t)per);
    chunk = std::min<int64_t>(chunk, 65535);
    LELPH_TRY(c->ip_out.ensure(per * std::min(chunk, std::max<int64_t>(nq, 1))));
    for (int64_t q0 = 0; q0 < nq; q0 += chunk)
    {
        const int64_t n = std::min(chunk, nq - q0);
        LELPH_TRY(ip_r2q(c, n, qpts + 3 * q0, c->ip_out.as<z64>()));
        HIP_TRY(hipMemcpyAsync(reinterpret_cast<char*>(dvscf_out) + per * q0, c->ip_out.p,
                               per * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return check_err(c);
}

int lelph_interp_set_dvscf(lelph_ctx* c, const double qpt[3], const double* eigVec, int kernel)
{
    if (!c || !qpt || !eigVec) return LELPH_ERR_ARG;
    if (kernel < LELPH_KERNEL_DFPT || kernel > LELPH_KERNEL_BARE_LOCAL) return LELPH_ERR_ARG;
    if (!c->ip_ready) return LELPH_ERR_STATE;
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    const int64_t ld = (int64_t)L.nmag * c->nf;
    const size_t bytes = sizeof(z64) * (size_t)nm * ld;
    c->dv_set = false;
    c->dv32_valid = false;
    LELPH_TRY(set_eig(c, eigVec));
    LELPH_TRY(c->dv64.ensure(bytes));
    if (kernel == LELPH_KERNEL_DFPT || kernel == LELPH_KERNEL_DFPT_LOCAL)
    {
        LELPH_TRY(c->ip_out.ensure(bytes));
        LELPH_TRY(ip_r2q(c, 1, qpt, c->ip_out.as<z64>()));
        // identity patterns: eig @ I^H = eig (read_dvscf.c:121-127)
        hipLaunchKernelGGL(k_mode_rotate, dim3((unsigned)cdiv((int64_t)nm * ld, 256)), dim3(256),
                           0, c->stream, c->ip_out.as<z64>(), c->eig_d.as<z64>(), nm, ld,
                           c->dv64.as<z64>());
        HIP_TRY(hipGetLastError());
    }
    else
        HIP_TRY(hipMemsetAsync(c->dv64.p, 0, bytes, c->stream));
    return finish_dv(c, qpt);
}

/* ------------------------------------------------------------------------ */
/* QE dvscf files (src/io/qe/read_dvscf.c, write_dyn_dvscf_qe.c)             */
/* ------------------------------------------------------------------------ */

// file sets (Nz, Ny, Nx) -> dst (device, (nsets, Nx, Ny, Nz)); pinned
// staging of up to 64 MiB per read, transposed on the device
static int read_qe_file(lelph_ctx* c, const char* path, z64* dst)
{
    const lelph_lattice& L = c->lat;
    const int64_t nf = c->nf, nsets = (int64_t)L.nmodes * L.nmag;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return LELPH_ERR_ARG;
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz < 0 || (int64_t)sz != nsets * nf * (int64_t)sizeof(z64))
    {
        std::fclose(fp);
        return LELPH_ERR_ARG;
    }
    const int64_t per = nf * (int64_t)sizeof(z64);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nsets, ((int64_t)64 << 20) / per));
    void* stage = nullptr;
    if (hipHostMalloc(&stage, (size_t)(per * chunk)) != hipSuccess)
    {
        std::fclose(fp);
        return LELPH_ERR_ALLOC;
    }
    int st = c->ip_W.ensure((size_t)(per * chunk));
    const unsigned tiles = (unsigned)(cdiv(L.fft_dims[0], 32) * cdiv(L.fft_dims[2], 32));
    for (int64_t s0 = 0; s0 < nsets && st == LELPH_OK; s0 += chunk)
    {
        const int64_t n = std::min(chunk, nsets - s0);
        if (std::fread(stage, (size_t)per, (size_t)n, fp) != (size_t)n)
        {
            st = LELPH_ERR_ARG;
            break;
        }
        if (hipMemcpyAsync(c->ip_W.p, stage, (size_t)(per * n), hipMemcpyHostToDevice,
                           c->stream) != hipSuccess)
        {
            st = LELPH_ERR_HIP;
            break;
        }
        hipLaunchKernelGGL(k_qe_transpose<true>, dim3(tiles, (unsigned)L.fft_dims[1], (unsigned)n),
                           dim3(256), 0, c->stream, c->ip_W.as<z64>(), (int)L.fft_dims[0],
                           (int)L.fft_dims[1], (int)L.fft_dims[2], dst + s0 * nf);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
            st = LELPH_ERR_HIP;
    }
    (void)hipHostFree(stage);
    std::fclose(fp);
    return st;
}

// eig @ pats^H (read_dvscf.c:121-127), FP64 on the host (nm x nm), uploaded
static int upload_eig_pat(lelph_ctx* c, const double* eigVec, const double* patterns)
{
    const int nm = c->lat.nmodes;
    std::vector<z64> ep((size_t)nm * nm);
    const z64* e = reinterpret_cast<const z64*>(eigVec);
    const z64* pat = reinterpret_cast<const z64*>(patterns);
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j)
        {
            z64 acc = {0, 0};
            for (int k = 0; k < nm; ++k)
                acc = cadd(acc, cmul(e[(size_t)i * nm + k], cconj(pat[(size_t)j * nm + k])));
            ep[(size_t)i * nm + j] = acc;
        }
    LELPH_TRY(c->ip_rot.ensure(sizeof(z64) * nm * nm));
    HIP_TRY(hipMemcpyAsync(c->ip_rot.p, ep.data(), sizeof(z64) * nm * nm, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LELPH_OK;
}

int lelph_read_dvscf_qe(lelph_ctx* c, const char* dvscf_file, const double* eigVec,
                        const double* patterns, double* dvscf_out)
{
    if (!c || !dvscf_file || !eigVec || !patterns || !dvscf_out) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    const int64_t ld = (int64_t)L.nmag * c->nf;
    const size_t bytes = sizeof(z64) * (size_t)nm * ld;
    LELPH_TRY(c->ip_tmp.ensure(bytes));
    LELPH_TRY(c->rot_tmp.ensure(bytes));
    LELPH_TRY(read_qe_file(c, dvscf_file, c->ip_tmp.as<z64>()));
    LELPH_TRY(upload_eig_pat(c, eigVec, patterns));
    hipLaunchKernelGGL(k_mode_rotate, dim3((unsigned)cdiv((int64_t)nm * ld, 256)), dim3(256), 0,
                       c->stream, c->ip_tmp.as<z64>(), c->ip_rot.as<z64>(), nm, ld,
                       c->rot_tmp.as<z64>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dvscf_out, c->rot_tmp.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LELPH_OK;
}

int lelph_set_dvscf_file(lelph_ctx* c, const double qpt[3], const double* eigVec,
                         const char* dvscf_file, const double* patterns, int kernel)
{
    if (!c || !qpt || !eigVec || !dvscf_file || !patterns) return LELPH_ERR_ARG;
    if (kernel != LELPH_KERNEL_DFPT && kernel != LELPH_KERNEL_DFPT_LOCAL) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int nm = L.nmodes;
    const int64_t ld = (int64_t)L.nmag * c->nf;
    const size_t bytes = sizeof(z64) * (size_t)nm * ld;
    c->dv_set = false;
    c->dv32_valid = false;
    LELPH_TRY(set_eig(c, eigVec));
    LELPH_TRY(c->dv64.ensure(bytes));
    LELPH_TRY(c->ip_tmp.ensure(bytes));
    LELPH_TRY(read_qe_file(c, dvscf_file, c->ip_tmp.as<z64>()));
    LELPH_TRY(upload_eig_pat(c, eigVec, patterns));
    hipLaunchKernelGGL(k_mode_rotate, dim3((unsigned)cdiv((int64_t)nm * ld, 256)), dim3(256), 0,
                       c->stream, c->ip_tmp.as<z64>(), c->ip_rot.as<z64>(), nm, ld,
                       c->dv64.as<z64>());
    HIP_TRY(hipGetLastError());
    return finish_dv(c, qpt);
}

int lelph_write_dvscf_qe(lelph_ctx* c, const char* dvscf_file, const double* dvscf_in)
{
    if (!c || !dvscf_file || !dvscf_in) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int64_t nf = c->nf, nsets = (int64_t)L.nmodes * L.nmag;
    const size_t bytes = sizeof(z64) * (size_t)(nsets * nf);
    LELPH_TRY(c->ip_tmp.ensure(bytes));
    LELPH_TRY(c->rot_tmp.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(c->ip_tmp.p, dvscf_in, bytes, hipMemcpyHostToDevice, c->stream));
    const unsigned tiles = (unsigned)(cdiv(L.fft_dims[0], 32) * cdiv(L.fft_dims[2], 32));
    hipLaunchKernelGGL(k_qe_transpose<false>, dim3(tiles, (unsigned)L.fft_dims[1], (unsigned)nsets),
                       dim3(256), 0, c->stream, c->ip_tmp.as<z64>(), (int)L.fft_dims[0],
                       (int)L.fft_dims[1], (int)L.fft_dims[2], c->rot_tmp.as<z64>());
    HIP_TRY(hipGetLastError());
    std::vector<z64> host((size_t)(nsets * nf));
    HIP_TRY(hipMemcpyAsync(host.data(), c->rot_tmp.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    FILE* fp = std::fopen(dvscf_file, "wb");
    if (!fp) return LELPH_ERR_ARG;
    const size_t w = std::fwrite(host.data(), sizeof(z64), host.size(), fp);
    const int cl = std::fclose(fp);
    return (w == host.size() && cl == 0) ? LELPH_OK : LELPH_ERR_ARG;
}

/* ------------------------------------------------------------------------ */
/* Fourier interpolation of dynamical matrices                               */
/* ------------------------------------------------------------------------ */

// long range of dyn (nq, nm, nm) on the device += sign * LR (+ ASR term)
// (add_ph_dyn_long_range, src/phonon/dyn_long_range.c:27-404)
static int dyn_lr_apply(lelph_ctx* c, const double* qs, int64_t nq, int sign, bool use_mass,
                        bool with_asr, z64* dyn)
{
    const lelph_lattice& L = c->lat;
    const int na = L.natom, nm = 3 * na;
    if (c->dy_eps.empty() || (c->dy_zb.empty() && c->dy_qp.empty())) return LELPH_OK;
    if (L.dimension == '2')
        for (int64_t i = 0; i < nq; ++i)
            if (std::fabs(qs[3 * i + 2]) > LELPH_EPS) return LELPH_ERR_ARG;
    DynLrGeom g{};
    std::memcpy(g.blat, L.blat_vec, sizeof(g.blat));
    const double* eps = c->dy_eps.data();
    std::memcpy(g.ea, eps, sizeof(g.ea));
    g.eps22 = eps[8];
    g.zlat = L.alat_vec[8];
    g.eta = c->dy_eta;
    g.d2 = L.dimension == '2';
    double factor = 4.0 * LELPH_RPI * LELPH_E2 / L.volume;
    if (sign < 0) factor = -factor;
    if (g.d2)
    {
        factor *= g.zlat / 2.0;
        for (double& x : g.ea) x = 0.0;
        g.ea[0] = 0.5 * g.zlat * (eps[0] - 1);
        g.ea[1] = 0.5 * g.zlat * eps[1];
        g.ea[3] = 0.5 * g.zlat * eps[3];
        g.ea[4] = 0.5 * g.zlat * (eps[4] - 1);
    }
    g.factor = factor;
    for (int i = 0; i < 3; ++i) g.Gg[i] = (int)c->dy_Gg[i];
    g.GZ = g.d2 ? 1 : g.Gg[2];
    g.natom = na;
    g.has_zb = !c->dy_zb.empty();
    g.has_qp = !c->dy_qp.empty();
    g.use_mass = use_mass ? 1 : 0;
    const int64_t nG = (int64_t)g.Gg[0] * g.Gg[1] * g.GZ;
    const int64_t per_q = nG * nm * (int64_t)sizeof(z64);
    int64_t chunk = std::max<int64_t>(1, ((int64_t)1 << 30) / per_q);
    chunk = std::min<int64_t>({chunk, nq, 65535});
    LELPH_TRY(c->dy_X1.ensure((size_t)(per_q * chunk)));
    LELPH_TRY(c->dy_X2.ensure((size_t)(per_q * chunk)));
    if (g.d2)
    {
        LELPH_TRY(c->dy_P1.ensure((size_t)(per_q * chunk)));
        LELPH_TRY(c->dy_P2.ensure((size_t)(per_q * chunk)));
    }
    LELPH_TRY(c->dy_qs.ensure(sizeof(double) * 3 * nq));
    HIP_TRY(hipMemcpyAsync(c->dy_qs.p, qs, sizeof(double) * 3 * nq, hipMemcpyHostToDevice,
                           c->stream));
    for (int64_t q0 = 0; q0 < nq; q0 += chunk)
    {
        const int64_t n = std::min(chunk, nq - q0);
        hipLaunchKernelGGL(k_dyn_lr_vec, dim3((unsigned)cdiv(nG, 256), (unsigned)n), dim3(256), 0,
                           c->stream, g, c->dy_qs.as<double>() + 3 * q0, c->d_apos,
                           c->dy_zbd.as<double>(), c->dy_qpd.as<double>(),
                           c->dy_massd.as<double>(), c->dy_X1.as<z64>(), c->dy_X2.as<z64>(),
                           g.d2 ? c->dy_P1.as<z64>() : nullptr, g.d2 ? c->dy_P2.as<z64>() : nullptr);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_dyn_lr_acc, dim3((unsigned)n), dim3(256), 0, c->stream,
                           c->dy_X1.as<z64>(), c->dy_X2.as<z64>(),
                           g.d2 ? c->dy_P1.as<z64>() : nullptr, g.d2 ? c->dy_P2.as<z64>() : nullptr,
                           nG, nm, factor, with_asr ? c->dy_asr.as<z64>() : nullptr,
                           sign < 0 ? 1.0 : -1.0, c->dy_massd.as<double>(), g.use_mass,
                           dyn + q0 * nm * nm);
        HIP_TRY(hipGetLastError());
    }
    return LELPH_OK;
}

int lelph_dyn_interp_begin(lelph_ctx* c, const int64_t qgrid[3], const double* masses,
                           const lelph_longrange* lr, double eta_ph, double EcutRy, int asr)
{
    if (!c || !qgrid || eta_ph <= 0 || asr < LELPH_ASR_NONE || asr > LELPH_ASR_ALL_HUANG)
        return LELPH_ERR_ARG;
    for (int i = 0; i < 3; ++i)
        if (qgrid[i] < 1 || qgrid[i] > 1024) return LELPH_ERR_ARG;
    const lelph_lattice& L = c->lat;
    const int na = L.natom, nm = 3 * na;
    if (nm > DYN_MAX_NM) return LELPH_ERR_UNSUPPORTED;
    if (lr && lr->Qpole && !lr->epsilon) return LELPH_ERR_ARG;
    c->dy_on = c->dy_ready = false;
    std::memcpy(c->dy_grid, qgrid, sizeof(c->dy_grid));
    c->dy_nq = qgrid[0] * qgrid[1] * qgrid[2];
    c->dy_eta = eta_ph;
    c->dy_asr_kind = asr;
    c->dy_mass.assign(masses ? masses : nullptr, masses ? masses + na : nullptr);
    auto cp = [](std::vector<double>& v, const double* p, size_t n) {
        v.assign(p ? p : nullptr, p ? p + n : nullptr);
    };
    cp(c->dy_eps, lr ? lr->epsilon : nullptr, 9);
    cp(c->dy_zb, lr ? lr->Zborn : nullptr, (size_t)9 * na);
    cp(c->dy_qp, lr ? lr->Qpole : nullptr, (size_t)27 * na);
    lr_fft_box(EcutRy, L.blat_vec, c->dy_Gg); // Ggrid_phonon (interpolation_driver.c:420-421)
    auto up = [&](DBuf& b, const std::vector<double>& v) -> int {
        if (v.empty()) return LELPH_OK;
        LELPH_TRY(b.ensure(sizeof(double) * v.size()));
        HIP_TRY(hipMemcpy(b.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        return LELPH_OK;
    };
    LELPH_TRY(up(c->dy_zbd, c->dy_zb));
    LELPH_TRY(up(c->dy_qpd, c->dy_qp));
    LELPH_TRY(up(c->dy_massd, c->dy_mass));
    // Wigner-Seitz vectors of every (R, atom, atom) (interpolation_driver.c:204-207)
    const int64_t nR = c->dy_nq;
    std::vector<int64_t> degen((size_t)nR * na * na);
    int64_t cap = 16 * nR * na * na, nws = 0;
    std::vector<int64_t> ws;
    for (;;)
    {
        ws.resize((size_t)3 * cap);
        const int st = lelph_wigner_seitz(qgrid, L.alat_vec, LELPH_EPS6, c->atomic_pos.data(), na,
                                          c->atomic_pos.data(), na, cap, ws.data(), degen.data(),
                                          &nws);
        if (st == LELPH_OK) break;
        if (st != LELPH_ERR_RANGE) return st;
        cap *= 4;
    }
    std::vector<int> wsi((size_t)3 * std::max<int64_t>(nws, 1));
    for (int64_t i = 0; i < 3 * nws; ++i) wsi[i] = (int)ws[i];
    c->dy_ws_h.assign(ws.begin(), ws.begin() + 3 * nws);
    c->dy_degen_h = degen;
    std::vector<int64_t> off((size_t)nR * na * na + 1, 0);
    for (int64_t i = 0; i < nR * na * na; ++i) off[i + 1] = off[i] + degen[i];
    LELPH_TRY(c->dy_ws.ensure(sizeof(int) * wsi.size()));
    LELPH_TRY(c->dy_off.ensure(sizeof(int64_t) * off.size()));
    HIP_TRY(hipMemcpy(c->dy_ws.p, wsi.data(), sizeof(int) * wsi.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->dy_off.p, off.data(), sizeof(int64_t) * off.size(),
                      hipMemcpyHostToDevice));
    LELPH_TRY(c->dy_dyn.ensure(sizeof(z64) * c->dy_nq * nm * nm));
    // the rotational rules act on the full force constants: the long-range
    // part of each coarse matrix is kept to be added in R space and removed
    // again (interpolation_driver.c:432-442,499-511,533-545)
    const bool keep_lr = asr >= LELPH_ASR_ALL && !c->dy_eps.empty() &&
                         (!c->dy_zb.empty() || !c->dy_qp.empty());
    c->dy_keep_lr = keep_lr;
    if (keep_lr)
    {
        LELPH_TRY(c->dy_lr.ensure(sizeof(z64) * c->dy_nq * nm * nm));
        HIP_TRY(hipMemsetAsync(c->dy_lr.p, 0, sizeof(z64) * c->dy_nq * nm * nm, c->stream));
    }
    // long-range ASR correction (compute_dyn_lr_asr_correction,
    // dyn_long_range.c:70-123): the q = 0 term without masses summed over b
    std::vector<z64> asrv((size_t)9 * na, z64{0, 0});
    if (!c->dy_eps.empty() && (!c->dy_zb.empty() || !c->dy_qp.empty()))
    {
        LELPH_TRY(c->dy_out.ensure(sizeof(z64) * nm * nm));
        HIP_TRY(hipMemsetAsync(c->dy_out.p, 0, sizeof(z64) * nm * nm, c->stream));
        const double q0[3] = {0, 0, 0};
        LELPH_TRY(dyn_lr_apply(c, q0, 1, 1, false, false, c->dy_out.as<z64>()));
        std::vector<z64> t((size_t)nm * nm);
        HIP_TRY(hipMemcpyAsync(t.data(), c->dy_out.p, sizeof(z64) * nm * nm,
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < na; ++b)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        asrv[(size_t)a * 9 + 3 * i + j] =
                            cadd(asrv[(size_t)a * 9 + 3 * i + j], t[(size_t)(3 * a + i) * nm + 3 * b + j]);
    }
    LELPH_TRY(c->dy_asr.ensure(sizeof(z64) * 9 * na));
    HIP_TRY(hipMemcpy(c->dy_asr.p, asrv.data(), sizeof(z64) * 9 * na, hipMemcpyHostToDevice));
    c->dy_have.assign((size_t)c->dy_nq, 0);
    c->dy_on = true;
    return LELPH_OK;
}

int lelph_dyn_interp_add_coarse(lelph_ctx* c, const double qpt[3], const double* omega,
                                const double* pol_vecs)
{
    if (!c || !qpt || !omega || !pol_vecs) return LELPH_ERR_ARG;
    if (!c->dy_on || c->dy_ready) return LELPH_ERR_STATE;
    const int na = c->lat.natom, nm = 3 * na;
    int64_t id[3];
    for (int i = 0; i < 3; ++i)
    {
        const double v = (qpt[i] - std::floor(qpt[i])) * c->dy_grid[i];
        const double r = std::rint(v);
        if (std::fabs(v - r) > 1e-4) return LELPH_ERR_ARG;
        id[i] = ((int64_t)r) % c->dy_grid[i];
    }
    const int64_t slot = (id[0] * c->dy_grid[1] + id[1]) * c->dy_grid[2] + id[2];
    if (c->dy_have[slot]) return LELPH_ERR_ARG;
    // pol_vecs_to_dyn (src/phonon/pol_vecs_to_dyn.c:8-40), nm x nm on the host
    const z64* pv = reinterpret_cast<const z64*>(pol_vecs);
    std::vector<z64> e((size_t)nm * nm), d((size_t)nm * nm);
    for (int v = 0; v < nm; ++v)
        for (int j = 0; j < nm; ++j)
            e[(size_t)v * nm + j] =
                cscale(pv[(size_t)v * nm + j], c->dy_mass.empty() ? 1.0 : std::pow(c->dy_mass[j / 3], 0.5));
    for (int a = 0; a < nm; ++a)
        for (int b = 0; b < nm; ++b)
        {
            z64 acc = {0, 0};
            for (int i = 0; i < nm; ++i)
            {
                const double f = (omega[i] > 0 ? 1.0 : -1.0) * omega[i] * omega[i];
                acc = cadd(acc, cmul(e[(size_t)i * nm + a], cscale(cconj(e[(size_t)i * nm + b]), f)));
            }
            d[(size_t)a * nm + b] = acc;
        }
    z64* slot_p = c->dy_dyn.as<z64>() + slot * nm * nm;
    HIP_TRY(hipMemcpyAsync(slot_p, d.data(), sizeof(z64) * nm * nm, hipMemcpyHostToDevice,
                           c->stream));
    LELPH_TRY(dyn_lr_apply(c, qpt, 1, -1, !c->dy_mass.empty(), true, slot_p));
    if (c->dy_keep_lr)
        LELPH_TRY(dyn_lr_apply(c, qpt, 1, +1, !c->dy_mass.empty(), true,
                               c->dy_lr.as<z64>() + slot * nm * nm));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->dy_have[slot] = 1;
    return LELPH_OK;
}

int lelph_dyn_interp_finish(lelph_ctx* c)
{
    if (!c) return LELPH_ERR_ARG;
    if (!c->dy_on || c->dy_ready) return LELPH_ERR_STATE;
    for (char h : c->dy_have)
        if (!h) return LELPH_ERR_STATE;
    const int na = c->lat.natom, nm = 3 * na;
    const int64_t nq = c->dy_nq, cols = (int64_t)nm * nm;
    LELPH_TRY(c->dy_P.ensure(sizeof(z64) * nq * nq));
    LELPH_TRY(c->dy_fc.ensure(sizeof(z64) * nq * cols));
    hipLaunchKernelGGL(k_q2r_phases, dim3((unsigned)cdiv(nq * nq, 256)), dim3(256), 0, c->stream,
                       (int)c->dy_grid[0], (int)c->dy_grid[1], (int)c->dy_grid[2],
                       c->dy_P.as<z64>());
    HIP_TRY(hipGetLastError());
    constexpr int MB = 16;
    hipLaunchKernelGGL((k_phase_gemm<MB>), dim3((unsigned)cdiv(cols, 256), (unsigned)cdiv(nq, MB)),
                       dim3(256), 0, c->stream, c->dy_P.as<z64>(), (int)nq, (int)nq, 1,
                       c->dy_dyn.as<z64>(), 1, cols, c->dy_fc.as<z64>(), nullptr, 0, 0, 0);
    HIP_TRY(hipGetLastError());
    if (c->dy_asr_kind <= LELPH_ASR_SIMPLE)
    {
        hipLaunchKernelGGL(k_fc_asr, dim3((unsigned)cdiv(9 * na, 64)), dim3(64), 0, c->stream,
                           c->dy_fc.as<z64>(), (int)nq, na, c->dy_massd.as<double>(),
                           c->dy_mass.empty() ? 0 : 1, c->dy_asr_kind);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->dy_ready = true;
        return LELPH_OK;
    }
    // projected rules (apply_acoustic_sum_rule_fc, src/phonon/asr.c:320-512):
    // host code on the force constants without mass factors
    const bool lr = c->dy_keep_lr;
    std::vector<z64> fc((size_t)(nq * cols)), fl;
    if (lr)
    {
        LELPH_TRY(c->dy_fclr.ensure(sizeof(z64) * nq * cols));
        hipLaunchKernelGGL((k_phase_gemm<MB>),
                           dim3((unsigned)cdiv(cols, 256), (unsigned)cdiv(nq, MB)), dim3(256), 0,
                           c->stream, c->dy_P.as<z64>(), (int)nq, (int)nq, 1, c->dy_lr.as<z64>(), 1,
                           cols, c->dy_fclr.as<z64>(), nullptr, 0, 0, 0);
        HIP_TRY(hipGetLastError());
        fl.resize((size_t)(nq * cols));
        HIP_TRY(hipMemcpyAsync(fl.data(), c->dy_fclr.p, sizeof(z64) * nq * cols,
                               hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(fc.data(), c->dy_fc.p, sizeof(z64) * nq * cols, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (lr)
        for (size_t i = 0; i < fc.size(); ++i) fc[i] = cadd(fc[i], fl[i]);
    // mass_normalize_force_constants (src/phonon/mass_normalize_phonons.c:36-71)
    auto mass_scale = [&](std::vector<z64>& f, double power) {
        if (c->dy_mass.empty()) return;
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < na; ++b)
            {
                const double m = std::pow(c->dy_mass[a], power) * std::pow(c->dy_mass[b], power);
                for (int64_t r = 0; r < nq; ++r)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                        {
                            z64& v = f[(size_t)(r * cols + (int64_t)(3 * a + i) * nm + 3 * b + j)];
                            v = cscale(v, m);
                        }
            }
    };
    mass_scale(fc, 0.5);
    LELPH_TRY(lelph_asr_fc(c->dy_asr_kind, c->dy_grid, na, reinterpret_cast<double*>(fc.data()),
                           c->atomic_pos.data(), c->lat.alat_vec, c->dy_ws_h.data(),
                           (int64_t)c->dy_ws_h.size() / 3, c->dy_degen_h.data()));
    mass_scale(fc, -0.5);
    if (lr)
        for (size_t i = 0; i < fc.size(); ++i) fc[i] = {fc[i].x - fl[i].x, fc[i].y - fl[i].y};
    HIP_TRY(hipMemcpy(c->dy_fc.p, fc.data(), sizeof(z64) * nq * cols, hipMemcpyHostToDevice));
    c->dy_ready = true;
    return LELPH_OK;
}

int lelph_dyn_interp(lelph_ctx* c, int64_t nq, const double* qpts, double* dyn_out,
                     double* omega_out, double* eig_out)
{
    if (!c || nq < 0 || (nq && !qpts)) return LELPH_ERR_ARG;
    if (!c->dy_ready) return LELPH_ERR_STATE;
    const int na = c->lat.natom, nm = 3 * na;
    const int64_t nn = (int64_t)nm * nm;
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(nq, 1), 65535);
    LELPH_TRY(c->dy_out.ensure(sizeof(z64) * nn * chunk));
    LELPH_TRY(c->dy_eig.ensure(sizeof(z64) * nn * chunk));
    LELPH_TRY(c->dy_freq.ensure(sizeof(double) * nm * chunk));
    LELPH_TRY(c->dy_qs2.ensure(sizeof(double) * 3 * chunk));
    for (int64_t q0 = 0; q0 < nq; q0 += chunk)
    {
        const int64_t n = std::min(chunk, nq - q0);
        HIP_TRY(hipMemcpyAsync(c->dy_qs2.p, qpts + 3 * q0, sizeof(double) * 3 * n,
                               hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_dyn_r2q, dim3((unsigned)cdiv(n * nn, 256)), dim3(256), 0, c->stream,
                           c->dy_qs2.as<double>(), (int)n, (int)c->dy_grid[0], (int)c->dy_grid[1],
                           (int)c->dy_grid[2], na, c->dy_off.as<int64_t>(), c->dy_ws.as<int>(),
                           c->dy_fc.as<z64>(), c->dy_out.as<z64>());
        HIP_TRY(hipGetLastError());
        LELPH_TRY(dyn_lr_apply(c, qpts + 3 * q0, n, 1, !c->dy_mass.empty(), true,
                               c->dy_out.as<z64>()));
        if (dyn_out)
            HIP_TRY(hipMemcpyAsync(dyn_out + 2 * q0 * nn, c->dy_out.p, sizeof(z64) * nn * n,
                                   hipMemcpyDeviceToHost, c->stream));
        if (omega_out || eig_out)
        {
            hipLaunchKernelGGL(k_herm_jacobi, dim3((unsigned)n), dim3(64), 0, c->stream,
                               c->dy_out.as<z64>(), nm, c->dy_massd.as<double>(),
                               c->dy_mass.empty() ? 0 : 1, c->dy_freq.as<double>(),
                               c->dy_eig.as<z64>(), 40);
            HIP_TRY(hipGetLastError());
            if (omega_out)
                HIP_TRY(hipMemcpyAsync(omega_out + q0 * nm, c->dy_freq.p, sizeof(double) * nm * n,
                                       hipMemcpyDeviceToHost, c->stream));
            if (eig_out)
                HIP_TRY(hipMemcpyAsync(eig_out + 2 * q0 * nn, c->dy_eig.p, sizeof(z64) * nn * n,
                                       hipMemcpyDeviceToHost, c->stream));
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return LELPH_OK;
}

int lelph_qe_parse_dyn(const char* path, int natom, double* qpt, double* masses, double* dynd);

int lelph_read_dyn_qe(lelph_ctx* c, const char* dyn_file, double* qpt, double* omega,
                      double* pol_vecs, double* amass)
{
    if (!c || !dyn_file || !qpt || !omega || !pol_vecs) return LELPH_ERR_ARG;
    const int na = c->lat.natom, nm = 3 * na;
    if (nm > DYN_MAX_NM) return LELPH_ERR_UNSUPPORTED;
    std::vector<double> m((size_t)na), d((size_t)2 * nm * nm);
    LELPH_TRY(lelph_qe_parse_dyn(dyn_file, na, qpt, m.data(), d.data()));
    LELPH_TRY(c->dy_out.ensure(sizeof(z64) * nm * nm));
    LELPH_TRY(c->dy_eig.ensure(sizeof(z64) * nm * nm));
    LELPH_TRY(c->dy_freq.ensure(sizeof(double) * nm));
    LELPH_TRY(c->dy_rmass.ensure(sizeof(double) * na));
    HIP_TRY(hipMemcpyAsync(c->dy_out.p, d.data(), sizeof(z64) * nm * nm, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->dy_rmass.p, m.data(), sizeof(double) * na, hipMemcpyHostToDevice,
                           c->stream));
    // zheev of the Hermitian part (read_dyn.c:236-271) on the device
    hipLaunchKernelGGL(k_herm_jacobi, dim3(1), dim3(64), 0, c->stream, c->dy_out.as<z64>(), nm,
                       c->dy_rmass.as<double>(), 1, c->dy_freq.as<double>(), c->dy_eig.as<z64>(),
                       40);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(omega, c->dy_freq.p, sizeof(double) * nm, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(pol_vecs, c->dy_eig.p, sizeof(z64) * nm * nm, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (amass) std::memcpy(amass, m.data(), sizeof(double) * na);
    return LELPH_OK;
}

int lelph_internal_grid(const lelph_ctx* c, int64_t* nbz, const double** kcrys, const int** kmap)
{
    if (!c || !nbz || !kcrys || !kmap) return LELPH_ERR_ARG;
    if (c->kpt_bz.empty()) return LELPH_ERR_STATE;
    *nbz = c->lat.nkpts_BZ;
    *kcrys = c->kpt_bz.data();
    *kmap = c->kmap.data();
    return LELPH_OK;
}

int lelph_interp_timings(lelph_ctx* c, double* ms, int n)
{
    if (!c || !ms || n < 0) return LELPH_ERR_ARG;
    for (int i = 0; i < n && i < 5; ++i) ms[i] = c->ip_ms[i];
    return LELPH_OK;
}

int lelph_last_timings(lelph_ctx* c, double* ms, int n)
{
    if (!c || !ms || n < 0) return LELPH_ERR_ARG;
    for (int i = 0; i < n && i < 8; ++i) ms[i] = c->timings[i];
    return LELPH_OK;
}

} // extern "C"
