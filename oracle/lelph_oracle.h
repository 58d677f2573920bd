/*
 * lelph_oracle.h -- FP64 CPU restatement of LetzElPhC's electron-phonon hot
 * path.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the
 * bench's cpu_baseline leg as the checker; never by the product path.
 *
 * Uses the structs of include/lelph_hip.h (the boundary types).  Every
 * function names the reference file:line it restates.
 */
#ifndef LELPH_ORACLE_H
#define LELPH_ORACLE_H

#include <stdint.h>

#include "../include/lelph_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* one iBZ wavefunction, struct WFC (src/common/dtypes.h:252-266) */
typedef struct orc_wfc
{
    const double* wfc;  /* (nspin,nbnds,nspinor,npw) complex interleaved */
    const double* gvec; /* (npw,3) cart, no 2pi                        */
    const double* Fk;   /* (nltimesj,ntype,npw) or NULL                 */
    int64_t npw;
} orc_wfc;

/* spline table of V_loc(|G|), struct Vloc_table (dtypes.h:190-210) */
typedef struct orc_vloc_table
{
    int64_t npts;
    double dg;
    double* g_co;  /* (npts)        */
    double* vlocg; /* (ntype,npts)  */
    double* vploc; /* (ntype,npts)  */
} orc_vloc_table;

/* ---- numerical helpers (src/common/numerical_func.c, spline.c) ---- */
double orc_legendre(int l, int m, double x);
double orc_Ylm(int l, int m, const double* vec);
double orc_simpson(const double* f, const double* dx, int64_t n);
void orc_prepare_spline(int64_t n, const double* x, const double* y,
                        double* dy);
double orc_spline_interpolate(double x, int64_t inear, const double* xi,
                              const double* yi, const double* dy);
void orc_reciprocal_vecs(const double* lat, double* blat);
int orc_SU2mat(const double* sym, int nspinor, int invert, int time_rev,
               double* su2 /* 4 complex */);

/* ---- FFT (own mixed radix; FFTW sign conventions) ---- */
/* in-place 3-D DFT of a (n0,n1,n2) row-major box; sign -1 forward, +1
 * backward, unnormalised (FFTW_FORWARD / FFTW_BACKWARD). */
int orc_fft3d(double* box, const int64_t* dims, int sign);
/* 1-D transform of n points (for tests). */
int orc_fft1d(double* data, int64_t n, int sign);

/* ---- V_loc table (src/dvloc/vlocg_table.c, VlocG_atom.c) ---- */
double orc_Vloc_Gspace(double* work, char cutoff, double Gnorm, int64_t ngrid,
                       const double* Vloc, const double* r, const double* rab,
                       double Zval, double eta, double volume);
int orc_create_vlocg_table(const lelph_lattice* lat, const lelph_pseudo* pp,
                           orc_vloc_table* tab);
void orc_free_vloc_table(orc_vloc_table* tab);

/* ---- dVlocq / add_dvscf (src/dvloc/dVbare_loc.c, add_dvscf.c) ---- */
int orc_dVlocq(const double* qpt, const lelph_lattice* lat,
               const lelph_pseudo* pp, const orc_vloc_table* tab,
               const double* eigVec, double* Vlocr);
void orc_add_dvscf_qe(double* dVscf, const double* dVloc,
                      const lelph_lattice* lat);
/* pattern -> mode rotation of src/io/qe/read_dvscf.c:107-158 */
void orc_dvscf_pattern_to_mode(double* dvscf, const double* eig,
                               const double* pats, const lelph_lattice* lat);

/* ---- the hot path ---- */
int orc_elphLocal(const double* qpt, const orc_wfc* wfcs,
                  const lelph_lattice* lat, int ikq, int ik, int kqsym,
                  int ksym, const double* dVlocr, double* elph_kq);
int orc_add_elphNonLocal(const orc_wfc* wfcs, const lelph_lattice* lat,
                         const lelph_pseudo* pp, int ikq, int ik, int kqsym,
                         int ksym, const double* eigVec, double* elph_kq);
void orc_elph_q_rotate(const double* Dl, const double* g, const double* Dr,
                       int nmodes, int nspin, int nb, int tim_rev,
                       double* gSq);
int orc_get_KplusQ_idxs(int64_t nbz, const double* kpts, const double* q,
                        int* idx);
/* the k loop of compute_and_write_elphq; OpenMP over k; g_out
 * (k_count,nmodes,nspin,nb,nb) for full-BZ k in [k_first, k_first+k_count). */
int orc_compute_elphq(const double* qpt, const orc_wfc* wfcs,
                      const lelph_lattice* lat, const lelph_pseudo* pp,
                      const double* eigVec, const double* dVscf, int non_loc,
                      int kminusq, int64_t k_first, int64_t k_count,
                      double* g_out);
/* sets the number of OpenMP threads used by orc_compute_elphq */
void orc_set_threads(int n);
/* threads used inside one orc_elphLocal (bands, GEMM rows) */
void orc_set_inner_threads(int n);
int orc_get_threads(void);

/* static helpers, exported for the parity tests */
void orc_su2rotate(int nspinor, int64_t npw, const double* su2, double* wfc);
void orc_apply_trans_wfc(const double* tr, const double* kvec, int64_t nsets,
                         int64_t npw, const double* gvecs, double* wfc,
                         int conjugate);
void orc_sum_VNL_KKp(const double* K, const double* Kp, const double* fcoeff,
                     int64_t nspin, int64_t nb, int64_t nspinor, double* out);

/* G-vector matching (src/wfc/gsort.c:49-95) */
void orc_find_gvecs_idxs(int64_t nsearch, const double* search, int64_t npw,
                         const double* gvecs, int64_t* out);
/* electronic_reps (src/symmetries/elec_reps.c:16-413): Dkmn (nspin,nb,nb) */
int orc_electronic_reps(const orc_wfc* wfcs, const lelph_lattice* lat,
                        const double* Rsym, const double* tauR, int tim_revR,
                        int64_t ikBZ, double* Dkmn);

/* ---- Fourier interpolation of dV_q (interp_oracle.c) ---- */
void orc_get_fft_box(double EcutRy, const double* blat, int64_t* box);
/* returns the number of vectors, -1 if cap is too small, -2 if none found */
int64_t orc_ws_vectors(const int64_t* grid, const double* lat, double eps,
                       const double* rvec_m, int64_t nrvec_m,
                       const double* rvec_n, int64_t nrvec_n, int64_t cap,
                       int64_t* ws_vecs, int64_t* ws_degen);
void orc_dVlong_range_kernel(const double* qpt, const double* gvecs, int64_t npw,
                             const double* Zvals, const double* eps, const double* Zeu,
                             const double* Qpole, int64_t natom, const double* atom_pos,
                             char dimension, double volume, double zlat, double EcutRy,
                             double eta_bare, double eta_ind, double* out);
int orc_dV_add_longrange(const double* qpt, const lelph_lattice* lat,
                         const lelph_longrange* lr, const double* eigVec, int sign,
                         double* dvscf);
void orc_multiply_eikr(double* pot, const double* q, const int64_t* N, int64_t nsets,
                       int sign);
void orc_dvscf_change_basis(double* dvscf, const double* rot, int64_t nsets, int64_t nm,
                            int64_t nmag, int64_t nf, char op);
int orc_rotate_dvscf(const double* qpt, const double* dvin, const lelph_symmetry* sym,
                     const lelph_lattice* lat, int composite, double* dvout);
void orc_fft_q2R(double* data, const int64_t* qgrid, int64_t nsets);
int orc_fft_R2q_dvscf(const double* dataR, const double* q, const int64_t* qgrid,
                      int64_t natom, int64_t nsets, const int64_t* ws, int64_t nws,
                      const int64_t* degen, double* dataq);
int orc_find_qpt_grid(int64_t nq, const double* q, int64_t* grid);
void orc_sorted_qpts_idxs(int64_t nq, const double* q, int64_t* idx);
/* dynamical matrices (src/phonon/, interpolation_utilities.c:78-163) */
void orc_pol_vecs_to_dyn(const double* omega, int64_t natom, const double* masses, double* pol);
void orc_dyn_lr_asr(const lelph_lattice* lat, const double* eps, const double* Zb,
                    const double* Qp, const int64_t* Gg, double eta, double* out);
int orc_add_ph_dyn_long_range(const double* qpt, const lelph_lattice* lat, const double* eps,
                              const double* Zb, const double* Qp, const int64_t* Gg, int sign,
                              const double* masses, const double* asr, double eta, double* dyn);
int orc_fft_R2q_dyn(const double* dataR, const double* q, const int64_t* qgrid, int64_t natom,
                    const int64_t* ws, int64_t nws, const int64_t* degen, double* dataq);
void orc_mass_normalize_fc(const double* masses, int64_t nsets, int64_t natom, double power,
                           double* frc);
void orc_asr_simple_fc(const int64_t* qgrid, int64_t natom, double* frc);
/* QE dvscf files (src/io/qe/read_dvscf.c, write_dyn_dvscf_qe.c) */
int orc_read_dvscf_qe(const char* path, const lelph_lattice* lat, const double* eig,
                      const double* pats, double* out);
int orc_write_dvscf_qe(const char* path, const lelph_lattice* lat, const double* dvscf);

#ifdef __cplusplus
}
#endif

#endif
