/*
 * lelph_hip.h -- C-ABI of the MI355X (gfx950) electron-phonon matrix-element
 * path  g^nu_mn(k,q) = < psi_m,k+q | dV_q,nu | psi_n,k >.
 *
 * Plain C, no MPI types, no torch types.  Every entry point returns an int
 * status (LELPH_OK == 0); lelph_strerror() turns it into text.  Complex
 * buffers are interleaved (re, im) doubles, i.e. the "re_im" trailing axis of
 * the reference's netCDF variables.
 *
 * The structs mirror the reference's struct Lattice / struct Pseudo / struct
 * WFC (src/common/dtypes.h:60-266) field by field, keeping the reference's
 * names, units and array layouts:
 *   - k points of the irreducible BZ in cartesian units without 2*pi,
 *   - G vectors (npw,3) cartesian without 2*pi,
 *   - blat_vec includes 2*pi (computed with the reference's ELPH_PI),
 *   - wavefunctions (nspin, nbnds, nspinor, npw),
 *   - dV (nmodes, nmag, Nx, Ny, Nz), x slowest, z fastest,
 *   - g (nmodes, nspin, nb_k, nb_kq): row = initial band (k), column = final
 *     band (k+q)  (src/dvloc/compute_Vbare_loc.c:323-332).
 *
 * One context per host thread.  Calls on one context are serialised on the
 * context's own HIP stream.
 */
#ifndef LELPH_HIP_H
#define LELPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- status codes ---------------- */
enum lelph_status
{
    LELPH_OK = 0,
    LELPH_ERR_ARG = 1,         /* bad argument / inconsistent sizes          */
    LELPH_ERR_ALLOC = 2,       /* device or host allocation failed           */
    LELPH_ERR_HIP = 3,         /* a HIP runtime call failed                  */
    LELPH_ERR_RANGE = 4,       /* G vector outside the FFT box, table range  */
    LELPH_ERR_UNSUPPORTED = 5, /* combination not supported by this build    */
    LELPH_ERR_STATE = 6,       /* call order violated (e.g. no dV set)       */
    LELPH_ERR_NOKQ = 7,        /* k+q not on the k grid (incommensurate)     */
    LELPH_ERR_SYM = 8,         /* symmetry matrix is not orthogonal          */
    LELPH_ERR_SCRATCH = 9      /* size exceeds a no-allocation scratch       */
};

/* ---------------- precision of the contractions ---------------- */
enum lelph_precision
{
    LELPH_FP64 = 0,   /* FP64 FFT, FP64 MFMA contraction                      */
    LELPH_FP32 = 1,   /* FP32 FFT (FP64 twiddles), FP32 MFMA, FP64 finalise   */
    LELPH_BF16X3 = 2, /* FP32 FFT, BF16 2-piece split, 3 MFMA products        */
    LELPH_BF16X6 = 3  /* FP32 FFT, BF16 3-piece split, 6 MFMA products        */
};

/* ---------------- kernel flags (src/elph/set_kernel.c:18-52) ------------ */
enum lelph_kernel
{
    LELPH_KERNEL_DFPT = 0,       /* dvscf + bare local + non local          */
    LELPH_KERNEL_DFPT_LOCAL = 1, /* dvscf + bare local                      */
    LELPH_KERNEL_BARE = 2,       /* bare local + non local (no screening)   */
    LELPH_KERNEL_BARE_LOCAL = 3  /* bare local only                         */
};

/* struct symmetry, src/common/dtypes.h:60-68 */
typedef struct lelph_symmetry
{
    double Rmat[9]; /* cartesian rotation, row major                          */
    double tau[3];  /* fractional translation, cartesian                      */
    int time_rev;   /* 1 if the operation includes time reversal              */
} lelph_symmetry;

/* struct Lattice, src/common/dtypes.h:80-140 */
typedef struct lelph_lattice
{
    int natom;
    int nsym;
    int timerev;
    int nspin;
    int nspinor;
    int nbnds;
    int nmodes;
    int nmag;
    char dimension; /* '3', '2' or '1' (Coulomb cut-off in dVlocq)            */
    int is_soc_present;
    int64_t fft_dims[3];
    double alat_vec[9]; /* a[i] = alat_vec[:,i]                               */
    double blat_vec[9]; /* b[i] = blat_vec[:,i], includes 2*pi                */
    double volume;
    const double* atomic_pos; /* (natom,3) cartesian                          */
    const int* atom_type;     /* (natom)                                      */
    int64_t nkpts_iBZ;
    const double* kpt_iredBZ; /* (nkpts_iBZ,3) cartesian, no 2*pi            */
    int64_t nkpts_BZ;
    const double* kpt_fullBZ_crys; /* (nkpts_BZ,3) crystal                   */
    const int* kmap;               /* (nkpts_BZ,2): (ik_iBZ, isym)           */
    const lelph_symmetry* syms;    /* (nsym)                                  */
} lelph_lattice;

/* struct local_pseudo, src/common/dtypes.h:212-224 */
typedef struct lelph_local_pseudo
{
    const double* Vloc_atomic; /* (ngrid) local pseudopotential V(r), Ry    */
    const double* r_grid;      /* (ngrid)                                    */
    const double* rab_grid;    /* (ngrid) dr/di                              */
    int64_t ngrid;
    double Zval;
} lelph_local_pseudo;

/* struct Pseudo, src/common/dtypes.h:226-250 */
typedef struct lelph_pseudo
{
    int ntype;
    int lmax;
    int64_t nltimesj;
    const double* PP_table; /* (nltimesj, ntype, 3): (l+1, 2j, pp_spin)     */
    const double* Fsign;    /* (nltimesj, ntype)                            */
    /* SOC f-coefficients: NULL, or (nltimesj*ntype) pointers, entry
     * [lidx*ntype+itype] -> complex (2l+1, 2l+1, nspinor, nspinor)         */
    const double* const* fCoeff;
    const lelph_local_pseudo* loc_pseudo; /* (ntype)                        */
    double qmax_abs; /* max |q_i| (crystal) the V_loc(G) table must cover    */
} lelph_pseudo;

/* one (k, k+q) work item of a batch: the arguments of elphLocal */
typedef struct lelph_pair
{
    int ikq, ik, kqsym, ksym;
} lelph_pair;

typedef struct lelph_ctx lelph_ctx;

/* ------------------------------------------------------------------------
 * Context.  Owns the device copies of the iBZ wavefunctions, G lists, KB
 * projectors, the V_loc(G) table and the current dV_q.
 * Replaces the allocation done by read_and_alloc_save_data
 * (src/io/get_save_data.c:51-638) for the data the hot path reads.
 * ---------------------------------------------------------------------- */
int lelph_create(const lelph_lattice* lattice, const lelph_pseudo* pseudo,
                 int precision, lelph_ctx** out);
void lelph_destroy(lelph_ctx* ctx);
const char* lelph_strerror(int status);
int lelph_set_precision(lelph_ctx* ctx, int precision);
/* Launch shapes of a context (no counterpart in the reference, whose FFTW
 * plans are measured per call, src/fft/planner.c:226-312): chosen per grid
 * by the project's autotuner (letzelphc_amd/autotune.py).  fft_lines and
 * fp32_form give bit-identical g; split_target moves the boundaries of the
 * fixed-order FP64 sum (g changes at the rounding level).  Zero fields
 * take the defaults. */
typedef struct lelph_tuning
{
    int split_target; /* blocks aimed at by the FP64/FP32 contraction's K
                         splits (default 2048)                               */
    int fft_lines;    /* lines per block of the compiled line FFTs: 8, 16,
                         32 where they fit in LDS, 0 = per-length default    */
    int fp32_form;    /* FP32 contraction: 0 dV and psi_k staged separately,
                         1 element-wise A = dV psi_k                          */
    int ctile;        /* FP64 / FP32 contraction block tile: 0 chosen from
                         the mode and band counts, 1 two modes x 32 x 64,
                         2 four modes x 32 x 32, 3 four modes x 16 x 48,
                         4 four modes x 16 x 16 (bit-identical g)             */
    int nl_form;      /* non-local projector GEMMs of bf16x3 / bf16x6
                         contexts: 0 FP32 operands on FP32 MFMA, 1 bf16
                         pieces (3 pieces, 6 products); changes g at the
                         1e-7 level                                           */
} lelph_tuning;
int lelph_set_tuning(lelph_ctx* ctx, const lelph_tuning* tuning);
int lelph_get_tuning(lelph_ctx* ctx, lelph_tuning* tuning);

/* The precision requested with lelph_create / lelph_set_precision and the
 * one the contractions run in on this context's shape: LELPH_BF16X3 runs as
 * LELPH_BF16X6 unless nspinor = 1, Nx*Ny*Nz is a multiple of 32 and
 * nbnds >= 64 (the shapes whose bf16x3 error is held to 1e-6).  No
 * counterpart in the reference (it has one precision per build,
 * src/elphC.h:16-32).  Either pointer may be NULL. */
int lelph_get_precision(lelph_ctx* ctx, int* requested, int* effective);
int lelph_get_device_count(int* count);

/* iBZ wavefunction ik_ibz: wfc (nspin,nbnds,nspinor,npw) complex, gvec
 * (npw,3) cartesian without 2*pi, Fk (nltimesj,ntype,npw) or NULL.
 * Replaces the struct WFC fill of src/io/get_save_data.c:516-609. */
int lelph_upload_wfc(lelph_ctx* ctx, int64_t ik_ibz, const double* wfc,
                     const double* gvec, const double* Fk, int64_t npw);

/* V_loc(|G|) spline table: create_vlocg_table (src/dvloc/vlocg_table.c:16-154).
 * Optional host copies of the result (each may be NULL): g_co (npts),
 * vlocg (ntype,npts), vploc (ntype,npts); *npts_out receives npts. */
int lelph_create_vlocg_table(lelph_ctx* ctx, int64_t* npts_out, double* g_co,
                             double* vlocg, double* vploc);

/* dVlocq (src/dvloc/dVbare_loc.c:16-231): bare local dV of q (crystal) in
 * the mode basis eigVec (nmodes, 3*natom).  Vlocr (nmodes,Nx,Ny,Nz) is an
 * optional host copy of the result (may be NULL). */
int lelph_dvlocq(lelph_ctx* ctx, const double qpt[3], const double* eigVec,
                 double* Vlocr);

/* Per-q set-up of the potential the local part contracts with; the caller
 * sequence of src/elph/elph_driver.c:210-245:
 *   dvscf   : (nmodes,nmag,Nx,Ny,Nz) complex or NULL (no screening),
 *   patterns: NULL if dvscf is already in the mode basis, else the
 *             (nmodes,nmodes) displacement patterns; dvscf is then in the
 *             pattern basis and is rotated as eig @ pats^H @ dvscf
 *             (src/io/qe/read_dvscf.c:107-158),
 *   kernel  : lelph_kernel; bare local dV is added (dVlocq + add_dvscf_qe,
 *             src/dvloc/add_dvscf.c:12-92).
 * The eigVec is kept for the non-local part of this q. */
int lelph_set_dvscf(lelph_ctx* ctx, const double qpt[3], const double* eigVec,
                    const double* dvscf, const double* patterns, int kernel);

/* Optional host copy of the current dV (nmodes,nmag,Nx,Ny,Nz). */
int lelph_get_dvscf(lelph_ctx* ctx, double* dvscf);

/* elphLocal (src/dvloc/compute_Vbare_loc.c:17-353).  Zeroes and fills
 * elph_kq (nmodes,nspin,nbnds,nbnds) with the local part. */
int lelph_elph_local(lelph_ctx* ctx, const double qpt[3], int ikq, int ik,
                     int kqsym, int ksym, double* elph_kq);

/* add_elphNonLocal (src/nonloc/Vbare_nloc.c:35-577).  ADDS the
 * Kleinman-Bylander part to elph_kq (the caller runs elph_local first, as
 * src/elph/compute_elph_matq.c:135-143 requires). */
int lelph_add_elph_nonlocal(lelph_ctx* ctx, int ikq, int ik, int kqsym,
                            int ksym, const double* eigVec, double* elph_kq);

/* elphLocal + add_elphNonLocal in one device round trip (non_loc = 0/1). */
int lelph_gkq(lelph_ctx* ctx, const double qpt[3], int ikq, int ik, int kqsym,
              int ksym, int non_loc, double* elph_kq);

/* A batch of (k,k+q) pairs for the current q; g_out (npairs, nmodes, nspin,
 * nbnds, nbnds).  Device-resident between pairs, one copy back at the end.
 * The pairs run along k -> k+q chains: a pair's psi_k box is the psi_k+q box
 * of the pair before it, and the last pair of a closed cycle takes its
 * psi_k+q box from the chain head's psi_k box, both with the shift theorem
 * where an umklapp vector separates them (within rounding of transforming
 * every side, as lelph_gkq does). */
int lelph_gkq_batch(lelph_ctx* ctx, const double qpt[3], int64_t npairs,
                    const lelph_pair* pairs, int non_loc, double* g_out);

/* Uneven batches over several contexts (cells of different grids, plane
 * wave, band and mode counts, FP64 or FP32 each): the k loops of
 * compute_and_write_elphq (src/elph/compute_elph_matq.c:116-223) of every
 * cell as one unit.  lelph_multi_create plans it: every pair gets its own
 * psi_k+q box and K-split partials, and its psi_k box is the psi_k+q box of
 * the pair before it on a k -> k+q chain (shifted where that side has an
 * umklapp vector), the last pair of a closed cycle shifting the head's psi_k
 * box into its psi_k+q box, as in lelph_gkq_batch, the contractions of all pairs of
 * one precision become the tile list of one persistent work-queue launch
 * (largest tiles first), and the whole sequence (boxes, contractions,
 * fixed-order reductions, non-local parts, copy-out) runs once eagerly and
 * is captured in a HIP graph.  lelph_multi_run replays the graph; g_out[i]
 * receives cell i's (npairs[i], nmodes, nspin, nbnds, nbnds).  g is
 * bit-identical to lelph_gkq_batch on each cell's pairs (and within
 * rounding of lelph_gkq per pair).  The contexts' dV and precision
 * must stay as they were at creation; bf16 split precisions are not
 * supported here (LELPH_ERR_UNSUPPORTED).  No counterpart in the
 * reference.  Timings (ms): boxes, contractions, reductions + non-local +
 * copy of the eager pass, the last replay in total, then the item counts
 * and grids of the FP64 and FP32 queues. */
typedef struct lelph_multi lelph_multi;
int lelph_multi_create(int ncells, lelph_ctx* const* ctxs, const double* qpts,
                       const int64_t* npairs, const lelph_pair* const* pairs, int non_loc,
                       lelph_multi** out);
int lelph_multi_run(lelph_multi* multi, double* const* g_out);
int lelph_multi_timings(lelph_multi* multi, double* ms /* 8 */);
void lelph_multi_destroy(lelph_multi* multi);

/* The k loop of compute_and_write_elphq (src/elph/compute_elph_matq.c:20-234)
 * for the iBZ q point qpt (crystal): k+q lookup on the full-BZ grid,
 * kminusq ("yambo" convention) swap, local + non-local per k.  g_out is the
 * (nkpts_BZ, nmodes, nspin, nbnds, nbnds) slab elph_mat[qpos, :, ...]. */
int lelph_compute_elphq(lelph_ctx* ctx, const double qpt[3], int non_loc,
                        int kminusq, double* g_out);

/* elph_q_rotate (src/symmetries/elphq_expand.c:16-73): g(Sk,Sq) from g(k,q)
 * and the electronic D matrices (nspin,nb,nb) of the k+q (Dmats_l) and k
 * (Dmats_r) points.  Host-side FP64. */
int lelph_elph_q_rotate(const double* Dmats_l, const double* elph_mat_q,
                        const double* Dmats_r, int nmodes, int nspin,
                        int nbnds, int tim_rev, double* elph_mat_Sq);

/* electronic_reps (src/symmetries/elec_reps.c:16-413): representation
 * matrices Dkmn (nspin, nbnds, nbnds) = <R k, b | U(R) | k, a> of the
 * symmetry (Rsym cartesian row major, tauR cartesian fractional translation,
 * tim_revR) at full-BZ point ikBZ.  R k must be on the k grid
 * (LELPH_ERR_SYM otherwise). */
int lelph_electronic_reps(lelph_ctx* ctx, const double Rsym[9], const double tauR[3],
                          int tim_revR, int64_t ikBZ, double* Dkmn);

/* elph_q_rotate (src/symmetries/elphq_expand.c:16-73) on the device, same
 * arguments and result as lelph_elph_q_rotate: g(Sk,Sq) (nmodes, nspin, nb,
 * nb) from g(k,q) and the D-matrices of the two sides. */
int lelph_elph_q_rotate_gpu(lelph_ctx* ctx, const double* Dmats_l, const double* elph_mat_q,
                            const double* Dmats_r, int nmodes, int nspin, int nb, int tim_rev,
                            double* elph_mat_Sq);

/* No counterpart in the reference: lelph_elph_q_rotate_gpu for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  All pointers are DEVICE pointers, laid out as in
 * lelph_elph_q_rotate_gpu; the kernels run on `stream` (a hipStream_t, NULL
 * for the default stream) and the call returns without waiting.  The
 * intermediate product lives in a static per-device array of
 * LELPH_NOALLOC_MAX_ELEMS complex128 (64 MiB), held until the process ends:
 * nmodes * nspin * nb^2 above that returns LELPH_ERR_SCRATCH.  No hipMalloc,
 * hipFree or allocating library call is made.  Calls that overlap on one
 * GPU share that array and race; order them on one stream. */
#define LELPH_NOALLOC_MAX_ELEMS (1 << 22)
int lelph_elph_q_rotate_noalloc(const double* d_Dmats_l, const double* d_elph_mat_q,
                                const double* d_Dmats_r, int nmodes, int nspin, int nb,
                                int tim_rev, double* d_elph_mat_Sq, void* stream);

/* get_KplusQ_idxs (src/common/numerical_func.c:471-524). */
int lelph_get_kplusq_idxs(int64_t nbz, const double* kpoints_crys,
                          const double qpt_crys[3], int* kplusq_idxs);

/* ------------------------------------------------------------------------
 * QE dvscf files: complex128 sets (nmodes * nmag) of (Nz, Ny, Nx), x fastest.
 * ---------------------------------------------------------------------- */

/* read_dvscf_qe (src/io/qe/read_dvscf.c:17-172): reads the file (pinned
 * staging), transposes to (Nx, Ny, Nz) on the device and rotates patterns
 * to modes with eigVec @ patterns^H; dvscf_out (nmodes, nmag, Nx, Ny, Nz). */
int lelph_read_dvscf_qe(lelph_ctx* ctx, const char* dvscf_file, const double* eigVec,
                        const double* patterns, double* dvscf_out);

/* lelph_set_dvscf with the potential read from a QE dvscf file (kernel
 * LELPH_KERNEL_DFPT or LELPH_KERNEL_DFPT_LOCAL), without a host copy of
 * the transposed potential (src/elph/elph_driver.c:226-245). */
int lelph_set_dvscf_file(lelph_ctx* ctx, const double qpt[3], const double* eigVec,
                         const char* dvscf_file, const double* patterns, int kernel);

/* write_dvscf_qe (src/io/qe/write_dyn_dvscf_qe.c:99-185): dvscf_in
 * (nmodes, nmag, Nx, Ny, Nz) written in the file layout above. */
int lelph_write_dvscf_qe(lelph_ctx* ctx, const char* dvscf_file, const double* dvscf_in);

/* ------------------------------------------------------------------------
 * Fourier interpolation of dV_q onto arbitrary q
 * (src/interpolation/interpolation_driver.c:190-420 coarse side,
 * :790-835 fine side).
 * ---------------------------------------------------------------------- */

/* Parameters of dV_add_longrange (src/dvloc/dVpot_add_long_range.c:18-241)
 * and dVlong_range_kernel (src/dvloc/dVlong_range_kernel.c:31-458).  A NULL
 * array switches its term off, as in the reference. */
typedef struct lelph_longrange
{
    const double* epsilon; /* (3,3) dielectric tensor, or NULL             */
    const double* Zborn;   /* (natom,3,3) Born charges, or NULL            */
    const double* Qpole;   /* (natom,3,3,3) quadrupoles, or NULL           */
    const double* Zvals;   /* (natom) valence charges (monopole), or NULL  */
    int only_induced;      /* 1: drop the monopole term                    */
    double eta_bare;       /* Ewald parameter of the bare (monopole) term  */
    double eta_induced;    /* Ewald parameter of the induced terms         */
    double EcutRy;         /* G box cut-off of the long-range sum, Ry      */
    int nmags_add[4];      /* add/subtract into dV[:, imag] if nonzero      */
} lelph_longrange;

/* dV_add_longrange (src/dvloc/dVpot_add_long_range.c:18-241) on the device:
 * dvscf (nmodes, nmag, Nx, Ny, Nz) += sign * long-range potential at qpt
 * (crystal), in the mode basis of eigVec (nmodes, 3*natom), or in the
 * cartesian basis (mode = 3*atom + x) when eigVec is NULL. */
int lelph_dv_longrange(lelph_ctx* ctx, const double qpt[3], const lelph_longrange* lr,
                       const double* eigVec, int sign, double* dvscf);

/* rotate_dvscf (src/symmetries/rotate_dvscf.c:22-368) on the device:
 * dV_Sq(r) = e^{-i Sq.tau} dV_q(S^-1 (r - tau)), SU(2) sandwich for nmag 4,
 * conjugate for time reversal.  composite_form as in the reference. */
int lelph_rotate_dvscf(lelph_ctx* ctx, const double qpt[3], const double* dvscf_in,
                       const lelph_symmetry* sym, int composite_form, double* dvscf_out);

/* build_wigner_seitz_vectors (src/interpolation/wigner_seitz.c:27-202), host
 * code: the superlattice vectors T that minimise |r_m - r_n - R - T| for
 * every R of the grid (x slowest, miller order of get_miller_idx) and every
 * (m, n) pair.  rvec_m / rvec_n NULL means one vector at the origin.
 * ws_degen (nR, nrvec_m, nrvec_n) gets the degeneracies; ws_vecs (cap, 3)
 * the vectors in crystal units of the unit cell.  *nws = sum of ws_degen;
 * LELPH_ERR_RANGE if that exceeds cap. */
int lelph_wigner_seitz(const int64_t grid[3], const double lat_vecs[9], double eps,
                       const double* rvec_m, int64_t nrvec_m, const double* rvec_n,
                       int64_t nrvec_n, int64_t cap, int64_t* ws_vecs, int64_t* ws_degen,
                       int64_t* nws);

/* Starts an interpolation on the context's lattice: coarse Gamma-centred
 * q grid qgrid, atomic masses (natom, or NULL: no mass normalisation, as
 * mass_normalize_pol_vecs with NULL), long-range parameters (NULL: none).
 * Device memory: nq * nmodes * nmag * N complex doubles. */
int lelph_interp_begin(lelph_ctx* ctx, const int64_t qgrid[3], const double* masses,
                       const lelph_longrange* lr);

/* One coarse q (crystal; any point of the full coarse grid, each exactly
 * once): dvscf (nmodes, nmag, N) in the mode basis of eigVec, as
 * get_dvscf_dyn_qe returns it.  Subtracts the long-range part, changes to the
 * cartesian basis (dVscf_change_basis 'C' with the mass-unnormalised
 * eigenvectors) and multiplies by e^{iqr} (interpolation_driver.c:316-320,
 * 380-412).  The long-range-free mode-basis potential stays on the device
 * for lelph_interp_add_star. */
int lelph_interp_add_coarse(lelph_ctx* ctx, const double qpt[3], const double* eigVec,
                            const double* dvscf);

/* A star member S q of the last coarse q added: rotate_dvscf of its
 * long-range-free potential with sym, then the same basis change with
 * eig_Sq (rotate_eig_vecs of the caller) and e^{iSq.r}
 * (interpolation_driver.c:329-363, 380-412).  qpt_Sq is S q in crystal
 * units. */
int lelph_interp_add_star(lelph_ctx* ctx, const lelph_symmetry* sym, const double qpt_Sq[3],
                          const double* eig_Sq);

/* q -> R: fft_q2R (src/interpolation/interpolation_utilities.c:25-76) over
 * the coarse grid once every coarse point is in. */
int lelph_interp_finish(lelph_ctx* ctx);

/* R -> q for nq fine points (crystal): fft_R2q_dvscf
 * (interpolation_utilities.c:165-238) with the Wigner-Seitz weights,
 * e^{-iqr}, long range added back in the cartesian basis
 * (interpolation_driver.c:822-835).  dvscf_out (nq, nmodes, nmag, N),
 * cartesian basis (mode = 3*atom + x), as write_dvscf_qe writes it. */
int lelph_interp_dvscf(lelph_ctx* ctx, int64_t nq, const double* qpts, double* dvscf_out);

/* The interpolated potential at qpt made the context's current dV_q
 * without leaving the device: lelph_interp_dvscf, then lelph_set_dvscf's
 * mode rotation with eigVec (identity patterns) and the bare local part. */
int lelph_interp_set_dvscf(lelph_ctx* ctx, const double qpt[3], const double* eigVec,
                           int kernel);

/* ------------------------------------------------------------------------
 * Fourier interpolation of dynamical matrices
 * (src/interpolation/interpolation_driver.c:190-530 coarse side, :855-1000
 * fine side; src/phonon/).  3-D and 2-D long range (dipole, quadrupole),
 * acoustic sum rule of enum lelph_asr_kind (src/phonon/asr.c).
 * ---------------------------------------------------------------------- */

/* Starts it on the context's lattice: coarse grid qgrid, atomic masses
 * (natom, or NULL), lr (epsilon, Zborn, Qpole used; NULL: no long range),
 * eta_ph, the cut-off of the long-range G sum (EcutRy, Ry; the reference
 * uses 30) and the sum rule.  3 natom <= 64. */
int lelph_dyn_interp_begin(lelph_ctx* ctx, const int64_t qgrid[3], const double* masses,
                           const lelph_longrange* lr, double eta_ph, double EcutRy, int asr);

/* One coarse q (crystal): frequencies omega (nmodes) and polarisation
 * vectors (nmodes, 3 natom) as read from the dyn file; pol_vecs_to_dyn
 * (src/phonon/pol_vecs_to_dyn.c:8-40) and the long range removed
 * (interpolation_driver.c:446-481). */
int lelph_dyn_interp_add_coarse(lelph_ctx* ctx, const double qpt[3], const double* omega,
                                const double* pol_vecs);

/* q -> R, mass factors and the sum rule (interpolation_driver.c:483-532). */
int lelph_dyn_interp_finish(lelph_ctx* ctx);

/* nq fine q (crystal): fft_R2q_dyn (interpolation_utilities.c:78-163), long
 * range added back, and (when omega_out or eig_out is given) the Hermitian
 * part diagonalised on the device (interpolation_driver.c:945-992):
 * dyn_out (nq, nmodes, nmodes) mass-normalised dynamical matrices,
 * omega_out (nq, nmodes) sign(w) sqrt|w| ascending, eig_out (nq, nmodes,
 * 3 natom) eigenvector rows / sqrt(M).  Each may be NULL.  Eigenvectors are
 * unique up to a phase per mode (a unitary mix within degenerate modes);
 * this solver does not reproduce LAPACK's choice. */
int lelph_dyn_interp(lelph_ctx* ctx, int64_t nq, const double* qpts, double* dyn_out,
                     double* omega_out, double* eig_out);

/* acoustic sum rule kinds (enum asr_kind, src/common/dtypes.h:71-78) */
enum lelph_asr_kind
{
    LELPH_ASR_NONE = 0,
    LELPH_ASR_SIMPLE = 1,
    LELPH_ASR_CRYSTAL = 2,
    LELPH_ASR_ALL = 3,
    LELPH_ASR_ALL_HUANG = 4
};

/* apply_acoustic_sum_rule_fc (src/phonon/asr.c:269-512), host code: frc
 * (ngrid, 3 natom, 3 natom) complex R-space force constants (without mass
 * factors) on the q grid, in place.  simple: the on-site block takes minus
 * the row sums.  crystal / all / all_huang: the real parts projected
 * orthogonally onto the translational (+ rotational, + Huang) conditions,
 * -R blocks the transposes and R = 0 symmetric, imaginary parts zero.
 * ws_vecs / ws_degen as lelph_wigner_seitz gives them for (atoms, atoms);
 * ws_vecs and lat_vecs are only read by all / all_huang. */
int lelph_asr_fc(int mode, const int64_t qgrid[3], int natom, double* frc,
                 const double* atomic_pos, const double* lat_vecs, const int64_t* ws_vecs,
                 int64_t n_ws_vecs, const int64_t* ws_degen);

/* apply_acoustic_sum_rule_born_charges (src/phonon/asr.c:155-267), host
 * code: Zborn (natom, 3, 3) in place with sum_a Z_a = 0 (the mean removed;
 * every mode but none gives this, the orthogonal projection onto the nine
 * conditions). */
int lelph_asr_born(int mode, double* Zborn, int natom);

/* quadrupole_3d_to_2d (src/phonon/quadrupole_3Dto2D.c:18-117), host code:
 * the 3-D dynamical quadrupoles Qpole (natom, 3, 3, 3) of a slab turned in
 * place into the 2-D ones the 2-D long range takes (Phys. Rev. X 11,
 * 041027, App. B), from epsilon (3, 3), the cartesian atomic positions and
 * Zborn (natom, 3, 3, or NULL). */
int lelph_quadrupole_3d_to_2d(int natom, const double* epsilon, const double* atomic_pos,
                              const double* Zborn, double* Qpole);

/* LO-TO direction (src/interpolation/interpolation_driver.c:734-786), host
 * code: q_crys = A^T (1e-5 dir / |dir|), the point that stands for Gamma
 * approached along the cartesian direction dir; LELPH_ERR_ARG when
 * |dir| < 1e-6. */
int lelph_loto_qpoint(const double lat_vecs[9], const double dir[3], double q_crys[3]);

/* ------------------------------------------------------------------------
 * QE phonon files (ph_save)
 * ---------------------------------------------------------------------- */

/* read_dyn_qe (src/io/qe/read_dyn.c:38-548, the old text format or the
 * XML one, first matrix of the file): qpt (3) as written, omega (nmodes) sign(w) sqrt|w|
 * ascending, pol_vecs (nmodes, 3 natom) eigenvector rows / sqrt(M), amass
 * (natom, or NULL) the masses of the file.  The Hermitian part is
 * diagonalised on the device (eigenvector phases are this solver's own). */
int lelph_read_dyn_qe(lelph_ctx* ctx, const char* dyn_file, double* qpt, double* omega,
                      double* pol_vecs, double* amass);

/* read_pattern_qe (src/io/qe/read_patterns.c:18-125), host only: pat_vecs
 * (nmodes, nmodes) complex from patterns.N.xml (identity for the
 * DVSCF_IS_IN_CARTESIAN_BASIS placeholder). */
int lelph_read_patterns_qe(const char* pat_file, int natom, double* pat_vecs);

/* data-file-schema.xml contents (parse_qexml, src/io/qe/read_qexml.c:20-643) */
typedef struct lelph_qexml_info
{
    int64_t natom, ntype;
    int64_t nsym;        /* crystal symmetries listed (before time-reversal doubling) */
    int64_t fft_dims[3]; /* basis_set/fft_grid nr1 nr2 nr3 */
    double lat_vec[9];   /* a[:, i] = a_i, bohr */
    double alat[3];      /* atomic_structure alat, three times */
    char dimension;      /* '2' when assume_isolated starts with "2D", else '3' */
    int is_soc_present;  /* magnetization/spinorbit */
    int nmag;            /* 1, 2 (lsda) or 4 (noncolin + do_magnetization) */
    int nspinor;         /* 2 when noncolin */
    int ph_mag_symm;     /* symmetries carry time_reversal="true" (magnetic group) */
    int ph_tim_rev;      /* time reversal usable for phonons */
    char pseudo_dir[1024];
} lelph_qexml_info;

/* parse_qexml (src/io/qe/read_qexml.c:20-643), host only.  Fills info;
 * with non-NULL arrays also atom_type (natom, 0-based species index),
 * atomic_pos (natom, 3; cartesian bohr as written), pseudo_files (ntype x
 * 256 chars), sym_mats (nsym, 3, 3; crystal rotations as written),
 * sym_tau (nsym, 3; crystal fractional translations, QE's x' = Sx - v
 * sign) and trevs (nsym).  For a magnetic group the operations without
 * time reversal come first (read_qexml.c:556-613).  Call once with NULL
 * arrays to learn natom, ntype and nsym. */
int lelph_parse_qexml(const char* xml_file, lelph_qexml_info* info, int* atom_type,
                      double* atomic_pos, char* pseudo_files, double* sym_mats, double* sym_tau,
                      int* trevs);

/* parse_upf (src/io/qe/parse_upf.c:28-336), host only: UPF-1 or UPF-2,
 * norm-conserving (LELPH_ERR_UNSUPPORTED otherwise).  ngrid and zval
 * always; r, rab, vloc (ngrid each, Ry) when not NULL. */
int lelph_parse_upf(const char* upf_file, int64_t* ngrid, double* zval, double* r, double* rab,
                    double* vloc);

/* read_ph_tensors_qe + read_quadrupole_fmt (src/io/qe/read_tensors.c),
 * host only: epsilon (3, 3), zborn (natom, 3, 3), qpole (natom, 3, 3, 3);
 * present bit 0 / 1 / 2 says which were found.  Absent files are not an
 * error; either path may be NULL. */
int lelph_read_ph_tensors_qe(const char* tensors_xml, const char* quadrupole_fmt, int natom,
                             double* epsilon, double* zborn, double* qpole, int* present);

/* read_qpts_qe (src/io/qe/read_dyn.c:551-609), host only: q grid product
 * nq_bz, iBZ count nq_ibz and, when qpts is not NULL, the iBZ q points
 * (nq_ibz, 3) of dyn0 (cartesian, 2 pi / alat). */
int lelph_read_qpts_qe(const char* dyn0_file, int64_t* nq_ibz, int64_t* nq_bz, double* qpts);

/* write_dyn_qe (src/io/qe/write_dyn_dvscf_qe.c:18-97), host only: one q,
 * dyn (nmodes, nmodes) mass normalised, written times sqrt(Ma Mb). */
int lelph_write_dyn_qe(const char* dyn_file, int natom, const double qpt[3], const double* dyn,
                       const double* masses);

/* Device time of the last lelph_interp_dvscf, summed over its q chunks,
 * milliseconds: [0] Wigner-Seitz weights, [1] R -> q phase GEMM with
 * e^{-iqr}, [2] long range added back, [3] total, [4] number of q.
 * No counterpart in the reference (timing of this library's kernels). */
int lelph_interp_timings(lelph_ctx* ctx, double* ms, int n);

/* Device time of the last elph_local / nonlocal / gkq / batch call, summed
 * over its pairs (HIP events), milliseconds:
 *   [0] sphere scatter + backward FFTs of psi_k and psi_k+q
 *   [1] local contraction kernel (the MFMA GEMM)
 *   [2] split-K reduction and 1/N
 *   [3] non-local part; in a batch, where it runs on a second stream beside
 *       the local parts, the time from the end of the pair's local part to
 *       the end of its non-local part
 *   [4] total device time per pair (start of its local part to the end of
 *       both parts; overlapping pairs of a batch count their overlap twice)
 *   [5] number of (k,k+q) pairs
 *   [6] number of psi sides transformed (sphere scatter + backward FFT); a
 *       batch takes the others over along k -> k+q chains with the shift
 *       theorem: a pair's psi_k box is the psi_k+q box of the pair before
 *       it, and the last pair of a closed cycle (k + nq = k + G) takes its
 *       psi_k+q box from the first pair's psi_k box
 * n: number of entries wanted (<= 8). */
int lelph_last_timings(lelph_ctx* ctx, double* ms, int n);

/* ------------------------------------------------------------------------
 * Learned stand-in for the Fourier interpolation of dV_q onto fine q
 * (src/interpolation/interpolation_driver.c:813-856, fft_R2q_dvscf
 * src/interpolation/interpolation_utilities.c:165-238); no counterpart
 * among the reference's functions.  A 3-D convolutional network on the FFT
 * grid (3x3x3 taps, periodic; 128 input channels, width 64, depth + 2
 * layers, SiLU), BF16 operands with FP32 accumulation on MFMA, FP32 master
 * weights, Adam, trained on the device.  It predicts the short-range
 * periodic part u_q(r) = e^{-2 pi i q.x} (dV_q(r) - dV_q^LR(r)) of nfields
 * complex fields (cartesian basis) at fine q as a correction to the
 * trilinear blend of the 8 coarse q around it; fields are laid out
 * (q, point, 2 nfields) (channel fastest, re/im interleaved per field) on
 * the host side of every call.  Nx*Ny*Nz must be a multiple of 64.
 * ---------------------------------------------------------------------- */
typedef struct lelph_nn lelph_nn;
int lelph_nn_create(const int64_t grid[3], int nfields, int depth, uint64_t seed,
                    lelph_nn** out);
void lelph_nn_destroy(lelph_nn* nn);
/* number of parameters (padding included), and the FP32 masters as
 * [layer][W (27 cin, 64) then b (64)] */
int lelph_nn_nparams(lelph_nn* nn, int64_t* n);
int lelph_nn_get_params(lelph_nn* nn, float* params);
int lelph_nn_set_params(lelph_nn* nn, const float* params);
/* the coarse grid's short-range fields (nqc, nfields, N) complex128 in the
 * cartesian basis (NOT multiplied by e^{-iqr}; the features do that per
 * neighbour), and the scale every field is divided by inside the network */
int lelph_nn_set_coarse(lelph_nn* nn, const int64_t qgrid[3], const double* fields,
                        double scale);
/* the network's input features X0 (S, N, 128) and trilinear base
 * (S, N, 2 nfields) of S fine q (crystal), scaled; either output may be NULL */
int lelph_nn_features(lelph_nn* nn, int S, const double* qpts, float* X0, float* base);
/* forward only on given inputs (X0 rounded to BF16): out = base + net(X0),
 * (S, N, 2 nfields), in the scaled units of the inputs */
int lelph_nn_forward(lelph_nn* nn, int S, const float* X0, const float* base, float* out);
/* mean-square loss of base + net(X0) against target (scaled units) and the
 * gradient of every parameter (layout of lelph_nn_get_params; may be NULL) */
int lelph_nn_loss_grad(lelph_nn* nn, int S, const float* X0, const float* base,
                       const float* target, double* loss, float* grad);
/* training data: (nq, N, 2 nfields) targets u_q, unscaled */
int lelph_nn_set_targets(lelph_nn* nn, int64_t nq, const float* targets);
/* Adam over the targets' q (crystal, nq rows matching lelph_nn_set_targets)
 * in shuffled batches; loss_hist (epochs) receives each epoch's mean loss */
int lelph_nn_train(lelph_nn* nn, int64_t nq, const double* qpts, int epochs, int batch,
                   double lr, uint64_t seed, double* loss_hist);
/* u_q of nq fine q, (nq, N, 2 nfields), unscaled */
int lelph_nn_predict(lelph_nn* nn, int64_t nq, const double* qpts, float* out);

#ifdef __cplusplus
}
#endif

#endif /* LELPH_HIP_H */
