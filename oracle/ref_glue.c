/*
 * ref_glue.c -- flat entry points over the reference's own interpolation
 * helpers (compiled from /root/reference/src into oracle/_ref), so the test
 * fixtures can be written with plain ctypes.  TEST INFRASTRUCTURE ONLY.
 *
 * Only fills the reference's structs and forwards; no arithmetic here.
 */
#include <mpi.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "common/ELPH_timers.h"
#include "common/dtypes.h"
#include "dvloc/dvloc.h"
#include "elphC.h"
#include "interpolation/wigner_seitz.h"
#include "symmetries/symmetries.h"

/* MPI (singleton) and the timer table the helpers record into */
int rg_init(void)
{
    int f = 0;
    MPI_Initialized(&f);
    if (!f && MPI_Init(NULL, NULL) != MPI_SUCCESS) return 1;
    init_ELPH_clocks();
    return 0;
}

/* build_wigner_seitz_vectors (src/interpolation/wigner_seitz.c:27) */
long long rg_ws(const long long* grid, const double* lat, double eps, const double* rm,
                long long nm, const double* rn, long long nn, long long cap, long long* vecs,
                long long* degen)
{
    ND_int* v = NULL;
    ND_int* d = NULL;
    ND_int n = build_wigner_seitz_vectors(grid, lat, eps, rm, nm, rn, nn, &v, &d);
    const ND_int nR = grid[0] * grid[1] * grid[2] * (rm ? nm : 1) * (rn ? nn : 1);
    if (n > cap)
    {
        free(v);
        free(d);
        return -1;
    }
    memcpy(vecs, v, sizeof(ND_int) * 3 * n);
    memcpy(degen, d, sizeof(ND_int) * nR);
    free(v);
    free(d);
    return n;
}

/* multiply_eikr (src/dvloc/multiply_eikr.c:14) on a full (one-rank) grid */
void rg_multiply_eikr(double* pot, const double* q, const long long* dims, long long nsets,
                      int sign)
{
    struct Lattice L;
    memset(&L, 0, sizeof(L));
    for (int i = 0; i < 3; ++i) L.fft_dims[i] = dims[i];
    L.nfftz_loc = dims[2];
    L.nfftz_loc_shift = 0;
    multiply_eikr((ELPH_cmplx*)pot, q, &L, nsets, sign);
}

/* rotate_dvscf (src/symmetries/rotate_dvscf.c:22) on MPI_COMM_SELF */
void rg_rotate_dvscf(const double* q, const double* in, const double* R, const double* tau,
                     int tr, const long long* dims, long long nmodes, long long nmag,
                     int nspinor, const double* alat, const double* blat, int composite,
                     double* out)
{
    struct Lattice L;
    memset(&L, 0, sizeof(L));
    for (int i = 0; i < 3; ++i) L.fft_dims[i] = dims[i];
    L.nfftz_loc = dims[2];
    L.nfftz_loc_shift = 0;
    L.nmodes = nmodes;
    L.nmag = nmag;
    L.nspinor = nspinor;
    memcpy(L.alat_vec, alat, sizeof(L.alat_vec));
    memcpy(L.blat_vec, blat, sizeof(L.blat_vec));
    struct symmetry s;
    memcpy(s.Rmat, R, sizeof(s.Rmat));
    memcpy(s.tau, tau, sizeof(s.tau));
    s.time_rev = tr != 0;
    rotate_dvscf(q, (const ELPH_cmplx*)in, &s, &L, composite != 0, (ELPH_cmplx*)out,
                 MPI_COMM_SELF);
}

/* alloc_and_Compute_f_Coeff (src/nonloc/fcoeff.c:24) for one channel (l, 2j)
 * of one atom type; out (2l+1, 2l+1, 2, 2) complex */
#include "nonloc/fcoeff.h"
int rg_fcoeff(int l, int twoj, double* out)
{
    struct Lattice L;
    memset(&L, 0, sizeof(L));
    L.nspinor = 2;
    L.is_soc_present = true;
    struct Pseudo P;
    memset(&P, 0, sizeof(P));
    double pp[3] = {l + 1.0, (double)twoj, 1.0};
    P.PP_table = pp;
    P.nltimesj = 1;
    P.ntype = 1;
    alloc_and_Compute_f_Coeff(&L, &P);
    if (!P.fCoeff || !P.fCoeff[0]) return 1;
    const int n = (2 * l + 1) * (2 * l + 1) * 4;
    memcpy(out, P.fCoeff[0], sizeof(ELPH_cmplx) * n);
    free_f_Coeff(&P);
    return 0;
}

/* quadrupole_3d_to_2d (src/phonon/quadrupole_3Dto2D.c:18) on a 2-D
 * lattice; Q (natom, 3, 3, 3) in place */
void quadrupole_3d_to_2d(struct Lattice* lattice, struct Phonon* phonon);
void rg_quad2d(long long natom, double* eps, double* pos, double* zb, double* Q)
{
    struct Lattice L;
    struct Phonon P;
    memset(&L, 0, sizeof(L));
    memset(&P, 0, sizeof(P));
    L.dimension = '2';
    L.natom = natom;
    L.atomic_pos = pos;
    P.epsilon = eps;
    P.Zborn = zb;
    P.Qpole = Q;
    quadrupole_3d_to_2d(&L, &P);
}
