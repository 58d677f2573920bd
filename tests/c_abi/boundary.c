/* The drop-in sequence of a C caller, compiled against include/lelph_hip.h
 * and linked to liblelph_hip.so: lelph_create -> lelph_upload_wfc per iBZ k
 * -> lelph_set_dvscf -> lelph_gkq per pair (the calls that replace
 * elphLocal + add_elphNonLocal at src/elph/compute_elph_matq.c:135-143 and
 * dVlocq + add_dvscf_qe at src/elph/elph_driver.c:239-245), then
 * lelph_gkq_batch over the same pairs.
 *
 * Usage: lelph_boundary IN OUT.  IN is a sequence of records
 *   char name[16] | int64 nbytes | nbytes of data
 * (written by tests/test_gpu_boundary.py from a seeded system); OUT gets
 * the g of every pair from lelph_gkq, then from lelph_gkq_batch, complex128
 * (npairs, nmodes, nspin, nbnds, nbnds) each.  Test code. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lelph_hip.h"

typedef struct
{
    char name[17];
    int64_t n;
    void* p;
} rec;

static rec recs[4096];
static int nrec;

static void* get(const char* name, int64_t* nbytes)
{
    for (int i = 0; i < nrec; ++i)
        if (!strcmp(recs[i].name, name))
        {
            if (nbytes) *nbytes = recs[i].n;
            return recs[i].p;
        }
    fprintf(stderr, "boundary: record %s missing\n", name);
    exit(3);
}

static double getd(const char* name) { return *(const double*)get(name, NULL); }

#define CHECK(call)                                                                      \
    do                                                                                   \
    {                                                                                    \
        int st_ = (call);                                                                \
        if (st_ != LELPH_OK)                                                             \
        {                                                                                \
            fprintf(stderr, "boundary: %s -> %d (%s)\n", #call, st_, lelph_strerror(st_)); \
            return 4;                                                                    \
        }                                                                                \
    } while (0)

int main(int argc, char** argv)
{
    if (argc != 3)
    {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    for (;;)
    {
        rec r;
        memset(&r, 0, sizeof r);
        if (fread(r.name, 1, 16, f) != 16) break;
        if (fread(&r.n, 8, 1, f) != 1) return 2;
        r.p = malloc(r.n > 0 ? (size_t)r.n : 1);
        if (r.n > 0 && fread(r.p, 1, (size_t)r.n, f) != (size_t)r.n) return 2;
        if (nrec == 4096) return 2;
        recs[nrec++] = r;
    }
    fclose(f);

    lelph_lattice lat;
    memset(&lat, 0, sizeof lat);
    const int64_t* li = get("lat_ints", NULL);
    lat.natom = (int)li[0];
    lat.nsym = (int)li[1];
    lat.timerev = (int)li[2];
    lat.nspin = (int)li[3];
    lat.nspinor = (int)li[4];
    lat.nbnds = (int)li[5];
    lat.nmodes = (int)li[6];
    lat.nmag = (int)li[7];
    lat.dimension = (char)li[8];
    lat.is_soc_present = (int)li[9];
    memcpy(lat.fft_dims, get("fft_dims", NULL), sizeof lat.fft_dims);
    memcpy(lat.alat_vec, get("alat_vec", NULL), sizeof lat.alat_vec);
    memcpy(lat.blat_vec, get("blat_vec", NULL), sizeof lat.blat_vec);
    lat.volume = getd("volume");
    lat.atomic_pos = get("atomic_pos", NULL);
    lat.atom_type = get("atom_type", NULL);
    int64_t nb;
    lat.kpt_iredBZ = get("kpt_iredBZ", &nb);
    lat.nkpts_iBZ = nb / 24;
    lat.kpt_fullBZ_crys = get("kpt_fullBZ", &nb);
    lat.nkpts_BZ = nb / 24;
    lat.kmap = get("kmap", NULL);
    const double* sd = get("syms", NULL); /* per op: Rmat 9, tau 3, time_rev 1 */
    lelph_symmetry* syms = calloc((size_t)lat.nsym, sizeof *syms);
    for (int i = 0; i < lat.nsym; ++i)
    {
        memcpy(syms[i].Rmat, sd + 13 * i, 9 * sizeof(double));
        memcpy(syms[i].tau, sd + 13 * i + 9, 3 * sizeof(double));
        syms[i].time_rev = (int)sd[13 * i + 12];
    }
    lat.syms = syms;

    lelph_pseudo pp;
    memset(&pp, 0, sizeof pp);
    const int64_t* pi = get("pp_ints", NULL);
    pp.ntype = (int)pi[0];
    pp.lmax = (int)pi[1];
    pp.nltimesj = pi[2];
    pp.PP_table = get("PP_table", NULL);
    pp.Fsign = get("Fsign", NULL);
    pp.fCoeff = NULL;
    pp.qmax_abs = getd("qmax_abs");
    lelph_local_pseudo* loc = calloc((size_t)pp.ntype, sizeof *loc);
    for (int t = 0; t < pp.ntype; ++t)
    {
        char nm[32];
        snprintf(nm, sizeof nm, "loc%d_v", t);
        loc[t].Vloc_atomic = get(nm, &nb);
        loc[t].ngrid = nb / 8;
        snprintf(nm, sizeof nm, "loc%d_r", t);
        loc[t].r_grid = get(nm, NULL);
        snprintf(nm, sizeof nm, "loc%d_rab", t);
        loc[t].rab_grid = get(nm, NULL);
        snprintf(nm, sizeof nm, "loc%d_z", t);
        loc[t].Zval = getd(nm);
    }
    pp.loc_pseudo = loc;

    /* section 1: the context and the iBZ wavefunctions */
    lelph_ctx* ctx = NULL;
    CHECK(lelph_create(&lat, &pp, LELPH_FP64, &ctx));
    for (int64_t ik = 0; ik < lat.nkpts_iBZ; ++ik)
    {
        char nm[32];
        snprintf(nm, sizeof nm, "wfc%ld", (long)ik);
        const double* wfc = get(nm, NULL);
        snprintf(nm, sizeof nm, "gvec%ld", (long)ik);
        const double* gvec = get(nm, &nb);
        snprintf(nm, sizeof nm, "Fk%ld", (long)ik);
        const double* Fk = get(nm, NULL);
        CHECK(lelph_upload_wfc(ctx, ik, wfc, gvec, Fk, nb / 24));
    }
    /* section 2: dV of this q */
    const double* qpt = get("qpt", NULL);
    CHECK(lelph_set_dvscf(ctx, qpt, get("eigvec", NULL), get("dvscf", NULL), NULL,
                          LELPH_KERNEL_DFPT));
    /* section 3: the pairs, one call each, then one batch */
    const int* pr = get("pairs", &nb);
    const int64_t npairs = nb / 16;
    const int64_t gsz = (int64_t)lat.nmodes * lat.nspin * lat.nbnds * lat.nbnds;
    double* g = calloc((size_t)(2 * npairs * gsz), 2 * sizeof(double));
    for (int64_t p = 0; p < npairs; ++p)
        CHECK(lelph_gkq(ctx, qpt, pr[4 * p], pr[4 * p + 1], pr[4 * p + 2], pr[4 * p + 3], 1,
                        g + 2 * p * gsz));
    lelph_pair* pairs = calloc((size_t)npairs, sizeof *pairs);
    for (int64_t p = 0; p < npairs; ++p)
    {
        pairs[p].ikq = pr[4 * p];
        pairs[p].ik = pr[4 * p + 1];
        pairs[p].kqsym = pr[4 * p + 2];
        pairs[p].ksym = pr[4 * p + 3];
    }
    CHECK(lelph_gkq_batch(ctx, qpt, npairs, pairs, 1, g + 2 * npairs * gsz));
    /* a call out of order is reported, not aborted */
    int st = lelph_gkq(ctx, qpt, 99999, 0, 0, 0, 1, g);
    if (st != LELPH_ERR_ARG)
    {
        fprintf(stderr, "boundary: bad k index gave %d\n", st);
        return 5;
    }
    lelph_destroy(ctx);

    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    fwrite(g, 2 * sizeof(double), (size_t)(2 * npairs * gsz), o);
    fclose(o);
    printf("boundary: %ld pairs, %d modes, %d bands: ok\n", (long)npairs, lat.nmodes, lat.nbnds);
    return 0;
}
