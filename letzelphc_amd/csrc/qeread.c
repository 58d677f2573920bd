/*
 * qeread.c -- host readers of QuantumESPRESSO phonon files (ph_save):
 * dynamical-matrix files (src/io/qe/read_dyn.c:67-293, old text format),
 * displacement patterns (src/io/qe/read_patterns.c:18-125) and the dyn
 * file writer of the interpolation (src/io/qe/write_dyn_dvscf_qe.c:18-97).
 * The dyn matrix read here is diagonalised on the device by
 * lelph_read_dyn_qe (lelph_hip.hip); this file only parses and formats.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lelph_hip.h"

#define QE_LINE 10000
#define LELPH_EPS_DYN 1e-5 /* ELPH_EPS, src/common/constants.h:9 */

/* every number strtod finds in s, at most n stored; returns the count
 * (parse_floats_from_string, src/common/string_func.c:48-88) */
static int64_t qe_floats(const char* s, double* out, int64_t n)
{
    const char* p = s;
    char* end;
    int64_t c = 0;
    while (*p)
    {
        const double v = strtod(p, &end);
        if (end == p)
        {
            ++p;
            continue;
        }
        if (out && c < n) out[c] = v;
        ++c;
        p = end;
    }
    return c;
}

static int qe_line(FILE* fp, char* buf)
{
    return fgets(buf, QE_LINE, fp) != NULL;
}

/* Parses the first dynamical matrix of an old-format dyn file:
 * qpt (3), masses (natom) and dyn (nm x nm, row major, D[(a,x)][(b,y)]
 * divided by sqrt(Ma Mb)).  Returns LELPH_OK or LELPH_ERR_ARG. */
int lelph_qe_parse_dyn(const char* path, int natom, double* qpt, double* masses, double* dynd)
{
    FILE* fp = fopen(path, "r");
    if (!fp) return LELPH_ERR_ARG;
    char* buf = malloc(QE_LINE);
    double f[16];
    double* tm = NULL;
    int st = LELPH_ERR_ARG;
    const int nm = 3 * natom;
    if (!buf) goto done;
    /* two comment lines, then ntype natom ibrav celldm(6) */
    if (!qe_line(fp, buf) || !qe_line(fp, buf) || !qe_line(fp, buf)) goto done;
    if (qe_floats(buf, f, 16) != 9) goto done;
    const int ntype = (int)rint(f[0]), na = (int)rint(f[1]), ibrav = (int)rint(f[2]);
    if (na != natom || ntype < 1) goto done;
    if (ibrav == 0)
        for (int i = 0; i < 4; ++i)
            if (!qe_line(fp, buf)) goto done;
    tm = malloc(sizeof(double) * ntype);
    if (!tm) goto done;
    for (int t = 0; t < ntype; ++t)
    {
        int it;
        char name[128];
        double m;
        if (!qe_line(fp, buf) || sscanf(buf, "%d '%127[^']' %lf", &it, name, &m) != 3) goto done;
        if (fabs(m) < LELPH_EPS_DYN) goto done;
        tm[t] = m;
    }
    for (int a = 0; a < natom; ++a)
    {
        if (!qe_line(fp, buf) || qe_floats(buf, f, 16) != 5) goto done;
        const int it = (int)rint(f[1]);
        if (it < 1 || it > ntype) goto done;
        masses[a] = tm[it - 1];
    }
    /* blank, "Dynamical ...", blank, q = ( ... ), blank */
    if (!qe_line(fp, buf) || !qe_line(fp, buf)) goto done;
    {
        const char* p = buf;
        while (*p && isspace((unsigned char)*p)) ++p;
        if (strncmp(p, "Dynamical", 9)) goto done;
    }
    if (!qe_line(fp, buf) || !qe_line(fp, buf)) goto done;
    if (qe_floats(buf, qpt, 3) != 3) goto done;
    if (!qe_line(fp, buf)) goto done;
    for (int a = 0; a < natom; ++a)
        for (int b = 0; b < natom; ++b)
        {
            int i1, i2;
            if (!qe_line(fp, buf) || sscanf(buf, "%d %d", &i1, &i2) != 2) goto done;
            if (i1 != a + 1 || i2 != b + 1) goto done;
            const double inv = 1.0 / sqrt(masses[a] * masses[b]);
            for (int x = 0; x < 3; ++x)
            {
                if (!qe_line(fp, buf) || qe_floats(buf, f, 16) != 6) goto done;
                for (int y = 0; y < 3; ++y)
                {
                    const int64_t o = 2 * ((int64_t)(3 * a + x) * nm + 3 * b + y);
                    dynd[o] = f[2 * y] * inv;
                    dynd[o + 1] = f[2 * y + 1] * inv;
                }
            }
        }
    st = LELPH_OK;
done:
    free(tm);
    free(buf);
    fclose(fp);
    return st;
}

/* write_dyn_qe (src/io/qe/write_dyn_dvscf_qe.c:18-97): one q, dyn
 * (nm x nm row major, mass normalised) written times sqrt(Ma Mb) */
int lelph_write_dyn_qe(const char* path, int natom, const double qpt[3], const double* dynd,
                       const double* masses)
{
    if (!path || natom < 1 || !qpt || !dynd || !masses) return LELPH_ERR_ARG;
    FILE* fp = fopen(path, "w");
    if (!fp) return LELPH_ERR_ARG;
    const int nm = 3 * natom;
    fprintf(fp, "Dynamical matrix file\n");
    fprintf(fp, "%s\n", "This is fake dyn file created by LetzElphC, only to be read by "
                        "LetzElphC.");
    fprintf(fp, "%3lld%5lld%4lld%12.7f%12.7f%12.7f%12.7f%12.7f%12.7f\n", (long long)natom,
            (long long)natom, -1LL, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    for (int a = 0; a < natom; ++a)
        fprintf(fp, " %4lld 'X%lld' %14.6f\n", (long long)(a + 1), (long long)(a + 1), masses[a]);
    for (int a = 0; a < natom; ++a)
        fprintf(fp, "%5lld%5lld%18.10f%18.10f%18.10f\n", (long long)(a + 1), (long long)(a + 1),
                0.0, 0.0, 0.0);
    fprintf(fp, "\n     Dynamical  Matrix in cartesian axes\n\n");
    fprintf(fp, "     q = ( %14.9f%14.9f%14.9f ) \n\n", qpt[0], qpt[1], qpt[2]);
    for (int a = 0; a < natom; ++a)
        for (int b = 0; b < natom; ++b)
        {
            fprintf(fp, "%5lld%5lld\n", (long long)(a + 1), (long long)(b + 1));
            const double ms = sqrt(masses[a] * masses[b]);
            for (int x = 0; x < 3; ++x)
            {
                for (int y = 0; y < 3; ++y)
                {
                    const int64_t o = 2 * ((int64_t)(3 * a + x) * nm + 3 * b + y);
                    fprintf(fp, "%12.8f %12.8f   ", dynd[o] * ms, dynd[o + 1] * ms);
                }
                fprintf(fp, "\n");
            }
        }
    fprintf(fp, "\n");
    return fclose(fp) == 0 ? LELPH_OK : LELPH_ERR_ARG;
}

/* inner text [*ib, *ie) of the first <tag ...>...</tag> in [b, e) */
static int xml_elem(const char* b, const char* e, const char* tag, const char** ib,
                    const char** ie)
{
    const size_t n = strlen(tag);
    for (const char* p = b; p + n + 1 < e; ++p)
    {
        if (p[0] != '<' || strncmp(p + 1, tag, n)) continue;
        const char c = p[1 + n];
        if (c != '>' && !isspace((unsigned char)c) && c != '/') continue;
        const char* gt = memchr(p, '>', (size_t)(e - p));
        if (!gt) return 0;
        if (gt[-1] == '/')
        {
            *ib = *ie = gt + 1;
            return 1;
        }
        char close[160];
        snprintf(close, sizeof(close), "</%s>", tag);
        const size_t cn = strlen(close);
        for (const char* q = gt + 1; q + cn <= e; ++q)
            if (!strncmp(q, close, cn))
            {
                *ib = gt + 1;
                *ie = q;
                return 1;
            }
        return 0;
    }
    return 0;
}

/* read_pattern_qe (src/io/qe/read_patterns.c:18-125): pat_vecs (nm, nm)
 * complex, rows the displacement patterns in irrep / perturbation order;
 * a DVSCF_IS_IN_CARTESIAN_BASIS root means the identity. */
int lelph_read_patterns_qe(const char* path, int natom, double* pat_vecs)
{
    if (!path || natom < 1 || !pat_vecs) return LELPH_ERR_ARG;
    FILE* fp = fopen(path, "rb");
    if (!fp) return LELPH_ERR_ARG;
    fseek(fp, 0, SEEK_END);
    const long sz = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    char* s = malloc((size_t)sz + 1);
    if (!s || fread(s, 1, (size_t)sz, fp) != (size_t)sz)
    {
        free(s);
        fclose(fp);
        return LELPH_ERR_ARG;
    }
    fclose(fp);
    s[sz] = 0;
    const int nm = 3 * natom;
    const char* e = s + sz;
    const char *ib, *ie;
    int st = LELPH_ERR_ARG;
    double* tmp = malloc(sizeof(double) * 2 * nm);
    if (!tmp) goto done;
    if (strstr(s, "<DVSCF_IS_IN_CARTESIAN_BASIS"))
    {
        for (int i = 0; i < 2 * nm * nm; ++i) pat_vecs[i] = 0.0;
        for (int i = 0; i < nm; ++i) pat_vecs[2 * (i * (nm + 1))] = 1.0;
        st = LELPH_OK;
        goto done;
    }
    const char *rb, *re;
    if (!xml_elem(s, e, "IRREPS_INFO", &rb, &re)) goto done;
    if (!xml_elem(rb, re, "NUMBER_IRR_REP", &ib, &ie)) goto done;
    const int nirr = atoi(ib);
    int imode = 0;
    for (int r = 0; r < nirr; ++r)
    {
        char tag[64];
        snprintf(tag, sizeof(tag), "REPRESENTION.%d", r + 1);
        const char *pb, *pe;
        if (!xml_elem(rb, re, tag, &pb, &pe)) goto done;
        if (!xml_elem(pb, pe, "NUMBER_OF_PERTURBATIONS", &ib, &ie)) goto done;
        const int np = atoi(ib);
        for (int p = 0; p < np; ++p)
        {
            snprintf(tag, sizeof(tag), "PERTURBATION.%d", p + 1);
            const char *qb, *qe;
            if (!xml_elem(pb, pe, tag, &qb, &qe)) goto done;
            if (!xml_elem(qb, qe, "DISPLACEMENT_PATTERN", &ib, &ie)) goto done;
            const size_t len = (size_t)(ie - ib);
            char* txt = malloc(len + 1);
            if (!txt) goto done;
            memcpy(txt, ib, len);
            txt[len] = 0;
            const int64_t nf = qe_floats(txt, tmp, 2 * nm);
            free(txt);
            if (nf != 2 * nm || imode >= nm) goto done;
            for (int i = 0; i < 2 * nm; ++i) pat_vecs[2 * (int64_t)imode * nm + i] = tmp[i];
            ++imode;
        }
    }
    if (imode == nm) st = LELPH_OK;
done:
    free(tmp);
    free(s);
    return st;
}
