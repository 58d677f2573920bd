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

static int parse_dyn_xml(const char* path, int natom, double* qpt, double* masses,
                         double* dynd);
static char* slurp(const char* path, long* size);

/* first non-blank line holds "<?xml" (is_dyn_xml, src/io/qe/read_dyn.c:611-641) */
static int is_dyn_xml(FILE* fp)
{
    char line[1024];
    int xml = 0;
    while (fgets(line, sizeof(line), fp))
    {
        const char* p = line;
        while (*p && isspace((unsigned char)*p)) ++p;
        if (!*p) continue;
        xml = strstr(p, "<?xml") != NULL;
        break;
    }
    rewind(fp);
    return xml;
}

/* Parses the first dynamical matrix of a dyn file, the old text format or
 * the XML one: qpt (3), masses (natom) and dyn (nm x nm, row major,
 * D[(a,x)][(b,y)] divided by sqrt(Ma Mb)).  Returns LELPH_OK or
 * LELPH_ERR_ARG. */
int lelph_qe_parse_dyn(const char* path, int natom, double* qpt, double* masses, double* dynd)
{
    if (!path || natom < 1 || !qpt || !masses || !dynd) return LELPH_ERR_ARG;
    FILE* fp = fopen(path, "r");
    if (!fp) return LELPH_ERR_ARG;
    if (is_dyn_xml(fp))
    {
        fclose(fp);
        return parse_dyn_xml(path, natom, qpt, masses, dynd);
    }
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
    long sz = 0;
    char* s = slurp(path, &sz);
    if (!s) return LELPH_ERR_ARG;
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

/* ------------------------------------------------------------------------
 * Minimal XML access for QE's data-file-schema.xml, tensors.xml and UPF:
 * elements are found by name among the descendants of a range, siblings
 * counted in document order (a match's own content is skipped when
 * counting), attributes read from the start tag.  Comments and
 * processing instructions are skipped.
 * ---------------------------------------------------------------------- */

struct xspan
{
    const char *tb, *te; /* start tag: '<' .. '>' inclusive of '<', te at '>' */
    const char *b, *e;   /* content */
};

static int name_end(char c)
{
    return c == '>' || c == '/' || isspace((unsigned char)c);
}

static const char* skip_special(const char* p, const char* e)
{
    if (p + 4 <= e && !strncmp(p, "<!--", 4))
    {
        for (const char* q = p + 4; q + 3 <= e; ++q)
            if (!strncmp(q, "-->", 3)) return q + 3;
        return e;
    }
    if (p + 2 <= e && (p[1] == '?' || p[1] == '!'))
    {
        const char* q = memchr(p, '>', (size_t)(e - p));
        return q ? q + 1 : e;
    }
    return NULL;
}

/* end of the element whose start tag ends at gt (not self-closing):
 * the '<' of the matching close tag, nesting of the same name counted */
static const char* close_of(const char* gt, const char* e, const char* tag, size_t n)
{
    int depth = 1;
    for (const char* p = gt + 1; p < e; ++p)
    {
        if (*p != '<') continue;
        const char* sk = skip_special(p, e);
        if (sk)
        {
            p = sk - 1;
            continue;
        }
        if (p[1] == '/' && p + 2 + n <= e && !strncmp(p + 2, tag, n) &&
            (p + 2 + n == e || p[2 + n] == '>' || isspace((unsigned char)p[2 + n])))
        {
            if (--depth == 0) return p;
        }
        else if (p + 1 + n <= e && !strncmp(p + 1, tag, n) && p + 1 + n < e && name_end(p[1 + n]))
        {
            const char* g2 = memchr(p, '>', (size_t)(e - p));
            if (!g2) return NULL;
            if (g2[-1] != '/') ++depth;
            p = g2;
        }
    }
    return NULL;
}

/* the idx-th element named tag in [b, e) */
static int xml_nth(const char* b, const char* e, const char* tag, int idx, struct xspan* out)
{
    const size_t n = strlen(tag);
    int seen = 0;
    for (const char* p = b; p < e; ++p)
    {
        if (*p != '<') continue;
        const char* sk = skip_special(p, e);
        if (sk)
        {
            p = sk - 1;
            continue;
        }
        if (p + 1 + n >= e || strncmp(p + 1, tag, n) || !name_end(p[1 + n])) continue;
        const char* gt = memchr(p, '>', (size_t)(e - p));
        if (!gt) return 0;
        const char *cb = gt + 1, *ce = gt + 1, *after = gt + 1;
        if (gt[-1] != '/')
        {
            ce = close_of(gt, e, tag, n);
            if (!ce) return 0;
            after = ce + n + 3;
        }
        if (seen++ == idx)
        {
            out->tb = p;
            out->te = gt;
            out->b = cb;
            out->e = ce;
            return 1;
        }
        p = after - 1;
    }
    return 0;
}

/* chain of names, each the first match inside the previous one; a name
 * "tag#k" takes the k-th match */
static int xml_path(const char* b, const char* e, struct xspan* out, const char* const* names,
                    int nn)
{
    struct xspan s = {b, b, b, e};
    for (int i = 0; i < nn; ++i)
    {
        char tag[96];
        int k = 0;
        const char* h = strchr(names[i], '#');
        size_t len = h ? (size_t)(h - names[i]) : strlen(names[i]);
        if (len >= sizeof(tag)) return 0;
        memcpy(tag, names[i], len);
        tag[len] = 0;
        if (h) k = atoi(h + 1);
        if (!xml_nth(s.b, s.e, tag, k, &s)) return 0;
    }
    *out = s;
    return 1;
}

#define XPATH(b, e, out, ...)                                                                  \
    xml_path((b), (e), (out), (const char* const[]){__VA_ARGS__},                              \
             (int)(sizeof((const char* const[]){__VA_ARGS__}) / sizeof(const char*)))

/* attribute name of the start tag of s, copied to out (NUL-terminated) */
static int xml_attr(const struct xspan* s, const char* name, char* out, size_t cap)
{
    const size_t n = strlen(name);
    for (const char* p = s->tb; p + n < s->te; ++p)
    {
        if (!isspace((unsigned char)p[0]) || strncmp(p + 1, name, n)) continue;
        const char* q = p + 1 + n;
        while (q < s->te && isspace((unsigned char)*q)) ++q;
        if (q >= s->te || *q != '=') continue;
        ++q;
        while (q < s->te && isspace((unsigned char)*q)) ++q;
        if (q >= s->te || (*q != '"' && *q != '\'')) continue;
        const char quote = *q++;
        const char* z = memchr(q, quote, (size_t)(s->te - q));
        if (!z || (size_t)(z - q) >= cap) return 0;
        memcpy(out, q, (size_t)(z - q));
        out[z - q] = 0;
        return 1;
    }
    return 0;
}

/* content of s as a NUL-terminated heap string */
static char* xml_text(const struct xspan* s)
{
    const size_t len = (size_t)(s->e - s->b);
    char* t = malloc(len + 1);
    if (!t) return NULL;
    memcpy(t, s->b, len);
    t[len] = 0;
    return t;
}

/* numbers of the content: count found (parse_floats_from_string) */
static int64_t xml_floats(const struct xspan* s, double* out, int64_t n)
{
    char* t = xml_text(s);
    if (!t) return -1;
    const int64_t c = qe_floats(t, out, n);
    free(t);
    return c;
}

/* content contains "true", case-insensitively */
static int xml_true(const struct xspan* s)
{
    for (const char* p = s->b; p + 4 <= s->e; ++p)
        if (tolower((unsigned char)p[0]) == 't' && tolower((unsigned char)p[1]) == 'r' &&
            tolower((unsigned char)p[2]) == 'u' && tolower((unsigned char)p[3]) == 'e')
            return 1;
    return 0;
}

static char* slurp(const char* path, long* size)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) return NULL;
    if (fseek(fp, 0, SEEK_END) != 0)
    {
        fclose(fp);
        return NULL;
    }
    const long sz = ftell(fp);
    if (sz < 0 || fseek(fp, 0, SEEK_SET) != 0)
    {
        fclose(fp);
        return NULL;
    }
    char* s = malloc((size_t)sz + 1);
    if (s && fread(s, 1, (size_t)sz, fp) != (size_t)sz)
    {
        free(s);
        s = NULL;
    }
    fclose(fp);
    if (s)
    {
        s[sz] = 0;
        *size = sz;
    }
    return s;
}

/* parse_upf (src/io/qe/parse_upf.c:28-336): UPF-1 (first line <PP_INFO>)
 * or UPF-2 (<UPF version ...>), norm-conserving only; z_valence and
 * mesh_size from the header, PP_MESH/PP_R, PP_MESH/PP_RAB and PP_LOCAL
 * with exactly mesh_size numbers each. */
int lelph_parse_upf(const char* path, int64_t* ngrid, double* zval, double* r, double* rab,
                    double* vloc)
{
    if (!path || !ngrid || !zval) return LELPH_ERR_ARG;
    long sz = 0;
    char* s = slurp(path, &sz);
    if (!s) return LELPH_ERR_ARG;
    const char* e = s + sz;
    int st = LELPH_ERR_ARG;
    const char* nl = memchr(s, '\n', (size_t)sz);
    const size_t l1 = nl ? (size_t)(nl - s) : (size_t)sz;
    char first[128];
    snprintf(first, sizeof(first), "%.*s", (int)(l1 < 127 ? l1 : 127), s);
    int64_t n = -1;
    double z = 0.0;
    if (strstr(first, "<UPF version"))
    {
        struct xspan h;
        char buf[256];
        if (!XPATH(s, e, &h, "PP_HEADER")) goto done;
        if (!xml_attr(&h, "z_valence", buf, sizeof(buf))) goto done;
        z = atof(buf);
        if (!xml_attr(&h, "mesh_size", buf, sizeof(buf))) goto done;
        n = atoll(buf);
        if (!xml_attr(&h, "pseudo_type", buf, sizeof(buf)) || strcmp(buf, "NC"))
        {
            st = LELPH_ERR_UNSUPPORTED;
            goto done;
        }
    }
    else if (strstr(first, "<PP_INFO>"))
    {
        /* header lines after <PP_HEADER>: version, element, type, nlcc,
         * functional, z_valence, total energy, cutoffs, lmax, mesh */
        const char* p = strstr(s, "<PP_HEADER>");
        if (!p) goto done;
        p = strchr(p, '\n');
        char line[1024];
        for (int k = 0; k < 10; ++k)
        {
            if (!p) goto done;
            ++p;
            const char* q = strchr(p, '\n');
            const size_t len = q ? (size_t)(q - p) : strlen(p);
            snprintf(line, sizeof(line), "%.*s", (int)(len < 1023 ? len : 1023), p);
            if (k == 2 && !strstr(line, "NC"))
            {
                st = LELPH_ERR_UNSUPPORTED;
                goto done;
            }
            if (k == 5) z = strtod(line, NULL);
            if (k == 9) n = strtoll(line, NULL, 10);
            p = q;
        }
    }
    else
    {
        st = LELPH_ERR_UNSUPPORTED;
        goto done;
    }
    if (n < 1) goto done;
    *ngrid = n;
    *zval = z;
    if (r || rab || vloc)
    {
        struct xspan m;
        if (r && (!XPATH(s, e, &m, "PP_MESH", "PP_R") || xml_floats(&m, r, n) != n)) goto done;
        if (rab && (!XPATH(s, e, &m, "PP_MESH", "PP_RAB") || xml_floats(&m, rab, n) != n))
            goto done;
        if (vloc && (!XPATH(s, e, &m, "PP_LOCAL") || xml_floats(&m, vloc, n) != n)) goto done;
    }
    st = LELPH_OK;
done:
    free(s);
    return st;
}

/* parse_qexml (src/io/qe/read_qexml.c:20-643) */
int lelph_parse_qexml(const char* path, lelph_qexml_info* info, int* atom_type,
                      double* atomic_pos, char* pseudo_files, double* sym_mats, double* sym_tau,
                      int* trevs)
{
    if (!path || !info) return LELPH_ERR_ARG;
    long sz = 0;
    char* s = slurp(path, &sz);
    if (!s) return LELPH_ERR_ARG;
    const char* e = s + sz;
    int st = LELPH_ERR_ARG;
    struct xspan x, sp;
    char buf[1024];
    double* tmp = NULL;
    lelph_qexml_info I;
    memset(&I, 0, sizeof(I));
    I.dimension = '3';
    I.nmag = 1;
    I.nspinor = 1;
    if (XPATH(s, e, &x, "output", "boundary_conditions", "assume_isolated"))
    {
        const char* p = x.b;
        while (p < x.e && isspace((unsigned char)*p)) ++p;
        if (p + 2 <= x.e && !strncmp(p, "2D", 2)) I.dimension = '2';
    }
    if (!XPATH(s, e, &x, "input", "control_variables", "pseudo_dir")) goto done;
    snprintf(I.pseudo_dir, sizeof(I.pseudo_dir), "%.*s", (int)(x.e - x.b), x.b);
    if (!XPATH(s, e, &sp, "output", "atomic_species") || !xml_attr(&sp, "ntyp", buf, 64)) goto done;
    I.ntype = atoll(buf);
    if (I.ntype < 1) goto done;
    char(*names)[16] = calloc((size_t)I.ntype, 16);
    if (!names) goto done;
    for (int t = 0; t < I.ntype; ++t)
    {
        struct xspan spc;
        if (!xml_nth(sp.b, sp.e, "species", t, &spc) || !xml_attr(&spc, "name", buf, sizeof(buf)))
        {
            free(names);
            goto done;
        }
        const char* nm = buf;
        while (isspace((unsigned char)*nm)) ++nm;
        if (strlen(nm) > 15)
        {
            free(names);
            goto done;
        }
        strcpy(names[t], nm);
        if (!xml_nth(spc.b, spc.e, "pseudo_file", 0, &x))
        {
            free(names);
            goto done;
        }
        if (pseudo_files)
            snprintf(pseudo_files + 256 * t, 256, "%.*s", (int)(x.e - x.b), x.b);
    }
    struct xspan as;
    if (!XPATH(s, e, &as, "output", "atomic_structure") || !xml_attr(&as, "nat", buf, 64))
    {
        free(names);
        goto done;
    }
    I.natom = atoll(buf);
    if (I.natom < 1 || !xml_attr(&as, "alat", buf, 64))
    {
        free(names);
        goto done;
    }
    I.alat[0] = I.alat[1] = I.alat[2] = atof(buf);
    for (int a = 0; a < I.natom; ++a)
    {
        struct xspan ap, at;
        if (!xml_nth(as.b, as.e, "atomic_positions", 0, &ap) || !xml_nth(ap.b, ap.e, "atom", a, &at) ||
            !xml_attr(&at, "name", buf, sizeof(buf)))
        {
            free(names);
            goto done;
        }
        const char* nm = buf;
        while (isspace((unsigned char)*nm)) ++nm;
        int it = -1;
        for (int t = 0; t < I.ntype; ++t)
            if (!strcmp(names[t], nm))
            {
                it = t;
                break;
            }
        if (it < 0)
        {
            free(names);
            goto done;
        }
        if (atom_type) atom_type[a] = it;
        if (atomic_pos && xml_floats(&at, atomic_pos + 3 * a, 3) != 3)
        {
            free(names);
            goto done;
        }
    }
    free(names);
    if (!XPATH(s, e, &x, "output", "basis_set", "fft_grid")) goto done;
    const char* nr[3] = {"nr1", "nr2", "nr3"};
    for (int i = 0; i < 3; ++i)
    {
        if (!xml_attr(&x, nr[i], buf, 64)) goto done;
        I.fft_dims[i] = atoll(buf);
    }
    const char* an[3] = {"a1", "a2", "a3"};
    for (int i = 0; i < 3; ++i)
    {
        double v[3];
        if (!XPATH(as.b, as.e, &x, "cell", an[i]) || xml_floats(&x, v, 3) != 3) goto done;
        for (int k = 0; k < 3; ++k) I.lat_vec[3 * k + i] = v[k]; /* a[:, i] */
    }
    if (!XPATH(s, e, &x, "output", "magnetization", "spinorbit")) goto done;
    I.is_soc_present = xml_true(&x);
    if (!XPATH(s, e, &x, "output", "magnetization", "lsda")) goto done;
    const int lsda = xml_true(&x);
    I.ph_tim_rev = 1;
    if (!XPATH(s, e, &x, "input", "symmetry_flags", "noinv")) goto done;
    if (xml_true(&x)) I.ph_tim_rev = 0;
    if (lsda)
        I.nmag = 2;
    else
    {
        if (!XPATH(s, e, &x, "output", "magnetization", "noncolin")) goto done;
        int mag = 0;
        if (xml_true(&x))
        {
            I.nspinor = 2;
            if (!XPATH(s, e, &x, "output", "magnetization", "do_magnetization")) goto done;
            mag = xml_true(&x);
        }
        I.nmag = mag ? 4 : 1;
        if (mag) I.ph_tim_rev = 0;
    }
    struct xspan sy;
    if (!XPATH(s, e, &sy, "output", "symmetries") || !xml_nth(sy.b, sy.e, "nsym", 0, &x)) goto done;
    {
        char* t = xml_text(&x);
        if (!t) goto done;
        I.nsym = atoi(t);
        free(t);
    }
    if (I.nsym < 1) goto done;
    const int64_t ns = I.nsym;
    tmp = malloc(sizeof(double) * 13 * (size_t)ns);
    if (!tmp) goto done;
    double *tm = tmp, *tt = tmp + 9 * ns, *tr = tmp + 12 * ns;
    int inversion = 0, magsym = 0;
    for (int64_t k = 0; k < ns; ++k)
    {
        struct xspan y;
        if (!xml_nth(sy.b, sy.e, "symmetry", (int)k, &y)) goto done;
        if (!xml_nth(y.b, y.e, "rotation", 0, &x) || xml_floats(&x, tm + 9 * k, 9) != 9) goto done;
        if (!xml_nth(y.b, y.e, "fractional_translation", 0, &x) ||
            xml_floats(&x, tt + 3 * k, 3) != 3)
            goto done;
        if (!xml_nth(y.b, y.e, "info", 0, &x)) goto done;
        tr[k] = 0;
        if (xml_attr(&x, "time_reversal", buf, 64))
        {
            for (char* p = buf; *p; ++p) *p = (char)tolower((unsigned char)*p);
            if (strstr(buf, "true"))
            {
                magsym = 1;
                I.ph_tim_rev = 1;
                tr[k] = 1;
            }
        }
        double sum = 0;
        for (int i = 0; i < 9; ++i)
        {
            const double d = (i % 4 == 0 ? 1.0 : 0.0) + tm[9 * k + i];
            sum += d * d;
        }
        if (sqrt(sum) < LELPH_EPS_DYN) inversion = 1;
    }
    /* magnetic groups: operations without time reversal first, then those
     * with it, each half in file order */
    int64_t nt = 0, nn = 0;
    if (magsym)
    {
        if (ns % 2) goto done;
        for (int64_t k = 0; k < ns; ++k)
            if (tr[k])
                ++nt;
            else
                ++nn;
        if (nt != nn) goto done;
    }
    if (sym_mats || sym_tau || trevs)
    {
        int64_t it = 0, in = 0;
        for (int64_t k = 0; k < ns; ++k)
        {
            const int64_t d = magsym ? (tr[k] ? ns / 2 + it++ : in++) : k;
            if (sym_mats) memcpy(sym_mats + 9 * d, tm + 9 * k, 9 * sizeof(double));
            if (sym_tau) memcpy(sym_tau + 3 * d, tt + 3 * k, 3 * sizeof(double));
            if (trevs) trevs[d] = tr[k] != 0;
        }
    }
    I.ph_mag_symm = magsym;
    if (inversion && !magsym) I.ph_tim_rev = 0;
    *info = I;
    st = LELPH_OK;
done:
    free(tmp);
    free(s);
    return st;
}

/* read_ph_tensors_qe + read_quadrupole_fmt (src/io/qe/read_tensors.c):
 * present bit 0 epsilon (3x3), bit 1 Born charges (natom, 3, 3), bit 2
 * quadrupoles (natom, 3, 3, 3); both stored transposed in tensors.xml.  A
 * missing file is not an error (nothing present). */
int lelph_read_ph_tensors_qe(const char* tensors_xml, const char* quadrupole_fmt, int natom,
                             double* epsilon, double* zborn, double* qpole, int* present)
{
    if (!present || natom < 1) return LELPH_ERR_ARG;
    *present = 0;
    long sz = 0;
    char* s = tensors_xml ? slurp(tensors_xml, &sz) : NULL;
    if (s)
    {
        const char* e = s + sz;
        struct xspan x;
        int st = LELPH_ERR_ARG;
        if (!XPATH(s, e, &x, "EF_TENSORS", "DONE_EFFECTIVE_CHARGE_EU")) goto tdone;
        const char* p = x.b;
        while (p < x.e && isspace((unsigned char)*p)) ++p;
        const int zeu = p < x.e && tolower((unsigned char)*p) == 't';
        if (!XPATH(s, e, &x, "EF_TENSORS", "DONE_ELECTRIC_FIELD")) goto tdone;
        p = x.b;
        while (p < x.e && isspace((unsigned char)*p)) ++p;
        const int eps = p < x.e && tolower((unsigned char)*p) == 't';
        if (eps)
        {
            double t[9];
            if (!XPATH(s, e, &x, "EF_TENSORS", "DIELECTRIC_CONSTANT") || xml_floats(&x, t, 9) != 9)
                goto tdone;
            if (epsilon)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) epsilon[3 * i + j] = t[3 * j + i];
            *present |= 1;
            if (zeu)
            {
                double* z = malloc(sizeof(double) * 9 * (size_t)natom);
                if (!z) goto tdone;
                if (!XPATH(s, e, &x, "EF_TENSORS", "EFFECTIVE_CHARGES_EU") ||
                    xml_floats(&x, z, 9 * (int64_t)natom) != 9 * (int64_t)natom)
                {
                    free(z);
                    goto tdone;
                }
                if (zborn)
                    for (int a = 0; a < natom; ++a)
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                zborn[9 * a + 3 * i + j] = z[9 * a + 3 * j + i];
                free(z);
                *present |= 2;
            }
        }
        st = LELPH_OK;
    tdone:
        free(s);
        if (st != LELPH_OK) return st;
    }
    FILE* fp = quadrupole_fmt ? fopen(quadrupole_fmt, "r") : NULL;
    if (!fp) return LELPH_OK;
    char line[1024];
    int st = LELPH_ERR_ARG;
    int64_t nd = 0;
    double* q = calloc(27 * (size_t)natom, sizeof(double));
    if (!q || !fgets(line, sizeof(line), fp)) goto qdone;
    while (fgets(line, sizeof(line), fp))
    {
        const char* p = line;
        while (*p && isspace((unsigned char)*p)) ++p;
        if (!*p) continue;
        if (nd >= 3 * (int64_t)natom) goto qdone;
        long long ia, id;
        double xx, yy, zz, yz, xz, xy;
        if (sscanf(line, "%lld %lld %lf %lf %lf %lf %lf %lf", &ia, &id, &xx, &yy, &zz, &yz, &xz,
                   &xy) != 8)
            goto qdone;
        --ia;
        --id;
        if (ia < 0 || ia >= natom || id < 0 || id > 2) goto qdone;
        double* Q = q + 27 * ia + id; /* Q[a][x][y][dir] */
        Q[0] = xx;
        Q[12] = yy;
        Q[24] = zz;
        Q[3] = Q[9] = xy;
        Q[6] = Q[18] = xz;
        Q[15] = Q[21] = yz;
        ++nd;
    }
    if (nd != 3 * (int64_t)natom) goto qdone;
    if (qpole) memcpy(qpole, q, sizeof(double) * 27 * (size_t)natom);
    *present |= 4;
    st = LELPH_OK;
qdone:
    free(q);
    fclose(fp);
    return st;
}

/* read_qpts_qe (src/io/qe/read_dyn.c:551-609): q grid and iBZ q points of
 * dyn0 (cartesian, 2 pi / alat units); qpts NULL returns the counts. */
int lelph_read_qpts_qe(const char* dyn0, int64_t* nq_ibz, int64_t* nq_bz, double* qpts)
{
    if (!dyn0 || !nq_ibz || !nq_bz) return LELPH_ERR_ARG;
    FILE* fp = fopen(dyn0, "r");
    if (!fp) return LELPH_ERR_ARG;
    char line[1024];
    long long g[3], n;
    int st = LELPH_ERR_ARG;
    if (!fgets(line, sizeof(line), fp) || sscanf(line, "%lld %lld %lld", g, g + 1, g + 2) != 3)
        goto done;
    if (g[0] < 1 || g[1] < 1 || g[2] < 1) goto done;
    if (!fgets(line, sizeof(line), fp) || sscanf(line, "%lld", &n) != 1 || n < 1) goto done;
    *nq_ibz = n;
    *nq_bz = g[0] * g[1] * g[2];
    for (long long i = 0; qpts && i < n; ++i)
    {
        if (!fgets(line, sizeof(line), fp) ||
            sscanf(line, "%lf %lf %lf", qpts + 3 * i, qpts + 3 * i + 1, qpts + 3 * i + 2) != 3)
            goto done;
    }
    st = LELPH_OK;
done:
    fclose(fp);
    return st;
}

/* read_dyn_xml (src/io/qe/read_dyn.c:306-548): GEOMETRY_INFO's
 * NUMBER_OF_TYPES, NUMBER_OF_ATOMS, MASS.t (amu, times 911.444243096 to the
 * Rydberg mass unit) and ATOM.i's INDEX (type), then DYNAMICAL_MAT_.1's
 * Q_POINT and PHI.a.b: 18 numbers with the first cartesian index fastest,
 * element (x, y) = phi[2x + 6y] + i phi[2x + 6y + 1]. */
static int parse_dyn_xml(const char* path, int natom, double* qpt, double* masses,
                         double* dynd)
{
    long sz = 0;
    char* s = slurp(path, &sz);
    if (!s) return LELPH_ERR_ARG;
    const char* e = s + sz;
    int st = LELPH_ERR_ARG;
    double* tm = NULL;
    struct xspan g, x;
    char tag[64], buf[64];
    const int nm = 3 * natom;
    if (!XPATH(s, e, &g, "GEOMETRY_INFO")) goto done;
    if (!xml_nth(g.b, g.e, "NUMBER_OF_TYPES", 0, &x)) goto done;
    {
        double v;
        if (xml_floats(&x, &v, 1) != 1) goto done;
        const int nt = (int)rint(v);
        if (!xml_nth(g.b, g.e, "NUMBER_OF_ATOMS", 0, &x) || xml_floats(&x, &v, 1) != 1) goto done;
        if ((int)rint(v) != natom || nt < 1) goto done;
        tm = malloc(sizeof(double) * (size_t)nt);
        if (!tm) goto done;
        for (int t = 0; t < nt; ++t)
        {
            snprintf(tag, sizeof(tag), "MASS.%d", t + 1);
            if (!xml_nth(g.b, g.e, tag, 0, &x) || xml_floats(&x, &v, 1) != 1) goto done;
            tm[t] = v * 911.444243096;
        }
        for (int a = 0; a < natom; ++a)
        {
            snprintf(tag, sizeof(tag), "ATOM.%d", a + 1);
            if (!xml_nth(g.b, g.e, tag, 0, &x) || !xml_attr(&x, "INDEX", buf, sizeof(buf)))
                goto done;
            const int it = atoi(buf) - 1;
            if (it < 0 || it >= nt) goto done;
            masses[a] = tm[it];
        }
    }
    struct xspan dm;
    if (!XPATH(s, e, &dm, "DYNAMICAL_MAT_.1")) goto done;
    if (!xml_nth(dm.b, dm.e, "Q_POINT", 0, &x) || xml_floats(&x, qpt, 3) != 3) goto done;
    for (int a = 0; a < natom; ++a)
        for (int b = 0; b < natom; ++b)
        {
            double phi[18];
            snprintf(tag, sizeof(tag), "PHI.%d.%d", a + 1, b + 1);
            if (!xml_nth(dm.b, dm.e, tag, 0, &x) || xml_floats(&x, phi, 18) != 18) goto done;
            const double inv = 1.0 / sqrt(masses[a] * masses[b]);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                {
                    const int64_t o = 2 * ((int64_t)(3 * a + i) * nm + 3 * b + j);
                    dynd[o] = phi[2 * i + 6 * j] * inv;
                    dynd[o + 1] = phi[2 * i + 6 * j + 1] * inv;
                }
        }
    st = LELPH_OK;
done:
    free(tm);
    free(s);
    return st;
}
