/*
 * nc4w.c -- netCDF-4 layout written with HDF5 (see include/lelph_nc4.h).
 *
 * netCDF-4 stores a dimension that has no coordinate variable as an HDF5
 * dataset of the dimension's name and length, made a dimension scale whose
 * NAME is "This is a netCDF dimension but not a netCDF variable.%10d" and
 * which carries the int attribute _Netcdf4Dimid; a variable is a dataset
 * attached to one scale per axis (DIMENSION_LIST) with the int array
 * attribute _Netcdf4Coordinates of its dimension ids.  These are the
 * conventions a netCDF library uses to read
 * a file back; def_ncVar's behaviour (src/io/nc4_def_var.c:10-98) is kept:
 * dimensions are created on first use and checked for length afterwards,
 * storage is contiguous unless chunk sizes are given, no fill.
 */
#include <hdf5.h>
#include <hdf5_hl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lelph_hip.h"
#include "../../include/lelph_nc4.h"

#define LELPH_TRY_C(expr)         \
    do                            \
    {                             \
        int _s = (expr);          \
        if (_s != LELPH_OK)       \
            return _s;            \
    } while (0)

#define NC4_MAX_DIMS 64
#define NC4_MAX_VARS 256

struct nc4_dim
{
    char name[128];
    int64_t len;
    hid_t ds;
};

struct nc4_var
{
    char name[128];
    int xtype, rank;
    hid_t ds;
};

struct lelph_nc4
{
    hid_t file;
    int ndims, nvars;
    struct nc4_dim dims[NC4_MAX_DIMS];
    struct nc4_var vars[NC4_MAX_VARS];
};

static hid_t file_type(int xtype)
{
    switch (xtype)
    {
        case LELPH_NC_DOUBLE: return H5Tcopy(H5T_IEEE_F64LE);
        case LELPH_NC_FLOAT: return H5Tcopy(H5T_IEEE_F32LE);
        case LELPH_NC_INT: return H5Tcopy(H5T_STD_I32LE);
        case LELPH_NC_CHAR:
        {
            hid_t t = H5Tcopy(H5T_C_S1);
            H5Tset_size(t, 1);
            H5Tset_strpad(t, H5T_STR_NULLTERM);
            H5Tset_cset(t, H5T_CSET_ASCII);
            return t;
        }
        default: return -1;
    }
}

static hid_t mem_type(int xtype)
{
    switch (xtype)
    {
        case LELPH_NC_DOUBLE: return H5Tcopy(H5T_NATIVE_DOUBLE);
        case LELPH_NC_FLOAT: return H5Tcopy(H5T_NATIVE_FLOAT);
        case LELPH_NC_INT: return H5Tcopy(H5T_NATIVE_INT);
        case LELPH_NC_CHAR: return file_type(LELPH_NC_CHAR);
        default: return -1;
    }
}

static int put_int_attr(hid_t obj, const char* name, int n, const int* v)
{
    hsize_t d = (hsize_t)n;
    hid_t sp = n == 1 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &d, NULL);
    hid_t a = H5Acreate2(obj, name, H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, H5T_NATIVE_INT, v);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return e < 0 ? LELPH_ERR_ARG : LELPH_OK;
}

int lelph_nc4_create(const char* path, lelph_nc4** out)
{
    if (!path || !out) return LELPH_ERR_ARG;
    *out = NULL;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
    H5Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    H5Pset_attr_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
    H5Pset_fclose_degree(fapl, H5F_CLOSE_STRONG);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, fapl);
    H5Pclose(fcpl);
    H5Pclose(fapl);
    if (f < 0) return LELPH_ERR_ARG;
    lelph_nc4* nc = calloc(1, sizeof(*nc));
    if (!nc)
    {
        H5Fclose(f);
        return LELPH_ERR_ALLOC;
    }
    nc->file = f;
    *out = nc;
    return LELPH_OK;
}

static int find_dim(const lelph_nc4* nc, const char* name)
{
    for (int i = 0; i < nc->ndims; ++i)
        if (!strcmp(nc->dims[i].name, name)) return i;
    return -1;
}

static int find_var(const lelph_nc4* nc, const char* name)
{
    for (int i = 0; i < nc->nvars; ++i)
        if (!strcmp(nc->vars[i].name, name)) return i;
    return -1;
}

static int def_dim(lelph_nc4* nc, const char* name, int64_t len, int* id)
{
    if (nc->ndims >= NC4_MAX_DIMS || strlen(name) >= 128 || len < 1) return LELPH_ERR_ARG;
    hsize_t d = (hsize_t)len;
    hid_t sp = H5Screate_simple(1, &d, NULL);
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_attr_creation_order(dcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    H5Pset_fill_time(dcpl, H5D_FILL_TIME_NEVER);
    hid_t ds = H5Dcreate2(nc->file, name, H5T_IEEE_F32BE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Pclose(dcpl);
    H5Sclose(sp);
    if (ds < 0) return LELPH_ERR_ARG;
    char sname[96];
    snprintf(sname, sizeof(sname), "This is a netCDF dimension but not a netCDF variable.%10lld",
             (long long)len);
    const int did = nc->ndims;
    if (H5DSset_scale(ds, sname) < 0 || put_int_attr(ds, "_Netcdf4Dimid", 1, &did))
    {
        H5Dclose(ds);
        return LELPH_ERR_ARG;
    }
    struct nc4_dim* dm = &nc->dims[nc->ndims++];
    strcpy(dm->name, name);
    dm->len = len;
    dm->ds = ds;
    *id = did;
    return LELPH_OK;
}

int lelph_nc4_def_var(lelph_nc4* nc, const char* name, int xtype, int rank,
                      const int64_t* dims, const char* const* dim_names, const int64_t* chunks)
{
    if (!nc || !name || rank < 1 || rank > 32 || !dims || !dim_names) return LELPH_ERR_ARG;
    if (strlen(name) >= 128 || nc->nvars >= NC4_MAX_VARS || find_var(nc, name) >= 0 ||
        find_dim(nc, name) >= 0)
        return LELPH_ERR_ARG;
    int ids[32];
    hsize_t hd[32], hc[32];
    for (int i = 0; i < rank; ++i)
    {
        if (!dim_names[i] || dims[i] < 1) return LELPH_ERR_ARG;
        int id = find_dim(nc, dim_names[i]);
        if (id < 0)
            LELPH_TRY_C(def_dim(nc, dim_names[i], dims[i], &id));
        else if (nc->dims[id].len != dims[i])
            return LELPH_ERR_ARG; /* nc4_def_var.c:41-46 */
        ids[i] = id;
        hd[i] = (hsize_t)dims[i];
        if (chunks)
        {
            if (chunks[i] < 1 || chunks[i] > dims[i]) return LELPH_ERR_ARG;
            hc[i] = (hsize_t)chunks[i];
        }
    }
    hid_t ft = file_type(xtype);
    if (ft < 0) return LELPH_ERR_ARG;
    hid_t sp = H5Screate_simple(rank, hd, NULL);
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_attr_creation_order(dcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    H5Pset_fill_time(dcpl, H5D_FILL_TIME_NEVER);
    if (chunks)
        H5Pset_chunk(dcpl, rank, hc);
    else
        H5Pset_layout(dcpl, H5D_CONTIGUOUS);
    hid_t ds = H5Dcreate2(nc->file, name, ft, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Pclose(dcpl);
    H5Sclose(sp);
    H5Tclose(ft);
    if (ds < 0) return LELPH_ERR_ARG;
    for (int i = 0; i < rank; ++i)
        if (H5DSattach_scale(ds, nc->dims[ids[i]].ds, (unsigned)i) < 0)
        {
            H5Dclose(ds);
            return LELPH_ERR_ARG;
        }
    if (rank > 1 && put_int_attr(ds, "_Netcdf4Coordinates", rank, ids))
    {
        H5Dclose(ds);
        return LELPH_ERR_ARG;
    }
    struct nc4_var* v = &nc->vars[nc->nvars++];
    strcpy(v->name, name);
    v->xtype = xtype;
    v->rank = rank;
    v->ds = ds;
    return LELPH_OK;
}

int lelph_nc4_put_vara(lelph_nc4* nc, const char* name, const int64_t* start,
                       const int64_t* count, const void* data)
{
    if (!nc || !name || (!start) != (!count)) return LELPH_ERR_ARG;
    const int iv = find_var(nc, name);
    if (iv < 0) return LELPH_ERR_ARG;
    struct nc4_var* v = &nc->vars[iv];
    hid_t fs = H5Dget_space(v->ds);
    hsize_t dims[32];
    H5Sget_simple_extent_dims(fs, dims, NULL);
    hsize_t st[32], ct[32];
    hsize_t n = 1;
    for (int i = 0; i < v->rank; ++i)
    {
        st[i] = start ? (hsize_t)start[i] : 0;
        ct[i] = count ? (hsize_t)count[i] : dims[i];
        if (start && (start[i] < 0 || count[i] < 0 || st[i] + ct[i] > dims[i]))
        {
            H5Sclose(fs);
            return LELPH_ERR_ARG;
        }
        n *= ct[i];
    }
    if (n == 0)
    {
        H5Sclose(fs);
        return LELPH_OK; /* the no-op write of nc4_def_var.c:84-95 */
    }
    if (!data)
    {
        H5Sclose(fs);
        return LELPH_ERR_ARG;
    }
    H5Sselect_hyperslab(fs, H5S_SELECT_SET, st, NULL, ct, NULL);
    hid_t ms = H5Screate_simple(v->rank, ct, NULL);
    hid_t mt = mem_type(v->xtype);
    herr_t e = H5Dwrite(v->ds, mt, ms, fs, H5P_DEFAULT, data);
    H5Tclose(mt);
    H5Sclose(ms);
    H5Sclose(fs);
    return e < 0 ? LELPH_ERR_ARG : LELPH_OK;
}

int lelph_nc4_close(lelph_nc4* nc)
{
    if (!nc) return LELPH_ERR_ARG;
    int st = LELPH_OK;
    for (int i = 0; i < nc->nvars; ++i) H5Dclose(nc->vars[i].ds);
    for (int i = 0; i < nc->ndims; ++i) H5Dclose(nc->dims[i].ds);
    if (H5Fclose(nc->file) < 0) st = LELPH_ERR_ARG;
    free(nc);
    return st;
}

/* ------------------------------------------------------------------------
 * Reading (nc_open / nc_inq_var / nc_get_vara of a netCDF-4 file): a
 * variable is the HDF5 dataset of its name; values convert to the asked
 * memory type as netCDF converts them.
 * ---------------------------------------------------------------------- */

int lelph_nc4_open(const char* path, lelph_nc4** out)
{
    if (!path || !out) return LELPH_ERR_ARG;
    *out = NULL;
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return LELPH_ERR_ARG;
    lelph_nc4* nc = calloc(1, sizeof(*nc));
    if (!nc)
    {
        H5Fclose(f);
        return LELPH_ERR_ALLOC;
    }
    nc->file = f;
    *out = nc;
    return LELPH_OK;
}

int lelph_nc4_inq_var(lelph_nc4* nc, const char* name, int* xtype, int* rank, int64_t* dims)
{
    if (!nc || !name) return LELPH_ERR_ARG;
    if (H5Lexists(nc->file, name, H5P_DEFAULT) <= 0) return LELPH_ERR_ARG;
    hid_t ds = H5Dopen2(nc->file, name, H5P_DEFAULT);
    if (ds < 0) return LELPH_ERR_ARG;
    hid_t t = H5Dget_type(ds);
    int xt = -1;
    const size_t sz = H5Tget_size(t);
    switch (H5Tget_class(t))
    {
        case H5T_FLOAT: xt = sz == 8 ? LELPH_NC_DOUBLE : LELPH_NC_FLOAT; break;
        case H5T_INTEGER: xt = sz == 4 ? LELPH_NC_INT : -1; break;
        case H5T_STRING: xt = LELPH_NC_CHAR; break;
        default: break;
    }
    H5Tclose(t);
    hid_t sp = H5Dget_space(ds);
    const int r = H5Sget_simple_extent_ndims(sp);
    hsize_t d[32];
    int st = LELPH_OK;
    if (r < 0 || r > 32)
        st = LELPH_ERR_ARG;
    else
    {
        H5Sget_simple_extent_dims(sp, d, NULL);
        if (rank) *rank = r;
        if (dims)
            for (int i = 0; i < r; ++i) dims[i] = (int64_t)d[i];
        if (xtype) *xtype = xt;
    }
    H5Sclose(sp);
    H5Dclose(ds);
    return st;
}

int lelph_nc4_get_vara(lelph_nc4* nc, const char* name, const int64_t* start,
                       const int64_t* count, int mtype, void* out)
{
    if (!nc || !name || !out || (!start) != (!count)) return LELPH_ERR_ARG;
    if (mtype != LELPH_NC_DOUBLE && mtype != LELPH_NC_INT && mtype != LELPH_NC_FLOAT)
        return LELPH_ERR_ARG;
    if (H5Lexists(nc->file, name, H5P_DEFAULT) <= 0) return LELPH_ERR_ARG;
    hid_t ds = H5Dopen2(nc->file, name, H5P_DEFAULT);
    if (ds < 0) return LELPH_ERR_ARG;
    hid_t fs = H5Dget_space(ds);
    const int r = H5Sget_simple_extent_ndims(fs);
    int st = LELPH_ERR_ARG;
    hid_t ms = -1;
    if (r < 0 || r > 32) goto done;
    hsize_t d[32], st_[32], ct[32];
    H5Sget_simple_extent_dims(fs, d, NULL);
    for (int i = 0; i < r; ++i)
    {
        st_[i] = start ? (hsize_t)start[i] : 0;
        ct[i] = count ? (hsize_t)count[i] : d[i];
        if (start && (start[i] < 0 || count[i] < 1 || st_[i] + ct[i] > d[i])) goto done;
    }
    if (r > 0)
    {
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, st_, NULL, ct, NULL);
        ms = H5Screate_simple(r, ct, NULL);
    }
    else
        ms = H5Screate(H5S_SCALAR);
    hid_t mt = mem_type(mtype);
    const herr_t e = H5Dread(ds, mt, ms, fs, H5P_DEFAULT, out);
    H5Tclose(mt);
    st = e < 0 ? LELPH_ERR_ARG : LELPH_OK;
done:
    if (ms >= 0) H5Sclose(ms);
    H5Sclose(fs);
    H5Dclose(ds);
    return st;
}
