/*
 * lelph_nc4.h -- netCDF-4 writer for ndb.elph / ndb.Dmats without libnetcdf.
 *
 * The reference writes its databases through netCDF-4 (def_ncVar,
 * src/io/nc4_def_var.c:10-98, and nc_put_vara).  This host library writes
 * the same netCDF-4 layout directly with HDF5: one dimension-scale dataset
 * per named dimension (netCDF's "dimension without a variable" form, with
 * its _Netcdf4Dimid), variables attached to their scales (DIMENSION_LIST,
 * _Netcdf4Coordinates), link and attribute creation order tracked, no fill,
 * contiguous or chunked storage as def_ncVar asks.  A netCDF reader
 * (netCDF4-python, yambopy) sees the reference's variable names, dimension
 * names, shapes, types and chunking.
 *
 * It also reads netCDF-4 files (the Yambo SAVE databases) the same way.
 *
 * Plain C; every call returns LELPH_OK (0) or a lelph_status code.
 */
#ifndef LELPH_NC4_H
#define LELPH_NC4_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* netCDF external types used by the reference's databases */
enum lelph_nc_type
{
    LELPH_NC_CHAR = 2,  /* NC_CHAR   */
    LELPH_NC_INT = 4,   /* NC_INT    */
    LELPH_NC_FLOAT = 5, /* NC_FLOAT  */
    LELPH_NC_DOUBLE = 6 /* NC_DOUBLE */
};

typedef struct lelph_nc4 lelph_nc4;

/* nc_create(path, NC_NETCDF4 | NC_CLOBBER) + ncsetfill(NC_NOFILL)
 * (src/elph/elph_driver.c:140-153) */
int lelph_nc4_create(const char* path, lelph_nc4** out);

/* def_ncVar (src/io/nc4_def_var.c:10-98): defines the variable and, on
 * first use, each named dimension; a dimension used again must have the
 * same length.  chunks NULL: contiguous, else chunk sizes (rank). */
int lelph_nc4_def_var(lelph_nc4* nc, const char* name, int xtype, int rank,
                      const int64_t* dims, const char* const* dim_names,
                      const int64_t* chunks);

/* nc_put_vara: writes the hyperslab start/count (rank each; NULL start and
 * count write the whole variable) from data in memory order, C layout,
 * in the variable's own type (double, float, int or char). */
int lelph_nc4_put_vara(lelph_nc4* nc, const char* name, const int64_t* start,
                       const int64_t* count, const void* data);

/* nc_close (of a file created or opened) */
int lelph_nc4_close(lelph_nc4* nc);

/* nc_open(path, NC_NOWRITE) of a netCDF-4 file, as the Yambo SAVE reader
 * opens ns.db1, ndb.kindx, ns.kb_pp_pwscf* and ns.wf_fragments_*
 * (src/io/get_save_data.c:108-118,164-170,454-460,566-573,701-707). */
int lelph_nc4_open(const char* path, lelph_nc4** out);

/* nc_inq_varid + nc_inq_vartype + dimension lengths: xtype (lelph_nc_type,
 * -1 for other types), rank and dims (up to 32) of variable name. */
int lelph_nc4_inq_var(lelph_nc4* nc, const char* name, int* xtype, int* rank, int64_t* dims);

/* nc_get_vara / nc_get_var (quick_read, quick_read_sub,
 * src/io/get_save_data.c:729-770): the hyperslab start/count (NULL: the
 * whole variable) converted to mem_type (LELPH_NC_DOUBLE, LELPH_NC_FLOAT
 * or LELPH_NC_INT), C order. */
int lelph_nc4_get_vara(lelph_nc4* nc, const char* name, const int64_t* start,
                       const int64_t* count, int mem_type, void* out);

#ifdef __cplusplus
}
#endif

#endif /* LELPH_NC4_H */
