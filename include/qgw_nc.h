/*
 * qgw_nc.h -- netCDF classic-format output and restart input of the qgw
 * driver (libqgw_host.so), written in-house: the reference links libnetcdf
 * (src/netcdf_io.h:1), which this build does not need.
 *
 * Files are netCDF classic (CDF-1, big-endian, 32-bit offsets), the format
 * libnetcdf's nc_create(..., NC_CLOBBER, ...) writes, readable by any netCDF
 * reader.  The reader also accepts the 64-bit-offset variant (CDF-2).
 */
#ifndef QGW_NC_H
#define QGW_NC_H

#include "qgw_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qgw_nc qgw_nc;

/* create_nc (src/netcdf_io.h:38-160): dimensions t (unlimited), level (nl),
 * y, x; float coordinate variables t, y, x (i*Delta) and level (0..nl-1);
 * float record variables psi and q (t, level, y, x).  y and x span the full
 * storage grid of a bounded run (n+2, m+2 points, ring included) or the
 * interior of a periodic run (NX, NY).  Returns NULL with *status set on
 * failure. */
qgw_nc *qgwh_nc_create(const char *path, const qgw_params *p, int *status);

/* write_nc (src/netcdf_io.h:171-277): appends one record, time t and psi, q
 * (full storage arrays nl*(m+2)*(n+2), reference idx layout) as floats; the
 * record count in the header is updated and the file flushed. */
int qgwh_nc_write(qgw_nc *nc, double t, const double *psi, const double *q);

/* Records written so far. */
int qgwh_nc_records(qgw_nc *nc);

int qgwh_nc_close(qgw_nc *nc);

/* read_nc (src/netcdf_io.h:286-386) of a restart file: the variables named
 * "q" (nl levels) and "topo" (one level) at record `rec` (record variables) or
 * their single slab (fixed variables), float or double, into full storage
 * arrays (the ring of a periodic run is left as is; a bounded run's file holds
 * the ring).  Either output may be NULL; a variable absent from the file
 * leaves its output untouched.  *found (may be NULL) gets bit 0 for q and bit
 * 1 for topo. */
int qgwh_nc_read(const char *path, const qgw_params *p, int rec, double *q, double *topo, int *found);

/* read_nc of one named variable with `layers` levels (src/netcdf_io.h:286-386;
 * the 4-D forcing reads "q_forcing_3d" with nl levels this way,
 * src/forcing.h:235-236,262-263) at record `rec` into the storage array out,
 * same rules as qgwh_nc_read.  *found (may be NULL) is 1 when the variable
 * exists. */
int qgwh_nc_read_var(const char *path, const qgw_params *p, const char *name, int layers, int rec, double *out,
                     int *found);

#ifdef __cplusplus
}
#endif

#endif /* QGW_NC_H */
