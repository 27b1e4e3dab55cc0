/* gatt_flow_io.h -- flow files of the time loop (flow_%010d.mat) without the MATLAB API.
 *
 * The reference writes and reads them through MATLAB's MAT-file library
 * (source/Fortran/readWriteMat.F90:1-260, matOpen(..., 'w7.3')): a MAT 7.3 file
 * is an HDF5 file behind a 512-byte MATLAB userblock, with one dataset per
 * variable carrying the attribute MATLAB_class = "double".  These functions
 * read and write that layout with libhdf5 (host code, no GPU).  Arrays are
 * nx*ny*nz doubles, x fastest (the Fortran order of the state); a grid with
 * nz == 1 is stored as the 2-D matrix MATLAB would store.
 *
 * Return codes are those of gatt_dns.h (0 = ok, negative = error); the last
 * error message is gatt_flow_last_error(). */
#ifndef GATT_FLOW_IO_H
#define GATT_FLOW_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* gatt_flow_last_error(void);

/* writeFlow + the NaN marking of exportFlow.F90 (readWriteMat.F90:150-230,
 * exportFlow.F90): t and U,V,W,R,E (q[0..4]) into `path`; points where
 * inside_wall[n] != 0 are written as NaN in every field (inside_wall may be
 * NULL). */
int gatt_flow_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double t, const double* const* q,
                    const uint8_t* inside_wall);

/* readFlow + importFlow.F90:24-40: t and U,V,W,R,E from `path`.  Points whose U
 * is NaN are inside walls: they get U = V = W = 0, R = E = 1 and, when
 * inside_wall is not NULL, inside_wall[n] = 1 (0 elsewhere).  A file without W
 * (a 2-D flow) reads W as zeros.  The stored sizes must match nx, ny, nz. */
int gatt_flow_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* t, double* const* q,
                   uint8_t* inside_wall);

#ifdef __cplusplus
}
#endif

#endif
