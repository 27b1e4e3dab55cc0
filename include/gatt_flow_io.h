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

/* writeFlowDeriv (readWriteMat.F90:231-353, exportFlow.F90:33-70, saveDerivs):
 * t, U,V,W,R,E and their rates Udot..Edot (qdot[0..4], main.F90:450-456 forms
 * them as (new - old)/dt), NaN inside walls in all ten fields. */
int gatt_flow_write_deriv(const char* path, int32_t nx, int32_t ny, int32_t nz, double t, const double* const* q,
                          const double* const* qdot, const uint8_t* inside_wall);

/* writeMeanFlow + exportMeanFlow.F90 (readWriteMat.F90:427-497, main.F90:525-529):
 * the SFD mean U,V,W,R,E (no t) into `path` (the reference's ../meanflowSFD.mat),
 * NaN inside walls. */
int gatt_meanflow_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* const* q,
                        const uint8_t* inside_wall);

/* readMeanFlow + importMeanFlow.F90 (readWriteMat.F90:356-424, main.F90:156-158,
 * resumeMeanFlow): U,V,W,R,E of the mean; the flow's wall points get
 * U = V = W = 0, R = E = 1.  inside_wall is an input here (the flow file's
 * mask, as importMeanFlow uses it); with NULL the NaN points of U are taken. */
int gatt_meanflow_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* const* q,
                       const uint8_t* inside_wall);

/* readSFD (readWriteMat.F90:87-135, main.F90:119-146, SFD type 2): the damping
 * field SFD_X of `path` (the reference's bin/SFD.mat, written by
 * preprocessing.m:104 from calcSFDregion.m).  Reads MAT 7.3 files and MATLAB's
 * default format (level 5 MAT files, v7 with zlib-compressed variables, little
 * endian); SFD_X must be a real double array of nx*ny*nz values. */
int gatt_sfd_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* sfd_x);

/* The SFD.mat of the preprocessing (preprocessing.m:104) in the MAT 7.3 layout,
 * variable SFD_X. */
int gatt_sfd_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* sfd_x);

/* Any set of double or complex double arrays as one MAT 7.3 file (MATLAB's
 * save(..., '-v7.3') of the variables): variable v is names[v], of MATLAB size
 * dims[o_v .. o_v + ranks[v] - 1] (ranks 2..8, the ranks concatenated in
 * dims), column-major data re[v]; when im and im[v] are not NULL the variable
 * is complex, stored as MATLAB stores it (compound {real, imag}).  Used for the
 * results and modes files of calcInstability.m:419, 450-452. */
int gatt_mat_write(const char* path, int32_t nvars, const char* const* names, const int32_t* ranks,
                   const int64_t* dims, const double* const* re, const double* const* im);

/* log.txt.  The header line of runDNS.m:58-66 (column names, five per tracked
 * point) and one line of main.F90:467-506 with its Fortran edit descriptors:
 * I10 save number (0 between saves), I10 step, ES13.8E1 t, ES10.5E1 dt, F8.5 CFL,
 * 5 x ES10.4E2 max changes, then per tracked point U,V,W,R,E (tracked[5*i..]):
 * with tracked_norm != 0, R - 1 and E/tracked_norm - 1; a value of exactly 0 is
 * written "0", others ES15.8E2; fields tab-separated, the line ends " \n".
 * Both write a NUL-terminated line into buf (size bytes) and return its length,
 * or a negative code when buf is too small. */
int gatt_log_header(int32_t ntracked, char* buf, int64_t size);
int gatt_log_line(int32_t nsave, int32_t tstep, double t, double dt, double cfl, const double* max_change,
                  int32_t ntracked, const double* tracked, double tracked_norm, char* buf, int64_t size);

#ifdef __cplusplus
}
#endif

#endif
