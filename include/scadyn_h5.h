/*
 * scadyn_h5.h -- the on-disk formats at the solve's boundary (HDF5, C API).
 *
 * The reference keeps its HDF5 I/O in the Fortran host (src/h5io.f90); this
 * separate library (libscadyn_h5.so, host code only) lets a C, C++ or Python
 * host read the reference's mesh files and write T-matrix files that the
 * reference's read_T (src/h5io.f90:138-313) reads back.
 *
 * Mesh file (read_mesh, src/h5io.f90:16-134; writer other/scadyn_mesh.py:77-103):
 *   coord    float64 [Nnode][3]   (Fortran coord(3, Nnode))
 *   etopol   int32   [N][4]       1-based node ids (Fortran etopol(4, N))
 *   param_r, param_i  float64 [N], or two-dimensional for several wavelength
 *            bars (read as the reference does: the flat data fills
 *            param(N, bars) column by column)
 * T-matrix file (T_write2file :429-558, T_empty :562-678, singleT_write2file
 * :682-833):
 *   Taai_r, Taai_i, Tabi_r, Tabi_i, Tbai_r, Tbai_i, Tbbi_r, Tbbi_i
 *            float64 [T_size], T_size = sum over bars of nm_b^2,
 *            nm_b = nmax_b (nmax_b + 2); bar b's nm_b x nm_b block is stored
 *            column-major at offset sum_{c<b} nm_c^2
 *   T-ref-a  float64 [3] = Re eps, Im eps, a
 *   T-wavlens float64 [bars] = wavelengths (2 pi / k)
 *
 * Complex arrays are interleaved (re, im) doubles.  Every call returns 0 or
 * SCADYN_H5_ERR (message in scadyn_h5_last_error()).
 */
#ifndef SCADYN_H5_H
#define SCADYN_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCADYN_H5_OK  0
#define SCADYN_H5_ERR 1

const char *scadyn_h5_last_error(void);

/* Sizes of a mesh file (n_bars = columns of param_r). */
int scadyn_h5_mesh_sizes(const char *path, int32_t *n_node, int32_t *n_tet, int32_t *n_bars);

/* read_mesh: coord [3*n_node], etopol [4*n_tet] (1-based, as stored) and
 * param [bars][n_tet] complex.  With a > 0 the coordinates are scaled so that
 * the volume-equivalent radius of the mesh is a (src/h5io.f90:128-131); the
 * permittivity is param_r + i param_i. */
int scadyn_h5_read_mesh(const char *path, double a, double *coord, int32_t *etopol, double *param);

/* A mesh file in the reference's layout (one bar of param). */
int scadyn_h5_write_mesh(const char *path, int32_t n_node, int32_t n_tet, const double *coord,
                         const int32_t *etopol, const double *param);

/* T_size of the bars' expansion orders. */
int64_t scadyn_h5_T_size(int32_t bars, const int32_t *nmax);

/* T_write2file: all bars; Taa..Tbb are complex [T_size] in the file's layout. */
int scadyn_h5_write_T(const char *path, int32_t bars, const int32_t *nmax, const double *Taa, const double *Tab,
                      const double *Tba, const double *Tbb, const double ref_a[3], const double *wavelengths);

/* T_empty: a zero T-matrix file, created only if `path` does not exist. */
int scadyn_h5_T_empty(const char *path, int32_t bars, const int32_t *nmax, const double ref_a[3],
                      const double *wavelengths);

/* singleT_write2file: bar `bar` (1-based) of an existing file, from its four
 * nm x nm column-major blocks (what scadyn_tmat_solve returns). */
int scadyn_h5_write_T_bar(const char *path, int32_t bars, const int32_t *nmax, int32_t bar, const double *Taa,
                          const double *Tab, const double *Tba, const double *Tbb);

/* read_T: sizes, then the four complex [T_size] arrays, T-ref-a and T-wavlens
 * (the last two may be null). */
int scadyn_h5_T_sizes(const char *path, int64_t *t_size, int32_t *bars);
int scadyn_h5_read_T(const char *path, double *Taa, double *Tab, double *Tba, double *Tbb, double ref_a[3],
                     double *wavelengths);

#ifdef __cplusplus
}
#endif
#endif /* SCADYN_H5_H */
