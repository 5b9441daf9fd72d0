/*
 * h5io.c -- HDF5 mesh and T-matrix files in the reference's layouts
 * (include/scadyn_h5.h; src/h5io.f90 in the reference).  Host code, C, HDF5
 * C API; built as libscadyn_h5.so apart from the HIP library.
 */
#include "../../include/scadyn_h5.h"

#include <hdf5.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static __thread char g_err[512];

static int err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return SCADYN_H5_ERR;
}

const char *scadyn_h5_last_error(void) { return g_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* dims of a dataset (rank <= 2); returns rank or -1 */
static int dset_dims(hid_t f, const char *name, hsize_t dims[2]) {
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) return -1;
    hid_t s = H5Dget_space(d);
    int rank = H5Sget_simple_extent_ndims(s);
    if (rank < 1 || rank > 2) rank = -1;
    else H5Sget_simple_extent_dims(s, dims, NULL);
    H5Sclose(s);
    H5Dclose(d);
    return rank;
}

static int read_all(hid_t f, const char *name, hid_t memtype, void *buf) {
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) return err("dataset %s not found", name);
    herr_t e = H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    H5Dclose(d);
    return e < 0 ? err("reading dataset %s failed", name) : SCADYN_H5_OK;
}

static int write_new(hid_t f, const char *name, hid_t type, int rank, const hsize_t *dims, const void *buf) {
    hid_t s = H5Screate_simple(rank, dims, NULL);
    hid_t d = H5Dcreate2(f, name, type, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = d < 0 ? -1 : H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    if (d >= 0) H5Dclose(d);
    H5Sclose(s);
    return e < 0 ? err("writing dataset %s failed", name) : SCADYN_H5_OK;
}

int scadyn_h5_mesh_sizes(const char *path, int32_t *n_node, int32_t *n_tet, int32_t *n_bars) {
    quiet();
    if (!path || !n_node || !n_tet || !n_bars) return err("null argument");
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return err("cannot open %s", path);
    hsize_t c[2] = {0, 0}, e[2] = {0, 0}, p[2] = {0, 0};
    int rc = dset_dims(f, "coord", c), re = dset_dims(f, "etopol", e), rp = dset_dims(f, "param_r", p);
    H5Fclose(f);
    if (rc != 2 || c[1] != 3) return err("%s: coord is not [Nnode][3]", path);
    if (re != 2 || e[1] != 4) return err("%s: etopol is not [N][4]", path);
    if (rp < 1) return err("%s: param_r missing", path);
    *n_node = (int32_t)c[0];
    *n_tet = (int32_t)e[0];
    *n_bars = rp == 1 ? 1 : (int32_t)(p[0] * p[1] / e[0]);
    if (rp == 2 && p[0] * p[1] != (hsize_t)*n_bars * e[0]) return err("%s: param_r size is not a multiple of N", path);
    if (rp == 1 && p[0] != e[0]) return err("%s: param_r has %llu entries for %llu tets", path,
                                            (unsigned long long)p[0], (unsigned long long)e[0]);
    return SCADYN_H5_OK;
}

int scadyn_h5_read_mesh(const char *path, double a, double *coord, int32_t *etopol, double *param) {
    int32_t nn, nt, nb;
    if (scadyn_h5_mesh_sizes(path, &nn, &nt, &nb)) return SCADYN_H5_ERR;
    if (!coord || !etopol || !param) return err("null argument");
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return err("cannot open %s", path);
    const size_t np = (size_t)nt * nb;
    double *pr = (double *)malloc(sizeof(double) * np), *pi = (double *)malloc(sizeof(double) * np);
    int rc = read_all(f, "coord", H5T_NATIVE_DOUBLE, coord);
    if (!rc) rc = read_all(f, "etopol", H5T_NATIVE_INT32, etopol);
    if (!rc) rc = read_all(f, "param_r", H5T_NATIVE_DOUBLE, pr);
    if (!rc) rc = read_all(f, "param_i", H5T_NATIVE_DOUBLE, pi);
    H5Fclose(f);
    if (!rc) {
        /* the flat data fills param(N, bars) column by column, as the
           reference's h5dread_f into param_r(N, bars) does */
        for (size_t e = 0; e < np; ++e) {
            param[2 * e] = pr[e];
            param[2 * e + 1] = pi[e];
        }
        for (int32_t t = 0; t < nt && !rc; ++t)
            for (int v = 0; v < 4; ++v)
                if (etopol[4 * t + v] < 1 || etopol[4 * t + v] > nn)
                    rc = err("%s: etopol entry %d of tet %d is outside 1..%d", path, etopol[4 * t + v], t + 1, nn);
    }
    free(pr);
    free(pi);
    if (rc || !(a > 0.0)) return rc;
    /* volume-equivalent radius (get_tetra_vol, then src/h5io.f90:128-131) */
    double vol = 0.0;
    for (int32_t t = 0; t < nt; ++t) {
        const double *p0 = coord + 3 * (etopol[4 * t] - 1), *p1 = coord + 3 * (etopol[4 * t + 1] - 1);
        const double *p2 = coord + 3 * (etopol[4 * t + 2] - 1), *p3 = coord + 3 * (etopol[4 * t + 3] - 1);
        double u[3], v[3], w[3];
        for (int c = 0; c < 3; ++c) {
            u[c] = p1[c] - p0[c];
            v[c] = p2[c] - p0[c];
            w[c] = p3[c] - p0[c];
        }
        const double det = w[0] * (u[1] * v[2] - u[2] * v[1]) + w[1] * (u[2] * v[0] - u[0] * v[2]) +
                           w[2] * (u[0] * v[1] - u[1] * v[0]);
        vol += fabs(det) / 6.0;
    }
    const double a_eff = pow(3.0 * vol / 4.0 / 3.141592653589793, 1.0 / 3.0);
    for (size_t i = 0; i < 3 * (size_t)nn; ++i) coord[i] *= a / a_eff;
    return SCADYN_H5_OK;
}

int scadyn_h5_write_mesh(const char *path, int32_t n_node, int32_t n_tet, const double *coord,
                         const int32_t *etopol, const double *param) {
    quiet();
    if (!path || !coord || !etopol || !param || n_node < 4 || n_tet < 1) return err("bad argument");
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return err("cannot create %s", path);
    double *pr = (double *)malloc(sizeof(double) * n_tet), *pi = (double *)malloc(sizeof(double) * n_tet);
    for (int32_t t = 0; t < n_tet; ++t) {
        pr[t] = param[2 * t];
        pi[t] = param[2 * t + 1];
    }
    hsize_t dc[2] = {(hsize_t)n_node, 3}, de[2] = {(hsize_t)n_tet, 4}, dp[1] = {(hsize_t)n_tet};
    int rc = write_new(f, "coord", H5T_NATIVE_DOUBLE, 2, dc, coord);
    if (!rc) rc = write_new(f, "etopol", H5T_NATIVE_INT32, 2, de, etopol);
    if (!rc) rc = write_new(f, "param_r", H5T_NATIVE_DOUBLE, 1, dp, pr);
    if (!rc) rc = write_new(f, "param_i", H5T_NATIVE_DOUBLE, 1, dp, pi);
    free(pr);
    free(pi);
    H5Fclose(f);
    return rc;
}

int64_t scadyn_h5_T_size(int32_t bars, const int32_t *nmax) {
    int64_t s = 0;
    for (int32_t b = 0; b < bars; ++b) {
        const int64_t nm = (int64_t)(nmax[b] + 1) * (nmax[b] + 1) - 1;
        s += nm * nm;
    }
    return s;
}

static const char *TNAMES[8] = {"Taai_r", "Taai_i", "Tabi_r", "Tabi_i", "Tbai_r", "Tbai_i", "Tbbi_r", "Tbbi_i"};

/* the eight real datasets from four complex arrays (null: zeros) */
static int write_T_file(const char *path, int32_t bars, int64_t ts, const double *const T[4], const double ref_a[3],
                        const double *wavelengths) {
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return err("cannot create %s", path);
    double *buf = (double *)calloc((size_t)(ts > 0 ? ts : 1), sizeof(double));
    hsize_t d1[1] = {(hsize_t)ts}, d3[1] = {3}, dw[1] = {(hsize_t)bars};
    int rc = SCADYN_H5_OK;
    for (int q = 0; q < 8 && !rc; ++q) {
        const double *src = T[q / 2];
        for (int64_t e = 0; e < ts; ++e) buf[e] = src ? src[2 * e + (q % 2)] : 0.0;
        rc = write_new(f, TNAMES[q], H5T_NATIVE_DOUBLE, 1, d1, buf);
    }
    if (!rc) rc = write_new(f, "T-ref-a", H5T_NATIVE_DOUBLE, 1, d3, ref_a);
    if (!rc) rc = write_new(f, "T-wavlens", H5T_NATIVE_DOUBLE, 1, dw, wavelengths);
    free(buf);
    H5Fclose(f);
    return rc;
}

int scadyn_h5_write_T(const char *path, int32_t bars, const int32_t *nmax, const double *Taa, const double *Tab,
                      const double *Tba, const double *Tbb, const double ref_a[3], const double *wavelengths) {
    quiet();
    if (!path || !nmax || bars < 1 || !Taa || !Tab || !Tba || !Tbb || !ref_a || !wavelengths)
        return err("bad argument");
    const double *T[4] = {Taa, Tab, Tba, Tbb};
    return write_T_file(path, bars, scadyn_h5_T_size(bars, nmax), T, ref_a, wavelengths);
}

int scadyn_h5_T_empty(const char *path, int32_t bars, const int32_t *nmax, const double ref_a[3],
                      const double *wavelengths) {
    quiet();
    if (!path || !nmax || bars < 1 || !ref_a || !wavelengths) return err("bad argument");
    if (access(path, F_OK) == 0) return SCADYN_H5_OK;  /* an existing file is kept (src/h5io.f90:597-598) */
    const double *T[4] = {NULL, NULL, NULL, NULL};
    return write_T_file(path, bars, scadyn_h5_T_size(bars, nmax), T, ref_a, wavelengths);
}

int scadyn_h5_write_T_bar(const char *path, int32_t bars, const int32_t *nmax, int32_t bar, const double *Taa,
                          const double *Tab, const double *Tba, const double *Tbb) {
    quiet();
    if (!path || !nmax || bar < 1 || bar > bars || !Taa || !Tab || !Tba || !Tbb) return err("bad argument");
    const int64_t ts = scadyn_h5_T_size(bars, nmax);
    const int64_t off = scadyn_h5_T_size(bar - 1, nmax);
    const int64_t nm = (int64_t)(nmax[bar - 1] + 1) * (nmax[bar - 1] + 1) - 1;
    hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) return err("cannot open %s for writing", path);
    const double *T[4] = {Taa, Tab, Tba, Tbb};
    int rc = SCADYN_H5_OK;
    for (int q = 0; q < 8 && !rc; ++q) {
        hsize_t dims[2];
        if (dset_dims(f, TNAMES[q], dims) != 1 || (int64_t)dims[0] != ts) {
            rc = err("%s: %s does not hold %lld values", path, TNAMES[q], (long long)ts);
            break;
        }
        /* the bar's slice [off, off + nm^2) as a hyperslab of the 1-D dataset */
        double *buf = (double *)malloc(sizeof(double) * (size_t)(nm * nm));
        for (int64_t e = 0; e < nm * nm; ++e) buf[e] = T[q / 2][2 * e + (q % 2)];
        hid_t d = H5Dopen2(f, TNAMES[q], H5P_DEFAULT);
        hid_t fs = H5Dget_space(d);
        hsize_t start[1] = {(hsize_t)off}, count[1] = {(hsize_t)(nm * nm)};
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, count, NULL);
        hid_t ms = H5Screate_simple(1, count, NULL);
        if (H5Dwrite(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf) < 0) rc = err("writing %s failed", TNAMES[q]);
        H5Sclose(ms);
        H5Sclose(fs);
        H5Dclose(d);
        free(buf);
    }
    H5Fclose(f);
    return rc;
}

int scadyn_h5_T_sizes(const char *path, int64_t *t_size, int32_t *bars) {
    quiet();
    if (!path || !t_size || !bars) return err("null argument");
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return err("cannot open %s", path);
    hsize_t d[2] = {0, 0}, w[2] = {0, 0};
    const int r = dset_dims(f, "Taai_r", d), rw = dset_dims(f, "T-wavlens", w);
    H5Fclose(f);
    if (r != 1) return err("%s: Taai_r is not one-dimensional", path);
    *t_size = (int64_t)d[0];
    *bars = rw == 1 ? (int32_t)w[0] : 0;
    return SCADYN_H5_OK;
}

int scadyn_h5_read_T(const char *path, double *Taa, double *Tab, double *Tba, double *Tbb, double ref_a[3],
                     double *wavelengths) {
    int64_t ts;
    int32_t bars;
    if (scadyn_h5_T_sizes(path, &ts, &bars)) return SCADYN_H5_ERR;
    if (!Taa || !Tab || !Tba || !Tbb) return err("null argument");
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return err("cannot open %s", path);
    double *buf = (double *)malloc(sizeof(double) * (size_t)(ts > 0 ? ts : 1));
    double *T[4] = {Taa, Tab, Tba, Tbb};
    int rc = SCADYN_H5_OK;
    for (int q = 0; q < 8 && !rc; ++q) {
        rc = read_all(f, TNAMES[q], H5T_NATIVE_DOUBLE, buf);
        for (int64_t e = 0; e < ts && !rc; ++e) T[q / 2][2 * e + (q % 2)] = buf[e];
    }
    if (!rc && ref_a) rc = read_all(f, "T-ref-a", H5T_NATIVE_DOUBLE, ref_a);
    if (!rc && wavelengths && bars > 0) rc = read_all(f, "T-wavlens", H5T_NATIVE_DOUBLE, wavelengths);
    free(buf);
    H5Fclose(f);
    return rc;
}
