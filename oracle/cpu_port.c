/*
 * cpu_port.c -- multi-threaded CPU port of the reference's matvec pieces, the
 * CPU baseline that the benchmark times beside the GPU (test and measurement
 * infrastructure only: nothing in the product links it).
 *
 * The reference's own CPU path cannot be compiled in this image: possu.f90
 * includes FFTW's fftw3.f03 and h5io.f90 needs the HDF5 Fortran module, and
 * neither exists here.  This port keeps the reference's algorithm and its
 * threading layout on the host's cores:
 *   scatter    sparse_T_matmul (src/sparse.f90:13-43) into the Nx*Ny*Nz grid,
 *              one grid copy per thread then a reduction (the reference runs
 *              it serially);
 *   gather     sparse_eps_matmul (src/sparse.f90:47-77), OpenMP over tets;
 *   near zone  matmul_Acorr_const (src/sparse_mat.f90:96-123): complex64
 *              blocks promoted to complex128, y_T += A x_B and
 *              y_B += A^T x_T, per-thread copies of y reduced at the end
 *              (the reference's OMP array reduction).
 * The FFT convolutions run in cpu_port.py through scipy.fft (pocketfft) on
 * the same threads, as the reference calls FFTW.
 *
 * Layouts: complex numbers are interleaved doubles; S is [N][27] per channel;
 * indS [N][27] 0-based grid node ids; blocks [P][9] complex64 column-major 3x3
 * (sp_mat(3,3,P)); ind [P][2] 0-based (tet_T, tet_B); x, y [N][3].
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int port_threads(void) { return omp_get_max_threads(); }

/* grid[indS(j,t)] += S(j,t) * x_t  (x strided: x[t*xs]) */
void port_scatter(int64_t N, const double *S, const int32_t *indS, const double *x, int64_t xs, int64_t ng,
                  double *grid) {
    const int nt = omp_get_max_threads();
    double *part = (double *)calloc((size_t)nt * ng * 2, sizeof(double));
#pragma omp parallel
    {
        double *g = part + (size_t)omp_get_thread_num() * ng * 2;
#pragma omp for schedule(static)
        for (int64_t t = 0; t < N; ++t) {
            const double xr = x[2 * t * xs], xi = x[2 * t * xs + 1];
            for (int j = 0; j < 27; ++j) {
                const double sr = S[2 * (t * 27 + j)], si = S[2 * (t * 27 + j) + 1];
                const int64_t n = indS[t * 27 + j];
                g[2 * n] += sr * xr - si * xi;
                g[2 * n + 1] += sr * xi + si * xr;
            }
        }
#pragma omp for schedule(static)
        for (int64_t e = 0; e < 2 * ng; ++e) {
            double a = 0.0;
            for (int p = 0; p < nt; ++p) a += part[(size_t)p * ng * 2 + e];
            grid[e] = a;
        }
    }
    free(part);
}

/* y_t (strided) += alpha * sum_j S(j,t) grid[indS(j,t)] */
void port_gather(int64_t N, const double *S, const int32_t *indS, const double *grid, double ar, double ai,
                 double *y, int64_t ys) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < N; ++t) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < 27; ++j) {
            const double a = S[2 * (t * 27 + j)], b = S[2 * (t * 27 + j) + 1];
            const int64_t n = indS[t * 27 + j];
            const double gr = grid[2 * n], gi = grid[2 * n + 1];
            sr += a * gr - b * gi;
            si += a * gi + b * gr;
        }
        y[2 * t * ys] += ar * sr - ai * si;
        y[2 * t * ys + 1] += ar * si + ai * sr;
    }
}

/* y += A_near x over the stored blocks (both triangles) */
void port_near(int64_t N, int64_t P, const float *blk, const int32_t *ind, const double *x, double *y) {
    const int nt = omp_get_max_threads();
    double *part = (double *)calloc((size_t)nt * N * 6, sizeof(double));
#pragma omp parallel
    {
        double *yy = part + (size_t)omp_get_thread_num() * N * 6;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < P; ++p) {
            const int64_t T = ind[2 * p], B = ind[2 * p + 1];
            const float *a = blk + 18 * p;  /* a[2*(i + 3*j)] = A(i,j) */
            for (int i = 0; i < 3; ++i) {
                double r = 0.0, im = 0.0;
                for (int j = 0; j < 3; ++j) {
                    const double ar = a[2 * (i + 3 * j)], ai = a[2 * (i + 3 * j) + 1];
                    const double xr = x[2 * (3 * B + j)], xi = x[2 * (3 * B + j) + 1];
                    r += ar * xr - ai * xi;
                    im += ar * xi + ai * xr;
                }
                yy[2 * (3 * T + i)] += r;
                yy[2 * (3 * T + i) + 1] += im;
            }
            if (T != B)
                for (int i = 0; i < 3; ++i) {
                    double r = 0.0, im = 0.0;
                    for (int j = 0; j < 3; ++j) {
                        const double ar = a[2 * (j + 3 * i)], ai = a[2 * (j + 3 * i) + 1];
                        const double xr = x[2 * (3 * T + j)], xi = x[2 * (3 * T + j) + 1];
                        r += ar * xr - ai * xi;
                        im += ar * xi + ai * xr;
                    }
                    yy[2 * (3 * B + i)] += r;
                    yy[2 * (3 * B + i) + 1] += im;
                }
        }
#pragma omp for schedule(static)
        for (int64_t e = 0; e < 6 * N; ++e) {
            double s = 0.0;
            for (int q = 0; q < nt; ++q) s += part[(size_t)q * N * 6 + e];
            y[e] += s;
        }
    }
    free(part);
}
