/*
 * scadyn_amd.h -- C ABI of the MI355X JVIE T-matrix solve.
 *
 * The reference (scadyn, Fortran) has no FFI on this path: calc_T
 * (src/T_matrix.f90:18-82) calls the module procedures update_projections,
 * build_G, compute_near_zone_interactions_const and compute_T_matrix
 * (src/T_matrix.f90:56-72) on the global `mesh`/`matrices` state.  This
 * header is the boundary that replaces those four calls with one call per
 * wavelength bar, bound from Fortran with ISO_C_BINDING
 * (scadyn_amd/fortran/scadyn_amd_iface.f90).
 *
 * Conventions (identical to the reference's in-memory layouts):
 *   coord   double[3*n_node]   Fortran coord(3, n_node), metres, already
 *                              rescaled to the volume-equivalent radius
 *                              (src/h5io.f90:128-131)
 *   etopol  int32 [4*n_tet]    Fortran etopol(4, n_tet), 1-based node ids
 *   param   double[2*n_tet]    complex(dp) param(n_tet): relative permittivity
 *                              of each tet for this bar (src/T_matrix.f90:37-43)
 *   complex vectors/matrices   interleaved (re, im) doubles, i.e. Fortran
 *                              complex(c_double_complex); matrices column-major
 *   unknown vector x(3N)       3*(t-1)+c: component c of tet t
 *                              (src/gmres_module.f90:80-86)
 *
 * Every call is synchronous, copies its inputs to the device and its
 * results back, and keeps no host pointer after it returns.  All calls
 * return 0 on success or a SCADYN_ERR_* code; scadyn_last_error() gives
 * the message of the last failure in the calling thread.
 */
#ifndef SCADYN_AMD_H
#define SCADYN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCADYN_OK               0
#define SCADYN_ERR_ARG          1  /* bad argument or size                          */
#define SCADYN_ERR_ORIENTATION  2  /* detJ <= 0 (src/common.f90:466-469)           */
#define SCADYN_ERR_GRID         3  /* centroid outside the cube grid (geometry.f90) */
#define SCADYN_ERR_SINGULAR     4  /* singular least-squares / triangular system    */
#define SCADYN_ERR_NOCONV       5  /* GMRES did not reach tol (strict mode only)     */
#define SCADYN_ERR_MEMORY       6  /* device memory budget exceeded                  */
#define SCADYN_ERR_DEVICE       7  /* HIP runtime failure                            */
#define SCADYN_ERR_UNSUPPORTED  8  /* option the build does not implement            */

#define SCADYN_PREC_FP64   0  /* FP64 matvec, reference-faithful GMRES          */
#define SCADYN_PREC_MIXED  1  /* FP32 matvec (and, with the blocked MGS, a Krylov basis
                                 stored in reduced precision with FP64 arithmetic)
                                 inside FP64 iterative refinement                    */

/* storage of the mixed-precision Krylov basis (scadyn_opts.basis) */
#define SCADYN_BASIS_AUTO  0
#define SCADYN_BASIS_C64   1
#define SCADYN_BASIS_F16   2

typedef struct scadyn_mesh {
    int32_t n_node;
    int32_t n_tet;
    const double *coord;
    const int32_t *etopol;
    const double *param;
} scadyn_mesh;

/* Solver options; field meanings follow mesh_struct (src/common.f90:78-89). */
typedef struct scadyn_opts {
    double tol;          /* GMRES tolerance on |g(i+1)|/||b||   (mesh%tol, 1e-5)       */
    int32_t maxit;       /* Krylov dimension per cycle            (mesh%maxit, 50)      */
    int32_t restart;     /* number of restart cycles              (mesh%restart, 4)     */
    int32_t M_ex;        /* stencil order, only 2                 (mesh%M_ex)           */
    int32_t near_zone;   /* near-zone radius in cubes, only 1     (mesh%near_zone)      */
    double grid_size;    /* grid step factor                      (mesh%grid_size, 1.8) */
    int32_t order;       /* basis order, only 0                   (mesh%order)          */
    int32_t precision;   /* SCADYN_PREC_*                                               */
    int32_t batch;       /* RHS slots solved concurrently, 0 = automatic                */
    int32_t strict;      /* nonzero: non-convergence returns SCADYN_ERR_NOCONV          */
    int32_t device;      /* HIP device ordinal                                          */
    int32_t fft_pad;     /* 0: smallest 2,3,5,7-smooth length >= 2N-1 per axis;
                            1: the reference's 2N                                       */
    double refine_tol;   /* mixed precision: a system is done when its true (FP64)
                            residual at a restart is <= refine_tol ||b||; the Arnoldi
                            steps use the FP32 matvec and tol is then the residual
                            reduction that ends a restart cycle early               */
    int32_t ortho;       /* Arnoldi orthogonalisation: 0 modified Gram-Schmidt (the
                            reference's, src/gmres_module.f90:241-244), mgs_block basis
                            vectors per pass; 1 classical Gram-Schmidt with DGKS
                            re-orthogonalisation (two fused passes per iteration);
                            2 block MGS: the MGS coefficients 8 basis vectors per pass
                            with an in-pass Gram correction (same recurrence, about
                            half the HBM traffic of 0)                                  */
    int32_t fft_kernel;  /* line-FFT kernels: 0 automatic (register two-stage kernel where
                            the axis length is R1*R2 with R1, R2 <= 20, else LDS
                            Stockham), 1 LDS Stockham on every axis                     */
    int32_t mgs_block;   /* ortho 0: basis vectors whose MGS coefficients one pass over the
                            rows evaluates (in-pass Gram correction, same recurrence):
                            0 automatic (4), 1, 2, 4, 6 or 8; a pass moves about
                            2 + 2/mgs_block vectors of HBM traffic per coefficient      */
    int32_t gather_kernel; /* complex64 gather + near zone: 0 vector FMA (default), 1 the near-zone
                              product on the matrix cores (v_mfma_f32_16x16x4f32; measured about 2x
                              slower, its operand set-up costs as much as the FMAs it replaces) */
    int32_t basis;       /* mixed precision with the blocked MGS (ortho 0, mgs_block > 1):
                            storage of the Krylov basis.  SCADYN_BASIS_AUTO (binary16 for
                            systems of 2^20 to 2^26 rows, else complex64),
                            SCADYN_BASIS_C64, or SCADYN_BASIS_F16: real and imaginary parts
                            in IEEE binary16, half the basis traffic of complex64.  The
                            arithmetic stays FP64 and every restart checks the FP64 true
                            residual, so the stored precision changes iteration counts,
                            not the accuracy of the result.  With binary16 a restart
                            cycle ends at a residual reduction of max(tol, 2^-10)      */
    int32_t reserved[3];
} scadyn_opts;

/* Defaults of the reference (src/common.f90:78-89). */
void scadyn_opts_default(scadyn_opts *opts);

/* Message of the last failure in this thread ("" if none). */
const char *scadyn_last_error(void);

/*
 * One wavelength bar: replaces update_projections + build_G +
 * compute_near_zone_interactions_const + compute_T_matrix
 * (src/T_matrix.f90:56-72).  k is the wavenumber (1/m), nmax the expansion
 * order; Taa, Tab, Tba, Tbb are caller-allocated nm x nm column-major
 * complex arrays, nm = nmax(nmax+2), returned already scaled by i k^3
 * (src/T_matrix.f90:270-275).  iters (optional, length 2*nm) receives
 * the GMRES iteration count of each column.
 */
int scadyn_tmat_solve(const scadyn_mesh *mesh, const scadyn_opts *opts,
                      double k, int32_t nmax,
                      double *Taa, double *Tab, double *Tba, double *Tbb,
                      int32_t *iters);

/* ---- operator-level API (the request loop, tests, benchmarks) -------- */

typedef struct scadyn_operator scadyn_operator;

/* Device-resident pFFT operator of one (mesh, k): grid, projections S*,
 * FFT(G), near-zone blocks (src/setup.f90:77-140, src/build_G.f90,
 * src/precorrection.f90:131-264). */
int scadyn_op_create(const scadyn_mesh *mesh, const scadyn_opts *opts, double k,
                     scadyn_operator **out);
void scadyn_op_destroy(scadyn_operator *op);

/* Sizes: info[0..15] = N_tet, Nx, Ny, Nz, N_cubes, N_tet_cube, n_blocks,
 * Lx, Ly, Lz, batch, device bytes, most restart cycles of a system in the last
 * solve, 1 if the near-zone rows are in cube order (per-cube gather), then the
 * tile widths of the five line-FFT passes of the current batch in FP64 and in
 * FP32 (8 bits each: x, y, z, inverse y, inverse x; tuned for large grids);
 * delta, box_delta and k in dinfo[0..2]; dinfo[3] = times the reference's
 * re-orthogonalisation test (src/gmres_module.f90:251) fired, summed over the
 * systems finished on this operator (the extra pass itself is not applied). */
int scadyn_op_info(const scadyn_operator *op, int64_t info[16], double dinfo[4]);

#define SCADYN_EXPORT_S       1  /* S,Sx,Sy,Sz: complex [4][N][27], reference tet order  */
#define SCADYN_EXPORT_BLOCKS  2  /* near-zone blocks: complex64 [n_blocks][9] col-major,
                                    reference order (src/precorrection.f90:182-258)     */
#define SCADYN_EXPORT_BLKIND  3  /* int32 [n_blocks][2] 1-based (tet_T, tet_B)           */
#define SCADYN_EXPORT_FG      4  /* FFT(G) octant, complex [Lz/2+1][Ly/2+1][Lx/2+1],
                                    scaled by 1/(Lx Ly Lz)                             */
/* Copy an internal array to host memory (nbytes must match exactly). */
int scadyn_op_export(const scadyn_operator *op, int32_t what, void *host, int64_t nbytes);

/* y = A x for nrhs vectors x[nrhs][3N] (compute_Ax2, src/gmres_module.f90:57-176). */
int scadyn_op_matvec(scadyn_operator *op, int32_t nrhs, const double *x, double *y);
/* The same with the matvec precision chosen explicitly (SCADYN_PREC_MIXED: S*,
 * grids and FFTs in complex64; x and y stay complex128). */
int scadyn_op_matvec_prec(scadyn_operator *op, int32_t precision, int32_t nrhs, const double *x, double *y);

/* Solve A x = rhs for nrhs right-hand sides [nrhs][3N] with restarted GMRES
 * (src/gmres_module.f90:180-325).  iters[nrhs] (optional) receives the
 * reference's iteration count `ite` of each system.  In mixed precision the
 * Arnoldi steps use the FP32 matvec and every restart the FP64 true residual;
 * a system is done when ||b - A x|| <= refine_tol ||b|| at a restart. */
int scadyn_op_solve(scadyn_operator *op, int32_t nrhs, const double *rhs, double *x,
                    int32_t *iters);

/* Regular-VSWF projection matrix mat(3N, 2nm) column-major, reference tet order
 * (vswf2constant, src/T_matrix.f90:280-305). */
int scadyn_op_vswf(scadyn_operator *op, int32_t nmax, double *mat);

/* Projection of given solutions onto the regular VSWFs, T_mat = mat^H X
 * (src/T_matrix.f90:266, before the i k^3 scale of :270-275): X is (3N, 2nm)
 * column-major in reference tet order, T (2nm, 2nm) column-major.  Runs the
 * FP64 matrix-core GEMM of scadyn_op_tmatrix. */
int scadyn_op_project(scadyn_operator *op, int32_t nmax, const double *X, double *T);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the projection GEMM of scadyn_op_project on device
 * pointers, T = mat^H X with mat (n x na) and X (n x nx) row-major complex128
 * (any row order, the same in both) and T (na x nx) row-major, on the
 * operator's device and stream, synchronous.  The GEMM's row-chunk partials
 * live in a fixed per-device scratch array of 4 Mi complex128 elements
 * (64 MiB, held by each device the library runs on until the process ends):
 * na * nx above that returns SCADYN_ERR_ARG.  Nothing is allocated.  Calls
 * that overlap on one device share the scratch and race. */
int scadyn_op_project_noalloc(scadyn_operator *op, int64_t n, int32_t na, int32_t nx, const double *mat,
                              const double *X, double *T);

/* T-matrix blocks of this operator (compute_T_matrix, src/T_matrix.f90:240-277). */
int scadyn_op_tmatrix(scadyn_operator *op, int32_t nmax,
                      double *Taa, double *Tab, double *Tba, double *Tbb,
                      int32_t *iters);

/* Selected columns of the T-matrix: only the systems of columns cols[0..ncols-1]
 * (0-based, M-waves 0..nm-1 then N-waves) are solved -- the body of
 * compute_T_matrix's column loop (src/T_matrix.f90:254-266) for those columns --
 * and T(:, cols[j]) over all 2nm rows, scaled by i k^3 (:270-275), is written
 * column-major into Tcols (2nm x ncols).  iters[ncols] (optional). */
int scadyn_op_tmatrix_cols(scadyn_operator *op, int32_t nmax, int32_t ncols, const int32_t *cols,
                           double *Tcols, int32_t *iters);

/* The same as a resumable job, for callers that take results as columns
 * converge (throughput measurement, streaming): the columns enter the
 * continuously refilled batch in the order given.  No counterpart in the
 * reference, whose column loop (src/T_matrix.f90:254-266) runs to the end.
 * advance runs solver steps until at least `want` more systems have finished
 * (or all have) and reports how many have finished in total; result writes
 * T(:, cols[j]) as scadyn_op_tmatrix_cols does for finished columns and zero
 * with iters[j] = -1 for the rest.  spread > 0 spreads the first fills of the
 * batch's B slots over that many solver steps (slot b starts at step
 * b * spread / B): with spread about the iterations of one system, columns
 * finish at an even rate from the first one on instead of in batch-wide
 * bursts.  The operator must outlive the job. */
typedef struct scadyn_tjob scadyn_tjob;
int scadyn_tjob_create(scadyn_operator *op, int32_t nmax, int32_t ncols, const int32_t *cols, int32_t spread,
                       scadyn_tjob **out);
int scadyn_tjob_advance(scadyn_tjob *job, int32_t want, int32_t *ndone);
int scadyn_tjob_result(scadyn_tjob *job, double *Tcols, int32_t *iters);
void scadyn_tjob_destroy(scadyn_tjob *job);

/* Kernel timing: when enabled, every launch of the hot kernels is bracketed
 * by HIP events on the operator's stream.  timings[i] = {total ms, launches,
 * algorithmic bytes} for kernel class i (see SCADYN_KCLASS_*); for the T-matrix
 * projection GEMM (SCADYN_KCLASS_TGEMM) the third number is floating-point
 * operations instead of bytes.  FP64 and FP32 matvec launches share a class. */
#define SCADYN_KCLASS_SCATTER  0
#define SCADYN_KCLASS_FFTX     1
#define SCADYN_KCLASS_FFTY     2
#define SCADYN_KCLASS_FFTZCONV 3
#define SCADYN_KCLASS_IFFTY    4
#define SCADYN_KCLASS_IFFTX    5
#define SCADYN_KCLASS_GATHER   6  /* gather + near-zone + combine */
#define SCADYN_KCLASS_KRYLOV   7
#define SCADYN_KCLASS_TGEMM    8  /* T = mat^H X on the FP64 matrix cores */
#define SCADYN_KCLASS_COUNT    9
int scadyn_op_profile(scadyn_operator *op, int32_t enable);
int scadyn_op_timings(scadyn_operator *op, double *timings /* [COUNT][3] */);
/* The same with the matvec classes (SCADYN_KCLASS_SCATTER .. GATHER) restricted
 * to the launches of one matvec precision (SCADYN_PREC_FP64: complex128 grids,
 * SCADYN_PREC_MIXED: complex64); the Krylov and GEMM classes are FP64 and
 * reported whole. */
int scadyn_op_timings_prec(scadyn_operator *op, int32_t precision, double *timings /* [COUNT][3] */);

/* ---- consumers of a T-matrix (host) ------------------------------------ */

/* Orientation-averaged cross sections of a bar's T (tr_T, src/mie.f90:728-759):
 * crs = {C_ext, C_sca, C_abs, Tr(Re T + T T^H)}, from the four nm x nm blocks
 * (column-major, as scadyn_tmat_solve returns them) and the wavenumber k. */
int scadyn_cross_sections(int32_t nmax, const double *Taa, const double *Tab, const double *Tba, const double *Tbb,
                          double k, double crs[4]);

/* Radiation force and torque along the incident direction for an x-polarised
 * plane wave of amplitude E (V/m) travelling along +z, the particle in its
 * reference orientation (forcetorque's z components, src/forces.f90:81-189,
 * F_z / T_z :215-257, planewave src/mie.f90:204-228): out = {F_z (N),
 * T_z (N m), Q_f,z, Q_t,z}, the efficiencies normalised as get_forces
 * (src/forces.f90:55-63) with the radius a (m).  The transverse components
 * need the VSWF rotations of the dynamics code and are not provided. */
int scadyn_force_torque_z(int32_t nmax, const double *Taa, const double *Tab, const double *Tba, const double *Tbb,
                          double k, double E, double a, double out[4]);

/* ---- request loop (serving) ------------------------------------------ */

/* Device bytes one T-matrix request needs at its peak (operator + the larger of
 * setup temporaries and solve workspace); host-side computation only.  No
 * counterpart in the reference (it has no server); the admission model of the
 * request loop.  The reference prints a similar estimate
 * (src/precorrection.f90:166-178). */
int scadyn_footprint(const scadyn_mesh *mesh, const scadyn_opts *opts, int32_t nmax, int64_t *bytes);

typedef struct scadyn_server scadyn_server;
typedef struct scadyn_server_opts {
    int32_t device;       /* HIP device ordinal                                   */
    int32_t max_active;   /* requests resident on the device at once (0: 8)       */
    int64_t hbm_budget;   /* device bytes the server may hold (0: 80% of free)    */
    int32_t host_only;    /* nonzero: no device is used; requests are validated,
                             admitted by footprint and retired after a stand-in
                             "solve" that computes nothing (T blocks zero, iteration
                             counts = the request id); tests the loop's admission
                             and completion logic on a host without a GPU        */
    int32_t reserved[7];
} scadyn_server_opts;

#define SCADYN_REQ_QUEUED  0
#define SCADYN_REQ_RUNNING 1
#define SCADYN_REQ_DONE    2
#define SCADYN_REQ_FAILED  3

/* A request loop on one device: T-matrix requests (mesh, k, nmax) are queued,
 * admitted FIFO while their footprint fits the budget, set up and solved with
 * their steps interleaved on separate streams (continuous batching).  No
 * counterpart in the reference. */
int scadyn_server_create(const scadyn_server_opts *opts, scadyn_server **out);
/* Copies the inputs; validates and sizes the request on the host (a request
 * larger than the whole budget fails with SCADYN_ERR_MEMORY); returns its id. */
int scadyn_server_submit(scadyn_server *sv, const scadyn_mesh *mesh, const scadyn_opts *opts, double k,
                         int32_t nmax, int64_t *id);
/* state and times {submit, admit, solve start, done} in seconds since creation */
int scadyn_server_status(scadyn_server *sv, int64_t id, int32_t *state, double times[4]);
/* Blocks until the request ends; copies T blocks (as scadyn_tmat_solve), the
 * iteration counts and the times; releases the request. Returns its error. */
int scadyn_server_wait(scadyn_server *sv, int64_t id, double *Taa, double *Tab, double *Tba, double *Tbb,
                       int32_t *iters, double times[4]);
/* Cancels a request: a queued one leaves the queue at once, an admitted one is
 * ended by the worker before its next solver step (its device memory and
 * footprint are released); either then waits as FAILED ("request cancelled").
 * A finished request is unaffected. */
int scadyn_server_cancel(scadyn_server *sv, int64_t id);
void scadyn_server_destroy(scadyn_server *sv);

#ifdef __cplusplus
}
#endif
#endif /* SCADYN_AMD_H */
