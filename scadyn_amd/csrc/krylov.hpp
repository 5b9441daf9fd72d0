// Batched restarted GMRES (src/gmres_module.f90:180-325) on RHS-innermost
// multi-vectors: every slot b runs its own GMRES(m) with the reference's
// semantics (x0 = rhs, modified Gram-Schmidt, complex Givens rotations,
// stopping on |g(i+1)|/||b|| < tol, `restart` cycles); slots advance one
// matvec per step and are refilled independently.
#pragma once

#include "internal.hpp"

namespace scadyn {

// PH_PARK: the slot's restart cycle has ended and it waits for the batch's
// next cycle boundary (aligned batches); every kernel treats it as idle.
enum SlotPhase : int { PH_IDLE = 0, PH_RESTART = 1, PH_ARNOLDI = 2, PH_DONE = 3, PH_PARK = 4 };
enum SlotAction : int { ACT_NONE = 0, ACT_V0 = 1, ACT_NEXT = 2, ACT_XUPD = 3 };

struct SlotState {
    int phase;
    int i;        // current Arnoldi column (0-based) while PH_ARNOLDI
    int cycle;    // restart cycle
    int ite;      // the reference's iteration counter
    int conv;     // 1 when |g(i+1)|/||b|| < tol was reached
    int act;      // SlotAction for the vector update
    int tgt;      // Krylov index written by ACT_V0 / ACT_NEXT
    int len;      // number of y entries for ACT_XUPD
    int reorth;   // times the reference's re-orthogonalisation test fired
    int col;      // RHS column held by the slot (-1 none)
    int pad[2];
    double bnorm, beta, hnorm, err;
};

struct KrylovDims {
    long long n;  // rows (3N)
    int B;        // slots
    int m;        // Krylov dimension (mesh%maxit)
    int nwg;      // workgroups of the reduction kernels
};

// V holds m+2 blocks of n*B: Krylov vectors 0..m and X at index m+1.
void launch_mgs_pass(const KrylovDims &d, int k, cd *W, const cd *V, const cd *Rhs, const SlotState *st,
                     const cd *H, cd *pdot, double *pnrm, double *pnav, hipStream_t s);
void launch_mgs_reduce(const KrylovDims &d, int k, const SlotState *st, cd *H, const cd *pdot,
                       const double *pnrm, const double *pnav, double *nrm2, double *nav2, hipStream_t s);
// verify = 1: restart residuals are true FP64 residuals and decide convergence
// against tol_true (mixed precision); an Arnoldi estimate below tol times the
// cycle's starting residual ends a cycle
// park = 1: a slot whose cycle ends goes to PH_PARK instead of PH_RESTART
void launch_slot_update(const KrylovDims &d, SlotState *st, cd *H, cd *cs, cd *sn, cd *g, cd *y,
                        const double *nrm2, const double *nav2, double tol, int restart, int verify, double tol_true,
                        int park, hipStream_t s);
// PH_PARK -> PH_RESTART for every slot (a cycle boundary of the batch)
void launch_unpark(const KrylovDims &d, SlotState *st, hipStream_t s);
// basis V stored as complex128 or (mixed precision) complex64; x in X (n x B)
void launch_vec_update(const KrylovDims &d, const SlotState *st, const cd *W, cd *V, cd *X, const cd *y,
                       hipStream_t s);
void launch_vec_update(const KrylovDims &d, const SlotState *st, const cf *W, cf *V, cd *X, const cd *y,
                       hipStream_t s);
void launch_vec_update(const KrylovDims &d, const SlotState *st, const cf *W, ch *V, cd *X, const cd *y,
                       hipStream_t s);
// the matvec inputs of a binary16 basis (block in_off[b] of V for active slots)
// converted to complex64 in xs (n x B)
void launch_stage_basis(const KrylovDims &d, const ch *V, const long long *in_off, const int *active, cf *xs,
                        hipStream_t s);
// load column col[b] (>= 0) of src (n x ld, row-major) into Rhs and X of slot b
void launch_load_rhs(const KrylovDims &d, const int *col, const cd *src, long long ld, cd *Rhs, cd *X,
                     double *pnrm, int x0zero, hipStream_t s);
void launch_load_finish(const KrylovDims &d, const int *col, SlotState *st, const double *pnrm,
                        hipStream_t s);
// store X of slots with col[b] >= 0 into column col[b] of dst (n x ld)
void launch_store_x(const KrylovDims &d, const int *col, const cd *X, cd *dst, long long ld, hipStream_t s);

// classical Gram-Schmidt + DGKS for column i of every Arnoldi slot (and the
// restart residual); kmax = max over slots of i.  pdot holds nwg*(m+1)*B.
void launch_cgs(const KrylovDims &d, cd *W, const cd *V, const cd *Rhs, const SlotState *st, cd *H, cd *pdot,
                double *pnrm, double *pnav, double *nrm2, double *nav2, int *flag, cd *cc, cd *hprev, int kmax,
                hipStream_t s);

// block MGS (MGS coefficients BS vectors per pass with in-pass Gram correction);
// kmax = max over Arnoldi slots of i.  pq holds nwg*bmgs_partials()*B, red
// B*(bmgs_partials()+2).
void launch_bmgs(const KrylovDims &d, cd *W, const cd *V, const cd *Rhs, const SlotState *st, cd *H, cd *pq,
                 double *pnrm, double *pnav, cd *red, double *nrm2, double *nav2, int kmax, hipStream_t s);
int bmgs_partials();

// Modified Gram-Schmidt evaluated bs (2, 4, 6 or 8) basis vectors per pass over the
// rows (mgsb_core.hpp): passes 0 .. ceil(K/bs) with K = max over Arnoldi slots
// of i+1 (0 when only restarting slots).  pdot holds nwg*mgsb_partials(bs)*B.
void launch_mgsb(const KrylovDims &d, int bs, int K, cd *W, const cd *V, const cd *Rhs, const SlotState *st, cd *H,
                 cd *pdot, double *pnrm, double *pnav, double *nrm2, double *nav2, hipStream_t s);
// mixed precision: complex64 basis, and w kept in the complex64 Wf once the
// first pass that changes it has run (read from the complex128 W before)
void launch_mgsb(const KrylovDims &d, int bs, int K, cd *W, cf *Wf, const cf *V, const cd *Rhs, const SlotState *st,
                 cd *H, cd *pdot, double *pnrm, double *pnav, double *nrm2, double *nav2, hipStream_t s);
// the same with the basis in binary16 (ch)
void launch_mgsb(const KrylovDims &d, int bs, int K, cd *W, cf *Wf, const ch *V, const cd *Rhs, const SlotState *st,
                 cd *H, cd *pdot, double *pnrm, double *pnav, double *nrm2, double *nav2, hipStream_t s);
int mgsb_partials(int bs);

// n x B block of the input array that slot s multiplies, and whether its
// result is used: a restarting slot multiplies x (block m+1 of the complex128
// V) unless its residual is taken elsewhere (then block 0, unused), an
// Arnoldi slot its basis vector i (0 <= i < m), others block 0 (unused).  The
// matvec reads every slot's input, so the block must exist: the complex64
// basis of the mixed-precision solve has m+1 blocks and no x.
__host__ __device__ inline long long input_block(const SlotState &s, int m, int restart_elsewhere, int &active) {
    active = 0;
    if (s.phase == PH_RESTART) {
        active = restart_elsewhere ? 0 : 1;
        return restart_elsewhere ? 0 : m + 1;
    }
    if (s.phase == PH_ARNOLDI) {
        active = 1;
        return s.i;
    }
    return 0;
}

// matvec input offsets / activity derived on the device from the slot phases;
// restart_elsewhere = 1 leaves restarting slots inactive (their residual
// matvec runs in the packed restart batch).  nblk: n x B blocks of the input
// array; *err is set (and the offset replaced by 0) for any offset outside it.
void launch_set_inputs(const KrylovDims &d, const SlotState *st, long long *in_off, int *active,
                       int restart_elsewhere, int nblk, int *err, hipStream_t s);
// columns slot[j] (j < cap, -1 none) of X (n x B) <-> a packed n x cap batch
void launch_pack_slots(const KrylovDims &d, const cd *X, const int *slot, int cap, cd *out, hipStream_t s);
void launch_unpack_slots(const KrylovDims &d, const cd *in, const int *slot, int cap, cd *W, hipStream_t s);
// mark slots with col[b] >= 0 idle
void launch_retire(const KrylovDims &d, const int *col, SlotState *st, hipStream_t s);

// T = mat^H X: mat is n x na, X n x nx (row-major); T is na x nx row-major (i, j).
// work holds the row-chunk partials (at least na*nx elements).
void launch_tgemm(const cd *mat, const cd *X, long long n, int na, int nx, cd *T, cd *work, size_t work_elems,
                  hipStream_t s);

}  // namespace scadyn
