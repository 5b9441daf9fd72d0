// Modified Gram-Schmidt of the batched GMRES evaluated BS basis vectors per
// pass over the rows: the per-thread row sweep and the per-slot coefficient
// step, host+device so that the host self-test runs the same code on the CPU.
//
// The MGS recurrence h_k = <v_k, w - sum_{l<k} h_l v_l>
// (src/gmres_module.f90:241-244) is linear in w, so for the vectors j of one
// block, with d_j = <v_j, w> taken from w before the block's own updates and
// G_jl = <v_j, v_l>,
//   h_j = d_j - sum_{l<j in block} G_jl h_l
// (forward substitution).  Pass p applies the deferred update
// w -= sum h_l v_l of block p-1 and takes the dots and Gram entries of block p
// from the same w: one read of w, one write, 2 BS basis vectors, for BS
// coefficients (one vector per pass: one read of w, one write, 2 vectors, for
// one coefficient).  The last pass (p = number of blocks) applies the last
// update and takes |w|^2; pass 0 also takes |w|^2 before orthogonalisation (the
// reference's re-orthogonalisation test) and turns a restarting slot's W into
// rhs - A x and its |.|^2.
//
// Buffers (m = Krylov dimension, n rows, B slots, nwg row-sweep workgroups):
//   V >= (m+1) x n x B (complex128 or complex64)    W, Rhs n x B    H B x (m+1) x m
//   partials: dots/Gram nwg x NQ x B, norms nwg x B each
#pragma once

#include "krylov.hpp"

namespace scadyn {
namespace mgsb {

#define MGSB_HD __host__ __device__ inline

template <int BS>
struct Blk {
    static constexpr int NQ = BS + BS * (BS - 1) / 2;  // dots + strictly lower Gram entries
    // index of G_jl (l < j) among the partials
    static MGSB_HD int g(int j, int l) { return BS + j * (j - 1) / 2 + l; }
    // blocks of vectors 0..i
    static MGSB_HD int nblocks(int i) { return (i + BS) / BS; }
};

// what a slot does in pass p
struct PassRole {
    bool rst;     // restarting slot, pass 0: W = rhs - W, |W|^2
    bool arn;     // Arnoldi slot taking part in this pass
    int ulo, nu;  // update with vectors ulo .. ulo+nu-1
    int dlo, nd;  // dots with vectors dlo .. dlo+nd-1 (nd = 0: final |w|^2)
};

template <int BS>
MGSB_HD PassRole pass_role(const SlotState &s, int p) {
    PassRole r;
    r.rst = s.phase == PH_RESTART && p == 0;
    const int nb = Blk<BS>::nblocks(s.i);
    r.arn = s.phase == PH_ARNOLDI && p <= nb;
    r.ulo = (p - 1) * BS;
    r.dlo = p * BS;
    r.nu = r.arn && p >= 1 ? (s.i + 1 - r.ulo < BS ? s.i + 1 - r.ulo : BS) : 0;
    r.nd = r.arn && p < nb ? (s.i + 1 - r.dlo < BS ? s.i + 1 - r.dlo : BS) : 0;
    return r;
}

// one thread's rows (row0, row0 + rstep, ...) of slot b: partial dots/Gram in
// acc[NQ], |w|^2 after the pass's update in nrm, before orthogonalisation in nav.
// The basis is stored as VT (complex128, or complex64 or binary16 complex for
// the mixed-precision solve); all arithmetic is FP64.  w is read from Win (WI) and, where the pass
// changes it, written to Wout (WO): the same complex128 array, or for the
// mixed-precision solve the complex128 matvec output in, a complex64 copy out
// from the first pass that writes w on (passes 0 and 1), then the copy alone.
template <int BS, typename VT, typename WI, typename WO>
MGSB_HD void sweep_rows(const KrylovDims &d, int p, const WI *Win, WO *Wout, const VT *V, const cd *Rhs,
                        const SlotState *st, const cd *H, int b, long long row0, long long rstep, cd *acc,
                        double &nrm, double &nav) {
    constexpr int NQ = Blk<BS>::NQ;
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = cmk(0.0, 0.0);
    nrm = nav = 0.0;
    const SlotState s = st[b];
    const PassRole r = pass_role<BS>(s, p);
    if (!r.rst && !r.arn) return;
    const size_t nB = (size_t)d.n * d.B;
    cd h[BS];
#pragma unroll
    for (int u = 0; u < BS; ++u)
        h[u] = u < r.nu ? H[(size_t)b * (d.m + 1) * d.m + (size_t)(d.m + 1) * s.i + (r.ulo + u)] : cmk(0.0, 0.0);
    for (long long row = row0; row < d.n; row += rstep) {
        const size_t e = (size_t)row * d.B + b;
        if (r.rst) {
            const cd w = csub(Rhs[e], Prec<cd>::from(Win[e]));
            Wout[e] = Prec<WO>::from(w);
            nrm += cabs2(w);
            continue;
        }
        // Operands are fetched before any is used: the basis indices are
        // clamped into the block (a vector beyond nu or nd is read again as
        // vector 0 of the block and not used), so no load waits on a branch
        // and they are in flight at once (loads issued at their use wait one
        // by one).  With the complex128 basis only the dot vectors go ahead
        // and the update vectors are loaded at their use: all of them up front
        // take 140 VGPRs (3 waves per SIMD instead of 4), which costs more.
        constexpr bool AHEAD = sizeof(VT) <= 8;
        const WI wraw = Win[e];
        VT vu[BS], vd[BS];
        if constexpr (AHEAD) {
            if (r.nu > 0) {
#pragma unroll
                for (int u = 0; u < BS; ++u) vu[u] = V[(size_t)(r.ulo + (u < r.nu ? u : 0)) * nB + e];
            }
        }
        if (r.nd > 0) {
#pragma unroll
            for (int u = 0; u < BS; ++u) vd[u] = V[(size_t)(r.dlo + (u < r.nd ? u : 0)) * nB + e];
        }
        cd w = Prec<cd>::from(wraw);
        if (p == 0) nav += cabs2(w);
        if (r.nu > 0) {
#pragma unroll
            for (int u = 0; u < BS; ++u)
                if (u < r.nu) {
                    if constexpr (!AHEAD) vu[u] = V[(size_t)(r.ulo + u) * nB + e];
                    w = csub(w, cmul(h[u], Prec<cd>::from(vu[u])));
                }
            Wout[e] = Prec<WO>::from(w);
        }
        if (r.nd > 0) {
            cd v[BS];
#pragma unroll
            for (int u = 0; u < BS; ++u) v[u] = u < r.nd ? Prec<cd>::from(vd[u]) : cmk(0.0, 0.0);
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                acc[j] = cadd(acc[j], cmulc(v[j], w));
#pragma unroll
                for (int l = 0; l < j; ++l) acc[Blk<BS>::g(j, l)] = cadd(acc[Blk<BS>::g(j, l)], cmulc(v[j], v[l]));
            }
        } else {
            nrm += cabs2(w);
        }
    }
}

// slot b after the partials of pass p are summed (red[NQ], nrm, nav): the
// block's MGS coefficients into H column i, or the norms
template <int BS>
MGSB_HD void slot_coef(const KrylovDims &d, int p, const SlotState *st, const cd *red, double nrm, double nav,
                       cd *Hall, double *nrm2, double *nav2, int b) {
    const SlotState s = st[b];
    const PassRole r = pass_role<BS>(s, p);
    if (r.rst) {
        nrm2[b] = nrm;
        return;
    }
    if (!r.arn) return;
    if (p == 0) nav2[b] = nav;
    if (r.nd == 0) {
        nrm2[b] = nrm;
        return;
    }
    cd *H = Hall + (size_t)b * (d.m + 1) * d.m + (size_t)(d.m + 1) * s.i;
    // all BS rows (the partials of vectors beyond the slot's column are zero),
    // stored for the nd that exist; fixed trip counts keep h and red in registers
    cd h[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) {
        cd a = red[j];
#pragma unroll
        for (int l = 0; l < j; ++l) a = csub(a, cmul(red[Blk<BS>::g(j, l)], h[l]));
        h[j] = a;
        if (j < r.nd) H[r.dlo + j] = a;
    }
}

}  // namespace mgsb
}  // namespace scadyn
