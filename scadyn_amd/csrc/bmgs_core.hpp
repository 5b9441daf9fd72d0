// Block modified Gram-Schmidt for the batched GMRES (opts.ortho = 2): the
// per-thread and per-slot bodies of the kernels, host+device so that the host
// self-test runs the same code on the CPU.
//
// The MGS recurrence h_k = <v_k, w - sum_{l<k} h_l v_l> (src/gmres_module.f90:
// 241-244) is evaluated BS basis vectors per pass: pass p applies the deferred
// update w -= sum h_l v_l of block p-1 and takes, from the same w, the dots
// d_j = <v_j, w> of block p together with the block's Gram entries
// G_jl = <v_j, v_l> (l < j); then
//   h_j = d_j - sum_{l<j in block} G_jl h_l
// is exactly the MGS coefficient (forward substitution, coef_slot).  A pass
// reads w once and 2 BS basis vectors, against 2 per coefficient for one MGS
// pass per vector.  Pass nb = ceil((i+1)/BS) applies the last update and takes
// |w|^2; pass 0 also takes |w|^2 before orthogonalisation (the reference's
// re-orthogonalisation test) and turns restarting slots' W into rhs - A x.
//
// Buffers (m = Krylov dimension, n rows, B slots, nwg row-sweep workgroups):
//   V  (m+2) x n x B      W, Rhs  n x B      H  B x (m+1) x m
//   pq nwg x BQ x B       pnrm, pnav  nwg x B      red  B x (BQ+2)
#pragma once

#include "krylov.hpp"

namespace scadyn {
namespace bmgs {

#define BMGS_HD __host__ __device__ inline

constexpr int BS = 8;
constexpr int BQ = BS + BS * (BS - 1) / 2;  // dots + strictly lower Gram entries
BMGS_HD int gram_idx(int j, int l) { return BS + j * (j - 1) / 2 + l; }  // l < j
BMGS_HD int red_index(int q, int b) { return b * (BQ + 2) + q; }

// 256 threads: SB = min(B, 256) slots x RW = 256 / SB rows per sweep step
struct RowSlot {
    int SB, RW, sl, rl, b;
    bool ok;
};
BMGS_HD RowSlot row_slot(int B, int tid, int by) {
    RowSlot r;
    r.SB = B < 256 ? B : 256;
    r.RW = 256 / r.SB;
    r.sl = tid % r.SB;
    r.rl = tid / r.SB;
    r.b = by * r.SB + r.sl;
    r.ok = r.rl < r.RW && r.b < B;
    return r;
}

// basis blocks of an Arnoldi slot at column i: vectors 0..i in ceil((i+1)/BS) blocks
BMGS_HD int nblocks(int i) { return (i + BS) / BS; }

// rows of one thread: accumulates dots/Gram (acc), |w|^2 after (nrm) and before (nav)
BMGS_HD void pass_rows(const KrylovDims &d, int p, cd *W, const cd *V, const cd *Rhs, const SlotState *st,
                       const cd *H, const RowSlot &m, long long bx, long long gx, cd *acc, double &nrm,
                       double &nav) {
    for (int q = 0; q < BQ; ++q) acc[q] = cmk(0.0, 0.0);
    nrm = nav = 0.0;
    if (!m.ok) return;
    const SlotState s = st[m.b];
    const size_t nB = (size_t)d.n * d.B;
    const bool rst = s.phase == PH_RESTART && p == 0;
    const bool arn = s.phase == PH_ARNOLDI && p <= nblocks(s.i);
    if (!rst && !arn) return;
    const int nb = nblocks(s.i);
    const int ulo = (p - 1) * BS, uhi = (arn && p >= 1) ? (p * BS < s.i + 1 ? p * BS : s.i + 1) : 0;
    const int dlo = p * BS, dhi = (arn && p < nb) ? ((p + 1) * BS < s.i + 1 ? (p + 1) * BS : s.i + 1) : 0;
    cd h[BS];
    for (int u = 0; u < BS; ++u) {
        h[u] = cmk(0.0, 0.0);
        if (ulo + u >= 0 && ulo + u < uhi)
            h[u] = H[(size_t)m.b * (d.m + 1) * d.m + (size_t)(d.m + 1) * s.i + (ulo + u)];
    }
    for (long long row = bx * m.RW + m.rl; row < d.n; row += gx * m.RW) {
        const size_t e = (size_t)row * d.B + m.b;
        cd w = W[e];
        if (rst) {
            w = csub(Rhs[e], w);
            W[e] = w;
            nrm += cabs2(w);
            continue;
        }
        if (p == 0) nav += cabs2(w);
        if (uhi > 0) {
            for (int u = 0; u < BS; ++u)
                if (ulo + u >= 0 && ulo + u < uhi) w = csub(w, cmul(h[u], V[(size_t)(ulo + u) * nB + e]));
            W[e] = w;
        }
        if (dhi > dlo) {
            cd v[BS];
            for (int u = 0; u < BS; ++u) {
                v[u] = cmk(0.0, 0.0);
                if (dlo + u < dhi) v[u] = V[(size_t)(dlo + u) * nB + e];
            }
            for (int j = 0; j < BS; ++j) {
                acc[j] = cadd(acc[j], cmulc(v[j], w));
                for (int l = 0; l < j; ++l) acc[gram_idx(j, l)] = cadd(acc[gram_idx(j, l)], cmulc(v[j], v[l]));
            }
        } else {
            nrm += cabs2(w);
        }
    }
}

// workgroup partial of the norms: rows rl of slot sl summed in order
BMGS_HD void store_norm_partials(const KrylovDims &d, const RowSlot &m, long long bx, const double *sn1,
                                 const double *sn2, double *pnrm, double *pnav) {
    double a = 0, b = 0;
    for (int r = 0; r < m.RW; ++r) {
        a += sn1[r * m.SB + m.sl];
        b += sn2[r * m.SB + m.sl];
    }
    pnrm[(size_t)bx * d.B + m.b] = a;
    pnav[(size_t)bx * d.B + m.b] = b;
}
BMGS_HD void store_dot_partial(const KrylovDims &d, const RowSlot &m, long long bx, int q, const cd *sq, cd *pq) {
    cd a = cmk(0.0, 0.0);
    for (int r = 0; r < m.RW; ++r) a = cadd(a, sq[r * m.SB + m.sl]);
    pq[((size_t)bx * BQ + q) * d.B + m.b] = a;
}

// what partial q of slot b sums: part[w * stride + off] for w < nwg
struct ReduceSrc {
    enum Kind { NONE = 0, DOT = 1, NORM = 2, ZERO = 3 } kind;
    const cd *z;
    const double *r;
    size_t stride, off;
};
BMGS_HD ReduceSrc reduce_src(const KrylovDims &d, int p, const SlotState *st, const cd *pq, const double *pnrm,
                             const double *pnav, int q, int b) {
    ReduceSrc src;
    src.kind = ReduceSrc::NONE;
    src.z = nullptr;
    src.r = nullptr;
    src.stride = src.off = 0;
    const SlotState s = st[b];
    const int nb = nblocks(s.i);
    const bool rst = s.phase == PH_RESTART && p == 0;
    const bool arn = s.phase == PH_ARNOLDI && p <= nb;
    if (!rst && !arn) return src;
    src.kind = ReduceSrc::ZERO;
    if (q < BQ) {
        if (arn && p < nb) {
            src.kind = ReduceSrc::DOT;
            src.z = pq;
            src.stride = (size_t)BQ * d.B;
            src.off = (size_t)q * d.B + b;
        }
    } else if (q == BQ) {
        src.kind = ReduceSrc::NORM;
        src.r = pnrm;
        src.stride = d.B;
        src.off = b;
    } else if (arn && p == 0) {
        src.kind = ReduceSrc::NORM;
        src.r = pnav;
        src.stride = d.B;
        src.off = b;
    }
    return src;
}

// forward substitution h_j = d_j - sum_{l<j} G_jl h_l into H column i; norms
BMGS_HD void coef_slot(const KrylovDims &d, int p, const SlotState *st, const cd *red, cd *Hall, double *nrm2,
                       double *nav2, int b) {
    const SlotState s = st[b];
    const cd *r = red + (size_t)red_index(0, b);
    if (s.phase == PH_RESTART && p == 0) {
        nrm2[b] = r[BQ].x;
        return;
    }
    const int nb = nblocks(s.i);
    if (s.phase != PH_ARNOLDI || p > nb) return;
    if (p == 0) nav2[b] = r[BQ + 1].x;
    if (p == nb) {
        nrm2[b] = r[BQ].x;
        return;
    }
    cd *H = Hall + (size_t)b * (d.m + 1) * d.m + (size_t)(d.m + 1) * s.i;
    const int lo = p * BS, hi = (p + 1) * BS < s.i + 1 ? (p + 1) * BS : s.i + 1;
    cd h[BS];
    for (int j = 0; j < hi - lo; ++j) {
        cd a = r[j];
        for (int l = 0; l < j; ++l) a = csub(a, cmul(r[gram_idx(j, l)], h[l]));
        h[j] = a;
        H[lo + j] = a;
    }
}

}  // namespace bmgs
}  // namespace scadyn
