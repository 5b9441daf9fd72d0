// Batched GMRES kernels and the T-matrix projection GEMM (FP64).
#include <algorithm>

#include "krylov.hpp"
#include "bmgs_core.hpp"
#include "mgsb_core.hpp"

namespace scadyn {

// ---- reductions over rows of an RHS-innermost multi-vector --------------------
// A workgroup of 256 threads covers SB = min(B,256) slots x RW = 256/SB rows per
// iteration; blockIdx.y selects the slot chunk.  Partial sums per workgroup are
// combined in a fixed order by the *_reduce kernels (deterministic).
struct RowSlotMap {
    int SB, RW, sl, rl, b;
    bool ok;
};
__device__ inline RowSlotMap row_slot_map(int B) {
    RowSlotMap r;
    r.SB = B < 256 ? B : 256;
    r.RW = 256 / r.SB;
    r.sl = threadIdx.x % r.SB;
    r.rl = threadIdx.x / r.SB;
    r.b = blockIdx.y * r.SB + r.sl;
    r.ok = r.rl < r.RW && r.b < B;
    return r;
}

// Deterministic sum over w < nwg of part[w * stride + off] by one workgroup of
// RED_T threads: fixed per-thread order, then a fixed LDS tree.  Every thread
// of the workgroup must call it; the result is returned to all of them.
constexpr int RED_T = 256;
template <typename T>
__device__ inline T zero_of();
template <>
__device__ inline double zero_of<double>() { return 0.0; }
template <>
__device__ inline cd zero_of<cd>() { return cmk(0.0, 0.0); }
__device__ inline double addv(double a, double b) { return a + b; }
__device__ inline cd addv(cd a, cd b) { return cadd(a, b); }

template <typename T>
__device__ T block_sum(const T *__restrict__ part, int nwg, size_t stride, size_t off) {
    __shared__ T sh[RED_T];
    T a = zero_of<T>();
    for (int w = threadIdx.x; w < nwg; w += RED_T) a = addv(a, part[(size_t)w * stride + off]);
    __syncthreads();  // sh may still be read by a previous call
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int h = RED_T / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = addv(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(256) void k_mgs_pass(KrylovDims d, int k, cd *__restrict__ W,
                                                  const cd *__restrict__ V, const cd *__restrict__ Rhs,
                                                  const SlotState *__restrict__ st, const cd *__restrict__ H,
                                                  cd *__restrict__ pdot, double *__restrict__ pnrm,
                                                  double *__restrict__ pnav) {
    __shared__ cd sdot[256];
    __shared__ double snrm[256], snav[256];
    RowSlotMap m = row_slot_map(d.B);
    cd dot = cmk(0, 0);
    double nrm = 0.0, nav = 0.0;
    if (m.ok) {
        const SlotState s = st[m.b];
        const size_t nB = (size_t)d.n * d.B;
        int mode = 0;  // 1: arnoldi pass, 2: restart residual
        if (s.phase == PH_ARNOLDI && k <= s.i + 1) mode = 1;
        if (s.phase == PH_RESTART && k == 0) mode = 2;
        if (mode) {
            cd h = cmk(0, 0);
            if (mode == 1 && k >= 1) h = H[(size_t)m.b * (d.m + 1) * d.m + (k - 1) + (size_t)(d.m + 1) * s.i];
            const cd *Vkm1 = V + (size_t)(k >= 1 ? k - 1 : 0) * nB;
            const cd *Vk = V + (size_t)k * nB;
            for (long long row = (long long)blockIdx.x * m.RW + m.rl; row < d.n; row += (long long)gridDim.x * m.RW) {
                const size_t e = (size_t)row * d.B + m.b;
                cd w = W[e];
                if (mode == 2) {
                    w = csub(Rhs[e], w);
                    W[e] = w;
                    nrm += cabs2(w);
                    continue;
                }
                if (k == 0) nav += cabs2(w);
                if (k >= 1) {
                    // w = w - h(k-1,i) v_{k-1}
                    const cd v = Vkm1[e];
                    w = csub(w, cmul(h, v));
                    W[e] = w;
                }
                if (k <= s.i) {
                    // h(k,i) = dot_product(v_k, w) = sum conj(v_k) w
                    const cd v = Vk[e];
                    dot = cadd(dot, cmulc(v, w));
                }
                if (k == s.i + 1) nrm += cabs2(w);
            }
        }
    }
    sdot[threadIdx.x] = dot;
    snrm[threadIdx.x] = nrm;
    snav[threadIdx.x] = nav;
    __syncthreads();
    if (m.ok && m.rl == 0) {
        cd a = cmk(0, 0);
        double b1 = 0, b2 = 0;
        for (int r = 0; r < m.RW; ++r) {
            a = cadd(a, sdot[r * m.SB + m.sl]);
            b1 += snrm[r * m.SB + m.sl];
            b2 += snav[r * m.SB + m.sl];
        }
        const size_t o = (size_t)blockIdx.x * d.B + m.b;
        pdot[o] = a;
        pnrm[o] = b1;
        pnav[o] = b2;
    }
}

// one workgroup per slot
__global__ __launch_bounds__(RED_T) void k_mgs_reduce(KrylovDims d, int k, const SlotState *__restrict__ st,
                                                      cd *__restrict__ H, const cd *__restrict__ pdot,
                                                      const double *__restrict__ pnrm,
                                                      const double *__restrict__ pnav, double *__restrict__ nrm2,
                                                      double *__restrict__ nav2) {
    const int b = blockIdx.x;
    const SlotState s = st[b];
    const bool arn = s.phase == PH_ARNOLDI && k <= s.i + 1;
    const bool rst = s.phase == PH_RESTART && k == 0;
    if (!arn && !rst) return;  // uniform over the workgroup
    const double x1 = block_sum<double>(pnrm, d.nwg, d.B, b);
    if (rst) {
        if (threadIdx.x == 0) nrm2[b] = x1;
        return;
    }
    const cd a = k <= s.i ? block_sum<cd>(pdot, d.nwg, d.B, b) : cmk(0.0, 0.0);
    const double x2 = k == 0 ? block_sum<double>(pnav, d.nwg, d.B, b) : 0.0;
    if (threadIdx.x == 0) {
        if (k <= s.i) H[(size_t)b * (d.m + 1) * d.m + k + (size_t)(d.m + 1) * s.i] = a;
        if (k == s.i + 1) nrm2[b] = x1;
        if (k == 0) nav2[b] = x2;
    }
}

// ---- MGS evaluated BS basis vectors per pass (opts.ortho = 0, mgs_block > 1) ----
// The per-thread row sweep and the per-slot coefficient step are in
// mgsb_core.hpp (shared with the host self-test).
template <int BS, typename VT, typename WI, typename WO>
__global__ __launch_bounds__(256) void k_mgsb_pass(KrylovDims d, int p, const WI *Win, WO *Wout,
                                                   const VT *__restrict__ V, const cd *__restrict__ Rhs,
                                                   const SlotState *__restrict__ st, const cd *__restrict__ H,
                                                   cd *__restrict__ pdot, double *__restrict__ pnrm,
                                                   double *__restrict__ pnav) {
    constexpr int NQ = mgsb::Blk<BS>::NQ;
    __shared__ cd sdot[256];
    __shared__ double snrm[256], snav[256];
    const RowSlotMap m = row_slot_map(d.B);
    cd acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = cmk(0.0, 0.0);
    double nrm = 0.0, nav = 0.0;
    if (m.ok)
        mgsb::sweep_rows<BS, VT, WI, WO>(d, p, Win, Wout, V, Rhs, st, H, m.b, (long long)blockIdx.x * m.RW + m.rl,
                             (long long)gridDim.x * m.RW, acc, nrm, nav);
    snrm[threadIdx.x] = nrm;
    snav[threadIdx.x] = nav;
    __syncthreads();
    if (m.ok && m.rl == 0) {
        double a = 0.0, c = 0.0;
        for (int r = 0; r < m.RW; ++r) {
            a += snrm[r * m.SB + m.sl];
            c += snav[r * m.SB + m.sl];
        }
        pnrm[(size_t)blockIdx.x * d.B + m.b] = a;
        pnav[(size_t)blockIdx.x * d.B + m.b] = c;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        __syncthreads();  // the previous partial's reads of sdot are done
        sdot[threadIdx.x] = acc[q];
        __syncthreads();
        if (m.ok && m.rl == 0) {
            cd a = cmk(0.0, 0.0);
            for (int r = 0; r < m.RW; ++r) a = cadd(a, sdot[r * m.SB + m.sl]);
            pdot[((size_t)blockIdx.x * NQ + q) * d.B + m.b] = a;
        }
    }
}

// one workgroup per slot: sums the pass's partials in a fixed order, then the
// block's coefficients (forward substitution) or the norms
template <int BS>
__global__ __launch_bounds__(RED_T) void k_mgsb_reduce(KrylovDims d, int p, const SlotState *__restrict__ st,
                                                       cd *__restrict__ H, const cd *__restrict__ pdot,
                                                       const double *__restrict__ pnrm,
                                                       const double *__restrict__ pnav, double *__restrict__ nrm2,
                                                       double *__restrict__ nav2) {
    constexpr int NQ = mgsb::Blk<BS>::NQ;
    const int b = blockIdx.x;
    const mgsb::PassRole r = mgsb::pass_role<BS>(st[b], p);
    if (!r.rst && !r.arn) return;  // uniform over the workgroup
    const double x1 = (r.rst || r.nd == 0) ? block_sum<double>(pnrm, d.nwg, d.B, b) : 0.0;
    const double x2 = (r.arn && p == 0) ? block_sum<double>(pnav, d.nwg, d.B, b) : 0.0;
    cd red[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        red[q] = r.nd > 0 ? block_sum<cd>(pdot, d.nwg, (size_t)NQ * d.B, (size_t)q * d.B + b) : cmk(0.0, 0.0);
    if (threadIdx.x == 0) mgsb::slot_coef<BS>(d, p, st, red, x1, x2, H, nrm2, nav2, b);
}

// passes 0 .. P; w in Wd (complex128) throughout, or (Wf given) read from Wd
// in passes 0 and 1 and kept in the complex64 Wf from the first write on
template <int BS, typename VT>
static void launch_mgsb_t(const KrylovDims &d, int K, cd *Wd, cf *Wf, const VT *V, const cd *Rhs,
                          const SlotState *st, cd *H, cd *pdot, double *pnrm, double *pnav, double *nrm2,
                          double *nav2, hipStream_t s) {
    const int P = (K + BS - 1) / BS;  // blocks of the vectors 0 .. K-1; pass P takes the last norm
    const int SB = d.B < 256 ? d.B : 256;
    const dim3 grid((unsigned)d.nwg, (unsigned)((d.B + SB - 1) / SB));
    for (int p = 0; p <= P; ++p) {
        if (!Wf)
            hipLaunchKernelGGL((k_mgsb_pass<BS, VT, cd, cd>), grid, dim3(256), 0, s, d, p, (const cd *)Wd, Wd, V, Rhs,
                               st, H, pdot, pnrm, pnav);
        else if (p <= 1)
            hipLaunchKernelGGL((k_mgsb_pass<BS, VT, cd, cf>), grid, dim3(256), 0, s, d, p, (const cd *)Wd, Wf, V, Rhs,
                               st, H, pdot, pnrm, pnav);
        else
            hipLaunchKernelGGL((k_mgsb_pass<BS, VT, cf, cf>), grid, dim3(256), 0, s, d, p, (const cf *)Wf, Wf, V, Rhs,
                               st, H, pdot, pnrm, pnav);
        hipLaunchKernelGGL((k_mgsb_reduce<BS>), dim3((unsigned)d.B), dim3(RED_T), 0, s, d, p, st, H, pdot, pnrm,
                           pnav, nrm2, nav2);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename VT>
static void launch_mgsb_v(const KrylovDims &d, int bs, int K, cd *Wd, cf *Wf, const VT *V, const cd *Rhs,
                          const SlotState *st, cd *H, cd *pdot, double *pnrm, double *pnav, double *nrm2,
                          double *nav2, hipStream_t s) {
    switch (bs) {
        case 2: launch_mgsb_t<2, VT>(d, K, Wd, Wf, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s); break;
        case 4: launch_mgsb_t<4, VT>(d, K, Wd, Wf, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s); break;
        case 6: launch_mgsb_t<6, VT>(d, K, Wd, Wf, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s); break;
        case 8: launch_mgsb_t<8, VT>(d, K, Wd, Wf, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s); break;
        default: fail(SCADYN_ERR_ARG, "MGS block size must be 2, 4, 6 or 8");
    }
}
void launch_mgsb(const KrylovDims &d, int bs, int K, cd *W, const cd *V, const cd *Rhs, const SlotState *st, cd *H,
                 cd *pdot, double *pnrm, double *pnav, double *nrm2, double *nav2, hipStream_t s) {
    launch_mgsb_v<cd>(d, bs, K, W, nullptr, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s);
}
void launch_mgsb(const KrylovDims &d, int bs, int K, cd *W, cf *Wf, const cf *V, const cd *Rhs, const SlotState *st,
                 cd *H, cd *pdot, double *pnrm, double *pnav, double *nrm2, double *nav2, hipStream_t s) {
    launch_mgsb_v<cf>(d, bs, K, W, Wf, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s);
}
void launch_mgsb(const KrylovDims &d, int bs, int K, cd *W, cf *Wf, const ch *V, const cd *Rhs, const SlotState *st,
                 cd *H, cd *pdot, double *pnrm, double *pnav, double *nrm2, double *nav2, hipStream_t s) {
    launch_mgsb_v<ch>(d, bs, K, W, Wf, V, Rhs, st, H, pdot, pnrm, pnav, nrm2, nav2, s);
}
int mgsb_partials(int bs) { return bs + bs * (bs - 1) / 2; }

// ---- block modified Gram-Schmidt (opts.ortho = 2) ------------------------------
// The per-thread / per-slot bodies are in bmgs_core.hpp (shared with the host
// self-test, which runs them on the CPU under AddressSanitizer).
using namespace bmgs;

__global__ __launch_bounds__(256) void k_bmgs_pass(KrylovDims d, int p, cd *__restrict__ W,
                                                   const cd *__restrict__ V, const cd *__restrict__ Rhs,
                                                   const SlotState *__restrict__ st, const cd *__restrict__ H,
                                                   cd *__restrict__ pq, double *__restrict__ pnrm,
                                                   double *__restrict__ pnav) {
    __shared__ cd sq[256];
    __shared__ double sn1[256], sn2[256];
    const RowSlot m = row_slot(d.B, threadIdx.x, blockIdx.y);
    cd acc[BQ];
    double nrm = 0.0, nav = 0.0;
    pass_rows(d, p, W, V, Rhs, st, H, m, blockIdx.x, gridDim.x, acc, nrm, nav);
    sn1[threadIdx.x] = nrm;
    sn2[threadIdx.x] = nav;
    __syncthreads();
    if (m.ok && m.rl == 0) store_norm_partials(d, m, blockIdx.x, sn1, sn2, pnrm, pnav);
    for (int q = 0; q < BQ; ++q) {
        __syncthreads();
        sq[threadIdx.x] = acc[q];
        __syncthreads();
        if (m.ok && m.rl == 0) store_dot_partial(d, m, blockIdx.x, q, sq, pq);
    }
}

// one workgroup per (partial q, slot): q < BQ dots/Gram, BQ |w|^2, BQ+1 |w0|^2
__global__ __launch_bounds__(RED_T) void k_bmgs_reduce(KrylovDims d, int p, const SlotState *__restrict__ st,
                                                       const cd *__restrict__ pq, const double *__restrict__ pnrm,
                                                       const double *__restrict__ pnav, cd *__restrict__ red) {
    const int q = blockIdx.x, b = blockIdx.y;
    const ReduceSrc src = reduce_src(d, p, st, pq, pnrm, pnav, q, b);
    if (src.kind == ReduceSrc::NONE) return;  // uniform over the workgroup
    cd a = cmk(0.0, 0.0);
    if (src.kind == ReduceSrc::DOT) a = block_sum<cd>(src.z, d.nwg, src.stride, src.off);
    if (src.kind == ReduceSrc::NORM) a.x = block_sum<double>(src.r, d.nwg, src.stride, src.off);
    if (threadIdx.x == 0) red[red_index(q, b)] = a;
}

__global__ void k_bmgs_coef(KrylovDims d, int p, const SlotState *__restrict__ st, const cd *__restrict__ red,
                            cd *__restrict__ Hall, double *__restrict__ nrm2, double *__restrict__ nav2) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    coef_slot(d, p, st, red, Hall, nrm2, nav2, b);
}

// ---- classical Gram-Schmidt with DGKS re-orthogonalisation (opts.ortho = 1) ----
// Pass "dots": h_k = <v_k, w> for the 16 basis vectors k in [16 z, 16 z + 16) of
// blockIdx.z (z = 0 also accumulates |w|^2 and turns restarting slots' W into
// the residual rhs - A x).  Pass "update": w -= sum_k h_k v_k and |w|^2.  Pass 2
// repeats both for slots whose norm dropped below 1/sqrt(2) of its value.
constexpr int CGS_KC = 16;

__global__ __launch_bounds__(256) void k_cgs_dots(KrylovDims d, int pass, cd *__restrict__ W,
                                                  const cd *__restrict__ V, const cd *__restrict__ Rhs,
                                                  const SlotState *__restrict__ st, const int *__restrict__ flag,
                                                  cd *__restrict__ pdot, double *__restrict__ pnav) {
    __shared__ cd sred[256];
    __shared__ double snav[256];
    RowSlotMap m = row_slot_map(d.B);
    const int k0 = blockIdx.z * CGS_KC;
    cd acc[CGS_KC];
#pragma unroll
    for (int q = 0; q < CGS_KC; ++q) acc[q] = cmk(0.0, 0.0);
    double nav = 0.0;
    int kmax = -1;  // highest basis index of this slot in this chunk
    bool rst = false;
    if (m.ok) {
        const SlotState s = st[m.b];
        const size_t nB = (size_t)d.n * d.B;
        const bool arn = s.phase == PH_ARNOLDI && (pass == 0 || flag[m.b]);
        rst = pass == 0 && s.phase == PH_RESTART && blockIdx.z == 0;
        if (arn) kmax = min(s.i, k0 + CGS_KC - 1);
        if (arn || rst)
            for (long long row = (long long)blockIdx.x * m.RW + m.rl; row < d.n; row += (long long)gridDim.x * m.RW) {
                const size_t e = (size_t)row * d.B + m.b;
                cd w = W[e];
                if (rst) {
                    w = csub(Rhs[e], w);
                    W[e] = w;
                    nav += cabs2(w);
                    continue;
                }
                if (blockIdx.z == 0) nav += cabs2(w);
#pragma unroll
                for (int q = 0; q < CGS_KC; ++q)
                    if (k0 + q <= kmax) acc[q] = cadd(acc[q], cmulc(V[(size_t)(k0 + q) * nB + e], w));
            }
    }
    snav[threadIdx.x] = nav;
    __syncthreads();
    if (m.ok && m.rl == 0 && blockIdx.z == 0) {
        double a = 0;
        for (int r = 0; r < m.RW; ++r) a += snav[r * m.SB + m.sl];
        pnav[(size_t)blockIdx.x * d.B + m.b] = a;
    }
    const int nk = d.m + 1;
    for (int q = 0; q < CGS_KC; ++q) {
        if (k0 + q >= nk) break;
        __syncthreads();
        sred[threadIdx.x] = acc[q];
        __syncthreads();
        if (m.ok && m.rl == 0) {
            cd a = cmk(0.0, 0.0);
            for (int r = 0; r < m.RW; ++r) a = cadd(a, sred[r * m.SB + m.sl]);
            pdot[((size_t)blockIdx.x * nk + (k0 + q)) * d.B + m.b] = a;
        }
    }
}

// sum partial dots into H column i (pass 0 stores, pass 1 adds); |w|^2 of
// restart slots.  One workgroup per slot.
__global__ __launch_bounds__(RED_T) void k_cgs_dots_reduce(KrylovDims d, int pass, const SlotState *__restrict__ st,
                                                           const int *__restrict__ flag, cd *__restrict__ Hall,
                                                           const cd *__restrict__ pdot,
                                                           const double *__restrict__ pnav,
                                                           double *__restrict__ nav2, double *__restrict__ nrm2) {
    const int nk = d.m + 1;
    const int b = blockIdx.x;
    const SlotState s = st[b];
    if (pass == 0 && (s.phase == PH_ARNOLDI || s.phase == PH_RESTART)) {
        const double a = block_sum<double>(pnav, d.nwg, d.B, b);
        if (threadIdx.x == 0) {
            if (s.phase == PH_ARNOLDI)
                nav2[b] = a;
            else
                nrm2[b] = a;
        }
    }
    if (s.phase != PH_ARNOLDI || (pass == 1 && !flag[b])) return;
    cd *H = Hall + (size_t)b * (d.m + 1) * d.m + (size_t)(d.m + 1) * s.i;
    for (int k = 0; k <= s.i; ++k) {
        const cd a = block_sum<cd>(pdot, d.nwg, (size_t)nk * d.B, (size_t)k * d.B + b);
        if (threadIdx.x == 0) H[k] = pass == 0 ? a : cadd(H[k], a);
    }
}

// w -= sum_{k<=i} c_k v_k with c = the dots of this pass (stored in H, or the
// pass-2 increments kept in cc); partial |w|^2
__global__ __launch_bounds__(256) void k_cgs_update(KrylovDims d, int pass, cd *__restrict__ W,
                                                    const cd *__restrict__ V, const SlotState *__restrict__ st,
                                                    const int *__restrict__ flag, const cd *__restrict__ cc,
                                                    double *__restrict__ pnrm) {
    __shared__ double snrm[256];
    RowSlotMap m = row_slot_map(d.B);
    double nrm = 0.0;
    if (m.ok) {
        const SlotState s = st[m.b];
        const size_t nB = (size_t)d.n * d.B;
        if (s.phase == PH_ARNOLDI && (pass == 0 || flag[m.b])) {
            const cd *c = cc + (size_t)m.b * (d.m + 1);
            for (long long row = (long long)blockIdx.x * m.RW + m.rl; row < d.n; row += (long long)gridDim.x * m.RW) {
                const size_t e = (size_t)row * d.B + m.b;
                cd w = W[e];
                for (int k = 0; k <= s.i; ++k) w = csub(w, cmul(c[k], V[(size_t)k * nB + e]));
                W[e] = w;
                nrm += cabs2(w);
            }
        }
    }
    snrm[threadIdx.x] = nrm;
    __syncthreads();
    if (m.ok && m.rl == 0) {
        double a = 0;
        for (int r = 0; r < m.RW; ++r) a += snrm[r * m.SB + m.sl];
        pnrm[(size_t)blockIdx.x * d.B + m.b] = a;
    }
}

// copy the coefficients of this pass (H column, or H - previous for pass 2),
// reduce |w|^2 and decide re-orthogonalisation (DGKS: |w| < |w_before| / sqrt 2)
__global__ void k_cgs_coef(KrylovDims d, int pass, const SlotState *__restrict__ st, const cd *__restrict__ Hall,
                           cd *__restrict__ cc, cd *__restrict__ hprev) {
    const int b = blockIdx.x;
    if (b >= d.B) return;
    const SlotState s = st[b];
    if (s.phase != PH_ARNOLDI) return;
    const cd *H = Hall + (size_t)b * (d.m + 1) * d.m + (size_t)(d.m + 1) * s.i;
    cd *c = cc + (size_t)b * (d.m + 1);
    cd *hp = hprev + (size_t)b * (d.m + 1);
    for (int k = threadIdx.x; k <= s.i; k += blockDim.x) {
        if (pass == 0) {
            c[k] = H[k];
            hp[k] = H[k];
        } else {
            c[k] = csub(H[k], hp[k]);
        }
    }
}

__global__ __launch_bounds__(RED_T) void k_cgs_norm(KrylovDims d, int pass, const SlotState *__restrict__ st,
                                                    const double *__restrict__ pnrm, const double *__restrict__ nav2,
                                                    int *__restrict__ flag, double *__restrict__ nrm2) {
    const int b = blockIdx.x;
    const SlotState s = st[b];
    if (s.phase != PH_ARNOLDI) {
        if (pass == 0 && threadIdx.x == 0) flag[b] = 0;
        return;
    }
    if (pass == 1 && !flag[b]) return;
    const double a = block_sum<double>(pnrm, d.nwg, d.B, b);
    if (threadIdx.x == 0) {
        nrm2[b] = a;
        if (pass == 0) flag[b] = (a < 0.5 * nav2[b]) ? 1 : 0;  // |w|^2 < |w0|^2 / 2
    }
}

// per-slot scalar work: Givens rotations, convergence test, least squares
// verify: restart residuals are true (FP64) residuals; a slot is done only when
// one is <= tol_true ||b||, and an Arnoldi estimate below tol times the cycle's
// starting residual ends the cycle early (x update, then a restart that checks
// the true residual)
__global__ void k_slot_update(KrylovDims d, SlotState *__restrict__ st, cd *__restrict__ Hall,
                              cd *__restrict__ csall, cd *__restrict__ snall, cd *__restrict__ gall,
                              cd *__restrict__ yall, const double *__restrict__ nrm2,
                              const double *__restrict__ nav2, double tol, int restart, int verify,
                              double tol_true, int park) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    SlotState s = st[b];
    const int m = d.m, ld = m + 1;
    cd *H = Hall + (size_t)b * ld * m;
    cd *cs = csall + (size_t)b * ld;
    cd *sn = snall + (size_t)b * ld;
    cd *g = gall + (size_t)b * ld;
    cd *y = yall + (size_t)b * m;
    s.act = ACT_NONE;
    if (s.phase == PH_RESTART) {
        const double beta = sqrt(nrm2[b]);
        for (int q = 0; q <= m; ++q) g[q] = cmk(0, 0);
        g[0] = cmk(beta, 0);
        s.beta = beta;
        if (beta == 0.0) {  // x is exact; the reference would divide by zero here
            s.phase = PH_DONE;
            s.conv = 1;
        } else if (verify && beta <= tol_true * s.bnorm) {
            s.phase = PH_DONE;
            s.conv = 1;
        } else {
            s.phase = PH_ARNOLDI;
            s.i = 0;
            s.act = ACT_V0;
            s.tgt = 0;
        }
    } else if (s.phase == PH_ARNOLDI) {
        const int i = s.i;
        const double hn = sqrt(nrm2[b]);
        const double nav = sqrt(nav2[b]);
        H[(i + 1) + (size_t)ld * i] = cmk(hn, 0);
        if (nav + 0.001 * hn == nav) s.reorth++;
        s.hnorm = hn;
        for (int kk = 0; kk < i; ++kk) {
            cd hk = H[kk + (size_t)ld * i], hk1 = H[kk + 1 + (size_t)ld * i];
            cd tmp = csub(cmul(cs[kk], hk), cmul(sn[kk], hk1));
            H[kk + 1 + (size_t)ld * i] = cadd(cmul(sn[kk], hk), cmul(cconj(cs[kk]), hk1));
            H[kk + (size_t)ld * i] = tmp;
        }
        cd hii = H[i + (size_t)ld * i], hi1 = H[i + 1 + (size_t)ld * i];
        const double nu = sqrt(cabs2(hii) + cabs2(hi1));
        if (nu != 0.0) {
            cs[i] = cconj(cmk(hii.x / nu, hii.y / nu));
            sn[i] = cmk(-hi1.x / nu, -hi1.y / nu);
            H[i + (size_t)ld * i] = csub(cmul(cs[i], hii), cmul(sn[i], hi1));
            H[i + 1 + (size_t)ld * i] = cmk(0, 0);
            cd t1 = g[i], t2 = g[i + 1];
            g[i] = csub(cmul(cs[i], t1), cmul(sn[i], t2));
            g[i + 1] = cadd(cmul(sn[i], t1), cmul(cconj(cs[i]), t2));
        }
        // the reference's |g(i+1)|/||b||; with verify, relative to the cycle's
        // starting (true) residual, so that every cycle is one correction step
        s.err = sqrt(cabs2(g[i + 1])) / (verify ? s.beta : s.bnorm);
        int len = 0;
        if (s.err < tol) {
            len = i + 1;
            if (verify) {
                s.cycle++;
                s.phase = (s.cycle >= restart) ? PH_DONE : (park ? PH_PARK : PH_RESTART);
            } else {
                s.conv = 1;
                s.phase = PH_DONE;
            }
        } else {
            s.ite++;
            if (i + 1 == m) {
                len = m;
                s.cycle++;
                s.phase = (s.cycle >= restart) ? PH_DONE : (park ? PH_PARK : PH_RESTART);
            } else {
                s.act = ACT_NEXT;
                s.tgt = i + 1;
                s.i = i + 1;
            }
        }
        if (len > 0) {
            // back substitution on the rotated (upper triangular) H(1:len,1:len)
            bool sing = false;
            for (int j = len - 1; j >= 0; --j) {
                cd acc = g[j];
                for (int l = j + 1; l < len; ++l) acc = csub(acc, cmul(H[j + (size_t)ld * l], y[l]));
                cd djj = H[j + (size_t)ld * j];
                if (djj.x == 0.0 && djj.y == 0.0) sing = true;
                y[j] = cdiv(acc, djj);
            }
            if (sing) {
                s.conv = -1;
                s.phase = PH_DONE;
            } else {
                s.act = ACT_XUPD;
                s.len = len;
            }
        }
    }
    st[b] = s;
}

// new basis vector v = w / h (stored as VT), or at a cycle's end x += V y
template <typename VT, typename WT>
__global__ __launch_bounds__(256) void k_vec_update(KrylovDims d, const SlotState *__restrict__ st,
                                                    const WT *__restrict__ W, VT *__restrict__ V,
                                                    cd *__restrict__ X, const cd *__restrict__ yall) {
    // one slot per thread (its state read once), rows strided as in the sweeps
    const RowSlotMap m = row_slot_map(d.B);
    if (!m.ok) return;
    const int b = m.b;
    const SlotState s = st[b];
    const size_t nB = (size_t)d.n * d.B;
    const long long r0 = (long long)blockIdx.x * m.RW + m.rl, rs = (long long)gridDim.x * m.RW;
    if (s.act == ACT_V0 || s.act == ACT_NEXT) {
        const double h = (s.act == ACT_V0) ? s.beta : s.hnorm;
        // a complex64 or binary16 basis takes w times one reciprocal per slot
        // (the quotient is rounded to the stored precision anyway, and two FP64
        // divisions per element made this sweep instruction-bound); the
        // complex128 basis divides, as the reference does
        const double hinv = h != 0.0 ? 1.0 / h : 1.0;
        VT *vt = V + (size_t)s.tgt * nB;
        for (long long row = r0; row < d.n; row += rs) {
            const size_t e = (size_t)row * d.B + b;
            cd w = Prec<cd>::from(W[e]);
            if constexpr (sizeof(VT) < sizeof(cd))
                w = cmk(w.x * hinv, w.y * hinv);
            else if (h != 0.0)
                w = cmk(w.x / h, w.y / h);
            vt[e] = Prec<VT>::from(w);
        }
    } else if (s.act == ACT_XUPD) {
        const cd *y = yall + (size_t)b * d.m;
        for (long long row = r0; row < d.n; row += rs) {
            const size_t e = (size_t)row * d.B + b;
            cd acc = cmk(0, 0);
            // unrolled so that the loads of several basis vectors are in flight
            // at once; the sum still runs in q order
#pragma unroll 8
            for (int q = 0; q < s.len; ++q) cfma(acc, Prec<cd>::from(V[(size_t)q * nB + e]), y[q]);
            X[e] = cadd(X[e], acc);
        }
    }
}

__global__ __launch_bounds__(256) void k_load_rhs(KrylovDims d, const int *__restrict__ col,
                                                  const cd *__restrict__ src, long long ld,
                                                  cd *__restrict__ Rhs, cd *__restrict__ X,
                                                  double *__restrict__ pnrm, int x0zero) {
    __shared__ double snrm[256];
    RowSlotMap m = row_slot_map(d.B);
    double nrm = 0.0;
    if (m.ok && col[m.b] >= 0) {
        const size_t nB = (size_t)d.n * d.B;
        const int c = col[m.b];
        for (long long row = (long long)blockIdx.x * m.RW + m.rl; row < d.n; row += (long long)gridDim.x * m.RW) {
            const size_t e = (size_t)row * d.B + m.b;
            cd v = src[(size_t)row * ld + c];
            Rhs[e] = v;
            // x0 = rhs as the reference (src/gmres_module.f90:198), or 0 for correction equations
            X[e] = x0zero ? cmk(0, 0) : v;
            nrm += cabs2(v);
        }
    }
    snrm[threadIdx.x] = nrm;
    __syncthreads();
    if (m.ok && m.rl == 0) {
        double a = 0;
        for (int r = 0; r < m.RW; ++r) a += snrm[r * m.SB + m.sl];
        pnrm[(size_t)blockIdx.x * d.B + m.b] = a;
    }
}

// one workgroup per slot
__global__ __launch_bounds__(RED_T) void k_load_finish(KrylovDims d, const int *__restrict__ col,
                                                       SlotState *__restrict__ st, const double *__restrict__ pnrm) {
    const int b = blockIdx.x;
    if (col[b] < 0) return;
    const double a = block_sum<double>(pnrm, d.nwg, d.B, b);
    if (threadIdx.x != 0) return;
    SlotState s;
    s.phase = PH_RESTART;
    s.i = 0;
    s.cycle = 0;
    s.ite = 0;
    s.conv = 0;
    s.act = ACT_NONE;
    s.tgt = 0;
    s.len = 0;
    s.reorth = 0;
    s.col = col[b];
    s.pad[0] = s.pad[1] = 0;
    s.bnorm = sqrt(a);
    s.beta = 0;
    s.hnorm = 0;
    s.err = 0;
    if (s.bnorm == 0.0) {  // zero right-hand side: x = 0
        s.phase = PH_DONE;
        s.conv = 1;
    }
    st[b] = s;
}

__global__ void k_store_x(KrylovDims d, const int *__restrict__ col, const cd *__restrict__ X,
                          cd *__restrict__ dst, long long ld) {
    const long long tot = d.n * d.B;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e % d.B);
        const int c = col[b];
        if (c < 0) continue;
        const long long row = e / d.B;
        dst[(size_t)row * ld + c] = X[e];
    }
}

// matvec input offsets and activity from the slot phases: a restarting slot
// multiplies x (block m+1 of the complex128 V), an Arnoldi slot its current
// basis vector.  The matvec kernels read every slot's input, active or not, so
// every offset must stay inside the input array: a restarting slot whose
// residual is taken elsewhere (the mixed-precision solve, whose complex64 basis
// has only m+1 blocks and no x) reads block 0.
// nblk: n x B blocks in the input array; an offset outside it (a state the
// solver never produces) is replaced by block 0 and reported in *err, so that
// no matvec reads outside the array
__global__ void k_set_inputs(KrylovDims d, const SlotState *__restrict__ st, long long *__restrict__ in_off,
                             int *__restrict__ active, int restart_elsewhere, int nblk, int *__restrict__ err) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    const long long nB = d.n * d.B;
    int a = 0;
    long long blk = input_block(st[b], d.m, restart_elsewhere, a);
    if (blk < 0 || blk >= nblk) {
        blk = 0;
        a = 0;
        *err = 1;
    }
    in_off[b] = blk * nB;
    active[b] = a;
}

// slots whose column was stored become idle
__global__ void k_unpark(KrylovDims d, SlotState *__restrict__ st) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < d.B && st[b].phase == PH_PARK) st[b].phase = PH_RESTART;
}

__global__ void k_retire(KrylovDims d, const int *__restrict__ col, SlotState *__restrict__ st) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B || col[b] < 0) return;
    st[b].phase = PH_IDLE;
    st[b].act = ACT_NONE;
}

// ---- T = mat^H X, split over row chunks -----------------------------------------
// T = mat^H X on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  Complex
// products as four real ones: Re T = Mr^T Xr + Mi^T Xi, Im T = Mr^T Xi - Mi^T Xr.
// A workgroup (4 waves) owns a 64 x 64 tile of T and one chunk of rows; each
// wave a 32 x 32 quarter as 2 x 2 blocks of 16 x 16.  Operand layout of the
// 16x16x4 f64 instruction: A[i][k] and B[k][j] in lane i + 16 k (resp. j + 16 k);
// D[i][j] in lane j + 16 (i % 4), element i / 4.  blockIdx.z (the row chunk)
// is the slowest grid index, so the 36-ish tiles of one chunk run together
// and read its rows of mat and X from HBM about once.
typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_tgemm(const cd *__restrict__ A, const cd *__restrict__ X, long long n,
                                               int na, int nx, long long chunk, cd *__restrict__ part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kk = lane >> 4, ii = lane & 15;
    const int i0 = blockIdx.x * 64 + (wv >> 1) * 32, j0 = blockIdx.y * 64 + (wv & 1) * 32;
    const long long r0 = blockIdx.z * chunk, r1 = std::min(n, r0 + chunk);
    v4d accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) accR[a][b] = accI[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int ca0 = i0 + ii, ca1 = i0 + 16 + ii, cb0 = j0 + ii, cb1 = j0 + 16 + ii;
#pragma unroll 2
    for (long long r = r0; r < r1; r += 4) {
        const long long row = r + kk;
        const bool rok = row < r1;
        const cd *ar = A + (size_t)row * na, *xr = X + (size_t)row * nx;
        const cd z = cmk(0.0, 0.0);
        const cd a0 = (rok && ca0 < na) ? ar[ca0] : z, a1 = (rok && ca1 < na) ? ar[ca1] : z;
        const cd b0 = (rok && cb0 < nx) ? xr[cb0] : z, b1 = (rok && cb1 < nx) ? xr[cb1] : z;
        const double am[2][2] = {{a0.x, a0.y}, {a1.x, a1.y}};
        const double bm[2][2] = {{b0.x, b0.y}, {b1.x, b1.y}};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                accR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[a][0], bm[b][0], accR[a][b], 0, 0, 0);
                accR[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[a][1], bm[b][1], accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(am[a][0], bm[b][1], accI[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-am[a][1], bm[b][0], accI[a][b], 0, 0, 0);
            }
    }
    cd *out = part + (size_t)blockIdx.z * na * nx;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = i0 + 16 * a + kk + 4 * e, j = j0 + 16 * b + ii;
                if (i < na && j < nx) out[(size_t)i * nx + j] = cmk(accR[a][b][e], accI[a][b][e]);
            }
}

__global__ void k_tgemm_reduce(const cd *__restrict__ part, int nz, size_t tot, cd *__restrict__ T) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        cd a = cmk(0, 0);
        for (int z = 0; z < nz; ++z) a = cadd(a, part[(size_t)z * tot + e]);
        T[e] = a;
    }
}

// restart batch: columns slot[j] (j < cap, -1 none) of X (n x B) into the
// packed n x cap batch, and the packed results back into W (n x B)
__global__ void k_pack_slots(KrylovDims d, const cd *__restrict__ X, const int *__restrict__ slot, int cap,
                             cd *__restrict__ out) {
    const long long tot = d.n * cap;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e % cap);
        const long long row = e / cap;
        const int b = slot[j];
        out[e] = b >= 0 ? X[(size_t)row * d.B + b] : cmk(0.0, 0.0);
    }
}
__global__ void k_unpack_slots(KrylovDims d, const cd *__restrict__ in, const int *__restrict__ slot, int cap,
                               cd *__restrict__ W) {
    const long long tot = d.n * cap;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e % cap);
        const int b = slot[j];
        if (b < 0) continue;
        const long long row = e / cap;
        W[(size_t)row * d.B + b] = in[e];
    }
}

// ---- launchers -------------------------------------------------------------------------
static dim3 rs_grid(const KrylovDims &d) {
    int SB = std::min(d.B, 256);
    return dim3((unsigned)d.nwg, (unsigned)((d.B + SB - 1) / SB));
}

void launch_mgs_pass(const KrylovDims &d, int k, cd *W, const cd *V, const cd *Rhs, const SlotState *st,
                     const cd *H, cd *pdot, double *pnrm, double *pnav, hipStream_t s) {
    hipLaunchKernelGGL(k_mgs_pass, rs_grid(d), dim3(256), 0, s, d, k, W, V, Rhs, st, H, pdot, pnrm, pnav);
    HIP_CHECK(hipGetLastError());
}
void launch_mgs_reduce(const KrylovDims &d, int k, const SlotState *st, cd *H, const cd *pdot,
                       const double *pnrm, const double *pnav, double *nrm2, double *nav2, hipStream_t s) {
    hipLaunchKernelGGL(k_mgs_reduce, dim3((unsigned)d.B), dim3(RED_T), 0, s, d, k, st, H, pdot, pnrm, pnav, nrm2,
                       nav2);
    HIP_CHECK(hipGetLastError());
}
void launch_slot_update(const KrylovDims &d, SlotState *st, cd *H, cd *cs, cd *sn, cd *g, cd *y,
                        const double *nrm2, const double *nav2, double tol, int restart, int verify, double tol_true,
                        int park, hipStream_t s) {
    hipLaunchKernelGGL(k_slot_update, dim3((d.B + 63) / 64), dim3(64), 0, s, d, st, H, cs, sn, g, y, nrm2, nav2,
                       tol, restart, verify, tol_true, park);
    HIP_CHECK(hipGetLastError());
}
void launch_unpark(const KrylovDims &d, SlotState *st, hipStream_t s) {
    hipLaunchKernelGGL(k_unpark, dim3((d.B + 63) / 64), dim3(64), 0, s, d, st);
    HIP_CHECK(hipGetLastError());
}
template <typename VT, typename WT>
static void launch_vec_update_t(const KrylovDims &d, const SlotState *st, const WT *W, VT *V, cd *X, const cd *y,
                                hipStream_t s) {
    const int SB = d.B < 256 ? d.B : 256;
    const dim3 grid((unsigned)d.nwg, (unsigned)((d.B + SB - 1) / SB));
    hipLaunchKernelGGL((k_vec_update<VT, WT>), grid, dim3(256), 0, s, d, st, W, V, X, y);
    HIP_CHECK(hipGetLastError());
}
void launch_vec_update(const KrylovDims &d, const SlotState *st, const cd *W, cd *V, cd *X, const cd *y,
                       hipStream_t s) {
    launch_vec_update_t<cd, cd>(d, st, W, V, X, y, s);
}
void launch_vec_update(const KrylovDims &d, const SlotState *st, const cf *W, cf *V, cd *X, const cd *y,
                       hipStream_t s) {
    launch_vec_update_t<cf, cf>(d, st, W, V, X, y, s);
}
void launch_vec_update(const KrylovDims &d, const SlotState *st, const cf *W, ch *V, cd *X, const cd *y,
                       hipStream_t s) {
    launch_vec_update_t<ch, cf>(d, st, W, V, X, y, s);
}

// xs[row B + b] = (complex64) V[in_off[b] + row B + b] for active slots b, the
// rows strided as in the sweeps
__global__ __launch_bounds__(256) void k_stage_basis(KrylovDims d, const ch *__restrict__ V,
                                                     const long long *__restrict__ in_off,
                                                     const int *__restrict__ active, cf *__restrict__ xs) {
    const RowSlotMap m = row_slot_map(d.B);
    if (!m.ok || !active[m.b]) return;
    const ch *v = V + in_off[m.b];
    for (long long row = (long long)blockIdx.x * m.RW + m.rl; row < d.n; row += (long long)gridDim.x * m.RW) {
        const size_t e = (size_t)row * d.B + m.b;
        xs[e] = Prec<cf>::from(v[e]);
    }
}
void launch_stage_basis(const KrylovDims &d, const ch *V, const long long *in_off, const int *active, cf *xs,
                        hipStream_t s) {
    const int SB = d.B < 256 ? d.B : 256;
    const dim3 grid((unsigned)d.nwg, (unsigned)((d.B + SB - 1) / SB));
    hipLaunchKernelGGL(k_stage_basis, grid, dim3(256), 0, s, d, V, in_off, active, xs);
    HIP_CHECK(hipGetLastError());
}
void launch_load_rhs(const KrylovDims &d, const int *col, const cd *src, long long ld, cd *Rhs, cd *X,
                     double *pnrm, int x0zero, hipStream_t s) {
    hipLaunchKernelGGL(k_load_rhs, rs_grid(d), dim3(256), 0, s, d, col, src, ld, Rhs, X, pnrm, x0zero);
    HIP_CHECK(hipGetLastError());
}
void launch_load_finish(const KrylovDims &d, const int *col, SlotState *st, const double *pnrm,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_load_finish, dim3((unsigned)d.B), dim3(RED_T), 0, s, d, col, st, pnrm);
    HIP_CHECK(hipGetLastError());
}
void launch_store_x(const KrylovDims &d, const int *col, const cd *X, cd *dst, long long ld, hipStream_t s) {
    long long tot = d.n * d.B;
    unsigned g = (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_store_x, dim3(g), dim3(256), 0, s, d, col, X, dst, ld);
    HIP_CHECK(hipGetLastError());
}
void launch_bmgs(const KrylovDims &d, cd *W, const cd *V, const cd *Rhs, const SlotState *st, cd *H, cd *pq,
                 double *pnrm, double *pnav, cd *red, double *nrm2, double *nav2, int kmax, hipStream_t s) {
    // passes 0 .. ceil((kmax+1)/BS) for the slot with the highest Arnoldi index kmax
    const int np = (kmax + BS) / BS;
    for (int p = 0; p <= np; ++p) {
        hipLaunchKernelGGL(k_bmgs_pass, rs_grid(d), dim3(256), 0, s, d, p, W, V, Rhs, st, H, pq, pnrm, pnav);
        hipLaunchKernelGGL(k_bmgs_reduce, dim3(BQ + 2, (unsigned)d.B), dim3(RED_T), 0, s, d, p, st, pq, pnrm, pnav,
                           red);
        hipLaunchKernelGGL(k_bmgs_coef, dim3((d.B + 63) / 64), dim3(64), 0, s, d, p, st, red, H, nrm2, nav2);
    }
    HIP_CHECK(hipGetLastError());
}
int bmgs_partials() { return BQ; }

void launch_cgs(const KrylovDims &d, cd *W, const cd *V, const cd *Rhs, const SlotState *st, cd *H, cd *pdot,
                double *pnrm, double *pnav, double *nrm2, double *nav2, int *flag, cd *cc, cd *hprev, int kmax,
                hipStream_t s) {
    const int SB = std::min(d.B, 256);
    const unsigned ys = (unsigned)((d.B + SB - 1) / SB);
    const unsigned zs = (unsigned)(kmax / CGS_KC + 1);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_cgs_dots, dim3((unsigned)d.nwg, ys, zs), dim3(256), 0, s, d, pass, W, V, Rhs, st, flag,
                           pdot, pnav);
        hipLaunchKernelGGL(k_cgs_dots_reduce, dim3((unsigned)d.B), dim3(RED_T), 0, s, d, pass, st, flag, H, pdot,
                           pnav, nav2, nrm2);
        hipLaunchKernelGGL(k_cgs_coef, dim3((unsigned)d.B), dim3(64), 0, s, d, pass, st, H, cc, hprev);
        hipLaunchKernelGGL(k_cgs_update, dim3((unsigned)d.nwg, ys), dim3(256), 0, s, d, pass, W, V, st, flag, cc,
                           pnrm);
        hipLaunchKernelGGL(k_cgs_norm, dim3((unsigned)d.B), dim3(RED_T), 0, s, d, pass, st, pnrm, nav2, flag, nrm2);
    }
    HIP_CHECK(hipGetLastError());
}

void launch_pack_slots(const KrylovDims &d, const cd *X, const int *slot, int cap, cd *out, hipStream_t s) {
    const long long tot = d.n * cap;
    const unsigned g = (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_pack_slots, dim3(g), dim3(256), 0, s, d, X, slot, cap, out);
    HIP_CHECK(hipGetLastError());
}
void launch_unpack_slots(const KrylovDims &d, const cd *in, const int *slot, int cap, cd *W, hipStream_t s) {
    const long long tot = d.n * cap;
    const unsigned g = (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_unpack_slots, dim3(g), dim3(256), 0, s, d, in, slot, cap, W);
    HIP_CHECK(hipGetLastError());
}
void launch_set_inputs(const KrylovDims &d, const SlotState *st, long long *in_off, int *active,
                       int restart_elsewhere, int nblk, int *err, hipStream_t s) {
    hipLaunchKernelGGL(k_set_inputs, dim3((d.B + 63) / 64), dim3(64), 0, s, d, st, in_off, active,
                       restart_elsewhere, nblk, err);
    HIP_CHECK(hipGetLastError());
}
void launch_retire(const KrylovDims &d, const int *col, SlotState *st, hipStream_t s) {
    hipLaunchKernelGGL(k_retire, dim3((d.B + 63) / 64), dim3(64), 0, s, d, col, st);
    HIP_CHECK(hipGetLastError());
}
void launch_tgemm(const cd *mat, const cd *X, long long n, int na, int nx, cd *T, cd *work, size_t work_elems,
                  hipStream_t s) {
    // 64 x 64 tiles; row chunks (multiples of 4 rows) so that about 2048
    // workgroups run, limited by the partial-sum workspace
    const int ta = (na + 63) / 64, tx = (nx + 63) / 64;
    const size_t tot = (size_t)na * nx;
    long long nz = std::max<long long>(1, 2048 / (ta * tx));
    nz = std::min<long long>(nz, std::max<long long>(1, (long long)(work_elems / tot)));
    nz = std::min<long long>(nz, std::max<long long>(1, n / 256));
    long long chunk = (n + nz - 1) / nz;
    chunk = (chunk + 3) / 4 * 4;
    nz = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_tgemm, dim3(ta, tx, (unsigned)nz), dim3(256), 0, s, mat, X, n, na, nx, chunk, work);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_tgemm_reduce, dim3(256), dim3(256), 0, s, work, (int)nz, tot, T);
    HIP_CHECK(hipGetLastError());
}

}  // namespace scadyn
