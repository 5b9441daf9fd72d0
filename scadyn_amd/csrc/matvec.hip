// Batched pFFT matvec y = A x (compute_Ax2, src/gmres_module.f90:57-176).
//
// Multi-vectors are RHS-innermost: element (row, slot b) of a batch of B
// vectors sits at row*B + b, rows being 3*t + c in the internal (cube-sorted)
// tet order.  Grid buffers are [z][y][x][4][B]: channel 0..2 carry S^T x_c
// (the three "Y" convolutions) and channel 3 carries Sx^T x_x + Sy^T x_y +
// Sz^T x_z; by linearity this one convolution replaces the reference's three
// divergence convolutions (:122-144), whose gathered sum it equals.
#include <type_traits>

#include "matvec.hpp"
#include "gather_core.hpp"

namespace scadyn {

// scatter (sparse_T_matmul, src/sparse.f90:13-43, + vec2arr2), node-centric:
// node p sums the stencil weights of the tets of the <= 8 cubes whose box holds p.
template <typename C>
__global__ __launch_bounds__(256) void k_scatter(const C *__restrict__ xin,
                                                 const long long *__restrict__ in_off, int B,
                                                 const int *__restrict__ cube_start,
                                                 const C *__restrict__ S4, int Nx, int Ny, int Nz,
                                                 int ncx, int ncy, int ncz, C *__restrict__ grid) {
    const long long tot = (long long)Nx * Ny * Nz * B;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e % B);
        const long long p = e / B;
        const int px = (int)(p % Nx), py = (int)((p / Nx) % Ny), pz = (int)(p / ((long long)Nx * Ny));
        const C *xb = xin + in_off[b] + b;
        C a0 = Prec<C>::zero(), a1 = a0, a2 = a0, a3 = a0;
        // cubes c with 2c <= p <= 2c+2 per axis
        const int cx0 = max(0, (px - 1) / 2), cx1 = min(ncx - 1, px / 2);
        const int cy0 = max(0, (py - 1) / 2), cy1 = min(ncy - 1, py / 2);
        const int cz0 = max(0, (pz - 1) / 2), cz1 = min(ncz - 1, pz / 2);
        for (int cz = cz0; cz <= cz1; ++cz)
            for (int cy = cy0; cy <= cy1; ++cy)
                for (int cx = cx0; cx <= cx1; ++cx) {
                    const int c = cx + ncx * (cy + ncy * cz);
                    const int j = (px - 2 * cx) + 3 * (py - 2 * cy) + 9 * (pz - 2 * cz);
                    for (int t = cube_start[c]; t < cube_start[c + 1]; ++t) {
                        const C *s = S4 + ((size_t)t * 27 + j) * 4;
                        const C x0 = Prec<C>::from(xb[(size_t)(3 * t) * B]),
                                x1 = Prec<C>::from(xb[(size_t)(3 * t + 1) * B]),
                                x2 = Prec<C>::from(xb[(size_t)(3 * t + 2) * B]);
                        const C s0 = s[0];
                        cfma(a0, s0, x0);
                        cfma(a1, s0, x1);
                        cfma(a2, s0, x2);
                        cfma(a3, s[1], x0);
                        cfma(a3, s[2], x1);
                        cfma(a3, s[3], x2);
                    }
                }
        C *gp = grid + (size_t)p * 4 * B + b;
        gp[0] = a0;
        gp[(size_t)B] = a1;
        gp[(size_t)2 * B] = a2;
        gp[(size_t)3 * B] = a3;
    }
}

// gather (sparse_eps_matmul, src/sparse.f90:47-77) + Y_aim = Y2 - k^2 Y (:148-153)
// + symmetric near-zone blocks (matmul_Acorr_const, src/sparse_mat.f90:96-123).
template <typename C>
__global__ __launch_bounds__(256) void k_gather_near(
    const C *__restrict__ grid, const C *__restrict__ xin, const long long *__restrict__ in_off, int B,
    int N, const int *__restrict__ tet_cube, const C *__restrict__ S4, int Nx, int Ny, int ncx, int ncy,
    const int64_t *__restrict__ row_start, const int *__restrict__ ent_col,
    const uint32_t *__restrict__ ent_blk, const cf *__restrict__ blocks, double k2,
    const int *__restrict__ slot_active, cd *__restrict__ y) {
    const long long tot = (long long)N * B;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e % B);
        const int t = (int)(e / B);
        cd *yo = y + (size_t)(3 * t) * B + b;
        if (slot_active && !slot_active[b]) continue;  // inactive slots are not written
        const int c = tet_cube[t];
        const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        const size_t base = (size_t)(2 * cx) + (size_t)Nx * ((size_t)(2 * cy) + (size_t)Ny * (2 * cz));
        const C *s = S4 + (size_t)t * 108;
        C Y0 = Prec<C>::zero(), Y1 = Y0, Y2 = Y0, Z0 = Y0, Z1 = Y0, Z2 = Y0;
        for (int jz = 0; jz < 3; ++jz)
            for (int jy = 0; jy < 3; ++jy)
                for (int jx = 0; jx < 3; ++jx) {
                    const int j = jx + 3 * jy + 9 * jz;
                    const size_t node = base + jx + (size_t)Nx * (jy + (size_t)Ny * jz);
                    const C *g = grid + node * 4 * B + b;
                    const C g0 = g[0], g1 = g[(size_t)B], g2 = g[(size_t)2 * B], g3 = g[(size_t)3 * B];
                    const C *sj = s + j * 4;
                    cfma(Y0, sj[0], g0);
                    cfma(Y1, sj[0], g1);
                    cfma(Y2, sj[0], g2);
                    cfma(Z0, sj[1], g3);
                    cfma(Z1, sj[2], g3);
                    cfma(Z2, sj[3], g3);
                }
        // near zone
        const C *xb = xin + in_off[b] + b;
        C n0 = Prec<C>::zero(), n1 = n0, n2 = n0;
        for (int64_t q = row_start[t]; q < row_start[t + 1]; ++q) {
            const int col = ent_col[q];
            const uint32_t eb = ent_blk[q];
            const cf *a = blocks + (size_t)(eb & 0x7fffffffu) * 9;
            const C x0 = Prec<C>::from(xb[(size_t)(3 * col) * B]),
                    x1 = Prec<C>::from(xb[(size_t)(3 * col + 1) * B]),
                    x2 = Prec<C>::from(xb[(size_t)(3 * col + 2) * B]);
            C m[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) m[i] = Prec<C>::from(a[i]);
            if (eb & 0x80000000u) {
                // transpose: y_i += sum_j A(j,i) x_j ; A(i,j) at i + 3j
                cfma(n0, m[0], x0); cfma(n0, m[1], x1); cfma(n0, m[2], x2);
                cfma(n1, m[3], x0); cfma(n1, m[4], x1); cfma(n1, m[5], x2);
                cfma(n2, m[6], x0); cfma(n2, m[7], x1); cfma(n2, m[8], x2);
            } else {
                cfma(n0, m[0], x0); cfma(n0, m[3], x1); cfma(n0, m[6], x2);
                cfma(n1, m[1], x0); cfma(n1, m[4], x1); cfma(n1, m[7], x2);
                cfma(n2, m[2], x0); cfma(n2, m[5], x1); cfma(n2, m[8], x2);
            }
        }
        typedef typename Prec<C>::real R;
        const R kk = (R)k2;
        yo[0] = Prec<cd>::from(cadd(n0, csub(Z0, cscale(Y0, kk))));
        yo[(size_t)B] = Prec<cd>::from(cadd(n1, csub(Z1, cscale(Y1, kk))));
        yo[(size_t)2 * B] = Prec<cd>::from(cadd(n2, csub(Z2, cscale(Y2, kk))));
    }
}

// ---- wave-uniform variants (slots in chunks of 64) ---------------------------
// One wave64 owns one grid node (scatter) or one tet (gather) for 64 slots:
// the node/tet index is wave-uniform (readfirstlane), so every per-node or
// per-tet operand (cube ranges, S*, near-zone entries and blocks) is a scalar
// load and only the slot-indexed x, grid and y accesses are vector loads,
// each 64 consecutive complex numbers.
template <typename C>
__global__ __launch_bounds__(256) void k_scatter_w64(const C *__restrict__ xin,
                                                     const long long *__restrict__ in_off, int B,
                                                     const int *__restrict__ cube_start,
                                                     const C *__restrict__ S4, int Nx, int Ny, int Nz,
                                                     int ncx, int ncy, int ncz, C *__restrict__ grid) {
    const int lane = threadIdx.x & 63;
    const int nch = (B + 63) >> 6;  // last chunk may be partial: lanes >= B load slot B-1, store nothing
    const long long nw = (long long)Nx * Ny * Nz * nch;
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (long long w = (long long)blockIdx.x * 4 + wib; w < nw; w += (long long)gridDim.x * 4) {
        const long long p = w / nch;
        const int braw = (int)(w - p * nch) * 64 + lane;
        const bool valid = braw < B;
        const int b = valid ? braw : B - 1;
        const int px = (int)(p % Nx), py = (int)((p / Nx) % Ny), pz = (int)(p / ((long long)Nx * Ny));
        const C *xb = xin + in_off[b] + b;
        C a0 = Prec<C>::zero(), a1 = a0, a2 = a0, a3 = a0;
        const int cx0 = max(0, (px - 1) / 2), cx1 = min(ncx - 1, px / 2);
        const int cy0 = max(0, (py - 1) / 2), cy1 = min(ncy - 1, py / 2);
        const int cz0 = max(0, (pz - 1) / 2), cz1 = min(ncz - 1, pz / 2);
        for (int cz = cz0; cz <= cz1; ++cz)
            for (int cy = cy0; cy <= cy1; ++cy)
                for (int cx = cx0; cx <= cx1; ++cx) {
                    const int c = cx + ncx * (cy + ncy * cz);
                    const int j = (px - 2 * cx) + 3 * (py - 2 * cy) + 9 * (pz - 2 * cz);
                    const int t1 = cube_start[c + 1];
                    for (int t = cube_start[c]; t < t1; ++t) {
                        const C *s = S4 + ((size_t)t * 27 + j) * 4;
                        const C x0 = Prec<C>::from(xb[(size_t)(3 * t) * B]),
                                x1 = Prec<C>::from(xb[(size_t)(3 * t + 1) * B]),
                                x2 = Prec<C>::from(xb[(size_t)(3 * t + 2) * B]);
                        const C s0 = s[0];
                        cfma(a0, s0, x0);
                        cfma(a1, s0, x1);
                        cfma(a2, s0, x2);
                        cfma(a3, s[1], x0);
                        cfma(a3, s[2], x1);
                        cfma(a3, s[3], x2);
                    }
                }
        if (valid) {
            C *gp = grid + (size_t)p * 4 * B + b;
            gp[0] = a0;
            gp[(size_t)B] = a1;
            gp[(size_t)2 * B] = a2;
            gp[(size_t)3 * B] = a3;
        }
    }
}

template <typename C>
__global__ __launch_bounds__(256) void k_gather_near_w64(
    const C *__restrict__ grid, const C *__restrict__ xin, const long long *__restrict__ in_off, int B,
    int N, const int *__restrict__ tet_cube, const C *__restrict__ S4, int Nx, int Ny, int ncx, int ncy,
    const int64_t *__restrict__ row_start, const int *__restrict__ ent_col,
    const uint32_t *__restrict__ ent_blk, const cf *__restrict__ blocks, double k2,
    const int *__restrict__ slot_active, cd *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int nch = (B + 63) >> 6;  // last chunk may be partial: lanes >= B load slot B-1, store nothing
    const long long nw = (long long)N * nch;
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (long long w = (long long)blockIdx.x * 4 + wib; w < nw; w += (long long)gridDim.x * 4) {
        const int t = (int)(w / nch);
        const int braw = (int)(w - (long long)t * nch) * 64 + lane;
        const bool valid = braw < B;
        const int b = valid ? braw : B - 1;
        if (!valid) continue;
        cd *yo = y + (size_t)(3 * t) * B + b;
        if (slot_active && !slot_active[b]) continue;  // inactive slots are not written
        const int c = tet_cube[t];
        const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        const size_t base = (size_t)(2 * cx) + (size_t)Nx * ((size_t)(2 * cy) + (size_t)Ny * (2 * cz));
        const C *s = S4 + (size_t)t * 108;
        C Y0 = Prec<C>::zero(), Y1 = Y0, Y2 = Y0, Z0 = Y0, Z1 = Y0, Z2 = Y0;
        for (int jz = 0; jz < 3; ++jz)
            for (int jy = 0; jy < 3; ++jy)
#pragma unroll
                for (int jx = 0; jx < 3; ++jx) {
                    const int j = jx + 3 * jy + 9 * jz;
                    const size_t node = base + jx + (size_t)Nx * (jy + (size_t)Ny * jz);
                    const C *g = grid + node * 4 * B + b;
                    const C g0 = g[0], g1 = g[(size_t)B], g2 = g[(size_t)2 * B], g3 = g[(size_t)3 * B];
                    const C *sj = s + j * 4;
                    cfma(Y0, sj[0], g0);
                    cfma(Y1, sj[0], g1);
                    cfma(Y2, sj[0], g2);
                    cfma(Z0, sj[1], g3);
                    cfma(Z1, sj[2], g3);
                    cfma(Z2, sj[3], g3);
                }
        const C *xb = xin + in_off[b] + b;
        C n0 = Prec<C>::zero(), n1 = n0, n2 = n0;
        const int64_t q1 = row_start[t + 1];
        for (int64_t q = row_start[t]; q < q1; ++q) {
            const int col = ent_col[q];
            const uint32_t eb = ent_blk[q];
            const cf *a = blocks + (size_t)(eb & 0x7fffffffu) * 9;
            const C x0 = Prec<C>::from(xb[(size_t)(3 * col) * B]),
                    x1 = Prec<C>::from(xb[(size_t)(3 * col + 1) * B]),
                    x2 = Prec<C>::from(xb[(size_t)(3 * col + 2) * B]);
            C m[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) m[i] = Prec<C>::from(a[i]);
            if (eb & 0x80000000u) {
                cfma(n0, m[0], x0); cfma(n0, m[1], x1); cfma(n0, m[2], x2);
                cfma(n1, m[3], x0); cfma(n1, m[4], x1); cfma(n1, m[5], x2);
                cfma(n2, m[6], x0); cfma(n2, m[7], x1); cfma(n2, m[8], x2);
            } else {
                cfma(n0, m[0], x0); cfma(n0, m[3], x1); cfma(n0, m[6], x2);
                cfma(n1, m[1], x0); cfma(n1, m[4], x1); cfma(n1, m[7], x2);
                cfma(n2, m[2], x0); cfma(n2, m[5], x1); cfma(n2, m[8], x2);
            }
        }
        typedef typename Prec<C>::real R;
        const R kk = (R)k2;
        yo[0] = Prec<cd>::from(cadd(n0, csub(Z0, cscale(Y0, kk))));
        yo[(size_t)B] = Prec<cd>::from(cadd(n1, csub(Z1, cscale(Y1, kk))));
        yo[(size_t)2 * B] = Prec<cd>::from(cadd(n2, csub(Z2, cscale(Y2, kk))));
    }
}

// ---- on-chip reuse variants (slots in chunks of 64) --------------------------
// Scatter, one wave per 2x2x2 block of grid nodes (2a..2a+1, 2b..2b+1, 2c..2c+1),
// ceil(N/2) blocks per axis:
// every node of the block lies in the boxes of the same <= 8 cubes (a-1..a per
// axis), so the wave loads each of their tets' x once for all 8 nodes, instead
// of once per node.  Per node the sum runs over cubes and tets in the same
// order as k_scatter.
template <typename C>
__global__ __launch_bounds__(256) void k_scatter_blk_w64(const C *__restrict__ xin,
                                                         const long long *__restrict__ in_off, int B,
                                                         const int *__restrict__ cube_start,
                                                         const C *__restrict__ S4, int Nx, int Ny, int Nz,
                                                         int ncx, int ncy, int ncz, C *__restrict__ grid) {
    const int lane = threadIdx.x & 63;
    const int nch = (B + 63) >> 6;  // last chunk may be partial: lanes >= B load slot B-1, store nothing
    // ceil(N/2) blocks per axis cover every node (nodes beyond the cubes' boxes get zeros)
    const int nbx = (Nx + 1) >> 1, nby = (Ny + 1) >> 1, nbz = (Nz + 1) >> 1;
    const long long nw = (long long)nbx * nby * nbz * nch;
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // items (node block, slot chunk) with the chunk fastest; a chunk-major order
    // with XCD-contiguous workgroup ranges (as in the gather) measured slower
    for (long long w = (long long)blockIdx.x * 4 + wib; w < nw; w += (long long)gridDim.x * 4) {
        const long long blk = w / nch;
        const int braw = (int)(w - blk * nch) * 64 + lane;
        const bool valid = braw < B;
        const int b = valid ? braw : B - 1;
        const int ba = (int)(blk % nbx), bb = (int)((blk / nbx) % nby), bc = (int)(blk / ((long long)nbx * nby));
        const C *xb = xin + in_off[b] + b;
        C acc[2][2][2][4];
#pragma unroll
        for (int iz = 0; iz < 2; ++iz)
#pragma unroll
            for (int iy = 0; iy < 2; ++iy)
#pragma unroll
                for (int ix = 0; ix < 2; ++ix)
#pragma unroll
                    for (int ch = 0; ch < 4; ++ch) acc[iz][iy][ix][ch] = Prec<C>::zero();
        // cube (ba-1+dx, bb-1+dy, bc-1+dz); node 2ba+ix is in its box when dx = 1
        // (local index ix) or dx = 0 and ix = 0 (local index 2)
#pragma unroll
        for (int dz = 0; dz < 2; ++dz) {
            const int cz = bc - 1 + dz;
            if (cz < 0 || cz >= ncz) continue;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                const int cy = bb - 1 + dy;
                if (cy < 0 || cy >= ncy) continue;
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                    const int cx = ba - 1 + dx;
                    if (cx < 0 || cx >= ncx) continue;
                    const int c = cx + ncx * (cy + ncy * cz);
                    const int t1 = cube_start[c + 1];
                    for (int t = cube_start[c]; t < t1; ++t) {
                        const C x0 = Prec<C>::from(xb[(size_t)(3 * t) * B]),
                                x1 = Prec<C>::from(xb[(size_t)(3 * t + 1) * B]),
                                x2 = Prec<C>::from(xb[(size_t)(3 * t + 2) * B]);
                        const C *st = S4 + (size_t)t * 108;
#pragma unroll
                        for (int iz = 0; iz < 2; ++iz) {
                            if (dz == 0 && iz == 1) continue;
                            const int jz = dz ? iz : 2;
#pragma unroll
                            for (int iy = 0; iy < 2; ++iy) {
                                if (dy == 0 && iy == 1) continue;
                                const int jy = dy ? iy : 2;
#pragma unroll
                                for (int ix = 0; ix < 2; ++ix) {
                                    if (dx == 0 && ix == 1) continue;
                                    const int jx = dx ? ix : 2;
                                    const C *s = st + (jx + 3 * jy + 9 * jz) * 4;
                                    C *a = acc[iz][iy][ix];
                                    const C s0 = s[0];
                                    cfma(a[0], s0, x0);
                                    cfma(a[1], s0, x1);
                                    cfma(a[2], s0, x2);
                                    cfma(a[3], s[1], x0);
                                    cfma(a[3], s[2], x1);
                                    cfma(a[3], s[3], x2);
                                }
                            }
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int iz = 0; iz < 2; ++iz)
#pragma unroll
            for (int iy = 0; iy < 2; ++iy)
#pragma unroll
                for (int ix = 0; ix < 2; ++ix) {
                    const int px = 2 * ba + ix, py = 2 * bb + iy, pz = 2 * bc + iz;
                    if (!valid || px >= Nx || py >= Ny || pz >= Nz) continue;
                    C *gp = grid + ((size_t)px + (size_t)Nx * ((size_t)py + (size_t)Ny * pz)) * 4 * B + b;
#pragma unroll
                    for (int ch = 0; ch < 4; ++ch) gp[(size_t)ch * B] = acc[iz][iy][ix][ch];
                }
    }
}

// Gather + near zone with the per-tet operands staged in LDS, one 64-thread
// workgroup per (non-empty cube, 64-slot chunk); the steps are in
// gather_core.hpp (shared with the host self-test).  Staging keeps many
// operand fetches in flight (wave-uniform operands read through scalar loads
// would wait on each other).
template <typename C, int NT, int GT, int VC>
__device__ __forceinline__ void gather_pass(const gcore::GatherArgs<C> &g, const gcore::Item &it, int tp, int lane,
                                            unsigned char *smraw) {
    C *sm = (C *)smraw;         // S* records of the pass's tets
    cf *smb = (cf *)smraw;      // then the near-zone blocks of each chunk
    using namespace gcore;
    C acc[NT][3];
    zero_acc<C, NT>(acc);
    __syncthreads();  // the previous pass's reads of sm are done
    stage_s4(g, tp, NT, lane, sm);
    __syncthreads();
    grid_part<C, NT>(g, it, sm, acc);
    const int64_t rsu = lane_row_start<C, GT, VC>(g, tp, NT, lane);
    int q = 0;
    for_near_cubes(g, it, [&](int c0, int c1) {
        for (int v0 = c0; v0 < c1; v0 += VC) {
            const int nv = min(VC, c1 - v0);
            __syncthreads();
            stage_blocks<C, GT, VC>(g, NT, nv, q, lane, rsu, smb);
            __syncthreads();
            chunk_part<C, NT, VC>(g, it, v0, nv, smb, acc);
            q += nv;
        }
    });
    write_out<C, NT>(g, it, tp, acc);
}

template <typename C, int GT, int VC>
__global__ __launch_bounds__(64, 3) void k_gather_lds(gcore::GatherArgs<C> g) {
    using namespace gcore;
    static_assert(GT == 8, "the pass switch below covers 1..8 row tets");
    __shared__ __align__(16) unsigned char sm[Shape<C, GT, VC>::BYTES];
    const int lane = threadIdx.x;
    const long long nw = (long long)g.n_ne * ((g.B + 63) >> 6);
    for (long long b = blockIdx.x; b < nw; b += gridDim.x) {
        // one item per workgroup when the grid covers all items (the launch
        // below); a grid-stride remainder keeps plain order
        const long long w = nw == gridDim.x ? xcd_remap(b, nw) : b;
        const Item it = make_item(g, w, lane);
        for (int tp = it.t0; tp < it.t1; tp += GT) {
            switch (min(GT, it.t1 - tp)) {  // wave-uniform
                case 1: gather_pass<C, 1, GT, VC>(g, it, tp, lane, sm); break;
                case 2: gather_pass<C, 2, GT, VC>(g, it, tp, lane, sm); break;
                case 3: gather_pass<C, 3, GT, VC>(g, it, tp, lane, sm); break;
                case 4: gather_pass<C, 4, GT, VC>(g, it, tp, lane, sm); break;
                case 5: gather_pass<C, 5, GT, VC>(g, it, tp, lane, sm); break;
                case 6: gather_pass<C, 6, GT, VC>(g, it, tp, lane, sm); break;
                case 7: gather_pass<C, 7, GT, VC>(g, it, tp, lane, sm); break;
                default: gather_pass<C, 8, GT, VC>(g, it, tp, lane, sm); break;
            }
        }
    }
}

// The complex64 gather with the near-zone product on the matrix cores
// (v_mfma_f32_16x16x4f32, operand algebra in gather_core.hpp): the grid part
// as k_gather_lds, then per near chunk the blocks staged in LDS as the A
// operand and the column tets' x as the B operand, the output tiles exchanged
// through LDS to the lanes that own the slots (opts.gather_kernel = 1).  FP32
// matrix instructions have twice the FP32 vector rate and the stored
// complex64 blocks need no conversion, but each k-step's operand set-up
// (index decode, LDS read, x load) costs about as many vector instructions as
// the FMAs it replaces: measured 23.0 vs 10.8 ms per launch at 824k tets,
// 59 slots, so the vector-FMA kernel stays the default.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int NT, int GT, int VC>
__device__ __forceinline__ void gather_pass_mf(const gcore::GatherArgs<cf> &g, const gcore::Item &it, int tp,
                                               int lane, unsigned char *smraw) {
    using namespace gcore;
    cf *sm = (cf *)smraw;
    cf *smb = (cf *)smraw;
    float *xb = (float *)smraw;
    // near part first (its output tiles in registers), then the tiles through
    // LDS into the lane's accumulators, then the grid part: the two phases'
    // accumulators are never live together
    constexpr int RT = mf_rows(NT);
    v4f D[RT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int st = 0; st < 4; ++st) D[rt][st] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
    const int lo = lane & 15, kq = lane >> 4;
    const int s0 = (it.b & ~63) + lo;  // slot of tile 0 for this lane's B operand; tile st adds 16 st
    long long off[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) off[st] = s0 + 16 * st < g.B ? g.in_off[s0 + 16 * st] : 0;
    const int64_t rsu = lane_row_start<cf, GT, VC>(g, tp, NT, lane);
    int q = 0;
    for_near_cubes(g, it, [&](int c0, int c1) {
        for (int v0 = c0; v0 < c1; v0 += VC) {
            const int nv = min(VC, c1 - v0);
            __syncthreads();  // the previous chunk's (or pass's) reads of the buffer are done
            stage_blocks<cf, GT, VC>(g, NT, nv, q, lane, rsu, smb);
            __syncthreads();
            const int S = (6 * nv + 3) / 4;
            // operands fetched one k-step ahead, so that the loads of step s + 1
            // are in flight while the matrix instructions of step s run
            float bn[4], an[RT];
#pragma unroll
            for (int st = 0; st < 4; ++st) bn[st] = mf_b(g, v0, nv, kq, s0 + 16 * st, off[st], s0 + 16 * st < g.B);
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) an[rt] = mf_a(smb, NT, nv, VC, 16 * rt + lo, kq);
            for (int s = 0; s < S; ++s) {
                float bo[4], ao[RT];
#pragma unroll
                for (int st = 0; st < 4; ++st) bo[st] = bn[st];
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) ao[rt] = an[rt];
                if (s + 1 < S) {
                    const int kp = 4 * (s + 1) + kq;
#pragma unroll
                    for (int st = 0; st < 4; ++st)
                        bn[st] = mf_b(g, v0, nv, kp, s0 + 16 * st, off[st], s0 + 16 * st < g.B);
#pragma unroll
                    for (int rt = 0; rt < RT; ++rt) an[rt] = mf_a(smb, NT, nv, VC, 16 * rt + lo, kp);
                }
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                    for (int st = 0; st < 4; ++st)
                        D[rt][st] = __builtin_amdgcn_mfma_f32_16x16x4f32(ao[rt], bo[st], D[rt][st], 0, 0, 0);
            }
            q += nv;
        }
    });
    __syncthreads();  // the last chunk's block reads are done
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
            for (int e = 0; e < 4; ++e) xb[mf_xidx(lane, e, rt, st)] = D[rt][st][e];
    __syncthreads();
    cf acc[NT][3];
    zero_acc<cf, NT>(acc);
    mf_combine<NT>(xb, lane, acc);
    __syncthreads();  // the exchange is read before the S* records overwrite it
    stage_s4(g, tp, NT, lane, sm);
    __syncthreads();
    grid_part<cf, NT>(g, it, sm, acc);
    write_out<cf, NT>(g, it, tp, acc);
}

template <int GT, int VC>
__global__ __launch_bounds__(64, 3) void k_gather_mf(gcore::GatherArgs<cf> g) {
    using namespace gcore;
    static_assert(GT == 8, "the pass switch below covers 1..8 row tets");
    constexpr int SB = Shape<cf, GT, VC>::BYTES > MF_XBYTES ? Shape<cf, GT, VC>::BYTES : MF_XBYTES;
    __shared__ __align__(16) unsigned char sm[SB];
    const int lane = threadIdx.x;
    const long long nw = (long long)g.n_ne * ((g.B + 63) >> 6);
    for (long long b = blockIdx.x; b < nw; b += gridDim.x) {
        const long long w = nw == gridDim.x ? xcd_remap(b, nw) : b;
        const Item it = make_item(g, w, lane);
        for (int tp = it.t0; tp < it.t1; tp += GT) {
            switch (min(GT, it.t1 - tp)) {  // wave-uniform
                case 1: gather_pass_mf<1, GT, VC>(g, it, tp, lane, sm); break;
                case 2: gather_pass_mf<2, GT, VC>(g, it, tp, lane, sm); break;
                case 3: gather_pass_mf<3, GT, VC>(g, it, tp, lane, sm); break;
                case 4: gather_pass_mf<4, GT, VC>(g, it, tp, lane, sm); break;
                case 5: gather_pass_mf<5, GT, VC>(g, it, tp, lane, sm); break;
                case 6: gather_pass_mf<6, GT, VC>(g, it, tp, lane, sm); break;
                case 7: gather_pass_mf<7, GT, VC>(g, it, tp, lane, sm); break;
                default: gather_pass_mf<8, GT, VC>(g, it, tp, lane, sm); break;
            }
        }
    }
}

// slot-chunk (wave64-per-item) kernels when the last 64-slot chunk is at least
// half full: even then they beat the slot-per-lane kernels, which share no
// operand loads between the tets of a cube or the nodes of a block
static bool wave_slots(int B) { return B % 64 == 0 || B % 64 >= 32; }

static unsigned grid_for(long long tot, int bs) {
    long long g = (tot + bs - 1) / bs;
    return (unsigned)std::min<long long>(std::max<long long>(g, 1), 1LL << 20);
}

template <typename C>
void launch_scatter(const MatvecArgs<C> &a, hipStream_t s) {
    if (wave_slots(a.B)) {
        const long long nw = (long long)((a.Nx + 1) / 2) * ((a.Ny + 1) / 2) * ((a.Nz + 1) / 2) * ((a.B + 63) / 64);
        hipLaunchKernelGGL((k_scatter_blk_w64<C>), dim3(grid_for(nw, 4)), dim3(256), 0, s, a.xin, a.in_off, a.B,
                           a.cube_start, a.S4, a.Nx, a.Ny, a.Nz, a.ncx, a.ncy, a.ncz, a.grid_in);
        HIP_CHECK(hipGetLastError());
        return;
    }
    long long tot = (long long)a.Nx * a.Ny * a.Nz * a.B;
    hipLaunchKernelGGL((k_scatter<C>), dim3(grid_for(tot, 256)), dim3(256), 0, s, a.xin, a.in_off, a.B,
                       a.cube_start, a.S4, a.Nx, a.Ny, a.Nz, a.ncx, a.ncy, a.ncz, a.grid_in);
    HIP_CHECK(hipGetLastError());
}

template <typename C>
void launch_gather_near(const MatvecArgs<C> &a, hipStream_t s) {
    if (wave_slots(a.B) && a.ne_cubes) {
        const long long nw = (long long)a.n_ne * ((a.B + 63) / 64);
        const gcore::GatherArgs<C> g = {a.grid_out, a.xin,     a.in_off,    a.B,      a.ne_cubes,
                                        a.n_ne,     a.cube_start, a.S4,      a.Nx,     a.Ny,
                                        a.ncx,      a.ncy,       a.ncz,       a.row_start, a.ent_blk,
                                        a.blocks,   a.k2,        a.slot_active, a.y};
        if constexpr (std::is_same<C, cf>::value) {
            if (a.gather_kernel == 1) {
                hipLaunchKernelGGL((k_gather_mf<gcore::GATHER_GT, gcore::GATHER_VC>), dim3(grid_for(nw, 1)), dim3(64),
                                   0, s, g);
                HIP_CHECK(hipGetLastError());
                return;
            }
        }
        hipLaunchKernelGGL((k_gather_lds<C, gcore::GATHER_GT, gcore::GATHER_VC>), dim3(grid_for(nw, 1)), dim3(64), 0, s,
                           g);
        HIP_CHECK(hipGetLastError());
        return;
    }
    if (wave_slots(a.B)) {
        const long long nw = (long long)a.N * ((a.B + 63) / 64);
        hipLaunchKernelGGL((k_gather_near_w64<C>), dim3(grid_for(nw, 4)), dim3(256), 0, s, a.grid_out, a.xin,
                           a.in_off, a.B, a.N, a.tet_cube, a.S4, a.Nx, a.Ny, a.ncx, a.ncy, a.row_start, a.ent_col,
                           a.ent_blk, a.blocks, a.k2, a.slot_active, a.y);
        HIP_CHECK(hipGetLastError());
        return;
    }
    long long tot = (long long)a.N * a.B;
    hipLaunchKernelGGL((k_gather_near<C>), dim3(grid_for(tot, 256)), dim3(256), 0, s, a.grid_out, a.xin,
                       a.in_off, a.B, a.N, a.tet_cube, a.S4, a.Nx, a.Ny, a.ncx, a.ncy, a.row_start, a.ent_col,
                       a.ent_blk, a.blocks, a.k2, a.slot_active, a.y);
    HIP_CHECK(hipGetLastError());
}

template void launch_scatter<cd>(const MatvecArgs<cd> &, hipStream_t);
template void launch_scatter<cf>(const MatvecArgs<cf> &, hipStream_t);
template void launch_gather_near<cd>(const MatvecArgs<cd> &, hipStream_t);
template void launch_gather_near<cf>(const MatvecArgs<cf> &, hipStream_t);

}  // namespace scadyn
