// Gather + near zone with LDS-staged operands, one wave per (non-empty cube,
// 64-slot chunk): the per-lane steps, host+device so that the host self-test
// runs the same code on the CPU.
//
//  * grid part: the S* records of up to GT tets of the cube (contiguous, 108
//    complex each) are copied to the staging buffer with coalesced loads, then
//    read as wave-uniform broadcasts while each lane streams its slot's 27 box
//    nodes (gather, src/sparse.f90:47-77, and Y2 - k^2 Y,
//    src/gmres_module.f90:148-153);
//  * near part (matmul_Acorr_const, src/sparse_mat.f90:96-123): for each near
//    cube, in chunks of VC column tets, lane (u, v) = (lane / VC, lane % VC)
//    fetches the block of row tet u and column tet v and stores it (complex64,
//    as stored) as a dense 3x3, transposed when the pair is stored as (v, u);
//    every lane then widens the chunk's blocks to the matvec precision and
//    applies them to its slot's x of those tets.
// The rows of a cube's tets list the near tets in the same order (near cube
// offset, then tet id), so entry q of every row of the cube is column q.
//
// Work order.  Item w is (slot chunk w / n_ne, cube ne_cubes[w % n_ne]): one
// chunk sweeps all cubes before the next, and the workgroups are renumbered so
// that each XCD runs one contiguous range of items (xcd_remap).  Neighbouring
// cubes then run at the same time on the same XCD, and the x of a near tet that
// one cube's wave fetches is still in that XCD's L2 when the waves of the
// cubes beside it along x ask for it.
//
// Buffers: grid [Nz][Ny][Nx][4][B], x [3N][B] in the matvec precision (at
// in_off[b]) and y [3N][B] complex128, S4
// [N][27][4], row_start [N+1], ent_blk [row_start[N]], blocks [P][9] complex64.
#pragma once

#include "internal.hpp"

namespace scadyn {
namespace gcore {

#define GC_HD __host__ __device__ inline

template <typename C>
struct GatherArgs {
    const C *grid;
    const C *xin;  // in the matvec precision
    const long long *in_off;
    int B;
    const int *ne_cubes;
    int n_ne;
    const int *cube_start;
    const C *S4;
    int Nx, Ny, ncx, ncy, ncz;
    const int64_t *row_start;
    const uint32_t *ent_blk;
    const cf *blocks;
    double k2;
    const int *slot_active;
    cd *y;
};

// row tets per pass (8 covers all tets of nearly every cube of a cube-split
// mesh, so the near-zone x of a cube is fetched once) and column tets per
// staged near chunk (GT x VC = 64: one block per lane)
constexpr int GATHER_GT = 8, GATHER_VC = 8;

// staging buffer: the S* records of GT tets (precision C) or, later in the
// pass, GT x VC near-zone blocks (complex64, as stored)
template <typename C, int GT, int VC>
struct Shape {
    static constexpr int S4_BYTES = GT * 108 * (int)sizeof(C);
    static constexpr int BLK_BYTES = GT * VC * 9 * (int)sizeof(cf);
    static constexpr int BYTES = S4_BYTES > BLK_BYTES ? S4_BYTES : BLK_BYTES;
};

// one wave's item: cube and slot
struct Item {
    int c, cx, cy, cz, t0, t1;
    int b;       // slot used for loads (B - 1 for idle lanes of a partial chunk)
    bool valid;  // lane holds a real slot
    bool act;    // slot active (inactive slots are not written)
    size_t base; // first box node of the cube
};

template <typename C>
GC_HD Item make_item(const GatherArgs<C> &g, long long w, int lane) {
    const int ci = (int)(w % g.n_ne);
    const int braw = (int)(w / g.n_ne) * 64 + lane;
    Item it;
    it.valid = braw < g.B;
    it.b = it.valid ? braw : g.B - 1;
    it.c = g.ne_cubes[ci];
    it.cx = it.c % g.ncx;
    it.cy = (it.c / g.ncx) % g.ncy;
    it.cz = it.c / (g.ncx * g.ncy);
    it.t0 = g.cube_start[it.c];
    it.t1 = g.cube_start[it.c + 1];
    it.act = g.slot_active ? g.slot_active[it.b] != 0 : true;
    it.base = (size_t)(2 * it.cx) + (size_t)g.Nx * ((size_t)(2 * it.cy) + (size_t)g.Ny * (2 * it.cz));
    return it;
}

// stage the S* records of tets tp .. tp+nt-1 (lane's share)
template <typename C>
GC_HD void stage_s4(const GatherArgs<C> &g, int tp, int nt, int lane, C *sm) {
    for (int e = lane; e < nt * 108; e += 64) sm[e] = g.S4[(size_t)tp * 108 + e];
}

// The steps below take the pass's row-tet count NT as a template argument
// (the kernel switches on it), so that the accumulators are exactly NT x 3
// registers and no loop over them carries a run-time bound.
template <typename C, int NT>
GC_HD void zero_acc(C (*acc)[3]) {
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[u][0] = acc[u][1] = acc[u][2] = Prec<C>::zero();
}

// y_c += S_{c+1}^T g3 - k^2 S^T g_c over the 27 box nodes
template <typename C, int NT>
GC_HD void grid_part(const GatherArgs<C> &g, const Item &it, const C *sm, C (*acc)[3]) {
    typedef typename Prec<C>::real R;
    const R kk = (R)g.k2;
    const int B = g.B;
    // box node j = jx + 3 jy + 9 jz.  In complex64 the 4 channels of the next
    // node are fetched while the current node is applied; complex128 has no
    // registers to spare for that at 3 waves per SIMD.
    constexpr bool PREFETCH = sizeof(C) == 8;
    auto node = [&](int j) {
        const int jx = j % 3, jy = (j / 3) % 3, jz = j / 9;
        return g.grid + (it.base + jx + (size_t)g.Nx * (jy + (size_t)g.Ny * jz)) * 4 * B + it.b;
    };
    C nx[4];
    if (PREFETCH) {
        const C *gp = node(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) nx[c] = gp[(size_t)c * B];
    }
#pragma unroll 1
    for (int j = 0; j < 27; ++j) {
        C g0, g1, g2, g3;
        if (PREFETCH) {
            g0 = nx[0], g1 = nx[1], g2 = nx[2], g3 = nx[3];
            if (j + 1 < 27) {
                const C *gp = node(j + 1);
#pragma unroll
                for (int c = 0; c < 4; ++c) nx[c] = gp[(size_t)c * B];
            }
        } else {
            const C *gp = node(j);
            g0 = gp[0], g1 = gp[(size_t)B], g2 = gp[(size_t)2 * B], g3 = gp[(size_t)3 * B];
        }
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const C *s = sm + u * 108 + j * 4;
            const C s0 = cscale(s[0], -kk);
            cfma(acc[u][0], s0, g0);
            cfma(acc[u][1], s0, g1);
            cfma(acc[u][2], s0, g2);
            cfma(acc[u][0], s[1], g3);
            cfma(acc[u][1], s[2], g3);
            cfma(acc[u][2], s[3], g3);
        }
    }
}

// lane's block of a near chunk: row tet tp + lane / VC, column v0 + lane % VC,
// entry q + lane % VC of the row (rsu = row_start of the lane's row tet).  The
// block stays complex64 (its stored precision) in LDS and is widened when used:
// half the LDS bytes of a complex128 copy for the FP64 matvec, same values.
template <typename C, int GT, int VC>
GC_HD void stage_blocks(const GatherArgs<C> &g, int nt, int nv, int q, int lane, int64_t rsu, cf *sm) {
    static_assert(GT * VC <= 64, "one block per lane");
    const int su = lane / VC, sv = lane % VC;
    if (su >= nt || sv >= nv) return;
    const uint32_t eb = g.ent_blk[rsu + q + sv];
    const cf *a = g.blocks + (size_t)(eb & 0x7fffffffu) * 9;
    cf *dst = sm + (su * VC + sv) * 9;
    if (eb & 0x80000000u) {
        for (int i = 0; i < 3; ++i)
            for (int jj = 0; jj < 3; ++jj) dst[i + 3 * jj] = a[jj + 3 * i];
    } else {
        for (int i = 0; i < 9; ++i) dst[i] = a[i];
    }
}

template <typename C, int GT, int VC>
GC_HD int64_t lane_row_start(const GatherArgs<C> &g, int tp, int nt, int lane) {
    const int su = lane / VC;
    return su < nt ? g.row_start[tp + su] : 0;
}

// apply the staged blocks of column tets v0 .. v0+nv-1
template <typename C, int NT, int VC>
GC_HD void chunk_part(const GatherArgs<C> &g, const Item &it, int v0, int nv, const cf *sm, C (*acc)[3]) {
    const int B = g.B;
    const C *xb = g.xin + g.in_off[it.b] + it.b;
    // x of the next column tet is fetched while the current one is applied
    C n0 = xb[(size_t)(3 * v0) * B], n1 = xb[(size_t)(3 * v0 + 1) * B], n2 = xb[(size_t)(3 * v0 + 2) * B];
#pragma unroll 1
    for (int v = 0; v < nv; ++v) {
        const C x0 = n0, x1 = n1, x2 = n2;
        if (v + 1 < nv) {
            const int col = v0 + v + 1;
            n0 = xb[(size_t)(3 * col) * B];
            n1 = xb[(size_t)(3 * col + 1) * B];
            n2 = xb[(size_t)(3 * col + 2) * B];
        }
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const cf *m = sm + (u * VC + v) * 9;  // column-major 3x3
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                cfma(acc[u][i], Prec<C>::from(m[i]), x0);
                cfma(acc[u][i], Prec<C>::from(m[i + 3]), x1);
                cfma(acc[u][i], Prec<C>::from(m[i + 6]), x2);
            }
        }
    }
}

template <typename C, int NT>
GC_HD void write_out(const GatherArgs<C> &g, const Item &it, int tp, C (*acc)[3]) {
    if (!it.valid || !it.act) return;  // inactive slots are not written
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        cd *yo = g.y + (size_t)(3 * (tp + u)) * g.B + it.b;
#pragma unroll
        for (int r = 0; r < 3; ++r) yo[(size_t)r * g.B] = Prec<cd>::from(acc[u][r]);
    }
}

// the near cubes of (cx, cy, cz) in row order; calls f(first tet, end tet)
template <typename C, typename F>
GC_HD void for_near_cubes(const GatherArgs<C> &g, const Item &it, F &&f) {
    for (int dz = -1; dz <= 1; ++dz) {
        const int zz = it.cz + dz;
        if (zz < 0 || zz >= g.ncz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int yy = it.cy + dy;
            if (yy < 0 || yy >= g.ncy) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = it.cx + dx;
                if (xx < 0 || xx >= g.ncx) continue;
                const int nc = xx + g.ncx * (yy + g.ncy * zz);
                f(g.cube_start[nc], g.cube_start[nc + 1]);
            }
        }
    }
}

// ---- near zone on the matrix cores (complex64 matvec) ----------------------
// For one chunk of nv column tets, the near-zone product of the pass's NT row
// tets is the real matrix product D = Am Bm over 64 slots:
//   D[n'][s]  rows n' = 2 (3u + i) + co: row tet u, component i, co = re / im
//   Bm[k'][s] = ci-part (re / im) of x_{3v+j}[s], k' = 2 (3v + j) + ci
//   Am[n'][k'] = Re A_uv(i, j) if co == ci, -Im A_uv(i, j) if (co, ci) = (re, im),
//                +Im A_uv(i, j) if (co, ci) = (im, re)
// (y_re += a x_re - b x_im, y_im += b x_re + a x_im).  One
// v_mfma_f32_16x16x4f32 per (k-step of 4, row tile of 16, slot tile of 16):
// lane l holds Am[16 rt + (l & 15)][4 s + (l >> 4)] and
// Bm[4 s + (l >> 4)][16 st + (l & 15)], and output element e of row
// mf_row(l, e), column l & 15 of its tile.  After the chunks, the tiles go
// through an LDS exchange buffer [16 RT][64] to the lanes that own the slots.
GC_HD constexpr int mf_rows(int NT) { return (6 * NT + 15) / 16; }  // row tiles of the pass
constexpr int MF_XBYTES = 16 * 3 * 64 * (int)sizeof(float);           // exchange buffer, NT <= 8
GC_HD int mf_row(int lane, int e) { return 4 * (lane >> 4) + e; }

// Am[np][kp] from the staged blocks (zero outside the pass's rows and the chunk)
GC_HD float mf_a(const cf *smb, int NT, int nv, int VC, int np, int kp) {
    const int u = np / 6, v = kp / 6;
    if (u >= NT || v >= nv) return 0.0f;
    const int i = (np % 6) >> 1, co = np & 1, j = (kp % 6) >> 1, ci = kp & 1;
    const cf a = smb[(u * VC + v) * 9 + i + 3 * j];
    return co == ci ? a.x : (co ? a.y : -a.y);
}

// Bm[kp][slot] (zero beyond the chunk or for a slot past the batch)
GC_HD float mf_b(const GatherArgs<cf> &g, int v0, int nv, int kp, int slot, long long off, bool ok) {
    const int v = kp / 6;
    if (!ok || v >= nv) return 0.0f;
    const int j = (kp % 6) >> 1;
    const cf x = g.xin[off + (size_t)(3 * (v0 + v) + j) * g.B + slot];
    return (kp & 1) ? x.y : x.x;
}

// exchange-buffer index of output element e of lane l in tile (rt, st)
GC_HD int mf_xidx(int lane, int e, int rt, int st) { return (16 * rt + mf_row(lane, e)) * 64 + 16 * st + (lane & 15); }

// the lane's slot: its near-zone outputs from the exchange buffer into acc
template <int NT>
GC_HD void mf_combine(const float *xb, int lane, cf (*acc)[3]) {
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            acc[u][i].x += xb[(2 * (3 * u + i)) * 64 + lane];
            acc[u][i].y += xb[(2 * (3 * u + i) + 1) * 64 + lane];
        }
}

}  // namespace gcore
}  // namespace scadyn
