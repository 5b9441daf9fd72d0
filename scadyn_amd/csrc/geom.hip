// Host-side setup: grid, cube boxes, tet -> cube map, near-zone pair lists,
// the 98x27 pseudo-inverse, and FFT length planning.  Integer/indexing work
// that the reference does serially on the host in src/geometry.f90,
// src/sparse_mat.f90:13-60 and src/precorrection.f90:182-258.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "fft.hpp"

namespace scadyn {

static inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -a[0] * b[2] + a[2] * b[0];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void build_geometry(const scadyn_mesh &mesh, const scadyn_opts &opts, HostGeometry &g) {
    if (mesh.n_tet <= 0 || mesh.n_node <= 0 || !mesh.coord || !mesh.etopol || !mesh.param)
        fail(SCADYN_ERR_ARG, "empty mesh or null array");
    if (opts.M_ex != 2) fail(SCADYN_ERR_UNSUPPORTED, "only stencil order M_ex = 2 is implemented");
    if (opts.near_zone != 1) fail(SCADYN_ERR_UNSUPPORTED, "only near_zone = 1 is implemented");
    if (opts.order != 0) fail(SCADYN_ERR_UNSUPPORTED, "only basis order 0 is implemented");
    const int N = mesh.n_tet, Nn = mesh.n_node;
    g.N = N;
    g.Nnode = Nn;
    const double *X = mesh.coord;
    std::vector<double> tc((size_t)N * 12), vol(N);
    double volsum = 0.0;
    for (int t = 0; t < N; ++t) {
        double p[4][3];
        for (int a = 0; a < 4; ++a) {
            int id = mesh.etopol[4 * (size_t)t + a];
            if (id < 1 || id > Nn)
                fail(SCADYN_ERR_ARG, "etopol entry out of range at tet " + std::to_string(t + 1));
            for (int c = 0; c < 3; ++c) p[a][c] = X[3 * (size_t)(id - 1) + c];
        }
        // gradshape determinant (src/common.f90:460-469)
        double J[3][3];
        for (int c = 0; c < 3; ++c) {
            J[c][0] = p[1][c] - p[0][c];
            J[c][1] = p[2][c] - p[0][c];
            J[c][2] = p[3][c] - p[0][c];
        }
        double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (!(det > 0.0))
            fail(SCADYN_ERR_ORIENTATION,
                 "Something is wrong (detJ <= 0) at tet " + std::to_string(t + 1));
        // tetra_volume (src/common.f90:255-267)
        double p41[3], p21[3], p31[3], cr[3];
        for (int c = 0; c < 3; ++c) {
            p41[c] = p[3][c] - p[0][c];
            p21[c] = p[1][c] - p[0][c];
            p31[c] = p[2][c] - p[0][c];
        }
        cross3(p21, p31, cr);
        double dt = p41[0] * cr[0] + p41[1] * cr[1] + p41[2] * cr[2];
        vol[t] = std::sqrt(dt * dt) / 6.0;
        volsum += vol[t];
        std::memcpy(&tc[12 * (size_t)t], p, sizeof(p));
    }
    double mn[3], mx[3];
    for (int c = 0; c < 3; ++c) mn[c] = mx[c] = X[c];
    for (int i = 1; i < Nn; ++i)
        for (int c = 0; c < 3; ++c) {
            mn[c] = std::min(mn[c], X[3 * (size_t)i + c]);
            mx[c] = std::max(mx[c], X[3 * (size_t)i + c]);
        }
    // delta with the default-real exponent 1.0/3.0 (src/geometry.f90:88)
    const double third32 = (double)(1.0f / 3.0f);
    double delta = opts.grid_size * 1.0 / opts.M_ex * std::pow(volsum / (double)N, third32);
    int n[3];
    for (int c = 0; c < 3; ++c) {
        n[c] = (int)std::ceil(std::fabs(mx[c] - mn[c]) / delta) + 1;
        n[c] = n[c] + (n[c] + 1) % 2 + 1;  // src/geometry.f90:94-98
    }
    g.Nx = n[0];
    g.Ny = n[1];
    g.Nz = n[2];
    g.delta = delta;
    g.box_delta = delta * opts.M_ex;
    for (int c = 0; c < 3; ++c) g.min_coord[c] = mn[c];
    g.ncx = (g.Nx - 2) / 2;
    g.ncy = (g.Ny - 2) / 2;
    g.ncz = (g.Nz - 2) / 2;
    g.N_cubes = g.ncx * g.ncy * g.ncz;
    // tet -> cube by centroid (src/geometry.f90:13-55, :342-353)
    std::vector<int> cube_of(N);
    for (int t = 0; t < N; ++t) {
        const double *p = &tc[12 * (size_t)t];
        int M[3];
        for (int c = 0; c < 3; ++c) {
            double cp = (p[c] + p[3 + c] + p[6 + c] + p[9 + c]) / 4;
            M[c] = (int)std::floor((cp - mn[c]) / g.box_delta);
        }
        if (M[0] < 0 || M[1] < 0 || M[2] < 0 || M[0] >= g.ncx || M[1] >= g.ncy || M[2] >= g.ncz)
            fail(SCADYN_ERR_GRID, "something is wrong (Point is not in any cube), tet " +
                                      std::to_string(t + 1));
        cube_of[t] = M[0] + g.ncx * M[1] + g.ncx * g.ncy * M[2];
    }
    // stable counting sort by cube: internal order
    g.cube_start.assign(g.N_cubes + 1, 0);
    for (int t = 0; t < N; ++t) g.cube_start[cube_of[t] + 1]++;
    g.N_tet_cube = 0;
    for (int c = 0; c < g.N_cubes; ++c) {
        g.N_tet_cube = std::max(g.N_tet_cube, g.cube_start[c + 1]);
        g.cube_start[c + 1] += g.cube_start[c];
    }
    std::vector<int> fill(g.cube_start.begin(), g.cube_start.end() - 1);
    g.perm.resize(N);
    g.inv.resize(N);
    for (int t = 0; t < N; ++t) {
        int i = fill[cube_of[t]]++;
        g.perm[i] = t;
        g.inv[t] = i;
    }
    g.tet_cube.resize(N);
    g.tcoord.resize((size_t)N * 12);
    g.vol.resize(N);
    g.param.resize(N);
    for (int i = 0; i < N; ++i) {
        int t = g.perm[i];
        g.tet_cube[i] = cube_of[t];
        std::memcpy(&g.tcoord[12 * (size_t)i], &tc[12 * (size_t)t], 12 * sizeof(double));
        g.vol[i] = vol[t];
        g.param[i] = cmk(mesh.param[2 * (size_t)t], mesh.param[2 * (size_t)t + 1]);
    }
}

void build_near_pairs(const HostGeometry &g, NearPairs &np) {
    // src/precorrection.f90:182-258: test cube, near cube (z,y,x order),
    // test tet, basis tet, keep reference ids tet_B >= tet_T.
    np.T.clear();
    np.B.clear();
    np.off.clear();
    for (int tc = 0; tc < g.N_cubes; ++tc) {
        int t0 = g.cube_start[tc], t1 = g.cube_start[tc + 1];
        if (t0 == t1) continue;
        int cx = tc % g.ncx, cy = (tc / g.ncx) % g.ncy, cz = tc / (g.ncx * g.ncy);
        for (int dz = -1; dz <= 1; ++dz) {
            int zz = cz + dz;
            if (zz < 0 || zz >= g.ncz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int yy = cy + dy;
                if (yy < 0 || yy >= g.ncy) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = cx + dx;
                    if (xx < 0 || xx >= g.ncx) continue;
                    int bc = xx + g.ncx * (yy + g.ncy * zz);
                    uint8_t off = (uint8_t)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1));
                    for (int T = t0; T < t1; ++T)
                        for (int B = g.cube_start[bc]; B < g.cube_start[bc + 1]; ++B)
                            if (g.perm[B] >= g.perm[T]) {
                                np.T.push_back(T);
                                np.B.push_back(B);
                                np.off.push_back(off);
                            }
                }
            }
        }
    }
    // row lists: row T gets (B, p), row B gets (T, p | transpose) when T != B
    const size_t P = np.T.size();
    if (P >= (size_t)BLK_TRANSPOSE) fail(SCADYN_ERR_UNSUPPORTED, "too many near-zone blocks");
    np.row_start.assign(g.N + 1, 0);
    for (size_t p = 0; p < P; ++p) {
        np.row_start[np.T[p] + 1]++;
        if (np.T[p] != np.B[p]) np.row_start[np.B[p] + 1]++;
    }
    for (int i = 0; i < g.N; ++i) np.row_start[i + 1] += np.row_start[i];
    np.ent_col.resize(np.row_start[g.N]);
    np.ent_blk.resize(np.row_start[g.N]);
    std::vector<int64_t> fill(np.row_start.begin(), np.row_start.end() - 1);
    for (size_t p = 0; p < P; ++p) {
        int64_t e = fill[np.T[p]]++;
        np.ent_col[e] = np.B[p];
        np.ent_blk[e] = (uint32_t)p;
        if (np.T[p] != np.B[p]) {
            e = fill[np.B[p]]++;
            np.ent_col[e] = np.T[p];
            np.ent_blk[e] = (uint32_t)p | BLK_TRANSPOSE;
        }
    }
    // order each row by (near-cube offset, column tet) and check that it holds
    // every tet of the near cubes exactly once
    np.cube_rows = true;
    std::vector<int64_t> key;
    for (int t = 0; t < g.N; ++t) {
        const int64_t r0 = np.row_start[t], r1 = np.row_start[t + 1];
        const int ct = g.tet_cube[t];
        const int tx = ct % g.ncx, ty = (ct / g.ncx) % g.ncy, tz = ct / (g.ncx * g.ncy);
        key.resize(r1 - r0);
        for (int64_t e = r0; e < r1; ++e) {
            const int cc = g.tet_cube[np.ent_col[e]];
            const int dx = cc % g.ncx - tx, dy = (cc / g.ncx) % g.ncy - ty, dz = cc / (g.ncx * g.ncy) - tz;
            key[e - r0] = (int64_t)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)) * g.N + np.ent_col[e];
        }
        // insertion sort: rows are short and close to sorted
        for (int64_t a = 1; a < r1 - r0; ++a) {
            const int64_t k = key[a];
            const int col = np.ent_col[r0 + a];
            const uint32_t blk = np.ent_blk[r0 + a];
            int64_t b = a - 1;
            while (b >= 0 && key[b] > k) {
                key[b + 1] = key[b];
                np.ent_col[r0 + b + 1] = np.ent_col[r0 + b];
                np.ent_blk[r0 + b + 1] = np.ent_blk[r0 + b];
                --b;
            }
            key[b + 1] = k;
            np.ent_col[r0 + b + 1] = col;
            np.ent_blk[r0 + b + 1] = blk;
        }
        if (!np.cube_rows) continue;
        int64_t e = r0;
        for (int dz = -1; dz <= 1 && np.cube_rows; ++dz)
            for (int dy = -1; dy <= 1 && np.cube_rows; ++dy)
                for (int dx = -1; dx <= 1 && np.cube_rows; ++dx) {
                    const int xx = tx + dx, yy = ty + dy, zz = tz + dz;
                    if (xx < 0 || xx >= g.ncx || yy < 0 || yy >= g.ncy || zz < 0 || zz >= g.ncz) continue;
                    const int bc = xx + g.ncx * (yy + g.ncy * zz);
                    for (int c = g.cube_start[bc]; c < g.cube_start[bc + 1]; ++c)
                        if (e >= r1 || np.ent_col[e++] != c) np.cube_rows = false;
                }
        if (e != r1) np.cube_rows = false;
    }
}

// Hestenes one-sided Jacobi SVD of A (m x n, column-major), m >= n.
// Replaces ZGESVD in pinv (src/possu.f90:369-422): pinv = V diag(1/s) U^H.
std::vector<cd> pinv_jacobi(const std::vector<cd> &A0, int m, int n, double thr) {
    std::vector<cd> A = A0, V((size_t)n * n, cmk(0, 0));
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = cmk(1, 0);
    auto col = [&](std::vector<cd> &M, int rows, int j) { return &M[(size_t)j * rows]; };
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                cd *ap = col(A, m, p), *aq = col(A, m, q);
                double alpha = 0, beta = 0;
                cd gam = cmk(0, 0);
                for (int i = 0; i < m; ++i) {
                    alpha += cabs2(ap[i]);
                    beta += cabs2(aq[i]);
                    gam = cadd(gam, cmulc(ap[i], aq[i]));
                }
                double ag = std::sqrt(cabs2(gam));
                if (ag == 0.0) continue;
                double rel = ag / std::sqrt(alpha * beta);
                off = std::max(off, rel);
                if (rel < 1e-17) continue;
                // phase so that ap^H (aq e^{-i phi}) is real positive
                cd ph = cmk(gam.x / ag, -gam.y / ag);
                double zeta = (beta - alpha) / (2.0 * ag);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < m; ++i) {
                    cd x = ap[i], y = cmul(aq[i], ph);
                    ap[i] = csub(cscale(x, c), cscale(y, s));
                    aq[i] = cadd(cscale(x, s), cscale(y, c));
                }
                cd *vp = col(V, n, p), *vq = col(V, n, q);
                for (int i = 0; i < n; ++i) {
                    cd x = vp[i], y = cmul(vq[i], ph);
                    vp[i] = csub(cscale(x, c), cscale(y, s));
                    vq[i] = cadd(cscale(x, s), cscale(y, c));
                }
            }
        if (off < 1e-16) break;
    }
    // pinv (n x m) = sum_j v_j (1/s_j) u_j^H, u_j = a_j / s_j
    std::vector<cd> P((size_t)n * m, cmk(0, 0));
    for (int j = 0; j < n; ++j) {
        cd *aj = col(A, m, j);
        double s2 = 0;
        for (int i = 0; i < m; ++i) s2 += cabs2(aj[i]);
        double s = std::sqrt(s2);
        if (!(s > thr)) continue;
        cd *vj = col(V, n, j);
        for (int i = 0; i < m; ++i) {
            cd u = cscale(cconj(aj[i]), 1.0 / (s * s));  // conj(u_i)/s
            for (int r = 0; r < n; ++r) P[(size_t)i * n + r] = cadd(P[(size_t)i * n + r], cmul(vj[r], u));
        }
    }
    return P;
}

static bool is_smooth(int L) {
    int r = L;
    for (int p : {2, 3, 5, 7})
        while (r % p == 0) r /= p;
    return r == 1;
}

// L = R1 R2 with both radices in the register FFT's set
static bool reg_splittable(int L) {
    for (int r1 = 2; r1 <= 20; ++r1)
        if (L % r1 == 0 && reg_radix_ok(r1) && reg_radix_ok(L / r1)) return true;
    return false;
}

// FFT length for a zero-padded axis of n >= 2N-1 points: the smallest
// 2,3,5,7-smooth length, or a longer one (at most n/8 more points) that the
// two-stage register FFT covers when the smallest does not
int smooth_length(int n) {
    int s0 = std::max(n, 1);
    while (!is_smooth(s0)) ++s0;
    if (reg_splittable(s0)) return s0;
    for (int L = s0 + 1; L <= n + n / 8; ++L)
        if (is_smooth(L) && reg_splittable(L)) return L;
    return s0;
}

std::vector<int> factor_radices(int L) {
    std::vector<int> rad;
    int r = L;
    while (r % 8 == 0) { rad.push_back(8); r /= 8; }
    while (r % 4 == 0) { rad.push_back(4); r /= 4; }
    while (r % 2 == 0) { rad.push_back(2); r /= 2; }
    for (int p : {7, 5, 3})
        while (r % p == 0) { rad.push_back(p); r /= p; }
    if (r != 1) {
        // a prime factor above 7 (reference 2N lengths): one generic stage
        rad.push_back(r);
    }
    return rad;
}

}  // namespace scadyn
