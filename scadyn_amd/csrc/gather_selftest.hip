// Host self-test of the LDS-staged per-cube gather (gather_core.hpp).
//
// Runs the kernel's own per-lane steps on the CPU, one wave at a time, with
// the workgroup barriers as phase boundaries, over exact-size buffers
// (AddressSanitizer on the host build reports any access outside them), and
// compares every output with a direct evaluation of
//   y_t = sum_j (S_{c+1}[t][j] g3[node_j] - k^2 S[t][j] g_c[node_j])
//       + sum_{near tets v} A(t, v) x_v
// on a synthetic cube lattice: 0 to 11 tets per cube (empty cubes, cubes with
// more tets than one staging pass holds), in the kernel's workgroup order
// (xcd_remap, itself checked to be a bijection), near-zone entries pointing at
// shuffled blocks with and without the transpose flag, inactive slots (left
// unwritten), and
// batch widths 64, 112 (partial second chunk), 56 and 40.
//
// The complex64 matrix-core pass (k_gather_mf) runs with the instruction
// emulated per tile.
//
// Exit status 0 when every case agrees to 1e-13 (complex128) / 1e-5 (complex64).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <random>
#include <type_traits>
#include <vector>

#include "gather_core.hpp"

using namespace scadyn;
using namespace scadyn::gcore;

static int failures = 0;

template <typename C>
static C to(cd a) {
    return Prec<C>::from(a);
}

// one pass of NT row tets, lane by lane, barriers as phase boundaries
template <typename C, int NT, int GT, int VC>
static void emulate_pass(const GatherArgs<C> &g, const std::vector<Item> &it, int tp, std::vector<cd> &smbuf) {
    // the kernel's staging buffer, viewed as the S* records (C) or the blocks (complex64)
    C *sm = (C *)smbuf.data();
    cf *smb = (cf *)smbuf.data();
    std::vector<std::array<std::array<C, 3>, NT>> acc(64);
    for (int l = 0; l < 64; ++l) zero_acc<C, NT>((C(*)[3])acc[l].data());
    for (int l = 0; l < 64; ++l) stage_s4(g, tp, NT, l, sm);
    for (int l = 0; l < 64; ++l) grid_part<C, NT>(g, it[l], sm, (C(*)[3])acc[l].data());
    std::vector<int64_t> rsu(64);
    for (int l = 0; l < 64; ++l) rsu[l] = lane_row_start<C, GT, VC>(g, tp, NT, l);
    int q = 0;
    for_near_cubes(g, it[0], [&](int c0, int c1) {
        for (int v0 = c0; v0 < c1; v0 += VC) {
            const int nv = std::min(VC, c1 - v0);
            for (int l = 0; l < 64; ++l) stage_blocks<C, GT, VC>(g, NT, nv, q, l, rsu[l], smb);
            for (int l = 0; l < 64; ++l) chunk_part<C, NT, VC>(g, it[l], v0, nv, smb, (C(*)[3])acc[l].data());
            q += nv;
        }
    });
    for (int l = 0; l < 64; ++l) write_out<C, NT>(g, it[l], tp, (C(*)[3])acc[l].data());
}

// The complex64 pass with the near zone on the matrix cores (k_gather_mf):
// the same operand functions per lane, v_mfma_f32_16x16x4f32 emulated per
// tile (D[m][n] += sum_k A[m][k] B[k][n], A[m][k] from lane m + 16 k, B[k][n]
// from lane n + 16 k, output (m, n) to lane n + 16 (m / 4)'s element m % 4 as
// mf_row states), the kernel's phase order and exchange buffer.
template <int NT, int GT, int VC>
static void emulate_pass_mf(const GatherArgs<cf> &g, const std::vector<Item> &it, int tp, std::vector<cd> &smbuf) {
    cf *sm = (cf *)smbuf.data();
    cf *smb = (cf *)smbuf.data();
    float *xb = (float *)smbuf.data();
    constexpr int RT = mf_rows(NT);
    std::vector<std::array<std::array<std::array<float, 4>, 4>, RT>> D(64);
    for (auto &d : D)
        for (auto &a : d)
            for (auto &b : a) b.fill(0.0f);
    std::vector<int64_t> rsu(64);
    std::vector<std::array<long long, 4>> off(64);
    for (int l = 0; l < 64; ++l) {
        rsu[l] = lane_row_start<cf, GT, VC>(g, tp, NT, l);
        const int s0 = (it[l].b & ~63) + (l & 15);
        for (int st = 0; st < 4; ++st) off[l][st] = s0 + 16 * st < g.B ? g.in_off[s0 + 16 * st] : 0;
    }
    int q = 0;
    for_near_cubes(g, it[0], [&](int c0, int c1) {
        for (int v0 = c0; v0 < c1; v0 += VC) {
            const int nv = std::min(VC, c1 - v0);
            for (int l = 0; l < 64; ++l) stage_blocks<cf, GT, VC>(g, NT, nv, q, l, rsu[l], smb);
            const int S = (6 * nv + 3) / 4;
            for (int s = 0; s < S; ++s) {
                float ao[64][RT], bo[64][4];
                for (int l = 0; l < 64; ++l) {
                    const int kp = 4 * s + (l >> 4), s0 = (it[l].b & ~63) + (l & 15);
                    for (int st = 0; st < 4; ++st)
                        bo[l][st] = mf_b(g, v0, nv, kp, s0 + 16 * st, off[l][st], s0 + 16 * st < g.B);
                    for (int rt = 0; rt < RT; ++rt) ao[l][rt] = mf_a(smb, NT, nv, VC, 16 * rt + (l & 15), kp);
                }
                for (int rt = 0; rt < RT; ++rt)
                    for (int st = 0; st < 4; ++st) {
                        float tile[16][16];
                        for (int m = 0; m < 16; ++m)
                            for (int n = 0; n < 16; ++n) {
                                float a = 0.0f;
                                for (int k = 0; k < 4; ++k) a += ao[m + 16 * k][rt] * bo[n + 16 * k][st];
                                tile[m][n] = a;
                            }
                        for (int l = 0; l < 64; ++l)
                            for (int e = 0; e < 4; ++e) D[l][rt][st][e] += tile[mf_row(l, e)][l & 15];
                    }
            }
            q += nv;
        }
    });
    for (int l = 0; l < 64; ++l)
        for (int rt = 0; rt < RT; ++rt)
            for (int st = 0; st < 4; ++st)
                for (int e = 0; e < 4; ++e) xb[mf_xidx(l, e, rt, st)] = D[l][rt][st][e];
    std::vector<std::array<std::array<cf, 3>, NT>> acc(64);
    for (int l = 0; l < 64; ++l) {
        zero_acc<cf, NT>((cf(*)[3])acc[l].data());
        mf_combine<NT>(xb, l, (cf(*)[3])acc[l].data());
    }
    for (int l = 0; l < 64; ++l) stage_s4(g, tp, NT, l, sm);
    for (int l = 0; l < 64; ++l) grid_part<cf, NT>(g, it[l], sm, (cf(*)[3])acc[l].data());
    for (int l = 0; l < 64; ++l) write_out<cf, NT>(g, it[l], tp, (cf(*)[3])acc[l].data());
}

template <int GT, int VC>
static void emulate_mf(const GatherArgs<cf> &g) {
    constexpr int SB = Shape<cf, GT, VC>::BYTES > MF_XBYTES ? Shape<cf, GT, VC>::BYTES : MF_XBYTES;
    std::vector<cd> sm((SB + 15) / 16);  // exact size, 16-byte aligned
    const long long nw = (long long)g.n_ne * ((g.B + 63) >> 6);
    for (long long b = 0; b < nw; ++b) {
        const long long w = xcd_remap(b, nw);
        std::vector<Item> it(64);
        for (int l = 0; l < 64; ++l) it[l] = make_item(g, w, l);
        for (int tp = it[0].t0; tp < it[0].t1; tp += GT) {
            switch (std::min(GT, it[0].t1 - tp)) {
                case 1: emulate_pass_mf<1, GT, VC>(g, it, tp, sm); break;
                case 2: emulate_pass_mf<2, GT, VC>(g, it, tp, sm); break;
                case 3: emulate_pass_mf<3, GT, VC>(g, it, tp, sm); break;
                case 4: emulate_pass_mf<4, GT, VC>(g, it, tp, sm); break;
                case 5: emulate_pass_mf<5, GT, VC>(g, it, tp, sm); break;
                case 6: emulate_pass_mf<6, GT, VC>(g, it, tp, sm); break;
                case 7: emulate_pass_mf<7, GT, VC>(g, it, tp, sm); break;
                default: emulate_pass_mf<8, GT, VC>(g, it, tp, sm); break;
            }
        }
    }
}

template <typename C, int GT, int VC>
static void emulate(const GatherArgs<C> &g) {
    static_assert(GT == 8, "pass switch covers 1..8");
    std::vector<cd> sm((Shape<C, GT, VC>::BYTES + 15) / 16);  // exact size, 16-byte aligned
    const long long nw = (long long)g.n_ne * ((g.B + 63) >> 6);
    for (long long b = 0; b < nw; ++b) {
        const long long w = xcd_remap(b, nw);  // the kernel's order (one item per workgroup)
        std::vector<Item> it(64);
        for (int l = 0; l < 64; ++l) it[l] = make_item(g, w, l);
        for (int tp = it[0].t0; tp < it[0].t1; tp += GT) {
            switch (std::min(GT, it[0].t1 - tp)) {
                case 1: emulate_pass<C, 1, GT, VC>(g, it, tp, sm); break;
                case 2: emulate_pass<C, 2, GT, VC>(g, it, tp, sm); break;
                case 3: emulate_pass<C, 3, GT, VC>(g, it, tp, sm); break;
                case 4: emulate_pass<C, 4, GT, VC>(g, it, tp, sm); break;
                case 5: emulate_pass<C, 5, GT, VC>(g, it, tp, sm); break;
                case 6: emulate_pass<C, 6, GT, VC>(g, it, tp, sm); break;
                case 7: emulate_pass<C, 7, GT, VC>(g, it, tp, sm); break;
                default: emulate_pass<C, 8, GT, VC>(g, it, tp, sm); break;
            }
        }
    }
}

template <typename C>
static void run_case(int B, std::mt19937_64 &rng, bool mf = false) {
    const int ncx = 4, ncy = 3, ncz = 3;
    const int Nx = 2 * ncx + 2, Ny = 2 * ncy + 2, Nz = 2 * ncz + 2;
    const int ncubes = ncx * ncy * ncz;
    std::uniform_int_distribution<int> cnt(0, 11);
    std::vector<int> cube_start(ncubes + 1, 0);
    for (int c = 0; c < ncubes; ++c) cube_start[c + 1] = cube_start[c] + (c % 5 == 0 ? 0 : cnt(rng));
    const int N = cube_start[ncubes];
    std::vector<int> tet_cube(N), ne;
    for (int c = 0; c < ncubes; ++c) {
        for (int t = cube_start[c]; t < cube_start[c + 1]; ++t) tet_cube[t] = c;
        if (cube_start[c + 1] > cube_start[c]) ne.push_back(c);
    }
    // rows in near-cube order; each (t, v) refers to a block, transposed for half of them
    std::vector<int64_t> row_start(N + 1, 0);
    std::vector<int> ent_col;
    for (int t = 0; t < N; ++t) {
        const int c = tet_cube[t], cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = cx + dx, yy = cy + dy, zz = cz + dz;
                    if (xx < 0 || xx >= ncx || yy < 0 || yy >= ncy || zz < 0 || zz >= ncz) continue;
                    const int nc = xx + ncx * (yy + ncy * zz);
                    for (int v = cube_start[nc]; v < cube_start[nc + 1]; ++v) ent_col.push_back(v);
                }
        row_start[t + 1] = (int64_t)ent_col.size();
    }
    const size_t E = ent_col.size();
    const size_t P = E / 2 + 1;
    std::normal_distribution<double> nd;
    std::uniform_int_distribution<size_t> pick(0, P - 1);
    std::vector<uint32_t> ent_blk(E);
    for (size_t e = 0; e < E; ++e) ent_blk[e] = (uint32_t)pick(rng) | ((rng() & 1) ? 0x80000000u : 0u);
    std::vector<cf> blocks(P * 9);
    for (auto &z : blocks) z = cmkf((float)nd(rng), (float)nd(rng));
    std::vector<C> S4((size_t)N * 108), grid((size_t)Nx * Ny * Nz * 4 * B);
    for (auto &z : S4) z = to<C>(cmk(nd(rng), nd(rng)));
    for (auto &z : grid) z = to<C>(cmk(nd(rng), nd(rng)));
    // x: slot b's vector at in_off[b] + row*B + b inside a buffer of two vector blocks
    std::vector<C> xin((size_t)2 * 3 * N * B);  // in the matvec precision
    for (auto &z : xin) z = to<C>(cmk(nd(rng), nd(rng)));
    std::vector<long long> in_off(B);
    std::vector<int> active(B);
    for (int b = 0; b < B; ++b) {
        in_off[b] = (b % 3 == 1) ? (long long)3 * N * B : 0;
        active[b] = (b % 7 != 3);
    }
    std::vector<cd> y((size_t)3 * N * B, cmk(-7.0, -7.0));
    const double k2 = 1.7;
    GatherArgs<C> g = {grid.data(), xin.data(), in_off.data(), B, ne.data(), (int)ne.size(), cube_start.data(),
                       S4.data(), Nx, Ny, ncx, ncy, ncz, row_start.data(), ent_blk.data(), blocks.data(), k2,
                       active.data(), y.data()};
    if constexpr (std::is_same<C, cf>::value) {
        if (mf)
            emulate_mf<GATHER_GT, GATHER_VC>(g);
        else
            emulate<C, GATHER_GT, GATHER_VC>(g);
    } else {
        emulate<C, GATHER_GT, GATHER_VC>(g);
    }
    // direct evaluation in double
    double num = 0.0, den = 0.0;
    for (int t = 0; t < N; ++t) {
        const int c = tet_cube[t], cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        for (int b = 0; b < B; ++b) {
            cd acc[3] = {cmk(0, 0), cmk(0, 0), cmk(0, 0)};
            if (!active[b]) acc[0] = acc[1] = acc[2] = cmk(-7.0, -7.0);  // not written
            if (active[b]) {
                for (int j = 0; j < 27; ++j) {
                    const int jx = j % 3, jy = (j / 3) % 3, jz = j / 9;
                    const size_t node = (size_t)(2 * cx + jx) + (size_t)Nx * ((2 * cy + jy) + (size_t)Ny * (2 * cz + jz));
                    for (int r = 0; r < 3; ++r) {
                        const cd gr = Prec<cd>::from(grid[(node * 4 + r) * B + b]);
                        const cd g3 = Prec<cd>::from(grid[(node * 4 + 3) * B + b]);
                        const cd s0 = Prec<cd>::from(S4[((size_t)t * 27 + j) * 4]);
                        const cd sr = Prec<cd>::from(S4[((size_t)t * 27 + j) * 4 + 1 + r]);
                        acc[r] = cadd(acc[r], csub(cmul(sr, g3), cscale(cmul(s0, gr), k2)));
                    }
                }
                for (int64_t e = row_start[t]; e < row_start[t + 1]; ++e) {
                    const int v = ent_col[e];
                    const uint32_t eb = ent_blk[e];
                    const cf *a = &blocks[(size_t)(eb & 0x7fffffffu) * 9];
                    for (int i = 0; i < 3; ++i)
                        for (int jj = 0; jj < 3; ++jj) {
                            const cf m = (eb & 0x80000000u) ? a[jj + 3 * i] : a[i + 3 * jj];
                            const cd xv = Prec<cd>::from(xin[in_off[b] + (size_t)(3 * v + jj) * B + b]);
                            acc[i] = cadd(acc[i], cmul(Prec<cd>::from(m), xv));
                        }
                }
            }
            for (int r = 0; r < 3; ++r) {
                const cd got = y[(size_t)(3 * t + r) * B + b];
                num += cabs2(csub(got, acc[r]));
                den += cabs2(acc[r]) + 1e-300;
            }
        }
    }
    const double err = std::sqrt(num / den);
    const double tol = sizeof(C) == 16 ? 1e-13 : 1e-5;
    const bool ok = err < tol;
    if (!ok) ++failures;
    std::printf("%s B=%3d N=%4d cubes %d (non-empty %zu) entries %zu  rel.err %.2e %s\n",
                sizeof(C) == 16 ? "c128" : (mf ? "c64m" : "c64 "), B, N, ncubes, ne.size(), E, err, ok ? "ok" : "FAIL");
}

// xcd_remap is a bijection of [0, G) that gives XCD b % 8 one contiguous range
static void check_remap() {
    std::vector<long long> Gs;
    for (long long G = 1; G <= 300; ++G) Gs.push_back(G);
    for (long long G : {71864LL, 71865LL, 1048575LL, 1048576LL}) Gs.push_back(G);
    int bad = 0;
    for (long long G : Gs) {
        std::vector<char> seen(G, 0);
        std::vector<long long> lo(8, G), hi(8, -1), n(8, 0);
        for (long long b = 0; b < G; ++b) {
            const long long w = xcd_remap(b, G);
            if (w < 0 || w >= G || seen[w]) {
                ++bad;
                break;
            }
            seen[w] = 1;
            const int x = (int)(b % 8);
            lo[x] = std::min(lo[x], w);
            hi[x] = std::max(hi[x], w);
            ++n[x];
        }
        for (int x = 0; x < 8; ++x)
            if (n[x] && hi[x] - lo[x] + 1 != n[x]) ++bad;  // contiguous
    }
    std::printf("xcd_remap over %zu grid sizes: %s\n", Gs.size(), bad ? "FAIL" : "ok");
    if (bad) ++failures;
}

int main() {
    check_remap();
    std::mt19937_64 rng(99);
    for (int B : {64, 112, 56, 40}) {
        run_case<cd>(B, rng);
        run_case<cf>(B, rng);
        run_case<cf>(B, rng, true);  // near zone on the (emulated) matrix cores
    }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
