// Host self-test of the block modified Gram-Schmidt (bmgs_core.hpp).
//
// Runs the kernels' own per-thread and per-slot code on the CPU: every
// workgroup of the row sweep thread by thread (with the workgroup reduction
// replayed through arrays in place of LDS), the partial-sum reduction and the
// forward substitution, over exact-size buffers (AddressSanitizer on the host
// build reports any access outside them).  Slots are set up in every phase
// (restart, Arnoldi at columns 0, BS-1, BS, 2BS-1, 2BS, m-1, idle, done) and the
// resulting H column, |w|^2 before and after, and the orthogonalised w are
// compared with plain one-vector-at-a-time MGS.  Batch widths cover one row per
// sweep step (B = 256), several (B = 5, 64, 112) and two slot chunks (B = 300).
//
// Exit status 0 when every case agrees to 1e-12 relative.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "bmgs_core.hpp"

using namespace scadyn;
using namespace scadyn::bmgs;

static int failures = 0;

static double crel(cd a, cd b, double scale) { return std::sqrt(cabs2(csub(a, b))) / scale; }

static void run_case(int B, int m, long long n, int nwg, std::mt19937_64 &rng) {
    KrylovDims d;
    d.n = n;
    d.B = B;
    d.m = m;
    d.nwg = nwg;
    const size_t nB = (size_t)n * B;
    std::normal_distribution<double> nd;
    auto rnd = [&]() { return cmk(nd(rng), nd(rng)); };
    std::vector<cd> V((size_t)(m + 2) * nB), W(nB), Rhs(nB), H((size_t)B * (m + 1) * m);
    for (auto &z : V) z = rnd();
    for (auto &z : W) z = rnd();
    for (auto &z : Rhs) z = rnd();
    for (auto &z : H) z = rnd();
    const std::vector<cd> W0 = W;
    std::vector<cd> pq((size_t)nwg * std::max(m + 1, BQ) * B), red((size_t)B * std::max(m + 1, BQ + 2));
    std::vector<double> pnrm((size_t)nwg * B), pnav((size_t)nwg * B), nrm2(B, -1.0), nav2(B, -1.0);
    std::vector<SlotState> st(B);
    const int cols[] = {0, BS - 1, BS, 2 * BS - 1, 2 * BS, m - 1};
    int kmax = 0;
    for (int b = 0; b < B; ++b) {
        SlotState s = {};
        const int kind = b % 9;
        if (kind < 6) {
            s.phase = PH_ARNOLDI;
            s.i = std::min(cols[kind], m - 1);
            kmax = std::max(kmax, s.i);
        } else if (kind == 6) {
            s.phase = PH_RESTART;
            s.i = m / 2;  // stale column of a previous cycle
        } else {
            s.phase = kind == 7 ? PH_IDLE : PH_DONE;
            s.i = m - 1;
        }
        st[b] = s;
    }
    // launch_bmgs, replayed
    const int SB = B < 256 ? B : 256;
    const int ys = (B + SB - 1) / SB;
    const int np = (kmax + BS) / BS;
    std::vector<cd> acc((size_t)256 * BQ), sq(256);
    std::vector<double> sn1(256), sn2(256);
    for (int p = 0; p <= np; ++p) {
        for (int by = 0; by < ys; ++by)
            for (int bx = 0; bx < nwg; ++bx) {
                for (int t = 0; t < 256; ++t) {
                    const RowSlot rs = row_slot(B, t, by);
                    pass_rows(d, p, W.data(), V.data(), Rhs.data(), st.data(), H.data(), rs, bx, nwg,
                              &acc[(size_t)t * BQ], sn1[t], sn2[t]);
                }
                for (int t = 0; t < 256; ++t) {
                    const RowSlot rs = row_slot(B, t, by);
                    if (rs.ok && rs.rl == 0)
                        store_norm_partials(d, rs, bx, sn1.data(), sn2.data(), pnrm.data(), pnav.data());
                }
                for (int q = 0; q < BQ; ++q) {
                    for (int t = 0; t < 256; ++t) sq[t] = acc[(size_t)t * BQ + q];
                    for (int t = 0; t < 256; ++t) {
                        const RowSlot rs = row_slot(B, t, by);
                        if (rs.ok && rs.rl == 0) store_dot_partial(d, rs, bx, q, sq.data(), pq.data());
                    }
                }
            }
        for (int b = 0; b < B; ++b)
            for (int q = 0; q < BQ + 2; ++q) {
                const ReduceSrc src = reduce_src(d, p, st.data(), pq.data(), pnrm.data(), pnav.data(), q, b);
                if (src.kind == ReduceSrc::NONE) continue;
                cd a = cmk(0.0, 0.0);
                for (int w = 0; w < nwg; ++w) {
                    if (src.kind == ReduceSrc::DOT) a = cadd(a, src.z[(size_t)w * src.stride + src.off]);
                    if (src.kind == ReduceSrc::NORM) a.x += src.r[(size_t)w * src.stride + src.off];
                }
                red[red_index(q, b)] = a;
            }
        for (int b = 0; b < B; ++b) coef_slot(d, p, st.data(), red.data(), H.data(), nrm2.data(), nav2.data(), b);
    }
    // plain MGS per slot
    double worst = 0.0;
    for (int b = 0; b < B; ++b) {
        const SlotState &s = st[b];
        std::vector<cd> w(n);
        for (long long r = 0; r < n; ++r) w[r] = W0[(size_t)r * B + b];
        if (s.phase == PH_ARNOLDI) {
            double w0 = 0.0;
            for (auto &z : w) w0 += cabs2(z);
            for (int k = 0; k <= s.i; ++k) {
                cd h = cmk(0.0, 0.0);
                for (long long r = 0; r < n; ++r) h = cadd(h, cmulc(V[(size_t)k * nB + (size_t)r * B + b], w[r]));
                for (long long r = 0; r < n; ++r) w[r] = csub(w[r], cmul(h, V[(size_t)k * nB + (size_t)r * B + b]));
                const cd got = H[(size_t)b * (m + 1) * m + (size_t)(m + 1) * s.i + k];
                worst = std::max(worst, crel(got, h, std::sqrt(cabs2(h)) + 1e-300));
            }
            double wn = 0.0, dw = 0.0;
            for (long long r = 0; r < n; ++r) {
                wn += cabs2(w[r]);
                dw = std::max(dw, std::sqrt(cabs2(csub(w[r], W[(size_t)r * B + b]))));
            }
            worst = std::max(worst, std::fabs(nrm2[b] - wn) / wn);
            worst = std::max(worst, std::fabs(nav2[b] - w0) / w0);
            worst = std::max(worst, dw / std::sqrt(wn / n));
        } else if (s.phase == PH_RESTART) {
            double wn = 0.0, dw = 0.0;
            for (long long r = 0; r < n; ++r) {
                const cd e = csub(Rhs[(size_t)r * B + b], W0[(size_t)r * B + b]);
                wn += cabs2(e);
                dw = std::max(dw, std::sqrt(cabs2(csub(e, W[(size_t)r * B + b]))));
            }
            worst = std::max(worst, std::fabs(nrm2[b] - wn) / wn);
            worst = std::max(worst, dw);
        } else {
            double dw = 0.0;
            for (long long r = 0; r < n; ++r)
                dw = std::max(dw, std::sqrt(cabs2(csub(W0[(size_t)r * B + b], W[(size_t)r * B + b]))));
            worst = std::max(worst, dw);  // untouched: exactly 0
        }
    }
    const bool ok = worst < 1e-12;
    if (!ok) ++failures;
    std::printf("B=%3d m=%2d n=%5lld nwg=%4d passes=%d  worst rel.err %.2e %s\n", B, m, n, nwg, np + 1, worst,
                ok ? "ok" : "FAIL");
}

int main() {
    std::mt19937_64 rng(7);
    const int Bs[] = {5, 64, 112, 256, 300};
    const int ms[] = {20, 50};
    for (int B : Bs)
        for (int m : ms) {
            const long long n = 600;
            const int rw = 256 / std::min(B, 256);
            const int nwg = (int)std::max(1LL, std::min(2048LL, (n + rw * 16 - 1) / (rw * 16)));
            run_case(B, m, n, nwg, rng);
            run_case(B, m, n, 7, rng);  // more workgroups than rows per sweep for the wide batches
        }
    // the launch geometry of the first large device run, which faulted (198k tets:
    // B 112, m 50, 2048 row-sweep workgroups), over 3 rows more than one sweep
    // step (2048 x 2 rows)
    run_case(112, 50, 4099, 2048, rng);
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
