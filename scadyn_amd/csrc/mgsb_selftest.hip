// Host self-test of the MGS evaluated several basis vectors per pass
// (mgsb_core.hpp).
//
// Replays the launch on the CPU: every workgroup of the row sweep thread by
// thread (256 threads, SB = min(B, 256) slots x 256/SB rows, the kernel's
// mapping), the per-workgroup sums in the kernel's order, the per-slot sums of
// the partials and the coefficient step, over exact-size buffers
// (AddressSanitizer on the host build reports any access outside them).  Slots
// are set up in every phase (Arnoldi at columns 0, BS-1, BS, 2BS-1, 2BS, m-1;
// restarting; idle; done) and the resulting H column, |w|^2 before and after,
// and the orthogonalised w are compared with plain one-vector-at-a-time MGS.
// Block sizes 2, 4, 6, 8; basis stored as complex128 and complex64 (FP64
// arithmetic on the stored values either way), w in complex128 or kept in
// complex64 from the first pass that writes it (1e-5 then); batch widths 5, 64, 112, 256,
// 300; Krylov dimensions 20 and 50; the default number of row-sweep workgroups
// and 7.
//
// It also checks that the matvec input block of every slot state lies inside
// the input array (input_block: m+2 blocks with the complex128 basis and x, m+1
// for the complex64 basis whose restarting slots take their residual
// elsewhere).
//
// Exit status 0 when every case agrees to 1e-12 relative (H entries relative
// to |v_k| |w|, the scale of the sum each is taken from: with a random,
// non-orthogonal basis a coefficient can be far smaller than its terms).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "mgsb_core.hpp"

using namespace scadyn;
using namespace scadyn::mgsb;

static int failures = 0;

static double crel(cd a, cd b, double scale) { return std::sqrt(cabs2(csub(a, b))) / scale; }

// wf: w kept in complex64 from the first pass that writes it (the mixed-
// precision solve), else one complex128 array
template <int BS, typename VT>
static void run_case(int B, int m, long long n, int nwg, std::mt19937_64 &rng, bool wf = false) {
    constexpr int NQ = Blk<BS>::NQ;
    KrylovDims d;
    d.n = n;
    d.B = B;
    d.m = m;
    d.nwg = nwg;
    const size_t nB = (size_t)n * B;
    std::normal_distribution<double> nd;
    auto rnd = [&]() { return cmk(nd(rng), nd(rng)); };
    std::vector<VT> V((size_t)(m + 1) * nB);  // the basis as stored (complex128, complex64 or binary16)
    std::vector<cd> W(nB), Rhs(nB), H((size_t)B * (m + 1) * m);
    std::vector<cf> Wf(wf ? nB : 0, cmkf(1e30f, 1e30f));  // written before it is read
    for (auto &z : V) z = Prec<VT>::from(rnd());
    // unit basis vectors per slot (not orthogonal): each MGS step is then a
    // projection and |w| can only shrink, so errors are bounded by |w0|
    for (int k = 0; k <= m; ++k)
        for (int b = 0; b < B; ++b) {
            double a = 0.0;
            for (long long r = 0; r < n; ++r) a += cabs2(Prec<cd>::from(V[(size_t)k * nB + (size_t)r * B + b]));
            const double f = 1.0 / std::sqrt(a);
            for (long long r = 0; r < n; ++r) {
                VT &z = V[(size_t)k * nB + (size_t)r * B + b];
                z = Prec<VT>::from(cscale(Prec<cd>::from(z), f));
            }
        }
    for (auto &z : W) z = rnd();
    for (auto &z : Rhs) z = rnd();
    for (auto &z : H) z = rnd();
    const std::vector<cd> W0 = W;
    std::vector<cd> pdot((size_t)nwg * NQ * B);
    std::vector<double> pnrm((size_t)nwg * B), pnav((size_t)nwg * B), nrm2(B, -1.0), nav2(B, -1.0);
    std::vector<SlotState> st(B);
    const int cols[] = {0, BS - 1, BS, 2 * BS - 1, 2 * BS, m - 1};
    int K = -1;
    for (int b = 0; b < B; ++b) {
        SlotState s = {};
        const int kind = b % 9;
        if (kind < 6) {
            s.phase = PH_ARNOLDI;
            s.i = std::min(cols[kind], m - 1);
            K = std::max(K, s.i + 1);
        } else if (kind == 6) {
            s.phase = PH_RESTART;
            s.i = m / 2;  // stale column of a previous cycle
            K = std::max(K, 0);
        } else {
            s.phase = kind == 7 ? PH_IDLE : PH_DONE;
            s.i = m - 1;
        }
        st[b] = s;
    }
    // launch_mgsb, replayed
    const int SB = B < 256 ? B : 256, RW = 256 / SB;
    const int ys = (B + SB - 1) / SB;
    const int P = (K + BS - 1) / BS;
    std::vector<cd> acc((size_t)256 * NQ);
    std::vector<double> snrm(256), snav(256);
    for (int p = 0; p <= P; ++p) {
        for (int by = 0; by < ys; ++by)
            for (int bx = 0; bx < nwg; ++bx) {
                for (int t = 0; t < 256; ++t) {
                    const int sl = t % SB, rl = t / SB, b = by * SB + sl;
                    const bool ok = rl < RW && b < B;
                    for (int q = 0; q < NQ; ++q) acc[(size_t)t * NQ + q] = cmk(0.0, 0.0);
                    snrm[t] = snav[t] = 0.0;
                    const long long r0 = (long long)bx * RW + rl, rs = (long long)nwg * RW;
                    cd *ac = &acc[(size_t)t * NQ];
                    if (ok && !wf)
                        sweep_rows<BS, VT, cd, cd>(d, p, W.data(), W.data(), V.data(), Rhs.data(), st.data(),
                                                   H.data(), b, r0, rs, ac, snrm[t], snav[t]);
                    else if (ok && p <= 1)
                        sweep_rows<BS, VT, cd, cf>(d, p, W.data(), Wf.data(), V.data(), Rhs.data(), st.data(),
                                                   H.data(), b, r0, rs, ac, snrm[t], snav[t]);
                    else if (ok)
                        sweep_rows<BS, VT, cf, cf>(d, p, Wf.data(), Wf.data(), V.data(), Rhs.data(), st.data(),
                                                   H.data(), b, r0, rs, ac, snrm[t], snav[t]);
                }
                for (int t = 0; t < 256; ++t) {
                    const int sl = t % SB, rl = t / SB, b = by * SB + sl;
                    if (!(rl < RW && b < B) || rl != 0) continue;
                    double a = 0.0, c = 0.0;
                    for (int r = 0; r < RW; ++r) {
                        a += snrm[r * SB + sl];
                        c += snav[r * SB + sl];
                    }
                    pnrm.at((size_t)bx * B + b) = a;
                    pnav.at((size_t)bx * B + b) = c;
                    for (int q = 0; q < NQ; ++q) {
                        cd z = cmk(0.0, 0.0);
                        for (int r = 0; r < RW; ++r) z = cadd(z, acc[(size_t)(r * SB + sl) * NQ + q]);
                        pdot.at(((size_t)bx * NQ + q) * B + b) = z;
                    }
                }
            }
        for (int b = 0; b < B; ++b) {
            const PassRole r = pass_role<BS>(st[b], p);
            if (!r.rst && !r.arn) continue;
            double x1 = 0.0, x2 = 0.0;
            cd red[NQ];
            for (int q = 0; q < NQ; ++q) red[q] = cmk(0.0, 0.0);
            for (int w = 0; w < nwg; ++w) {
                if (r.rst || r.nd == 0) x1 += pnrm.at((size_t)w * B + b);
                if (r.arn && p == 0) x2 += pnav.at((size_t)w * B + b);
                if (r.nd > 0)
                    for (int q = 0; q < NQ; ++q) red[q] = cadd(red[q], pdot.at((size_t)w * NQ * B + (size_t)q * B + b));
            }
            slot_coef<BS>(d, p, st.data(), red, x1, x2, H.data(), nrm2.data(), nav2.data(), b);
        }
    }
    // plain MGS per slot; the orthogonalised w as the solve leaves it
    auto wout = [&](long long r, int b) { return wf ? Prec<cd>::from(Wf[(size_t)r * B + b]) : W[(size_t)r * B + b]; };
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
                auto vk = [&](long long r) { return Prec<cd>::from(V[(size_t)k * nB + (size_t)r * B + b]); };
                double vn = 0.0, wn0 = 0.0;  // scale of the sum: |v_k| |w| (h itself may cancel)
                for (long long r = 0; r < n; ++r) {
                    vn += cabs2(vk(r));
                    wn0 += cabs2(w[r]);
                }
                for (long long r = 0; r < n; ++r) h = cadd(h, cmulc(vk(r), w[r]));
                for (long long r = 0; r < n; ++r) w[r] = csub(w[r], cmul(h, vk(r)));
                const cd got = H[(size_t)b * (m + 1) * m + (size_t)(m + 1) * s.i + k];
                worst = std::max(worst, crel(got, h, std::sqrt(vn * wn0) + 1e-300));
            }
            double wn = 0.0, dw = 0.0;
            for (long long r = 0; r < n; ++r) {
                wn += cabs2(w[r]);
                dw = std::max(dw, std::sqrt(cabs2(csub(w[r], wout(r, b)))));
            }
            // relative to the norm before orthogonalisation (the scale of the
            // rounding; w itself can shrink by orders of magnitude)
            worst = std::max(worst, std::fabs(nrm2[b] - wn) / w0);
            worst = std::max(worst, std::fabs(nav2[b] - w0) / w0);
            worst = std::max(worst, dw / std::sqrt(w0 / n));
        } else if (s.phase == PH_RESTART) {
            double wn = 0.0, dw = 0.0;
            for (long long r = 0; r < n; ++r) {
                const cd e = csub(Rhs[(size_t)r * B + b], W0[(size_t)r * B + b]);
                wn += cabs2(e);
                dw = std::max(dw, std::sqrt(cabs2(csub(e, wout(r, b)))));
            }
            worst = std::max(worst, std::fabs(nrm2[b] - wn) / wn);
            worst = std::max(worst, dw / std::sqrt(wn / n));
        } else {
            double dw = 0.0;
            for (long long r = 0; r < n; ++r)
                dw = std::max(dw, std::sqrt(cabs2(csub(W0[(size_t)r * B + b], W[(size_t)r * B + b]))));
            worst = std::max(worst, dw);  // untouched: exactly 0
            if (nrm2[b] != -1.0 || nav2[b] != -1.0) worst = 1.0;  // norms of idle slots untouched
        }
    }
    // complex64 w: rounded at every pass that writes it (relative 6e-8 each)
    const double tol = wf ? 1e-5 : 1e-12;
    const bool ok = worst < tol;
    if (!ok) ++failures;
    std::printf("BS=%d %s%s B=%3d m=%2d n=%5lld nwg=%4d passes=%2d  worst rel.err %.2e %s\n", BS,
                sizeof(VT) == 16 ? "c128" : sizeof(VT) == 8 ? "c64 " : "f16 ", wf ? " w64" : "    ", B, m, n, nwg, P + 1, worst,
                ok ? "ok" : "FAIL");
}

// matvec input blocks of every slot state fall inside the input array: m+2
// blocks (complex128 V with x) in the FP64 solve, m+1 (complex64 basis) when
// restarting slots take their residual elsewhere (the mixed-precision solve)
static void check_input_blocks() {
    int bad = 0;
    for (int m : {1, 2, 10, 50})
        for (int rest = 0; rest < 2; ++rest) {
            const int nblk = rest ? m + 1 : m + 2;
            for (int ph : {PH_IDLE, PH_RESTART, PH_ARNOLDI, PH_DONE})
                for (int i = 0; i < m; ++i) {
                    SlotState s = {};
                    s.phase = ph;
                    s.i = i;
                    int a = -1;
                    const long long blk = input_block(s, m, rest, a);
                    if (blk < 0 || blk >= nblk) ++bad;
                    if (a != (ph == PH_ARNOLDI || (ph == PH_RESTART && !rest) ? 1 : 0)) ++bad;
                }
        }
    std::printf("matvec input blocks: %s\n", bad ? "FAIL" : "ok");
    if (bad) ++failures;
}

int main() {
    check_input_blocks();
    std::mt19937_64 rng(11);
    const int Bs[] = {5, 64, 112, 256, 300};
    const int ms[] = {20, 50};
    for (int B : Bs)
        for (int m : ms) {
            const long long n = 600;
            const int rw = 256 / std::min(B, 256);
            const int nwg = (int)std::max(1LL, std::min(2048LL, (n + rw * 16 - 1) / (rw * 16)));
            run_case<2, cd>(B, m, n, nwg, rng);
            run_case<4, cd>(B, m, n, nwg, rng);
            run_case<8, cd>(B, m, n, nwg, rng);
            run_case<4, cd>(B, m, n, 7, rng);
            run_case<6, cd>(B, m, n, nwg, rng);
            run_case<4, cf>(B, m, n, nwg, rng);  // complex64 basis of the mixed-precision solve
            run_case<8, cf>(B, m, n, 7, rng);
            run_case<4, cf>(B, m, n, nwg, rng, true);  // ... with complex64 w
            run_case<2, cf>(B, m, n, 7, rng, true);
            run_case<4, ch>(B, m, n, nwg, rng, true);  // binary16 basis with complex64 w
            run_case<8, ch>(B, m, n, 7, rng, true);
        }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
