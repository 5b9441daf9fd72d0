// Host self-test of the two-stage register line FFT.
//
// Runs the kernel's own per-thread stage code (fft_reg_core.hpp) on the CPU,
// one workgroup at a time with the barriers as phase boundaries, and compares
// every output with a direct long-double DFT.  Covers every register split
// (R1, R2) the pass planner can choose, zero-padded inputs (n_in < L),
// cropped outputs (n_out < L), partial column tiles, forward and inverse
// transforms, the fused convolution pass, and complex128 / complex64.
// The pass geometry comes from the library's own planner (make_line_pass).
//
// Exit status 0 when every case is within tolerance.
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "fft_reg_core.hpp"

using namespace scadyn;
using namespace scadyn::regfft;
typedef std::complex<long double> cl;

template <typename C>
static cl to_cl(C a) {
    return cl((long double)a.x, (long double)a.y);
}
template <typename C>
static C from_cl(cl a) {
    C r;
    r.x = (typename Prec<C>::real)a.real();
    r.y = (typename Prec<C>::real)a.imag();
    return r;
}

static const long double PI_L = 3.141592653589793238462643383279502884L;

template <int MODE, typename C>
static void emulate(const LinePass &p, const C *in, C *out, const C *twg, int sign, const C *fg, const ConvInfo &ci) {
    const long long tiles = (p.inner + p.ti - 1) / p.ti, blocks = tiles * p.outer;
    std::vector<C> lds((size_t)p.L * p.ti), tw(twg, twg + p.L);
    std::vector<std::array<C, 20>> regs(p.threads);
    const int nph = MODE == 0 ? 2 : 4;
    const int rm = reg_rmax(p);  // the instance the launcher picks
    for (long long bi = 0; bi < blocks; ++bi) {
        const long long b = xcd_remap(bi, blocks);  // the kernel's workgroup -> tile order
        for (int ph = 0; ph < nph; ++ph)
            for (int t = 0; t < p.threads; ++t) {
                if (rm == 12)
                    line_reg_phase<MODE, C, 12>(ph, in, out, lds.data(), tw.data(), fg, ci, p, make_ctx(p, b, t),
                                                sign, regs[t].data());
                else if (rm == 14)
                    line_reg_phase<MODE, C, 14>(ph, in, out, lds.data(), tw.data(), fg, ci, p, make_ctx(p, b, t),
                                                sign, regs[t].data());
                else if (rm == 16)
                    line_reg_phase<MODE, C, 16>(ph, in, out, lds.data(), tw.data(), fg, ci, p, make_ctx(p, b, t),
                                                sign, regs[t].data());
                else if (rm == 18)
                    line_reg_phase<MODE, C, 18>(ph, in, out, lds.data(), tw.data(), fg, ci, p, make_ctx(p, b, t),
                                                sign, regs[t].data());
                else
                    line_reg_phase<MODE, C, 20>(ph, in, out, lds.data(), tw.data(), fg, ci, p, make_ctx(p, b, t),
                                                sign, regs[t].data());
            }
    }
}

template <typename C>
static std::vector<C> twiddles(int L) {
    std::vector<C> t(L);
    for (int m = 0; m < L; ++m) t[m] = from_cl<C>(std::polar(1.0L, -2.0L * PI_L * m / L));
    return t;
}

// direct DFT of one zero-padded line: X[k] = sum_{n < n_in} x[n] exp(sign 2 pi i n k / L)
static std::vector<cl> dft(const std::vector<cl> &x, int L, int sign) {
    std::vector<cl> X(L), w(L);
    for (int m = 0; m < L; ++m) w[m] = std::polar(1.0L, sign * 2.0L * PI_L * m / L);
    for (int k = 0; k < L; ++k) {
        cl acc = 0;
        for (size_t n = 0; n < x.size(); ++n) acc += x[n] * w[(n * (size_t)k) % L];
        X[k] = acc;
    }
    return X;
}

static int failures = 0;

template <typename C>
static void check(const char *what, int L, const LinePass &p, const std::vector<C> &got, const std::vector<cl> &ref,
                  double tol) {
    long double num = 0, den = 0;
    for (size_t e = 0; e < ref.size(); ++e) {
        num += std::norm(to_cl(got[e]) - ref[e]);
        den += std::norm(ref[e]);
    }
    const double err = (double)std::sqrt(num / den);
    const bool ok = err < tol;
    if (!ok) ++failures;
    std::printf("%-5s %-4s L=%3d R1=%2d R2=%2d ti=%2d thr=%3d  rel.err %.2e %s\n", what,
                sizeof(C) == 16 ? "c128" : "c64", L, p.rad[0], p.rad[1], p.ti, p.threads, err, ok ? "ok" : "FAIL");
}

template <typename C>
static void line_case(int L, int sign, std::mt19937_64 &rng, long long inner = 19, size_t lds = 64 * 1024) {
    const int n_in = sign < 0 ? L / 2 + 1 : L, n_out = sign < 0 ? L : L / 2 + 1;
    const long long outer = 2;
    LinePass p;
    make_line_pass(p, L, n_in, n_out, inner, outer, lds, sizeof(C), true);
    if (!p.reg) {
        std::printf("L=%d not on the register path\n", L);
        ++failures;
        return;
    }
    std::normal_distribution<double> nd;
    std::vector<C> in((size_t)outer * n_in * inner), out((size_t)outer * n_out * inner);
    for (auto &z : in) z = from_cl<C>(cl(nd(rng), nd(rng)));
    std::vector<C> tw = twiddles<C>(L);
    emulate<0, C>(p, in.data(), out.data(), tw.data(), sign, nullptr, ConvInfo());
    std::vector<cl> ref(out.size());
    for (long long o = 0; o < outer; ++o)
        for (long long i = 0; i < inner; ++i) {
            std::vector<cl> x(n_in);
            for (int n = 0; n < n_in; ++n) x[n] = to_cl(in[((size_t)o * n_in + n) * inner + i]);
            std::vector<cl> X = dft(x, L, sign);
            for (int k = 0; k < n_out; ++k) ref[((size_t)o * n_out + k) * inner + i] = X[k];
        }
    check<C>(sign < 0 ? "fwd" : "inv", L, p, out, ref, sizeof(C) == 16 ? 1e-14 : 2e-6);
}

// fused z-pass: lines along z of a [Lz][Ly][Lx][nb] grid, zero-padded to n_in = Nz,
// forward, x Fg(min(kx, Lx-kx), min(ky, Ly-ky), min(kz, Lz-kz)), inverse, cropped to Nz
template <typename C>
static void conv_case(int Lz, std::mt19937_64 &rng, size_t lds = 64 * 1024) {
    const int Nz = Lz / 2, Lx = 4, Ly = 3, nb = 3;
    const int Hx = Lx / 2 + 1, Hy = Ly / 2 + 1, Hz = Lz / 2 + 1;
    const long long inner = (long long)Lx * Ly * nb;
    LinePass p;
    make_line_pass(p, Lz, Nz, Nz, inner, 1, lds, sizeof(C), true);
    ConvInfo ci;
    ci.Lx = Lx;
    ci.Ly = Ly;
    ci.Lz = Lz;
    ci.Hx = Hx;
    ci.Hy = Hy;
    ci.nb = nb;
    std::normal_distribution<double> nd;
    std::vector<C> in((size_t)Nz * inner), out((size_t)Nz * inner), fg((size_t)Hx * Hy * Hz);
    for (auto &z : in) z = from_cl<C>(cl(nd(rng), nd(rng)));
    for (auto &z : fg) z = from_cl<C>(cl(nd(rng), nd(rng)));
    std::vector<C> tw = twiddles<C>(Lz);
    emulate<1, C>(p, in.data(), out.data(), tw.data(), -1, fg.data(), ci);
    std::vector<cl> ref(out.size());
    for (long long i = 0; i < inner; ++i) {
        const long long xy = i / nb;
        const int gx = (int)(xy % Lx), gy = (int)(xy / Lx);
        const int kx = std::min(gx, Lx - gx), ky = std::min(gy, Ly - gy);
        std::vector<cl> x(Nz);
        for (int n = 0; n < Nz; ++n) x[n] = to_cl(in[(size_t)n * inner + i]);
        std::vector<cl> X = dft(x, Lz, -1);
        for (int k = 0; k < Lz; ++k) X[k] *= to_cl(fg[((size_t)std::min(k, Lz - k) * Hy + ky) * Hx + kx]);
        std::vector<cl> y = dft(X, Lz, +1);
        for (int n = 0; n < Nz; ++n) ref[(size_t)n * inner + i] = y[n];
    }
    check<C>("conv", Lz, p, out, ref, sizeof(C) == 16 ? 1e-14 : 2e-6);
}

int main() {
    std::mt19937_64 rng(12345);
    const int rad[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20};
    std::vector<int> lengths;
    for (int a : rad)
        for (int b : rad) {
            const int L = a * b;
            bool seen = false;
            for (int l : lengths) seen |= l == L;
            if (!seen) lengths.push_back(L);
        }
    for (int L : lengths) {
        line_case<cd>(L, -1, rng);
        line_case<cd>(L, +1, rng);
        line_case<cf>(L, -1, rng);
        conv_case<cd>(L, rng);
        conv_case<cf>(L, rng);
        // the widest tiles (32 complex64 columns, up to 1024 threads; 16 complex128
        // columns) within the 160 KiB of LDS a workgroup may hold, partial last tile
        line_case<cf>(L, -1, rng, 37, 160 * 1024);
        line_case<cd>(L, +1, rng, 21, 160 * 1024);
        conv_case<cf>(L, rng, 160 * 1024);
        conv_case<cd>(L, rng, 160 * 1024);
    }
    std::printf("%zu lengths, %d failures\n", lengths.size(), failures);
    return failures ? 1 : 0;
}
