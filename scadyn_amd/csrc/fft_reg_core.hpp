// Device (and host self-test) core of the two-stage register FFT for line lengths L = R1 * R2 with R1, R2 <= 20.
//
// Each thread owns one radix-R sub-transform per stage and keeps it in VGPRs
// (compile-time unrolled in-register DFTs with compile-time twiddles); the
// only data exchange is one LDS transpose between the two stages (two for the
// fused convolution pass).  Stage 1 reads its operands straight from HBM and
// the last stage writes straight to HBM, both coalesced along the inner
// (contiguous) index: lanes of a wave cover `ti` consecutive inner columns.
//
// Index algebra (Stockham with two stages):
//   stage 1 (radix R1, item j < R2):  x[j + r R2], r < R1  ->  DFT_R1  ->  Y[j R1 + r]
//   stage 2 (radix R2, item j < R1):  Y[j + r R1] * W_L^(j r)  ->  DFT_R2  ->  X[j + r R1]
// The fused convolution continues from stage 2's registers: X[k] * Fg(k), the
// inverse's first stage (radix R2) on the same registers, one LDS exchange and
// the inverse's second stage (radix R1).
#pragma once

#include <type_traits>

#include "fft.hpp"

namespace scadyn {
namespace regfft {

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
#define HD __host__ __device__ __forceinline__

template <int B, int E, typename F>
HD void sfor(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}
#define CV(x) decltype(x)::value

// cos and sin of 2 pi num / den, evaluated by the compiler: reduction to
// [0, pi/4] by exact rational symmetries, then the Taylor series.
struct CS {
    double c, s;
};
constexpr double kTwoPi = 6.28318530717958647692528676655900577;
constexpr CS cs_small(double t) {
    double c = 1.0, s = t, tc = 1.0, ts = t;
    const double t2 = t * t;
    for (int k = 1; k < 12; ++k) {
        tc *= -t2 / (double)((2 * k - 1) * (2 * k));
        ts *= -t2 / (double)((2 * k) * (2 * k + 1));
        c += tc;
        s += ts;
    }
    return {c, s};
}
constexpr CS cs_frac(long long num, long long den) {
    num %= den;
    if (num < 0) num += den;
    bool neg_s = false, neg_c = false, swap = false;
    if (2 * num > den) {  // r -> 1 - r
        num = den - num;
        neg_s = true;
    }
    if (4 * num > den) {  // r -> 1/2 - r
        num = den - 2 * num;
        den *= 2;
        neg_c = true;
    }
    if (8 * num > den) {  // r -> 1/4 - r
        num = den - 4 * num;
        den *= 4;
        swap = true;
    }
    CS v = cs_small(kTwoPi * (double)num / (double)den);
    double c = swap ? v.s : v.c, s = swap ? v.c : v.s;
    return {neg_c ? -c : c, neg_s ? -s : s};
}

HD int imin(int a, int b) { return a < b ? a : b; }

template <typename C>
HD C czero() {
    return Prec<C>::zero();
}

// z * SIGN i
template <int SIGN, typename C>
HD C mul_i(C z) {
    C r;
    if constexpr (SIGN < 0) {
        r.x = z.y;
        r.y = -z.x;
    } else {
        r.x = -z.y;
        r.y = z.x;
    }
    return r;
}

// z * W_N^E, W_N = exp(SIGN 2 pi i / N), E and N compile-time
template <int N, int E, int SIGN, typename C>
HD C twiddle(C z) {
    typedef typename Prec<C>::real R;
    constexpr int e = ((E % N) + N) % N;
    if constexpr (e == 0) {
        return z;
    } else if constexpr (4 * e == N) {
        return mul_i<SIGN>(z);
    } else if constexpr (2 * e == N) {
        C r;
        r.x = -z.x;
        r.y = -z.y;
        return r;
    } else if constexpr (4 * e == 3 * N) {
        return mul_i<-SIGN>(z);
    } else {
        constexpr CS w = cs_frac(e, N);
        const R c = (R)w.c, s = (R)(SIGN * w.s);
        C r;
        r.x = z.x * c - z.y * s;
        r.y = z.x * s + z.y * c;
        return r;
    }
}

// In-register DFT of length N, natural order in and out:
//   v[k] <- sum_n v[n] exp(SIGN 2 pi i n k / N)
template <int N, int SIGN, typename C>
struct RDFT;

template <int SIGN, typename C>
struct RDFT<2, SIGN, C> {
    HD static void run(C *v) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int SIGN, typename C>
struct RDFT<4, SIGN, C> {
    HD static void run(C *v) {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C a2 = cadd(v[1], v[3]), a3 = mul_i<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    }
};

// odd prime N: pairs (n, N - n) share the cosine and sine factors
template <int N, int SIGN, typename C>
HD void dft_odd(C *v) {
    typedef typename Prec<C>::real R;
    constexpr int H = (N - 1) / 2;
    C t[H + 1], u[H + 1];
    sfor<1, H + 1>([&](auto nn) {
        constexpr int n = CV(nn);
        t[n] = cadd(v[n], v[N - n]);
        u[n] = csub(v[n], v[N - n]);
    });
    const C x0 = v[0];
    C sum = x0;
    sfor<1, H + 1>([&](auto nn) { sum = cadd(sum, t[CV(nn)]); });
    sfor<1, H + 1>([&](auto kk) {
        constexpr int k = CV(kk);
        C a = x0, b = czero<C>();
        sfor<1, H + 1>([&](auto nn) {
            constexpr int n = CV(nn);
            constexpr CS w = cs_frac((long long)n * k, N);
            a.x += (R)w.c * t[n].x;
            a.y += (R)w.c * t[n].y;
            b.x += (R)w.s * u[n].x;
            b.y += (R)w.s * u[n].y;
        });
        const C ib = mul_i<SIGN>(b);
        v[k] = cadd(a, ib);
        v[N - k] = csub(a, ib);
    });
    v[0] = sum;
}

template <int SIGN, typename C>
struct RDFT<3, SIGN, C> {
    HD static void run(C *v) { dft_odd<3, SIGN, C>(v); }
};
template <int SIGN, typename C>
struct RDFT<5, SIGN, C> {
    HD static void run(C *v) { dft_odd<5, SIGN, C>(v); }
};
template <int SIGN, typename C>
struct RDFT<7, SIGN, C> {
    HD static void run(C *v) { dft_odd<7, SIGN, C>(v); }
};

// N = A B: A-point DFTs over n1 for each n2 (n = B n1 + n2), twiddle
// W_N^(n2 k1), B-point DFTs over n2 for each k1; output index k1 + A k2.
template <int A, int B, int SIGN, typename C>
HD void dft_comp(C *v) {
    constexpr int N = A * B;
    C t[N];
    sfor<0, B>([&](auto n2c) {
        constexpr int n2 = CV(n2c);
        C u[A];
        sfor<0, A>([&](auto n1) { u[CV(n1)] = v[B * CV(n1) + n2]; });
        RDFT<A, SIGN, C>::run(u);
        sfor<0, A>([&](auto k1c) {
            constexpr int k1 = CV(k1c);
            t[k1 * B + n2] = twiddle<N, n2 * k1, SIGN>(u[k1]);
        });
    });
    sfor<0, A>([&](auto k1c) {
        constexpr int k1 = CV(k1c);
        C u[B];
        sfor<0, B>([&](auto n2) { u[CV(n2)] = t[k1 * B + CV(n2)]; });
        RDFT<B, SIGN, C>::run(u);
        sfor<0, B>([&](auto k2) { v[k1 + A * CV(k2)] = u[CV(k2)]; });
    });
}

#define COMPOSITE(N, A, B)                                                       \
    template <int SIGN, typename C>                                              \
    struct RDFT<N, SIGN, C> {                                                    \
        HD static void run(C *v) { dft_comp<A, B, SIGN, C>(v); } \
    };
COMPOSITE(6, 3, 2)
COMPOSITE(8, 4, 2)
COMPOSITE(9, 3, 3)
COMPOSITE(10, 5, 2)
COMPOSITE(12, 4, 3)
COMPOSITE(14, 7, 2)
COMPOSITE(15, 5, 3)
COMPOSITE(16, 4, 4)
COMPOSITE(18, 9, 2)
COMPOSITE(20, 5, 4)
#undef COMPOSITE

// radices the register path instantiates (must match reg_radix_ok)
#define REG_RADICES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12) X(14) X(15) X(16) X(18) X(20)

struct Ctx {
    long long o;   // outer slab
    long long i;   // inner column
    bool col_ok;
    int c;         // column within the tile
    int j;         // item index within the stage
    int sh;        // log2 of the tile width
    size_t in0;    // element (o, line point 0, i) of the input
    size_t out0;   // element (o, line point 0, i) of the output
};

// Addressing is incremental: every stage walks its R elements with a running
// pointer or index (one add each) instead of a product per element.

// stage 1: radix R, x[j + r (L/R)] from HBM (zero beyond n_in), DFT, LDS
template <int R, int SIGN, typename C>
HD void stage_first(const C *__restrict__ in, C *lds, const LinePass &p, const Ctx &x) {
    const int LR = p.L / R;
    if (x.j >= LR) return;
    C v[R];
    const size_t step = (size_t)LR * p.inner;
    const C *src = in + x.in0 + (size_t)x.j * p.inner;
    int n = x.j;
    sfor<0, R>([&](auto rr) {
        constexpr int r = CV(rr);
        v[r] = (x.col_ok && n < p.n_in) ? *src : czero<C>();
        src += step;
        n += LR;
    });
    RDFT<R, SIGN, C>::run(v);
    const int lb = ((x.j * R) << x.sh) + x.c;
    sfor<0, R>([&](auto rr) {
        constexpr int r = CV(rr);
        lds[lb + (r << x.sh)] = v[r];
    });
}

// stage 2 operand load: LDS[j + r (L/R)] * W_L^(SIGN j r), then DFT
template <int R, int SIGN, typename C>
HD void stage_second(C *v, const C *lds, const C *tw, const LinePass &p, const Ctx &x) {
    const int LR = p.L / R;
    const int ls = LR << x.sh;
    int li = (x.j << x.sh) + x.c;
    sfor<0, R>([&](auto rr) {
        constexpr int r = CV(rr);
        v[r] = lds[li];
        li += ls;
    });
    int ti = 0;  // j r
    sfor<1, R>([&](auto rr) {
        constexpr int r = CV(rr);
        ti += x.j;
        C w = tw[ti];
        if constexpr (SIGN > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
    });
    RDFT<R, SIGN, C>::run(v);
}

// stage 2 to HBM: outputs k = j + r (L/R), cropped to n_out
template <int R, int SIGN, typename C>
HD void stage_last(C *__restrict__ out, const C *lds, const C *tw, const LinePass &p,
                                           const Ctx &x) {
    const int LR = p.L / R;
    if (x.j >= LR) return;
    C v[R];
    stage_second<R, SIGN, C>(v, lds, tw, p, x);
    if (!x.col_ok) return;
    const size_t step = (size_t)LR * p.inner;
    C *dst = out + x.out0 + (size_t)x.j * p.inner;
    int k = x.j;
    sfor<0, R>([&](auto rr) {
        constexpr int r = CV(rr);
        if (k < p.n_out) *dst = v[r];
        dst += step;
        k += LR;
    });
}

// convolution middle, first half: forward stage 2 (radix R), x Fg, inverse
// stage 1 (radix R) on the same registers.  The inner index is below 2^31
// (make_line_pass keeps larger passes off the register kernel), so the grid
// point is found with 32-bit divisions.
template <int R, typename C>
HD void conv_compute(C *v, const C *lds, const C *tw, const C *__restrict__ fg, const ConvInfo &ci, const LinePass &p,
                     const Ctx &x) {
    const int LR = p.L / R;
    if (x.j >= LR) return;
    stage_second<R, -1, C>(v, lds, tw, p, x);
    if (x.col_ok) {
        const unsigned xy = (unsigned)x.i / (unsigned)ci.nb;
        const int gy = (int)(xy / (unsigned)ci.Lx), gx = (int)(xy - (unsigned)gy * (unsigned)ci.Lx);
        const int kx = imin(gx, ci.Lx - gx), ky = imin(gy, ci.Ly - gy);
        const C *f = fg + (size_t)ky * ci.Hx + kx;
        const int zs = ci.Hy * ci.Hx;
        int k = x.j;
        sfor<0, R>([&](auto rr) {
            constexpr int r = CV(rr);
            const int kz = imin(k, ci.Lz - k);
            v[r] = cmul(v[r], f[(size_t)kz * zs]);
            k += LR;
        });
    }
    RDFT<R, +1, C>::run(v);
}

// second half, after a barrier: the inverse stage-1 outputs to LDS
template <int R, typename C>
HD void conv_store(const C *v, C *lds, const LinePass &p, const Ctx &x) {
    if (x.j >= p.L / R) return;
    const int lb = ((x.j * R) << x.sh) + x.c;
    sfor<0, R>([&](auto rr) {
        constexpr int r = CV(rr);
        lds[lb + (r << x.sh)] = v[r];
    });
}

HD int log2_pow2(int t) {
    int s = 0;
    while ((1 << s) < t) ++s;
    return s;
}

HD Ctx make_ctx(const LinePass &p, long long block, int thread) {
    Ctx x;
    x.sh = log2_pow2(p.ti);  // the tile width is a power of two
    const long long tiles = (p.inner + p.ti - 1) >> x.sh;
    x.o = block / tiles;
    const long long i0 = (block - x.o * tiles) << x.sh;
    x.c = thread & (p.ti - 1);
    x.j = thread >> x.sh;
    x.i = i0 + x.c;
    x.col_ok = x.i < p.inner;
    x.in0 = (size_t)x.o * p.n_in * (size_t)p.inner + (size_t)x.i;
    x.out0 = (size_t)x.o * p.n_out * (size_t)p.inner + (size_t)x.i;
    return x;
}

// The phases of one workgroup, separated by barriers.  MODE 0 (one transform
// of sign s): phase 0 = stage 1, phase 1 = stage 2.  MODE 1 (forward, x Fg,
// inverse): phase 0 = forward stage 1, phase 1 = conv_compute (registers v),
// phase 2 = conv_store, phase 3 = inverse stage 2.
template <int MODE, typename C, int RMAX = 20>
HD void line_reg_phase(int phase, const C *__restrict__ in, C *__restrict__ out, C *lds, const C *tw,
                       const C *__restrict__ fg, const ConvInfo &ci, const LinePass &p, const Ctx &x, int sign,
                       C *v) {
    // RMAX: largest radix compiled into this instance (fewer VGPRs for small radices)
    const int R1 = p.rad[0], R2 = p.rad[1];
    const int s1 = (MODE == 0) ? sign : -1;
    if (phase == 0) {
#define FIRST(R)                                          \
    case R:                                               \
        if constexpr (R <= RMAX) {                        \
            if (s1 < 0)                                   \
                stage_first<R, -1, C>(in, lds, p, x);     \
            else                                          \
                stage_first<R, +1, C>(in, lds, p, x);     \
        }                                                 \
        break;
        switch (R1) { REG_RADICES(FIRST) }
#undef FIRST
    } else if (MODE == 0) {
#define LAST(R)                                                 \
    case R:                                                     \
        if constexpr (R <= RMAX) {                              \
            if (s1 < 0)                                         \
                stage_last<R, -1, C>(out, lds, tw, p, x);       \
            else                                                \
                stage_last<R, +1, C>(out, lds, tw, p, x);       \
        }                                                       \
        break;
        switch (R2) { REG_RADICES(LAST) }
#undef LAST
    } else if (phase == 1) {
#define MID(R)                                                                     \
    case R:                                                                        \
        if constexpr (R <= RMAX) conv_compute<R, C>(v, lds, tw, fg, ci, p, x);     \
        break;
        switch (R2) { REG_RADICES(MID) }
#undef MID
    } else if (phase == 2) {
#define STO(R)                                                   \
    case R:                                                      \
        if constexpr (R <= RMAX) conv_store<R, C>(v, lds, p, x); \
        break;
        switch (R2) { REG_RADICES(STO) }
#undef STO
    } else {
#define ILAST(R)                                                            \
    case R:                                                                 \
        if constexpr (R <= RMAX) stage_last<R, +1, C>(out, lds, tw, p, x);  \
        break;
        switch (R1) { REG_RADICES(ILAST) }
#undef ILAST
    }
}

// smallest compiled radix bound covering a pass (instances 12, 14, 16, 18 and
// 20: complex128 takes 106, 124, 139, 158 and 174 VGPRs, i.e. 4, 4, 3, 3 and 2
// waves per SIMD)
HD int reg_rmax(const LinePass &p) {
    const int r = p.rad[0] > p.rad[1] ? p.rad[0] : p.rad[1];
    return r <= 12 ? 12 : (r <= 14 ? 14 : (r <= 16 ? 16 : (r <= 18 ? 18 : 20)));
}

}  // namespace regfft
}  // namespace scadyn
