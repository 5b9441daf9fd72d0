// Mixed-radix Stockham line FFTs in LDS for gfx950 (wave64), complex128 or
// complex64 (the mixed-precision matvec).
#include <algorithm>
#include <cmath>

#include "fft.hpp"

namespace scadyn {

// multiply by sign*i
template <typename C>
__device__ inline C mul_si(C z, int sign) {
    C r;
    if (sign < 0) {
        r.x = z.y;
        r.y = -z.x;
    } else {
        r.x = -z.y;
        r.y = z.x;
    }
    return r;
}

template <typename C>
__device__ inline C mkc(double re, double im) {
    C r;
    r.x = (typename Prec<C>::real)re;
    r.y = (typename Prec<C>::real)im;
    return r;
}

template <int R, typename C>
struct Bfly;

template <typename C>
struct Bfly<2, C> {
    __device__ static inline void run(C *v, int, const C *, int) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename C>
struct Bfly<4, C> {
    __device__ static inline void run(C *v, int sign, const C *, int) {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C a2 = cadd(v[1], v[3]), a3 = mul_si(csub(v[1], v[3]), sign);
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    }
};

template <typename C>
struct Bfly<8, C> {
    __device__ static inline void run(C *v, int sign, const C *, int) {
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<4, C>::run(e, sign, nullptr, 0);
        Bfly<4, C>::run(o, sign, nullptr, 0);
        const double h = 0.70710678118654752440;
        // w8^k, k = 0..3 with w8 = (1 + sign i)/sqrt(2)
        C w1 = mkc<C>(h, sign * h), w3 = mkc<C>(-h, sign * h);
        C t0 = o[0], t1 = cmul(o[1], w1), t2 = mul_si(o[2], sign), t3 = cmul(o[3], w3);
        v[0] = cadd(e[0], t0);
        v[4] = csub(e[0], t0);
        v[1] = cadd(e[1], t1);
        v[5] = csub(e[1], t1);
        v[2] = cadd(e[2], t2);
        v[6] = csub(e[2], t2);
        v[3] = cadd(e[3], t3);
        v[7] = csub(e[3], t3);
    }
};

template <typename C>
struct Bfly<3, C> {
    __device__ static inline void run(C *v, int sign, const C *, int) {
        typedef typename Prec<C>::real R;
        const R s3 = (R)0.86602540378443864676;  // sqrt(3)/2
        C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        C m = csub(v[0], cscale(t1, (R)0.5));
        C r = mul_si(cscale(t2, s3), sign);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <typename C>
struct Bfly<5, C> {
    __device__ static inline void run(C *v, int sign, const C *, int) {
        typedef typename Prec<C>::real R;
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        C a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
        C a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
        C b1 = mul_si(cadd(cscale(t3, s1), cscale(t4, s2)), sign);
        C b2 = mul_si(csub(cscale(t3, s2), cscale(t4, s1)), sign);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    }
};

// radix 7 straight from the DFT definition with the length-L twiddle table:
// W_7^q = tw[(L/7) q]
template <typename C>
struct Bfly<7, C> {
    __device__ static inline void run(C *v, int sign, const C *tw, int L) {
        C out[7];
        const int st = L / 7;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            C acc = v[0];
#pragma unroll
            for (int r = 1; r < 7; ++r) {
                C w = tw[st * ((r * k) % 7)];
                if (sign > 0) w.y = -w.y;
                cfma(acc, v[r], w);
            }
            out[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) v[k] = out[k];
    }
};

// One Stockham stage of radix R on ti interleaved columns (LDS -> LDS).
template <int R, typename C>
__device__ void stage(const C *__restrict__ src, C *__restrict__ dst, int L, int Ns, int ti,
                      const C *__restrict__ tw, int sign) {
    const int LR = L / R;
    const int step = L / (Ns * R);
    const int c = threadIdx.x & (ti - 1);
    const int jstep = blockDim.x / ti;
    for (int j = threadIdx.x / ti; j < LR; j += jstep) {
        const int m = j % Ns;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[(j + r * LR) * ti + c];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C w = tw[m * r * step];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        Bfly<R, C>::run(v, sign, tw, L);
        const int idxD = (j / Ns) * Ns * R + m;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[(idxD + r * Ns) * ti + c] = v[r];
    }
}

// Stage for a radix above 8 (only with the reference's 2N lengths): O(R^2),
// inputs re-read from LDS for every output.
template <typename C>
__device__ void stage_generic(const C *__restrict__ src, C *__restrict__ dst, int L, int R, int Ns, int ti,
                              const C *__restrict__ tw, int sign) {
    const int LR = L / R;
    const int step = L / (Ns * R);
    const int c = threadIdx.x & (ti - 1);
    const int jstep = blockDim.x / ti;
    for (int j = threadIdx.x / ti; j < LR; j += jstep) {
        const int m = j % Ns;
        const int idxD = (j / Ns) * Ns * R + m;
        for (int k = 0; k < R; ++k) {
            C acc = Prec<C>::zero();
            for (int r = 0; r < R; ++r) {
                C x = src[(j + r * LR) * ti + c];
                C w = tw[m * r * step];
                C wr = tw[LR * ((r * k) % R)];
                if (sign > 0) {
                    w.y = -w.y;
                    wr.y = -wr.y;
                }
                cfma(acc, cmul(x, w), wr);
            }
            dst[(idxD + k * Ns) * ti + c] = acc;
        }
    }
}

template <typename C>
__device__ C *run_stages(C *a, C *b, const LinePass &p, const C *tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < p.nrad; ++s) {
        const int R = p.rad[s];
        switch (R) {
            case 2: stage<2, C>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 3: stage<3, C>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 4: stage<4, C>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 5: stage<5, C>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 7: stage<7, C>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 8: stage<8, C>(a, b, p.L, Ns, p.ti, tw, sign); break;
            default: stage_generic<C>(a, b, p.L, R, Ns, p.ti, tw, sign); break;
        }
        __syncthreads();
        C *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

template <int MODE, typename C>
__global__ __launch_bounds__(256) void k_line(const C *__restrict__ in, C *__restrict__ out, LinePass p,
                                              const C *__restrict__ twg, int sign, const C *__restrict__ fg,
                                              ConvInfo ci) {
    extern __shared__ char smraw[];
    C *sm = reinterpret_cast<C *>(smraw);
    C *tw = sm;
    C *b0 = sm + p.L;
    C *b1 = b0 + (size_t)p.L * p.ti;
    const long long tiles = (p.inner + p.ti - 1) / p.ti;
    const long long o = blockIdx.x / tiles;
    const long long i0 = (blockIdx.x - o * tiles) * p.ti;
    for (int m = threadIdx.x; m < p.L; m += blockDim.x) tw[m] = twg[m];
    const int c = threadIdx.x & (p.ti - 1);
    const int jstep = blockDim.x / p.ti;
    const long long i = i0 + c;
    const bool col_ok = i < p.inner;
    for (int j = threadIdx.x / p.ti; j < p.L; j += jstep) {
        C v = Prec<C>::zero();
        if (j < p.n_in && col_ok) v = in[((size_t)o * p.n_in + j) * p.inner + i];
        b0[j * p.ti + c] = v;
    }
    __syncthreads();
    C *res;
    if (MODE == 0) {
        res = run_stages<C>(b0, b1, p, tw, sign);
    } else {
        res = run_stages<C>(b0, b1, p, tw, -1);
        // x Fg(kx, ky, kz): the column fixes (kx, ky), the line index is kz
        if (col_ok) {
            long long xy = i / ci.nb;
            int x = (int)(xy % ci.Lx), y = (int)(xy / ci.Lx);
            int kx = min(x, ci.Lx - x), ky = min(y, ci.Ly - y);
            const C *f = fg + (size_t)ky * ci.Hx + kx;
            const size_t zs = (size_t)ci.Hy * ci.Hx;
            for (int j = threadIdx.x / p.ti; j < p.L; j += jstep) {
                int kz = min(j, ci.Lz - j);
                res[j * p.ti + c] = cmul(res[j * p.ti + c], f[kz * zs]);
            }
        }
        __syncthreads();
        C *other = (res == b0) ? b1 : b0;
        res = run_stages<C>(res, other, p, tw, +1);
    }
    if (col_ok)
        for (int j = threadIdx.x / p.ti; j < p.n_out; j += jstep)
            out[((size_t)o * p.n_out + j) * p.inner + i] = res[j * p.ti + c];
}

size_t line_pass_lds(const LinePass &p, size_t elem) {
    // twiddles + one buffer (register kernel) or two buffers (Stockham)
    return elem * ((size_t)p.L + (p.reg ? 1 : 2) * (size_t)p.L * p.ti);
}

// Register-kernel tile width ti: the largest power of two <= want that fits the
// inner extent, the LDS budget and the kernel's thread bound.
static int reg_tile(int L, int rmax, long long inner, size_t lds_budget, size_t elem, int want) {
    const int tmax = reg_max_threads(elem);
    int ti = want;
    while (ti > 1 && (ti > inner || elem * ((size_t)L + (size_t)L * ti) > lds_budget || ti * rmax > tmax)) ti /= 2;
    return ti;
}

bool set_line_tile(LinePass &p, int ti, size_t lds_budget, size_t elem) {
    if (!p.reg) return false;
    const int rmax = std::max(p.rad[0], p.rad[1]);
    const int t = reg_tile(p.L, rmax, p.inner, lds_budget, elem, ti);
    if (t != ti) return false;
    p.ti = ti;
    p.threads = ((ti * rmax + 63) / 64) * 64;
    return true;
}

void make_line_pass(LinePass &p, int L, int n_in, int n_out, long long inner, long long outer, size_t lds_budget,
                    size_t elem, bool allow_reg) {
    p.L = L;
    p.n_in = n_in;
    p.n_out = n_out;
    p.inner = inner;
    p.outer = outer;
    p.reg = 0;
    p.threads = 256;
    // the register kernel indexes inner columns in 32 bits (convolution pass)
    if (allow_reg && inner < (1LL << 31)) {
        // most balanced split L = R1 R2 (R1 >= R2) with both radices in the register set
        int best = 0;
        for (int r1 = 20; r1 >= 2; --r1)
            if (L % r1 == 0 && reg_radix_ok(r1) && reg_radix_ok(L / r1) && r1 >= L / r1) best = r1;
        if (best) {
            const int r1 = best, r2 = L / best, rmax = std::max(r1, r2);
            // tile: 16 columns of 16-byte or 32 of 8-byte elements (256 contiguous bytes per row)
            const int ti = reg_tile(L, rmax, inner, lds_budget, elem, elem >= 16 ? 16 : 32);
            p.reg = 1;
            p.nrad = 2;
            p.rad[0] = r1;
            p.rad[1] = r2;
            p.ti = ti;
            p.threads = ((ti * rmax + 63) / 64) * 64;
            return;
        }
    }
    std::vector<int> r = factor_radices(L);
    if ((int)r.size() > 16) fail(SCADYN_ERR_UNSUPPORTED, "FFT length with too many factors");
    p.nrad = (int)r.size();
    for (int s = 0; s < p.nrad; ++s) p.rad[s] = r[s];
    int ti = 64;
    while (ti > 1 && (ti > inner || elem * ((size_t)L + 2 * (size_t)L * ti) > lds_budget)) ti /= 2;
    p.ti = ti;
}

template <typename C>
static void launch(const LinePass &p, const C *in, C *out, const C *tw, int sign, const C *fg, const ConvInfo &ci,
                   int mode, hipStream_t s) {
    if (p.reg) {
        launch_line_reg<C>(p, in, out, tw, sign, fg, ci, mode, s);
        return;
    }
    long long tiles = (p.inner + p.ti - 1) / p.ti;
    long long blocks = tiles * p.outer;
    if (blocks <= 0) return;
    if (blocks > 0x7fffffffLL) fail(SCADYN_ERR_UNSUPPORTED, "FFT grid too large");
    size_t lds = line_pass_lds(p, sizeof(C));
    if (mode == 0)
        hipLaunchKernelGGL((k_line<0, C>), dim3((unsigned)blocks), dim3(256), lds, s, in, out, p, tw, sign, fg, ci);
    else
        hipLaunchKernelGGL((k_line<1, C>), dim3((unsigned)blocks), dim3(256), lds, s, in, out, p, tw, sign, fg, ci);
    HIP_CHECK(hipGetLastError());
}

template <typename C>
void launch_line_fft(const LinePass &p, const C *in, C *out, const C *tw, int sign, hipStream_t s) {
    launch<C>(p, in, out, tw, sign, nullptr, ConvInfo(), 0, s);
}

template <typename C>
void launch_line_conv(const LinePass &p, const C *in, C *out, const C *tw, const C *fg, const ConvInfo &ci,
                      hipStream_t s) {
    launch<C>(p, in, out, tw, -1, fg, ci, 1, s);
}

template void launch_line_fft<cd>(const LinePass &, const cd *, cd *, const cd *, int, hipStream_t);
template void launch_line_fft<cf>(const LinePass &, const cf *, cf *, const cf *, int, hipStream_t);
template void launch_line_conv<cd>(const LinePass &, const cd *, cd *, const cd *, const cd *, const ConvInfo &,
                                   hipStream_t);
template void launch_line_conv<cf>(const LinePass &, const cf *, cf *, const cf *, const cf *, const ConvInfo &,
                                   hipStream_t);

}  // namespace scadyn
