// Mixed-radix Stockham line FFTs in LDS for gfx950 (wave64).
#include <algorithm>
#include <cmath>

#include "fft.hpp"

namespace scadyn {

// multiply by sign*i
__device__ inline cd mul_si(cd z, int sign) { return sign < 0 ? cmk(z.y, -z.x) : cmk(-z.y, z.x); }

template <int R>
__device__ inline void butterfly(cd *v, int sign, const cd *tw, int L);

template <>
__device__ inline void butterfly<2>(cd *v, int, const cd *, int) {
    cd a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ inline void butterfly<4>(cd *v, int sign, const cd *, int) {
    cd a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cd a2 = cadd(v[1], v[3]), a3 = mul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <>
__device__ inline void butterfly<8>(cd *v, int sign, const cd *, int) {
    cd e[4] = {v[0], v[2], v[4], v[6]};
    cd o[4] = {v[1], v[3], v[5], v[7]};
    butterfly<4>(e, sign, nullptr, 0);
    butterfly<4>(o, sign, nullptr, 0);
    const double h = 0.70710678118654752440;
    // w8^k, k = 0..3 with w8 = (1 + sign i)/sqrt(2)
    cd w1 = cmk(h, sign * h), w3 = cmk(-h, sign * h);
    cd t0 = o[0], t1 = cmul(o[1], w1), t2 = mul_si(o[2], sign), t3 = cmul(o[3], w3);
    v[0] = cadd(e[0], t0);
    v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
}

template <>
__device__ inline void butterfly<3>(cd *v, int sign, const cd *, int) {
    const double s3 = 0.86602540378443864676;  // sqrt(3)/2
    cd t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    cd m = csub(v[0], cscale(t1, 0.5));
    cd r = mul_si(cscale(t2, s3), sign);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

template <>
__device__ inline void butterfly<5>(cd *v, int sign, const cd *, int) {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    cd t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    cd t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    cd a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    cd a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    cd b1 = mul_si(cadd(cscale(t3, s1), cscale(t4, s2)), sign);
    cd b2 = mul_si(csub(cscale(t3, s2), cscale(t4, s1)), sign);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

// radix 7 (and any small odd radix) straight from the DFT definition with
// the length-L twiddle table: W_R^q = tw[(L/R) q]
template <>
__device__ inline void butterfly<7>(cd *v, int sign, const cd *tw, int L) {
    cd out[7];
    const int st = L / 7;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        cd acc = v[0];
#pragma unroll
        for (int r = 1; r < 7; ++r) {
            cd w = tw[st * ((r * k) % 7)];
            if (sign > 0) w.y = -w.y;
            cfma(acc, v[r], w);
        }
        out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = out[k];
}

// One Stockham stage of radix R on ti interleaved columns (LDS -> LDS).
template <int R>
__device__ void stage(const cd *__restrict__ src, cd *__restrict__ dst, int L, int Ns, int ti,
                      const cd *__restrict__ tw, int sign) {
    const int LR = L / R;
    const int step = L / (Ns * R);
    const int c = threadIdx.x & (ti - 1);
    const int jstep = blockDim.x / ti;
    for (int j = threadIdx.x / ti; j < LR; j += jstep) {
        const int m = j % Ns;
        cd v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[(j + r * LR) * ti + c];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cd w = tw[m * r * step];
                if (sign > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        butterfly<R>(v, sign, tw, L);
        const int idxD = (j / Ns) * Ns * R + m;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[(idxD + r * Ns) * ti + c] = v[r];
    }
}

// Stage for a radix above 8 (only with the reference's 2N lengths): O(R^2),
// inputs re-read from LDS for every output.
__device__ void stage_generic(const cd *__restrict__ src, cd *__restrict__ dst, int L, int R,
                              int Ns, int ti, const cd *__restrict__ tw, int sign) {
    const int LR = L / R;
    const int step = L / (Ns * R);
    const int c = threadIdx.x & (ti - 1);
    const int jstep = blockDim.x / ti;
    for (int j = threadIdx.x / ti; j < LR; j += jstep) {
        const int m = j % Ns;
        const int idxD = (j / Ns) * Ns * R + m;
        for (int k = 0; k < R; ++k) {
            cd acc = cmk(0, 0);
            for (int r = 0; r < R; ++r) {
                cd x = src[(j + r * LR) * ti + c];
                cd w = tw[m * r * step];
                cd wr = tw[LR * ((r * k) % R)];
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

__device__ cd *run_stages(cd *a, cd *b, const LinePass &p, const cd *tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < p.nrad; ++s) {
        const int R = p.rad[s];
        switch (R) {
            case 2: stage<2>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 3: stage<3>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 4: stage<4>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 5: stage<5>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 7: stage<7>(a, b, p.L, Ns, p.ti, tw, sign); break;
            case 8: stage<8>(a, b, p.L, Ns, p.ti, tw, sign); break;
            default: stage_generic(a, b, p.L, R, Ns, p.ti, tw, sign); break;
        }
        __syncthreads();
        cd *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_line(const cd *__restrict__ in, cd *__restrict__ out,
                                              LinePass p, const cd *__restrict__ twg, int sign,
                                              ConvInfo ci) {
    extern __shared__ cd sm[];
    cd *tw = sm;
    cd *b0 = sm + p.L;
    cd *b1 = b0 + (size_t)p.L * p.ti;
    const long long tiles = (p.inner + p.ti - 1) / p.ti;
    const long long o = blockIdx.x / tiles;
    const long long i0 = (blockIdx.x - o * tiles) * p.ti;
    for (int m = threadIdx.x; m < p.L; m += blockDim.x) tw[m] = twg[m];
    const int c = threadIdx.x & (p.ti - 1);
    const int jstep = blockDim.x / p.ti;
    const long long i = i0 + c;
    const bool col_ok = i < p.inner;
    for (int j = threadIdx.x / p.ti; j < p.L; j += jstep) {
        cd v = cmk(0, 0);
        if (j < p.n_in && col_ok) v = in[((size_t)o * p.n_in + j) * p.inner + i];
        b0[j * p.ti + c] = v;
    }
    __syncthreads();
    cd *res;
    if (MODE == 0) {
        res = run_stages(b0, b1, p, tw, sign);
    } else {
        res = run_stages(b0, b1, p, tw, -1);
        // x Fg(kx, ky, kz): the column fixes (kx, ky), the line index is kz
        if (col_ok) {
            long long xy = i / ci.nb;
            int x = (int)(xy % ci.Lx), y = (int)(xy / ci.Lx);
            int kx = min(x, ci.Lx - x), ky = min(y, ci.Ly - y);
            const cd *f = ci.fg + (size_t)ky * ci.Hx + kx;
            const size_t zs = (size_t)ci.Hy * ci.Hx;
            for (int j = threadIdx.x / p.ti; j < p.L; j += jstep) {
                int kz = min(j, ci.Lz - j);
                res[j * p.ti + c] = cmul(res[j * p.ti + c], f[kz * zs]);
            }
        }
        __syncthreads();
        cd *other = (res == b0) ? b1 : b0;
        res = run_stages(res, other, p, tw, +1);
    }
    if (col_ok)
        for (int j = threadIdx.x / p.ti; j < p.n_out; j += jstep)
            out[((size_t)o * p.n_out + j) * p.inner + i] = res[j * p.ti + c];
}

size_t line_pass_lds(const LinePass &p) { return sizeof(cd) * ((size_t)p.L + 2 * (size_t)p.L * p.ti); }

void make_line_pass(LinePass &p, int L, int n_in, int n_out, long long inner, long long outer,
                    size_t lds_budget) {
    p.L = L;
    p.n_in = n_in;
    p.n_out = n_out;
    p.inner = inner;
    p.outer = outer;
    std::vector<int> r = factor_radices(L);
    if ((int)r.size() > 16) fail(SCADYN_ERR_UNSUPPORTED, "FFT length with too many factors");
    p.nrad = (int)r.size();
    for (int s = 0; s < p.nrad; ++s) p.rad[s] = r[s];
    int ti = 64;
    while (ti > 1 && (ti > inner || sizeof(cd) * ((size_t)L + 2 * (size_t)L * ti) > lds_budget)) ti /= 2;
    p.ti = ti;
}

static void launch(const LinePass &p, const cd *in, cd *out, const cd *tw, int sign,
                   const ConvInfo &ci, int mode, hipStream_t s) {
    long long tiles = (p.inner + p.ti - 1) / p.ti;
    long long blocks = tiles * p.outer;
    if (blocks <= 0) return;
    if (blocks > 0x7fffffffLL) fail(SCADYN_ERR_UNSUPPORTED, "FFT grid too large");
    size_t lds = line_pass_lds(p);
    if (mode == 0)
        hipLaunchKernelGGL(k_line<0>, dim3((unsigned)blocks), dim3(256), lds, s, in, out, p, tw, sign, ci);
    else
        hipLaunchKernelGGL(k_line<1>, dim3((unsigned)blocks), dim3(256), lds, s, in, out, p, tw, sign, ci);
    HIP_CHECK(hipGetLastError());
}

void launch_line_fft(const LinePass &p, const cd *in, cd *out, const cd *tw, int sign, hipStream_t s) {
    launch(p, in, out, tw, sign, ConvInfo(), 0, s);
}

void launch_line_conv(const LinePass &p, const cd *in, cd *out, const cd *tw, const ConvInfo &ci,
                      hipStream_t s) {
    launch(p, in, out, tw, -1, ci, 1, s);
}

}  // namespace scadyn
