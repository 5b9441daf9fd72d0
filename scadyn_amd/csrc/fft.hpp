// Batched strided line FFTs in LDS (Stockham autosort, mixed radix).
//
// One pass transforms lines along one axis of a grid laid out as
// [outer][axis][inner] (inner contiguous).  A workgroup owns a tile of `ti`
// consecutive inner columns of one outer slab and all L points of the line,
// so global loads/stores are coalesced along the inner (contiguous) index.
// The input line has n_in valid points (the rest are the zero padding of the
// pruned transform, src/possu.f90:140-204) and only the first n_out outputs
// are stored (the crop of src/possu.f90:225-289 / arr2vec).
#pragma once

#include "internal.hpp"

namespace scadyn {

struct LinePass {
    int L = 0, n_in = 0, n_out = 0;
    int nrad = 0;
    int rad[16] = {0};
    long long inner = 0;  // contiguous inner extent (elements)
    long long outer = 0;  // number of outer slabs
    int ti = 0;           // tile width (inner columns per workgroup)
    int reg = 0;          // 1: two-stage register kernel, rad = {R1, R2}
    int threads = 256;    // workgroup size
};

// Thread bound of the register line-FFT kernel: complex64 instances take at
// most 92 VGPRs, so they launch up to 1024 threads (4 waves per SIMD);
// complex128 instances (up to 174 VGPRs) stay at 512.
inline int reg_max_threads(size_t elem) { return elem <= 8 ? 1024 : 512; }

// Fg lookup for the fused z-pass: inner index i of the [Ly][Lx][nb] layout
// -> frequency (kx, ky); kz is the line index.  Fg holds the even octant.
struct ConvInfo {
    int Lx = 0, Ly = 0, Lz = 0, Hx = 0, Hy = 0;
    long long nb = 0;  // batch columns per grid point
};

// allow_reg: use the two-stage register kernel when L = R1 R2 with both
// radices in the register set
void make_line_pass(LinePass &p, int L, int n_in, int n_out, long long inner, long long outer, size_t lds_budget,
                    size_t elem, bool allow_reg = true);
bool reg_radix_ok(int r);
// register kernel: switch the tile width to ti (a power of two); false when ti
// does not fit the pass (then p is unchanged)
bool set_line_tile(LinePass &p, int ti, size_t lds_budget, size_t elem);
template <typename C>
void launch_line_reg(const LinePass &p, const C *in, C *out, const C *tw, int sign, const C *fg, const ConvInfo &ci,
                     int mode, hipStream_t s);
size_t line_pass_lds(const LinePass &p, size_t elem);

// forward (sign -1) or inverse (sign +1) transform; tw = exp(-2 pi i m / L)
template <typename C>
void launch_line_fft(const LinePass &p, const C *in, C *out, const C *tw, int sign, hipStream_t s);
// forward, multiply by the Fg octant, inverse (the z-pass of the convolution)
template <typename C>
void launch_line_conv(const LinePass &p, const C *in, C *out, const C *tw, const C *fg, const ConvInfo &ci,
                      hipStream_t s);

}  // namespace scadyn
