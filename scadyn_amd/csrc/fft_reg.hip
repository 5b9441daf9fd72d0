// Two-stage register line-FFT kernel (the algebra and the per-thread stages
// are in fft_reg_core.hpp, shared with the host self-test).
#include "fft_reg_core.hpp"

namespace scadyn {
namespace regfft {

template <int MODE, typename C, int RMAX>
__global__ __launch_bounds__(sizeof(C) <= 8 ? 1024 : 512) void k_line_reg(const C *__restrict__ in, C *__restrict__ out, LinePass p,
                                                  const C *__restrict__ twg, int sign, const C *__restrict__ fg,
                                                  ConvInfo ci) {
    extern __shared__ char smraw[];
    C *tw = reinterpret_cast<C *>(smraw);
    C *lds = tw + p.L;
    for (int m = threadIdx.x; m < p.L; m += blockDim.x) tw[m] = twg[m];
    // neighbouring tiles of a line row share cache lines (rows need not be
    // 128-byte aligned): give them to one XCD
    const Ctx x = make_ctx(p, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x);
    C v[RMAX];
    line_reg_phase<MODE, C, RMAX>(0, in, out, lds, tw, fg, ci, p, x, sign, v);
    __syncthreads();
    line_reg_phase<MODE, C, RMAX>(1, in, out, lds, tw, fg, ci, p, x, sign, v);
    if (MODE == 1) {
        __syncthreads();  // every stage-2 read of LDS is done
        line_reg_phase<MODE, C, RMAX>(2, in, out, lds, tw, fg, ci, p, x, sign, v);
        __syncthreads();
        line_reg_phase<MODE, C, RMAX>(3, in, out, lds, tw, fg, ci, p, x, sign, v);
    }
}

}  // namespace regfft

using namespace regfft;

bool reg_radix_ok(int r) {
    switch (r) {
#define OK(R) case R:
        REG_RADICES(OK)
#undef OK
        return true;
        default: return false;
    }
}

template <typename C>
void launch_line_reg(const LinePass &p, const C *in, C *out, const C *tw, int sign, const C *fg, const ConvInfo &ci,
                     int mode, hipStream_t s) {
    const long long tiles = (p.inner + p.ti - 1) / p.ti;
    const long long blocks = tiles * p.outer;
    if (blocks <= 0) return;
    if (blocks > 0x7fffffffLL) fail(SCADYN_ERR_UNSUPPORTED, "FFT grid too large");
    const size_t lds = line_pass_lds(p, sizeof(C));
    if (p.threads > reg_max_threads(sizeof(C))) fail(SCADYN_ERR_UNSUPPORTED, "FFT tile exceeds the thread bound");
    const dim3 grid((unsigned)blocks), block(p.threads);
    const int rm = reg_rmax(p);
    if (lds > 65536) {  // tiles beyond 64 KiB of LDS (MI355X allows 160 KiB per workgroup)
#define ATTR(M, R)                                                                                        \
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_line_reg<M, C, R>),                  \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (mode == 0) {
            if (rm == 12) { ATTR(0, 12) } else if (rm == 14) { ATTR(0, 14) } else if (rm == 16) { ATTR(0, 16) }
            else if (rm == 18) { ATTR(0, 18) } else { ATTR(0, 20) }
        } else {
            if (rm == 12) { ATTR(1, 12) } else if (rm == 14) { ATTR(1, 14) } else if (rm == 16) { ATTR(1, 16) }
            else if (rm == 18) { ATTR(1, 18) } else { ATTR(1, 20) }
        }
#undef ATTR
    }
    if (mode == 0) {
        if (rm == 12)
            hipLaunchKernelGGL((k_line_reg<0, C, 12>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else if (rm == 14)
            hipLaunchKernelGGL((k_line_reg<0, C, 14>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else if (rm == 16)
            hipLaunchKernelGGL((k_line_reg<0, C, 16>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else if (rm == 18)
            hipLaunchKernelGGL((k_line_reg<0, C, 18>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else
            hipLaunchKernelGGL((k_line_reg<0, C, 20>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
    } else {
        if (rm == 12)
            hipLaunchKernelGGL((k_line_reg<1, C, 12>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else if (rm == 14)
            hipLaunchKernelGGL((k_line_reg<1, C, 14>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else if (rm == 16)
            hipLaunchKernelGGL((k_line_reg<1, C, 16>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else if (rm == 18)
            hipLaunchKernelGGL((k_line_reg<1, C, 18>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
        else
            hipLaunchKernelGGL((k_line_reg<1, C, 20>), grid, block, lds, s, in, out, p, tw, sign, fg, ci);
    }
    HIP_CHECK(hipGetLastError());
}

template void launch_line_reg<cd>(const LinePass &, const cd *, cd *, const cd *, int, const cd *, const ConvInfo &,
                                  int, hipStream_t);
template void launch_line_reg<cf>(const LinePass &, const cf *, cf *, const cf *, int, const cf *, const ConvInfo &,
                                  int, hipStream_t);

}  // namespace scadyn
