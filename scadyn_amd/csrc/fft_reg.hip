// Two-stage register line-FFT kernel (the algebra and the per-thread stages
// are in fft_reg_core.hpp, shared with the host self-test).
#include "fft_reg_core.hpp"

namespace scadyn {
namespace regfft {

template <int MODE, typename C, int RMAX>
__global__ __launch_bounds__(512) void k_line_reg(const C *__restrict__ in, C *__restrict__ out, LinePass p,
                                                  const C *__restrict__ twg, int sign, const C *__restrict__ fg,
                                                  ConvInfo ci) {
    extern __shared__ char smraw[];
    C *tw = reinterpret_cast<C *>(smraw);
    C *lds = tw + p.L;
    for (int m = threadIdx.x; m < p.L; m += blockDim.x) tw[m] = twg[m];
    const Ctx x = make_ctx(p, blockIdx.x, threadIdx.x);
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
    const dim3 grid((unsigned)blocks), block(p.threads);
    const int rm = reg_rmax(p);
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
