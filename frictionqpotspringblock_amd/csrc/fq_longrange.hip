// Launchers of the batched long-range kernels (fq_longrange.hpp).
#include "fq_launch.hpp"
#include "fq_longrange.hpp"

namespace fq {

hipError_t launch_lr_gemm(const double* pref, const LrBuffers& lb, int64_t lr_m, hipStream_t stream)
{
    dim3 grid((unsigned)((lb.N + LR_BM - 1) / LR_BM), (unsigned)((lb.R + LR_BN - 1) / LR_BN));
    hipLaunchKernelGGL(k_lr_gemm, grid, dim3(LR_THREADS), 0, stream, pref, lb, lr_m);
    return hipGetLastError();
}

hipError_t launch_lr_elem(int ext, const SysView* views, const LrBuffers& lb, const RunArgs& ra, hipStream_t stream)
{
    if (ext) {
        hipLaunchKernelGGL((k_lr_elem<1>), dim3((unsigned)lb.R), dim3(1024), 0, stream, views, lb, ra);
    }
    else {
        hipLaunchKernelGGL((k_lr_elem<0>), dim3((unsigned)lb.R), dim3(1024), 0, stream, views, lb, ra);
    }
    return hipGetLastError();
}

hipError_t launch_lr_prologue(const SysView* views, const LrBuffers& lb, hipStream_t stream)
{
    hipLaunchKernelGGL(k_lr_prologue, dim3((unsigned)lb.R), dim3(1024), 0, stream, views, lb);
    return hipGetLastError();
}

} // namespace fq
