// Launchers of the batched long-range kernels (fq_longrange.hpp).
#include "fq_launch.hpp"
#include "fq_longrange.hpp"

namespace fq {

int lr_split_blocks()
{
    return LR_BK;
}

hipError_t launch_lr_gemm(const double* pref, const LrBuffers& lb, int64_t lr_m, hipStream_t stream)
{
    LrGeom geo;
    geo.tm = (int)((lb.N + LR_BM - 1) / LR_BM);
    geo.tn = (int)((lb.R + LR_BN - 1) / LR_BN);
    geo.nsplit = lb.nsplit;
    const int64_t nkb = (lb.N + LR_BK - 1) / LR_BK;
    geo.klen = ((nkb + lb.nsplit - 1) / lb.nsplit) * LR_BK;
    if ((int64_t)(geo.nsplit - 1) * geo.klen >= lb.N) {
        return hipErrorInvalidValue; // an empty summation range
    }
    const unsigned total = (unsigned)geo.tm * (unsigned)geo.tn * (unsigned)geo.nsplit;
    hipLaunchKernelGGL(k_lr_gemm, dim3(total), dim3(LR_THREADS), 0, stream, pref, lb, lr_m, geo);
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
