// Launchers of the adjoint kernels (fq_adjoint.hpp).
#include "fq_adjoint.hpp"
#include "fq_launch.hpp"

namespace fq {

// (particles per thread, threads per workgroup) variants; the host picks the
// first one that holds the system
#define FQ_ADJ_CASES(X) X(1, 1024) X(2, 1024) X(4, 1024) X(8, 512)

int adj_variants(int* ppt, int* maxt, int cap)
{
    const int P[] = {1, 2, 4, 8}, M[] = {1024, 1024, 1024, 512};
    int n = 0;
    for (int i = 0; i < 4 && n < cap; ++i, ++n) {
        ppt[n] = P[i];
        maxt[n] = M[i];
    }
    return n;
}

bool launch_adj_record(int ppt, int threads, const SysView* views, const AdjBuffers& ab, int windows, int flow,
                       double v_frame, hipStream_t stream)
{
    switch (ppt) {
#define FQ_REC(P, M)                                                                                           \
    case P:                                                                                                    \
        hipLaunchKernelGGL((k_adj_record<P, M>), dim3((unsigned)ab.R, (unsigned)windows), dim3(threads), 0,     \
                           stream, views, ab, flow, v_frame);                                                  \
        return true;
        FQ_ADJ_CASES(FQ_REC)
#undef FQ_REC
        default:
            return false;
    }
}

bool launch_adj_reverse(int ppt, int threads, const SysView* views, const AdjBuffers& ab, int steps,
                        hipStream_t stream)
{
    switch (ppt) {
#define FQ_REV(P, M)                                                                                           \
    case P:                                                                                                    \
        hipLaunchKernelGGL((k_adj_reverse<P, M>), dim3((unsigned)ab.R), dim3(threads), 0, stream, views, ab,    \
                           steps);                                                                             \
        return true;
        FQ_ADJ_CASES(FQ_REV)
#undef FQ_REV
        default:
            return false;
    }
}

} // namespace fq
