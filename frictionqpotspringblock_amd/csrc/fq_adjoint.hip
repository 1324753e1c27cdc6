// Launchers of the adjoint kernels (fq_adjoint.hpp).
#include "fq_adjoint.hpp"
#include "fq_launch.hpp"

namespace fq {

// record: (particles per thread, threads per workgroup) variants; the host
// picks the first one that holds the system
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
    (void)threads; // the kernel runs on exactly its variant's thread count
    const int64_t seg = ab.rnt == 1 ? ab.N : ab.rtp + 2 * ab.rhalo; // particles of a workgroup's segment
    const size_t lds = 2 * sizeof(double) * (size_t)seg;
    switch (ppt) {
#define FQ_REC(P, M)                                                                                           \
    case P:                                                                                                    \
        if (ab.rnt == 1) {                                                                                     \
            hipLaunchKernelGGL((k_adj_record<P, M, true>), dim3((unsigned)ab.R, (unsigned)windows), dim3(M),   \
                               lds, stream, views, ab, flow, v_frame);                                         \
        }                                                                                                      \
        else {                                                                                                 \
            hipLaunchKernelGGL((k_adj_record<P, M, false>), dim3((unsigned)(ab.R * ab.rnt), (unsigned)windows),\
                               dim3(M), lds, stream, views, ab, flow, v_frame);                                \
        }                                                                                                      \
        return true;
        FQ_ADJ_CASES(FQ_REC)
#undef FQ_REC
        default:
            return false;
    }
}

// reverse: particles of the extended segment per thread
#define FQ_ADJ_SLOTS(X) X(1) X(2) X(3) X(4) X(6) X(8)

int adj_reverse_slots(int64_t extended)
{
    const int S[] = {1, 2, 3, 4, 6, 8};
    for (int s : S) {
        if (extended <= (int64_t)s * ADJ_REV_THREADS) {
            return s;
        }
    }
    return 0;
}

bool launch_adj_reverse(const SysView* views, const AdjBuffers& ab, int slots, int steps, hipStream_t stream)
{
    switch (slots) {
#define FQ_REV(P)                                                                                              \
    case P:                                                                                                    \
        hipLaunchKernelGGL((k_adj_reverse<P>), dim3((unsigned)(ab.R * ab.nt)), dim3(ADJ_REV_THREADS), 0,       \
                           stream, views, ab, steps);                                                          \
        break;
        FQ_ADJ_SLOTS(FQ_REV)
#undef FQ_REV
        default:
            return false;
    }
    hipLaunchKernelGGL(k_adj_params, dim3((unsigned)((ab.R * 4 + 255) / 256)), dim3(256), 0, stream, ab);
    return true;
}

} // namespace fq
