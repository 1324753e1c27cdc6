// Instantiations of the streaming kernel (fq_stream.hpp).
#include "fq_launch.hpp"
#include "fq_stream.hpp"

namespace fq {

#define FQ_STREAM(P, I, E)                                                                                     \
    if (pot == P && inter == I && ext == E) {                                                                  \
        hipLaunchKernelGGL((k_stream<P, I, E>), dim3(nsys), dim3(STREAM_THREADS), 0, stream, views, ra);       \
        return true;                                                                                           \
    }

bool launch_stream(int pot, int inter, int ext, const SysView* views, int nsys, const RunArgs& ra,
                   hipStream_t stream)
{
    FQ_STREAM(POT_CUSPY, INT_LAPLACE1D, 0)
    FQ_STREAM(POT_CUSPY, INT_LAPLACE1D, 1)
    FQ_STREAM(POT_CUSPY, INT_QUARTIC1D, 0)
    FQ_STREAM(POT_CUSPY, INT_QUARTIC1D, 1)
    FQ_STREAM(POT_CUSPY, INT_QUARTICGRADIENT1D, 0)
    FQ_STREAM(POT_CUSPY, INT_NONE, 0)
    FQ_STREAM(POT_CUSPY, INT_NONE, 1)
    FQ_STREAM(POT_CUSPY, INT_LONGRANGE1D, 0)
    FQ_STREAM(POT_CUSPY, INT_LONGRANGE1D, 1)
    FQ_STREAM(POT_CUSPY, INT_LAPLACE2D, 0)
    FQ_STREAM(POT_CUSPY, INT_QUARTICGRADIENT2D, 0)
    FQ_STREAM(POT_SEMISMOOTH, INT_LAPLACE1D, 0)
    FQ_STREAM(POT_SEMISMOOTH, INT_QUARTIC1D, 0)
    FQ_STREAM(POT_SEMISMOOTH, INT_NONE, 0)
    FQ_STREAM(POT_SMOOTH, INT_LAPLACE1D, 0)
    FQ_STREAM(POT_SMOOTH, INT_QUARTIC1D, 0)
    FQ_STREAM(POT_SMOOTH, INT_NONE, 0)
    return false;
}

} // namespace fq
