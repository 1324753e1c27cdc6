// Launchers of the service kernels (fq_service.hpp).
#include "fq_launch.hpp"
#include "fq_service.hpp"

namespace fq {

hipError_t launch_landscape_init(const SysView* views, int nsys, int64_t maxN, hipStream_t stream)
{
    dim3 grid((unsigned)((maxN + SVC_THREADS - 1) / SVC_THREADS), (unsigned)nsys);
    hipLaunchKernelGGL(k_landscape_init, grid, dim3(SVC_THREADS), 0, stream, views);
    return hipGetLastError();
}

hipError_t launch_restore(const SysView* views, int64_t N, const uint64_t* state, const double* value,
                          const int64_t* index, hipStream_t stream)
{
    dim3 grid((unsigned)((N + SVC_THREADS - 1) / SVC_THREADS), 1);
    hipLaunchKernelGGL(k_restore, grid, dim3(SVC_THREADS), 0, stream, views, state, value, index);
    return hipGetLastError();
}

hipError_t launch_trigger(const SysView* views, int64_t p, double eps, int32_t direction, int32_t do_align,
                          hipStream_t stream)
{
    hipLaunchKernelGGL(k_trigger, dim3(1), dim3(1), 0, stream, views, p, eps, direction, do_align);
    return hipGetLastError();
}

hipError_t launch_trigger_each(const SysView* views, int nsys, const int64_t* ps, double eps, int32_t direction,
                               int32_t do_align, hipStream_t stream)
{
    hipLaunchKernelGGL(k_trigger_each, dim3(nsys), dim3(1), 0, stream, views, ps, eps, direction, do_align);
    return hipGetLastError();
}

hipError_t launch_refill(const SysView* views, int nsys, hipStream_t stream)
{
    hipLaunchKernelGGL(k_refill, dim3(nsys), dim3(SVC_THREADS), 0, stream, views);
    return hipGetLastError();
}

hipError_t launch_qs_record(const SysView* views, int nsys, const QsArgs& qa, int64_t e, int begin, hipStream_t stream)
{
    hipLaunchKernelGGL(k_qs_record, dim3(nsys), dim3(SVC_THREADS), 0, stream, views, qa, e, (int32_t)begin);
    return hipGetLastError();
}

hipError_t launch_set_iref(const SysView* views, int nsys, int64_t maxN, hipStream_t stream)
{
    dim3 grid((unsigned)((maxN + SVC_THREADS - 1) / SVC_THREADS), (unsigned)nsys);
    hipLaunchKernelGGL(k_set_iref, grid, dim3(SVC_THREADS), 0, stream, views);
    return hipGetLastError();
}

#define FQ_FORCES(P, I, E)                                                                                     \
    if (pot == P && inter == I && ext == E) {                                                                  \
        hipLaunchKernelGGL((k_forces<P, I, E>), dim3(nsys), dim3(SVC_THREADS), 0, stream, views, flags);       \
        return true;                                                                                           \
    }

static bool launch_forces_one(int pot, int inter, int ext, const SysView* views, int nsys, int32_t flags,
                              hipStream_t stream);

// Long-range systems: the interaction forces by k_lr_fint (many workgroups
// per system) between the updates that k_forces applies first (uniform
// displacement, quench, thermal draws) and the other forces.
bool launch_forces(int pot, int inter, int ext, const SysView* views, int nsys, int32_t flags, hipStream_t stream)
{
    if (inter != INT_LONGRANGE1D || !(flags & FF_INT)) {
        return launch_forces_one(pot, inter, ext, views, nsys, flags, stream);
    }
    const int32_t first = flags & (FF_ADD_U | FF_QUENCH | FF_THERMAL);
    if (first && !launch_forces_one(pot, inter, ext, views, nsys, first, stream)) {
        return false;
    }
    hipLaunchKernelGGL(k_lr_fint, dim3(64, (unsigned)nsys), dim3(SVC_THREADS), 0, stream, views);
    return launch_forces_one(pot, inter, ext, views, nsys, flags & ~(first | FF_INT), stream);
}

static bool launch_forces_one(int pot, int inter, int ext, const SysView* views, int nsys, int32_t flags,
                              hipStream_t stream)
{
    FQ_FORCES(POT_CUSPY, INT_LAPLACE1D, 0)
    FQ_FORCES(POT_CUSPY, INT_LAPLACE1D, 1)
    FQ_FORCES(POT_CUSPY, INT_QUARTIC1D, 0)
    FQ_FORCES(POT_CUSPY, INT_QUARTIC1D, 1)
    FQ_FORCES(POT_CUSPY, INT_QUARTICGRADIENT1D, 0)
    FQ_FORCES(POT_CUSPY, INT_NONE, 0)
    FQ_FORCES(POT_CUSPY, INT_NONE, 1)
    FQ_FORCES(POT_CUSPY, INT_LONGRANGE1D, 0)
    FQ_FORCES(POT_CUSPY, INT_LONGRANGE1D, 1)
    FQ_FORCES(POT_CUSPY, INT_LAPLACE2D, 0)
    FQ_FORCES(POT_CUSPY, INT_QUARTICGRADIENT2D, 0)
    FQ_FORCES(POT_SEMISMOOTH, INT_LAPLACE1D, 0)
    FQ_FORCES(POT_SEMISMOOTH, INT_QUARTIC1D, 0)
    FQ_FORCES(POT_SEMISMOOTH, INT_NONE, 0)
    FQ_FORCES(POT_SMOOTH, INT_LAPLACE1D, 0)
    FQ_FORCES(POT_SMOOTH, INT_QUARTIC1D, 0)
    FQ_FORCES(POT_SMOOTH, INT_NONE, 0)
    return false;
}

bool launch_max_uniform(int pot, const SysView* views, int nsys, int32_t direction, hipStream_t stream)
{
    if (pot == POT_CUSPY) {
        hipLaunchKernelGGL((k_max_uniform<POT_CUSPY>), dim3(nsys), dim3(SVC_THREADS), 0, stream, views, direction);
        return true;
    }
    if (pot == POT_SEMISMOOTH) {
        hipLaunchKernelGGL((k_max_uniform<POT_SEMISMOOTH>), dim3(nsys), dim3(SVC_THREADS), 0, stream, views,
                           direction);
        return true;
    }
    return false;
}

bool launch_nopassing(int ppt, int threads, size_t lds, const SysView* views, int nsys, const RunArgs& ra,
                      hipStream_t stream)
{
    switch (ppt) {
        case 1:
            hipLaunchKernelGGL((k_nopassing<1>), dim3(nsys), dim3(threads), lds, stream, views, ra);
            return true;
        case 2:
            hipLaunchKernelGGL((k_nopassing<2>), dim3(nsys), dim3(threads), lds, stream, views, ra);
            return true;
        case 4:
            hipLaunchKernelGGL((k_nopassing<4>), dim3(nsys), dim3(threads), lds, stream, views, ra);
            return true;
        case 8:
            hipLaunchKernelGGL((k_nopassing<8>), dim3(nsys), dim3(threads), lds, stream, views, ra);
            return true;
        case 16:
            hipLaunchKernelGGL((k_nopassing<16>), dim3(nsys), dim3(threads), lds, stream, views, ra);
            return true;
        default:
            return false;
    }
}

} // namespace fq
