// Launchers of the tiled per-step path (fq_tile.hpp).
#include "fq_launch.hpp"
#include "fq_tile.hpp"

namespace fq {

#define FQ_TILE_INTS(X, P)                                                                                     \
    X(P, INT_NONE) X(P, INT_LAPLACE1D) X(P, INT_QUARTIC1D) X(P, INT_QUARTICGRADIENT1D) X(P, INT_LAPLACE2D)    \
        X(P, INT_QUARTICGRADIENT2D)

// long range: the reference's class has the Cuspy potential only
#define FQ_TILE_LR(X) X(POT_CUSPY, INT_LONGRANGE1D)

bool tile_supported(int pot, int inter, int ext)
{
    return ext == 0 && (inter == INT_NONE || inter == INT_LAPLACE1D || inter == INT_QUARTIC1D ||
                        inter == INT_QUARTICGRADIENT1D || inter == INT_LAPLACE2D ||
                        inter == INT_QUARTICGRADIENT2D || (inter == INT_LONGRANGE1D && pot == POT_CUSPY));
}

hipError_t launch_tile_prologue(const SysView* views, const TileDesc* tiles, int ntiles, hipStream_t stream)
{
    hipLaunchKernelGGL(k_tile_prologue, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, tiles);
    return hipGetLastError();
}

bool launch_tile_step(int pot, int inter, const SysView* views, const TileDesc* tiles, int ntiles, size_t lds,
                      double* part, int32_t* tflag, const RunArgs& ra, hipStream_t stream)
{
#define FQ_STEP(P, I)                                                                                          \
    if (pot == P && inter == I) {                                                                              \
        hipLaunchKernelGGL((k_tile_step<P, I>), dim3((unsigned)ntiles), dim3(TILE_THREADS), lds, stream, views, \
                           tiles, part, tflag, ra);                                                            \
        return true;                                                                                           \
    }
    FQ_TILE_INTS(FQ_STEP, POT_CUSPY)
    FQ_TILE_INTS(FQ_STEP, POT_SEMISMOOTH)
    FQ_TILE_INTS(FQ_STEP, POT_SMOOTH)
    FQ_TILE_LR(FQ_STEP)
#undef FQ_STEP
    return false;
}

hipError_t launch_tile_decide(const SysView* views, int nsys, const int32_t* tile_first, const double* part,
                              int32_t* tflag, const RunArgs& ra, hipStream_t stream)
{
    hipLaunchKernelGGL(k_tile_decide, dim3((unsigned)((nsys + 63) / 64)), dim3(64), 0, stream, views, nsys,
                       tile_first, part, tflag, ra);
    return hipGetLastError();
}

bool tile_block_supported(int pot, int inter)
{
    (void)pot;
    return inter == INT_NONE || inter == INT_LAPLACE1D || inter == INT_QUARTIC1D || inter == INT_QUARTICGRADIENT1D ||
           inter == INT_LAPLACE2D || inter == INT_QUARTICGRADIENT2D;
}

bool launch_tile_block(int pot, int inter, const SysView* views, const TileDesc* tiles, int ntiles, double* part,
                       int32_t* tfail, const RunArgs& ra, int K, size_t lds, hipStream_t stream)
{
#define FQ_BLOCK(P, I)                                                                                         \
    if (pot == P && inter == I) {                                                                              \
        hipLaunchKernelGGL((k_tile_block<P, I>), dim3((unsigned)ntiles), dim3(TB_THREADS), 0, stream, views,  \
                           tiles, part, tfail, ra, K);                                                         \
        return true;                                                                                           \
    }
#define FQ_BLOCK2D(P, I)                                                                                       \
    if (pot == P && inter == I) {                                                                              \
        hipLaunchKernelGGL((k_tile_block2d<P, I>), dim3((unsigned)ntiles), dim3(TL_THREADS), lds, stream,      \
                           views, tiles, part, tfail, ra, K);                                                  \
        return true;                                                                                           \
    }
#define FQ_BLOCK_ALL(P) FQ_BLOCK(P, INT_NONE) FQ_BLOCK(P, INT_LAPLACE1D) FQ_BLOCK(P, INT_QUARTIC1D)            \
    FQ_BLOCK(P, INT_QUARTICGRADIENT1D) FQ_BLOCK2D(P, INT_LAPLACE2D) FQ_BLOCK2D(P, INT_QUARTICGRADIENT2D)
    FQ_BLOCK_ALL(POT_CUSPY)
    FQ_BLOCK_ALL(POT_SEMISMOOTH)
    FQ_BLOCK_ALL(POT_SMOOTH)
#undef FQ_BLOCK_ALL
#undef FQ_BLOCK2D
#undef FQ_BLOCK
    return false;
}

hipError_t launch_tile_block_decide(const SysView* views, int nsys, const int32_t* tile_first, const double* part,
                                    const int32_t* tfail, const RunArgs& ra, int K, hipStream_t stream)
{
    hipLaunchKernelGGL(k_tile_block_decide, dim3((unsigned)nsys), dim3(64), 0, stream, views, tile_first, part, tfail,
                       ra, K);
    return hipGetLastError();
}

bool launch_tile_finish(int pot, int inter, const SysView* views, const TileDesc* tiles, int ntiles,
                        hipStream_t stream)
{
    bool ok = false;
#define FQ_FIN(P, I)                                                                                           \
    if (!ok && pot == P && inter == I) {                                                                       \
        hipLaunchKernelGGL((k_tile_finish<P, I>), dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, \
                           tiles);                                                                             \
        ok = true;                                                                                             \
    }
    FQ_TILE_INTS(FQ_FIN, POT_CUSPY)
    FQ_TILE_INTS(FQ_FIN, POT_SEMISMOOTH)
    FQ_TILE_INTS(FQ_FIN, POT_SMOOTH)
    FQ_TILE_LR(FQ_FIN)
#undef FQ_FIN
    if (ok) {
        hipLaunchKernelGGL(k_tile_finish_u, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, tiles);
    }
    return ok;
}

// no-passing sweeps of chains over tiles (Cuspy + Laplace1d)
hipError_t launch_np_sweep(const SysView* views, const TileDesc* tiles, int ntiles, int nsys,
                           const int32_t* tile_first, double* part, const RunArgs& ra, hipStream_t stream)
{
    hipLaunchKernelGGL(k_np_sweep, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, tiles);
    hipLaunchKernelGGL(k_np_resid, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, tiles, part);
    hipLaunchKernelGGL(k_np_decide, dim3((unsigned)((nsys + 63) / 64)), dim3(64), 0, stream, views, nsys,
                       tile_first, part, ra);
    return hipGetLastError();
}

hipError_t launch_np_finish(const SysView* views, const TileDesc* tiles, int ntiles, hipStream_t stream)
{
    hipLaunchKernelGGL(k_np_finish, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, tiles);
    hipLaunchKernelGGL(k_tile_finish_u, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, stream, views, tiles);
    return hipGetLastError();
}

} // namespace fq
