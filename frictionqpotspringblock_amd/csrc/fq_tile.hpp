// Tiled per-step path for systems beyond a resident workgroup: every system
// is cut into tiles of ~TILE_PARTICLES particles (whole rows for lattices),
// one workgroup per tile, one launch per step for all tiles of all systems of
// a group, so that a large system runs on many CUs. Same expressions as
// k_resident / k_stream (fq_physics.hpp), so the trajectories are
// bit-identical; it implements
//   timeStep            detail.h:1539-1570
//   timeSteps/flowSteps detail.h:1577-1583, 1637-1645   (MODE_STEPS)
//   minimise (inertial) detail.h:1754-1785              (MODE_MINIMISE)
// for chains and lattices without external forcing.
//
// State: u, v, a double-buffered in HBM (RunState::cur says which buffer holds
// the state); the well (jloc and its two yield positions, cached in yl_a /
// yr_a) is updated in place by the particle's owner. One step is two kernels:
//   k_tile_step   : slips u + dt v + dt^2/2 a of the tile and of its halo
//                   (the neighbours' slips are recomputed from their state:
//                   the same expression, hence the same values), well walk
//                   (an unreachable well flags a refill for the system),
//                   forces, the three velocity passes into the other buffer,
//                   per-tile residual partial sums;
//   k_tile_decide : per system: commit the step (flip the buffer, counters)
//                   unless a refill is pending, StopList from the tile sums
//                   in a fixed order.
// A discarded step (refill) leaves the state buffer untouched; wells already
// moved by it stay correct for the state (a walk from any well of the window
// finds the same one). k_tile_finish writes the final state into the primary
// arrays with its forces (quenched on convergence).
// Roofline: HBM-bound, ~68 B per particle-step (u, v, a, jloc, y_l, y_r read;
// u, v, a written) plus the halo.
#pragma once

#include "fq_resident.hpp"

namespace fq {

constexpr int TILE_THREADS = 256;
constexpr int64_t TILE_PARTICLES = 4096;
constexpr int TILE_FLAG_REFILL = 1;
constexpr int TILE_FLAG_NAN = 2;

template <int INT>
constexpr bool is_lattice()
{
    return INT == INT_LAPLACE2D || INT == INT_QUARTICGRADIENT2D;
}

// LDS exchange of a tile: chains hold [halo, p0 .. p1-1, halo]; lattices the
// tile's rows with one halo row above and below (width `cols`).
__device__ __forceinline__ int64_t tile_lds_len(int64_t p0, int64_t p1, int64_t cols, bool lattice)
{
    return lattice ? (p1 - p0) + 2 * cols : (p1 - p0) + 2;
}

// Refresh the cached yield positions of every particle of the tile from its
// window (at the start of a run and after a refill).
__global__ void __launch_bounds__(TILE_THREADS) k_tile_prologue(const SysView* __restrict__ views,
                                                                 const TileDesc* __restrict__ tiles)
{
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const Landscape& L = S.ls;
    for (int64_t p = td.p0 + threadIdx.x; p < td.p1; p += blockDim.x) {
        const int32_t j = L.jloc[p];
        S.yl_a[p] = L.y[p * L.nchunk + j];
        S.yr_a[p] = L.y[p * L.nchunk + j + 1];
    }
}

template <int POT, int INT>
__global__ void __launch_bounds__(TILE_THREADS) k_tile_step(const SysView* __restrict__ views,
                                                             const TileDesc* __restrict__ tiles,
                                                             double* __restrict__ part, int32_t* __restrict__ tflag,
                                                             RunArgs ra)
{
    extern __shared__ double s_x[];
    __shared__ double s_red[2][TILE_THREADS / 64];
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    RunState* const rs = S.rs;
    const int32_t status = uniform_i(rs->status);
    const int64_t steps = uniform_ll(rs->steps), target = uniform_ll(rs->target);
    if ((status != ST_RUNNING && status != ST_BUDGET) || steps >= target) {
        return;
    }
    const int cb = uniform_i(rs->cur), nb = cb ^ 1;
    const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6, nw = T >> 6;
    const int64_t N = S.N, cols = S.cols, R = S.rows;
    const int64_t p0 = td.p0, p1 = td.p1;
    constexpr bool LAT = is_lattice<INT>();
    const Landscape& L = S.ls;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);
    const double* const Uc = cb ? S.u_b : S.u;
    const double* const Vc = cb ? S.v_b : S.v;
    const double* const Ac = cb ? S.a_b : S.a;
    double* const Un = nb ? S.u_b : S.u;
    double* const Vn = nb ? S.v_b : S.v;
    double* const An = nb ? S.a_b : S.a;

    const double dt = S.dt;
    const double c2 = 0.5 * dt * dt;
    const double hdt = 0.5 * dt;
    const double meta = -S.eta;
    const double inv_m = S.inv_m;
    const double kf = S.k_frame;
    const double uf = ra.flow ? uniform_d(rs->u_frame) + ra.v_frame * dt : uniform_d(rs->u_frame);

    // slips of the tile and its halo into LDS: local index k of global q
    const int64_t h = LAT ? cols : 1;  // halo width
    const int64_t n = p1 - p0;
    const int64_t len = n + 2 * h;
    for (int64_t k = t; k < len; k += T) {
        int64_t q = p0 - h + k;
        q = q < 0 ? q + N : (q >= N ? q - N : q);
        s_x[k] = Uc[q] + dt * Vc[q] + c2 * Ac[q];
    }
    __syncthreads();

    // well walk of the tile's own particles (in place)
    int flag = 0;
    for (int64_t k = t; k < n; k += T) {
        const int64_t p = p0 + k;
        const double x = s_x[h + k];
        flag |= __builtin_isnan(x) ? TILE_FLAG_NAN : 0;
        double lo = S.yl_a[p], hi = S.yr_a[p];
        if (!(lo <= x && x < hi) && __builtin_isfinite(x)) {
            const double* row = L.y + p * nch;
            int32_t j = L.jloc[p];
            while (x >= hi && j < jmax) {
                ++j;
                lo = hi;
                hi = row[j + 1];
            }
            while (x < lo && j > jmin) {
                --j;
                hi = lo;
                lo = row[j];
            }
            if (lo <= x && x < hi) {
                L.jloc[p] = j;
                S.yl_a[p] = lo;
                S.yr_a[p] = hi;
            }
            else {
                flag |= TILE_FLAG_REFILL;
            }
        }
    }

    // local accessor of the slips (lattices: global site q of a row in
    // [r0 - 1, r1] modulo the rows)
    const int64_t r0 = LAT ? p0 / cols : 0;
    auto X = [&](int64_t q) -> double {
        if constexpr (LAT) {
            int64_t d = q / cols - r0 + 1;
            d = d < 0 ? d + R : (d >= R ? d - R : d);
            return s_x[d * cols + q % cols];
        }
        else {
            int64_t d = q - p0 + 1;
            d = d < 0 ? d + N : (d >= N ? d - N : d);
            return s_x[d];
        }
    };

    double sf = 0.0, sff = 0.0;
    for (int64_t k = t; k < n; k += T) {
        const int64_t p = p0 + k;
        const double x = s_x[h + k];
        double fi;
        if constexpr (INT == INT_NONE) {
            fi = 0.0;
        }
        else if constexpr (LAT) {
            fi = lattice_force<INT>(S, X, p / cols, p % cols);
        }
        else {
            fi = chain_force<INT>(S, s_x[k], x, s_x[k + 2]);
        }
        const double fp = potential_force<POT>(S, x, S.yl_a[p], S.yr_a[p]);
        const double ff = kf * (uf - x);
        const double sum3 = ff + fp + fi;
        const double vn = Vc[p], an = Ac[p];
        double vv = vn + dt * an;
        double f = sum3 + meta * vv;
        double aa = f * inv_m;
        vv = vn + hdt * (an + aa);
        f = sum3 + meta * vv;
        aa = f * inv_m;
        vv = vn + hdt * (an + aa);
        f = sum3 + meta * vv;
        aa = f * inv_m;
        Un[p] = x;
        Vn[p] = vv;
        An[p] = aa;
        sf += f * f;
        sff += ff * ff;
    }
    if (flag) {
        atomicOr(tflag + td.sys, flag);
    }
    if (ra.mode == MODE_MINIMISE) {
        sf = wave_sum(sf);
        sff = wave_sum(sff);
        if (lane == 0) {
            s_red[0][wave] = sf;
            s_red[1][wave] = sff;
        }
        __syncthreads();
        if (t == 0) {
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < nw; ++w) {
                a0 += s_red[0][w];
                a1 += s_red[1][w];
            }
            part[2 * blockIdx.x] = a0;
            part[2 * blockIdx.x + 1] = a1;
        }
    }
}

// Commit or discard the step of every system (one thread per system).
__global__ void __launch_bounds__(64) k_tile_decide(const SysView* __restrict__ views, int nsys,
                                                     const int32_t* __restrict__ tile_first,
                                                     const double* __restrict__ part, int32_t* __restrict__ tflag,
                                                     RunArgs ra)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) {
        return;
    }
    const SysView& S = views[s];
    RunState* const rs = S.rs;
    if ((rs->status != ST_RUNNING && rs->status != ST_BUDGET) || rs->steps >= rs->target) {
        return;
    }
    const int32_t fl = tflag[s];
    tflag[s] = 0;
    if (fl & TILE_FLAG_REFILL) {
        rs->status = ST_REFILL;
        return;
    }
    rs->cur ^= 1;
    rs->steps += 1;
    rs->inc += 1;
    if (ra.flow) {
        rs->u_frame = rs->u_frame + ra.v_frame * S.dt;
    }
    if (fl & TILE_FLAG_NAN) {
        rs->status = ST_NAN;
        return;
    }
    if (ra.mode == MODE_MINIMISE) {
        double sf = 0.0, sff = 0.0;
        for (int k = tile_first[s]; k < tile_first[s + 1]; ++k) {
            sf += part[2 * k];
            sff += part[2 * k + 1];
        }
        const double rf = __builtin_sqrt(sf), rff = __builtin_sqrt(sff);
        const double r = (rff != 0.0) ? rf / rff : rf;
        rs->below = !(r >= ra.tol) ? rs->below + 1 : 0;
        rs->below2 = !(r >= ra.tol2) ? rs->below2 + 1 : 0;
        rs->dec = !(r >= rs->prev) ? rs->dec + 1 : 0;
        rs->prev = r;
        const int64_t nt = ra.niter_tol;
        if (nt <= 0 || (rs->dec >= nt - 1 && rs->below >= nt) || rs->below2 >= nt) {
            rs->status = ST_CONVERGED;
        }
    }
}

// The final state into the primary arrays and the forces of it (quenched if
// the StopList converged).
template <int POT, int INT>
__global__ void __launch_bounds__(TILE_THREADS) k_tile_finish(const SysView* __restrict__ views,
                                                               const TileDesc* __restrict__ tiles)
{
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const RunState* rs = S.rs;
    const int cb = uniform_i(rs->cur);
    const int32_t status = uniform_i(rs->status);
    const int64_t N = S.N, cols = S.cols;
    constexpr bool LAT = is_lattice<INT>();
    const double* const Uc = cb ? S.u_b : S.u;
    const double* const Vc = cb ? S.v_b : S.v;
    const double* const Ac = cb ? S.a_b : S.a;
    const double meta = -S.eta;
    const double kf = S.k_frame;
    const double uf = rs->u_frame;
    auto U = [&](int64_t q) -> double { return Uc[q]; };
    for (int64_t p = td.p0 + threadIdx.x; p < td.p1; p += blockDim.x) {
        const double x = Uc[p];
        double vv = Vc[p], aa = Ac[p];
        if (status == ST_CONVERGED) {
            vv = 0.0;
            aa = 0.0;
        }
        if (status != ST_NAN) {
            double fi;
            if constexpr (INT == INT_NONE) {
                fi = 0.0;
            }
            else if constexpr (LAT) {
                fi = lattice_force<INT>(S, U, p / cols, p % cols);
            }
            else {
                fi = chain_force<INT>(S, Uc[p == 0 ? N - 1 : p - 1], x, Uc[p == N - 1 ? 0 : p + 1]);
            }
            const double fp = potential_force<POT>(S, x, S.yl_a[p], S.yr_a[p]);
            const double ff = kf * (uf - x);
            const double fd = meta * vv;
            S.f_int[p] = fi;
            S.f_pot[p] = fp;
            S.f_frame[p] = ff;
            S.f_damp[p] = fd;
            S.f[p] = ff + fp + fi + fd;
        }
        S.v[p] = vv;
        S.a[p] = aa;
    }
}

// Second half of the finish: the slips into the primary array (after every
// tile of k_tile_finish read its neighbours from the state buffer).
__global__ void __launch_bounds__(TILE_THREADS) k_tile_finish_u(const SysView* __restrict__ views,
                                                                 const TileDesc* __restrict__ tiles)
{
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    if (S.rs->cur == 0) {
        return;
    }
    for (int64_t p = td.p0 + threadIdx.x; p < td.p1; p += blockDim.x) {
        S.u[p] = S.u_b[p];
    }
}

} // namespace fq
