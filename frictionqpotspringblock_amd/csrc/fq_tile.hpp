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
    // (long range: every slip of the system, at its own index)
    constexpr bool LR = INT == INT_LONGRANGE1D;
    const int64_t h = LR ? p0 : (LAT ? cols : 1); // halo width (long range: the slips before the tile)
    const int64_t n = p1 - p0;
    const int64_t len = LR ? N : n + 2 * h;
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
        else if constexpr (LR) {
            // the reference's sequential sum over all particles (detail.h:
            // 849-867), slips from LDS (one address per step of the sum:
            // a broadcast), prefactors from the system's table
            const double* __restrict__ pref = S.pref;
            fi = longrange_force(S, [&](int64_t i) { return s_x[i]; }, [&](int64_t d) { return pref[d]; }, p);
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

// ---------------------------------------------------------------------------
// Blocks of K steps per launch (chains). Each tile computes its particles
// [p0, p1) together with a halo of K particles on each side, all in
// registers, for up to K steps with one barrier per step (neighbour slips
// through LDS, as in k_resident); the halo is recomputed from its state with
// the same expressions as its owner, so the tile's particles stay exact for K
// steps (errors from the halo's open ends move inwards one particle per step).
// The tile records, per step, the residual partial sums of its own particles
// and the first step s at which one of them needed a refill or became NaN,
// as (s << 2) | TILE_FLAG_*; the halo's failures are recorded by their
// owners. k_tile_block_decide then commits the K steps, or, if the StopList
// stops earlier or a failure occurred, asks the next block to redo the first
// c steps from the unchanged state buffer (deterministic, hence identical)
// and commits those: c = s - 1 before a refill (then status REFILL, and the
// host refills as after k_tile_step), c = s through a NaN step (then status
// NAN, as k_tile_decide does).
// ---------------------------------------------------------------------------

constexpr int TB_THREADS = 256;
constexpr int TB_PPT = 5;                 // particles per thread of the extended segment
constexpr int TB_KMAX = (TB_THREADS * TB_PPT - TB_TILE_PARTICLES) / 2; // 128
static_assert(TB_KMAX == TB_KMAX_STEPS, "extended segment and block length disagree");
constexpr int32_t TB_NO_FAIL = 0x7fffffff;

// steps the block of system S takes (the redo count, or min(K, target - steps))
__device__ __forceinline__ int tb_steps(const RunState* rs, int K)
{
    const int32_t redo = rs->tb_redo;
    if (redo > 0) {
        return redo;
    }
    const int64_t room = rs->target - rs->steps;
    return (int)(room < K ? (room > 0 ? room : 0) : K);
}

template <int POT, int INT>
__global__ void __launch_bounds__(TB_THREADS) k_tile_block(const SysView* __restrict__ views,
                                                           const TileDesc* __restrict__ tiles,
                                                           double* __restrict__ part, int32_t* __restrict__ tfail,
                                                           RunArgs ra, int K)
{
    constexpr int PPT = TB_PPT;
    __shared__ double s_first[2][TB_THREADS];
    __shared__ double s_last[2][TB_THREADS];
    __shared__ double s_red[2][TB_THREADS / 64][2];
    __shared__ int s_fail;
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const SysView P = S; // parameters in registers (no reload after each barrier)
    const RunState* const rs = S.rs;
    const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6, nw = T >> 6;
    const int32_t status = uniform_i(rs->status);
    if (status != ST_RUNNING && status != ST_BUDGET) {
        return;
    }
    const int Ke = uniform_i(tb_steps(rs, K));
    const int cb = uniform_i(rs->cur), nb = cb ^ 1;
    const int64_t N = S.N, p0 = td.p0, n = td.p1 - td.p0;
    const int64_t next = n + 2 * (int64_t)K; // particles of the extended segment
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
    const double dv = ra.v_frame * dt;
    const bool mini = ra.mode == MODE_MINIMISE;

    // extended index i = t * PPT + k, global particle (p0 - K + i) mod N
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(next - base >= PPT ? PPT : (next - base > 0 ? next - base : 0));
    const int Tlast = (int)((next - 1) / PPT); // thread holding the last particle
    double u[PPT], v[PPT], a[PPT], yl[PPT], yr[PPT];
    // constants of the current well (well_consts), kept for SemiSmooth: they
    // save its two divisions per step (7-13 % of the kernel's time). Smooth's
    // one division would cost more in registers (occupancy) than it saves;
    // Cuspy's mid-point is a single addition and multiplication.
    constexpr bool WC = POT == POT_SEMISMOOTH;
    double wc0[PPT], wc1[PPT];
    int32_t jl[PPT];
    auto gidx = [&](int k) -> int64_t { // global index of local particle k (periodic)
        int64_t q = (p0 - K + base + k) % N;
        return q < 0 ? q + N : q;
    };
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        u[k] = v[k] = a[k] = yl[k] = yr[k] = wc0[k] = wc1[k] = 0.0;
        jl[k] = 0;
        if (k < cnt) {
            const int64_t q = gidx(k);
            u[k] = Uc[q];
            v[k] = Vc[q];
            a[k] = Ac[q];
            // the well from jloc and the window (another tile may be writing
            // the cached yield positions of its own particles meanwhile; jloc
            // is one word, either value is a well of the window, and the
            // first walk finds the particle's well from either)
            jl[k] = L.jloc[q];
            yl[k] = L.y[q * nch + jl[k]];
            yr[k] = L.y[q * nch + jl[k] + 1];
            if constexpr (WC) {
                well_consts<POT>(P, yl[k], yr[k], wc0[k], wc1[k]);
            }
        }
    }
    if (t == 0) {
        s_fail = TB_NO_FAIL;
    }
    double uf = rs->u_frame;
    __syncthreads();

    int s;
    for (s = 1; s <= Ke; ++s) {
        const int b = s & 1;
        double* const x = u; // position update in place: x = u + dt v + dt^2/2 a
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            x[k] = u[k] + dt * v[k] + c2 * a[k];
        }
        if (cnt > 0) {
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (k == 0) {
                    s_first[b][t] = x[0];
                }
                if (k == cnt - 1) {
                    s_last[b][t] = x[k];
                }
            }
        }
        __syncthreads();
        if (mini && s > 1 && t == 0) { // residual sums of step s - 1
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < nw; ++w) {
                a0 += s_red[b ^ 1][w][0];
                a1 += s_red[b ^ 1][w][1];
            }
            part[((int64_t)blockIdx.x * K + (s - 2)) * 2] = a0;
            part[((int64_t)blockIdx.x * K + (s - 2)) * 2 + 1] = a1;
        }
        if ((uniform_i(s_fail) >> 2) < s) {
            break; // a particle of the tile failed at an earlier step
        }
        if (ra.flow) {
            uf = uf + dv;
        }
        double xlast = x[0]; // (selected, not indexed: no scratch)
#pragma unroll
        for (int k = 1; k < PPT; ++k) {
            xlast = k == cnt - 1 ? x[k] : xlast;
        }
        // open ends of the extended segment: any finite value
        const double left = (t == 0) ? x[0] : s_last[b][t - 1];
        const double right = (t == Tlast) ? xlast : s_first[b][t + 1 < T ? t + 1 : t];
        double sf = 0.0, sff = 0.0;
        int fail = 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (k < cnt) {
                const int64_t i = base + k;
                const bool exact = i >= s - 1 && i < next - s + 1; // the state of step s - 1 is exact
                const bool own = i >= K && i < K + n;
                const double xk = x[k];
                if (own && __builtin_isnan(xk)) {
                    fail |= TILE_FLAG_NAN;
                }
                if (exact && !(yl[k] <= xk && xk < yr[k]) && __builtin_isfinite(xk)) {
                    const double* row = L.y + gidx(k) * nch;
                    int32_t j = jl[k];
                    double lo = yl[k], hi = yr[k];
                    while (xk >= hi && j < jmax) {
                        ++j;
                        lo = hi;
                        hi = row[j + 1];
                    }
                    while (xk < lo && j > jmin) {
                        --j;
                        hi = lo;
                        lo = row[j];
                    }
                    if (lo <= xk && xk < hi) {
                        jl[k] = j;
                        yl[k] = lo;
                        yr[k] = hi;
                        if constexpr (WC) {
                            well_consts<POT>(P, lo, hi, wc0[k], wc1[k]);
                        }
                    }
                    else if (own) {
                        fail |= TILE_FLAG_REFILL;
                    }
                }
                double fi;
                if constexpr (INT == INT_NONE) {
                    fi = 0.0;
                }
                else {
                    double um = 0.0, up = 0.0;
#pragma unroll
                    for (int q = 0; q < PPT; ++q) {
                        if (q == k) {
                            um = (q == 0) ? left : x[q > 0 ? q - 1 : 0];
                            up = (q == cnt - 1) ? right : x[q + 1 < PPT ? q + 1 : q];
                        }
                    }
                    fi = chain_force<INT>(P, um, xk, up);
                }
                const double fp = WC ? potential_force_wc<POT>(P, xk, yl[k], yr[k], wc0[k], wc1[k])
                                     : potential_force<POT>(P, xk, yl[k], yr[k]);
                const double ff = kf * (uf - xk);
                const double sum3 = ff + fp + fi;
                const double vn = v[k], an = a[k];
                double vv = vn + dt * an;
                double f = sum3 + meta * vv;
                double aa = f * inv_m;
                vv = vn + hdt * (an + aa);
                f = sum3 + meta * vv;
                aa = f * inv_m;
                vv = vn + hdt * (an + aa);
                f = sum3 + meta * vv;
                aa = f * inv_m;
                v[k] = vv;
                a[k] = aa;
                if (own) {
                    sf += f * f;
                    sff += ff * ff;
                }
            }
        }
        if (fail) { // a refill before a NaN of the same step
            atomicMin(&s_fail, (s << 2) | ((fail & TILE_FLAG_REFILL) ? TILE_FLAG_REFILL : TILE_FLAG_NAN));
        }
        if (mini) {
            sf = wave_sum(sf);
            sff = wave_sum(sff);
            if (lane == 0) {
                s_red[b][wave][0] = sf;
                s_red[b][wave][1] = sff;
            }
        }
    }
    const int done = s - 1;
    __syncthreads();
    if (t == 0) {
        if (mini && done >= 1 && done == Ke) { // sums of the last step
            const int lb = done & 1;
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < nw; ++w) {
                a0 += s_red[lb][w][0];
                a1 += s_red[lb][w][1];
            }
            part[((int64_t)blockIdx.x * K + (done - 1)) * 2] = a0;
            part[((int64_t)blockIdx.x * K + (done - 1)) * 2 + 1] = a1;
        }
        tfail[blockIdx.x] = s_fail;
    }
    // the tile's particles: state after the block into the other buffer (used
    // only if the decision commits all Ke steps), wells in place
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int64_t i = base + k;
        if (k < cnt && i >= K && i < K + n) {
            const int64_t q = gidx(k);
            Un[q] = u[k];
            Vn[q] = v[k];
            An[q] = a[k];
            L.jloc[q] = jl[k];
            S.yl_a[q] = yl[k];
            S.yr_a[q] = yr[k];
        }
    }
}

// ---------------------------------------------------------------------------
// k_tile_block2d: blocks of K steps over tiles of lattices (Laplace2d,
// QuarticGradient2d), the 2-d form of k_tile_block. A tile owns rows
// [r0, r0 + nr) x columns [c0, c0 + nc) and recomputes K halo rows and
// columns on every side (periodic; a halo may repeat sites of a small
// lattice, each copy is computed on its own): He x We = (nr + 2K) x (nc + 2K)
// extended sites, at most TL_SITES, site i = k * T + t row-major (a wavefront
// reads and writes runs of consecutive particles of a row). Velocity,
// acceleration and well stay in registers; the slip of each step goes to LDS
// (two buffers, one barrier per step), where the neighbours read it. The
// state of step s - 1 is exact on the extended rows and columns
// [s - 1, extent - s + 1): only those sites take step s (the others would only
// produce values that no exact site reads). Failure codes, residual partial
// sums, the decision and the redo are those of k_tile_block
// (k_tile_block_decide).
// ---------------------------------------------------------------------------

// the extended sites of a tile as the lattice lattice_force wraps around
// (opposite edges read each other: finite values, and edge sites are no
// longer exact after the first step)
struct TileRows {
    int64_t rows, cols;
    double k, k2, k4;
};

template <int POT, int INT>
__global__ void __launch_bounds__(TL_THREADS) k_tile_block2d(const SysView* __restrict__ views,
                                                             const TileDesc* __restrict__ tiles,
                                                             double* __restrict__ part,
                                                             int32_t* __restrict__ tfail, RunArgs ra, int K)
{
    constexpr int PPT = TL_PPT;
    constexpr int T = TL_THREADS;
    constexpr int nw = T / 64;
    extern __shared__ double s_x[]; // [2][He * We]: the slips after the even and the odd steps
    __shared__ double s_red[2][nw][2];
    __shared__ int s_fail;
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const SysView P = S; // parameters in registers (no reload after each barrier)
    const RunState* const rs = S.rs;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int32_t status = uniform_i(rs->status);
    if (status != ST_RUNNING && status != ST_BUDGET) {
        return;
    }
    const int Ke = uniform_i(tb_steps(rs, K));
    const int cb = uniform_i(rs->cur), nb = cb ^ 1;
    // tile geometry (uniform)
    const int R = uniform_i((int)S.rows), C = uniform_i((int)S.cols);
    const int nr = uniform_i(td.nr), nc = uniform_i(td.nc);
    const int He = nr + 2 * K, We = nc + 2 * K; // extended rows and columns
    const int next = He * We;                   // extended sites (<= TL_SITES, checked by the host)
    const int gr0 = uniform_i(((td.r0 - K) % R + R) % R); // global row and column of extended site 0
    const int gc0 = uniform_i(((td.c0 - K) % C + C) % C);
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
    const double dt = uniform_d(S.dt);
    const double c2 = 0.5 * dt * dt;
    const double hdt = 0.5 * dt;
    const double meta = -uniform_d(S.eta);
    const double inv_m = uniform_d(S.inv_m);
    const double kf = uniform_d(S.k_frame);
    const double dv = ra.v_frame * dt;
    const bool mini = ra.mode == MODE_MINIMISE;
    const TileRows geo{He, We, uniform_d(S.k), uniform_d(S.k2), uniform_d(S.k4)};

    // per site: velocity, acceleration, well, extended row << 16 | column
    // (He, We < 2^13); the slip lives in LDS only (the buffer of the last
    // step taken)
    double v[PPT], a[PPT], yl[PPT], yr[PPT];
    int32_t jl[PPT], ec[PPT];
    auto gidx = [&](int32_t rc) -> int64_t { // global particle of extended site rc
        const int g = (gr0 + (rc >> 16)) % R;
        const int c = (gc0 + (rc & 0xffff)) % C;
        return (int64_t)g * C + c;
    };
    // the sites taking step s: extended row and column in [s - 1, extent - s + 1)
    auto takes = [&](int32_t rc, int s) -> bool {
        const int e = rc >> 16, c = rc & 0xffff;
        return e >= s - 1 && e < He - s + 1 && c >= s - 1 && c < We - s + 1;
    };
    auto owns = [&](int32_t rc) -> bool {
        const int e = rc >> 16, c = rc & 0xffff;
        return e >= K && e < K + nr && c >= K && c < K + nc;
    };
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int i = k * T + t;
        v[k] = a[k] = yl[k] = yr[k] = 0.0;
        jl[k] = 0;
        ec[k] = -1; // no site
        if (i < next) {
            const int e = i / We;
            ec[k] = (e << 16) | (i - e * We);
            const int64_t q = gidx(ec[k]);
            const double u0 = Uc[q];
            v[k] = Vc[q];
            a[k] = Ac[q];
            // the well from jloc and the window (as in k_tile_block)
            jl[k] = L.jloc[q];
            yl[k] = L.y[q * nch + jl[k]];
            yr[k] = L.y[q * nch + jl[k] + 1];
            // both buffers start with finite slips: an edge site of the exact
            // region reads neighbours that no longer take steps
            s_x[i] = u0;
            s_x[next + i] = u0;
        }
    }
    if (t == 0) {
        s_fail = TB_NO_FAIL;
    }
    double uf = rs->u_frame;
    __syncthreads();

    int s;
    for (s = 1; s <= Ke; ++s) {
        const int b = s & 1;
        double* const xs = s_x + b * next;
        const double* const xp = s_x + (b ^ 1) * next;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int i = k * T + t;
            if (ec[k] >= 0 && takes(ec[k], s)) {
                xs[i] = xp[i] + dt * v[k] + c2 * a[k];
            }
        }
        __syncthreads();
        if (mini && s > 1 && t == 0) { // residual sums of step s - 1
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < nw; ++w) {
                a0 += s_red[b ^ 1][w][0];
                a1 += s_red[b ^ 1][w][1];
            }
            part[((int64_t)blockIdx.x * K + (s - 2)) * 2] = a0;
            part[((int64_t)blockIdx.x * K + (s - 2)) * 2 + 1] = a1;
        }
        if ((uniform_i(s_fail) >> 2) < s) {
            break; // a site of the tile failed at an earlier step
        }
        if (ra.flow) {
            uf = uf + dv;
        }
        double sf = 0.0, sff = 0.0;
        int fail = 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int i = k * T + t;
            const int32_t eck = ec[k];
            if (eck >= 0 && takes(eck, s)) {
                const bool own = owns(eck);
                const double xk = xs[i];
                if (own && __builtin_isnan(xk)) {
                    fail |= TILE_FLAG_NAN;
                }
                if (!(yl[k] <= xk && xk < yr[k]) && __builtin_isfinite(xk)) {
                    const double* row = L.y + gidx(eck) * nch;
                    int32_t j = jl[k];
                    double wl = yl[k], wr = yr[k];
                    while (xk >= wr && j < jmax) {
                        ++j;
                        wl = wr;
                        wr = row[j + 1];
                    }
                    while (xk < wl && j > jmin) {
                        --j;
                        wr = wl;
                        wl = row[j];
                    }
                    if (wl <= xk && xk < wr) {
                        jl[k] = j;
                        yl[k] = wl;
                        yr[k] = wr;
                    }
                    else if (own) {
                        fail |= TILE_FLAG_REFILL;
                    }
                }
                const double fi = lattice_force<INT>(geo, [&](int64_t q) { return xs[q]; }, eck >> 16, eck & 0xffff);
                const double fp = potential_force<POT>(P, xk, yl[k], yr[k]);
                const double ff = kf * (uf - xk);
                const double sum3 = ff + fp + fi;
                const double vn = v[k], an = a[k];
                double vv = vn + dt * an;
                double f = sum3 + meta * vv;
                double aa = f * inv_m;
                vv = vn + hdt * (an + aa);
                f = sum3 + meta * vv;
                aa = f * inv_m;
                vv = vn + hdt * (an + aa);
                f = sum3 + meta * vv;
                aa = f * inv_m;
                v[k] = vv;
                a[k] = aa;
                if (own) {
                    sf += f * f;
                    sff += ff * ff;
                }
            }
        }
        if (fail) { // a refill before a NaN of the same step
            atomicMin(&s_fail, (s << 2) | ((fail & TILE_FLAG_REFILL) ? TILE_FLAG_REFILL : TILE_FLAG_NAN));
        }
        if (mini) {
            sf = wave_sum(sf);
            sff = wave_sum(sff);
            if (lane == 0) {
                s_red[b][wave][0] = sf;
                s_red[b][wave][1] = sff;
            }
        }
    }
    const int done = s - 1;
    __syncthreads();
    if (t == 0) {
        if (mini && done >= 1 && done == Ke) { // sums of the last step
            const int lb = done & 1;
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < nw; ++w) {
                a0 += s_red[lb][w][0];
                a1 += s_red[lb][w][1];
            }
            part[((int64_t)blockIdx.x * K + (done - 1)) * 2] = a0;
            part[((int64_t)blockIdx.x * K + (done - 1)) * 2 + 1] = a1;
        }
        tfail[blockIdx.x] = s_fail;
    }
    // the tile's own sites: state after the block into the other buffer (used
    // only if the decision commits all Ke steps), wells in place
    const double* const xd = s_x + (done & 1) * next;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int i = k * T + t;
        if (ec[k] >= 0 && owns(ec[k])) {
            const int64_t q = gidx(ec[k]);
            Un[q] = xd[i];
            Vn[q] = v[k];
            An[q] = a[k];
            L.jloc[q] = jl[k];
            S.yl_a[q] = yl[k];
            S.yr_a[q] = yr[k];
        }
    }
}

// Commit (or ask for a redo of) the block of every system: one wavefront per
// system; lane l sums the tiles' partial norms of steps l, l + 64, ... (over
// the tiles in order), then lane 0 runs the StopList over the steps in order.
__global__ void __launch_bounds__(64) k_tile_block_decide(const SysView* __restrict__ views,
                                                           const int32_t* __restrict__ tile_first,
                                                           const double* __restrict__ part,
                                                           const int32_t* __restrict__ tfail, RunArgs ra, int K)
{
    __shared__ double s_r[TB_KMAX];
    const int sy = blockIdx.x, lane = threadIdx.x;
    const SysView& S = views[sy];
    RunState* const rs = S.rs;
    const int32_t status = uniform_i(rs->status);
    if (status != ST_RUNNING && status != ST_BUDGET) {
        return;
    }
    const int Ke = uniform_i(tb_steps(rs, K));
    const int k0 = tile_first[sy], k1 = tile_first[sy + 1];
    int32_t code = TB_NO_FAIL; // (first failing step << 2) | TILE_FLAG_* over the tiles
    for (int k = k0 + lane; k < k1; k += 64) {
        code = tfail[k] < code ? tfail[k] : code;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t o = __shfl_xor(code, off, 64);
        code = o < code ? o : code;
    }
    code = uniform_i(code);
    const int fs = code >> 2, ftype = code & 3;
    // steps that may be committed: those before a refill, through a NaN
    const int valid = ftype == TILE_FLAG_NAN ? (fs < Ke ? fs : Ke) : (fs - 1 < Ke ? fs - 1 : Ke);
    const int checked = ftype == TILE_FLAG_NAN ? valid - 1 : valid; // StopList steps (not the NaN step)
    const bool mini = ra.mode == MODE_MINIMISE;
    if (mini) {
        for (int st = lane; st < checked; st += 64) {
            double sf = 0.0, sff = 0.0;
            for (int k = k0; k < k1; ++k) {
                sf += part[((int64_t)k * K + st) * 2];
                sff += part[((int64_t)k * K + st) * 2 + 1];
            }
            const double rf = __builtin_sqrt(sf), rff = __builtin_sqrt(sff);
            s_r[st] = (rff != 0.0) ? rf / rff : rf;
        }
    }
    __syncthreads();
    if (lane != 0) {
        return;
    }
    int c = valid;
    bool conv = false;
    int32_t below = rs->below, below2 = rs->below2, dec = rs->dec;
    double prev = rs->prev;
    if (mini) {
        const int64_t nt = ra.niter_tol;
        for (int st = 0; st < checked; ++st) {
            const double r = s_r[st];
            below = !(r >= ra.tol) ? below + 1 : 0;
            below2 = !(r >= ra.tol2) ? below2 + 1 : 0;
            dec = !(r >= prev) ? dec + 1 : 0;
            prev = r;
            if (nt <= 0 || (dec >= nt - 1 && below >= nt) || below2 >= nt) {
                c = st + 1;
                conv = true;
                break;
            }
        }
    }
    // the failure to report after the committed steps (none if the StopList
    // stopped first)
    const int32_t pending = (!conv && fs <= Ke) ? ftype : 0;
    if (c < Ke) {
        if (c > 0) { // the next block redoes the first c steps
            rs->tb_redo = c;
            rs->tb_fail = pending;
            return;
        }
        // nothing to commit (a refill at the first step): report it now
        rs->tb_redo = 0;
        rs->tb_fail = 0;
        rs->status = pending == TILE_FLAG_NAN ? ST_NAN : ST_REFILL;
        return;
    }
    // commit the Ke steps
    if (Ke > 0) {
        rs->cur ^= 1;
    }
    rs->steps += Ke;
    rs->inc += Ke;
    if (ra.flow) {
        double uf = rs->u_frame;
        for (int st = 0; st < Ke; ++st) {
            uf = uf + ra.v_frame * S.dt;
        }
        rs->u_frame = uf;
    }
    rs->below = below;
    rs->below2 = below2;
    rs->dec = dec;
    rs->prev = prev;
    rs->tb_redo = 0;
    const int32_t report = rs->tb_fail ? rs->tb_fail : pending;
    rs->tb_fail = 0;
    if (conv) {
        rs->status = ST_CONVERGED;
    }
    else if (report == TILE_FLAG_NAN) {
        rs->status = ST_NAN;
    }
    else if (report == TILE_FLAG_REFILL) {
        rs->status = ST_REFILL;
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
            else if constexpr (INT == INT_LONGRANGE1D) {
                const double* __restrict__ pref = S.pref;
                fi = longrange_force(S, U, [&](int64_t d) { return pref[d]; }, p);
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

// ---------------------------------------------------------------------------
// No-passing minimiser (Jacobi sweep, Cuspy + Laplace1d, detail.h:1694-1753)
// for chains of any length, over tiles of TILE_PARTICLES particles, one
// launch per kernel per sweep:
//   k_np_sweep   : every particle of the tile jumps to the minimum of its
//                  quadratic well given the previous iterate of its
//                  neighbours (read from the state buffer, the reference's
//                  snapshot `m_v_n`), walking its well until it is stable,
//                  with the landscape window moved in place by its owner
//                  (align_particle: the same steps as the resident kernel);
//                  the new iterate goes into the other buffer;
//   k_np_resid   : per tile, the partial sums of |f|^2 and |f_frame|^2 of the
//                  new iterate (f = f_potential + f_interactions + f_frame,
//                  detail.h:1741);
//   k_np_decide  : per system, the StopList on the residual (tile sums in a
//                  fixed order), then the buffer flip.
// The well of each particle only depends on its own slip, so the in-place
// walk of a sweep never has to be undone: a sweep that the StopList stops
// after is the last one, and sweeps after a stop are not launched (the
// kernels return on a stopped system). k_np_finish writes the final iterate
// and its forces (quenched on convergence, as the resident kernel does).
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(TILE_THREADS) k_np_sweep(const SysView* __restrict__ views,
                                                            const TileDesc* __restrict__ tiles)
{
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const RunState* const rs = S.rs;
    const int32_t status = uniform_i(rs->status);
    const int64_t steps = uniform_ll(rs->steps), target = uniform_ll(rs->target);
    if ((status != ST_RUNNING && status != ST_BUDGET) || steps >= target) {
        return;
    }
    const int cb = uniform_i(rs->cur);
    const double* const Uc = cb ? S.u_b : S.u;
    double* const Un = cb ? S.u : S.u_b;
    const Landscape& L = S.ls;
    const int64_t N = S.N;
    const double k = S.k, kf = S.k_frame, mu = S.mu;
    const double uf = uniform_d(rs->u_frame);
    for (int64_t p = td.p0 + threadIdx.x; p < td.p1; p += blockDim.x) {
        const double uneigh = Uc[p == 0 ? N - 1 : p - 1] + Uc[p == N - 1 ? 0 : p + 1];
        double yl = S.yl_a[p], yr = S.yr_a[p];
        int32_t j = L.jloc[p];
        bool moved = false; // the well or the window changed
        double un;
        // the well sequence is monotone (the minimum moves with the well), so
        // this terminates; the cap only bounds the work
        for (int guard = 0; guard < (1 << 20); ++guard) {
            const double umin = 0.5 * (yl + yr);
            un = (k * uneigh + kf * uf + mu * umin) / (2 * k + kf + mu);
            const int32_t jold = j;
            moved |= align_particle(L, p, un, yl, yr, j) != 0;
            if (j == jold) {
                break;
            }
            moved = true;
        }
        Un[p] = un;
        if (moved) {
            L.jloc[p] = j;
            S.yl_a[p] = yl;
            S.yr_a[p] = yr;
        }
    }
}

// Forces of the iterate x at particle p (f_potential, f_frame, f_interactions
// as the reference computes them after a sweep, detail.h:1735-1741).
__device__ __forceinline__ void np_forces(const SysView& S, const double* __restrict__ x, int64_t p, double uf,
                                          double& fp, double& ff, double& fi)
{
    const int64_t N = S.N;
    const double u = x[p];
    const double umin = 0.5 * (S.yl_a[p] + S.yr_a[p]);
    ff = S.k_frame * (uf - u);
    fp = S.mu * (umin - u);
    fi = chain_force<INT_LAPLACE1D>(S, x[p == 0 ? N - 1 : p - 1], u, x[p == N - 1 ? 0 : p + 1]);
}

__global__ void __launch_bounds__(TILE_THREADS) k_np_resid(const SysView* __restrict__ views,
                                                            const TileDesc* __restrict__ tiles,
                                                            double* __restrict__ part)
{
    __shared__ double s_red[2][TILE_THREADS / 64];
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const RunState* const rs = S.rs;
    const int32_t status = uniform_i(rs->status);
    const int64_t steps = uniform_ll(rs->steps), target = uniform_ll(rs->target);
    if ((status != ST_RUNNING && status != ST_BUDGET) || steps >= target) {
        return;
    }
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
    const double* const Un = uniform_i(rs->cur) ? S.u : S.u_b; // the new iterate
    const double uf = uniform_d(rs->u_frame);
    double sf = 0.0, sff = 0.0;
    for (int64_t p = td.p0 + t; p < td.p1; p += blockDim.x) {
        double fp, ff, fi;
        np_forces(S, Un, p, uf, fp, ff, fi);
        const double f = fp + fi + ff;
        sf += f * f;
        sff += ff * ff;
    }
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

// One thread per system: commit the sweep (flip the buffer) and push its
// residual through the StopList (detail.h:1743-1750).
__global__ void __launch_bounds__(64) k_np_decide(const SysView* __restrict__ views, int nsys,
                                                  const int32_t* __restrict__ tile_first,
                                                  const double* __restrict__ part, RunArgs ra)
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
    rs->cur ^= 1;
    rs->steps += 1;
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

// The final iterate and its forces into the primary arrays (after at least
// one sweep; quenched on convergence: v = a = 0, f_damping = -eta * 0, f in
// the order of computeForce, detail.h:1321-1330).
__global__ void __launch_bounds__(TILE_THREADS) k_np_finish(const SysView* __restrict__ views,
                                                             const TileDesc* __restrict__ tiles)
{
    const TileDesc td = tiles[blockIdx.x];
    const SysView& S = views[td.sys];
    const RunState* rs = S.rs;
    if (uniform_ll(rs->steps) == 0) {
        return;
    }
    const double* const Uc = uniform_i(rs->cur) ? S.u_b : S.u;
    const bool conv = uniform_i(rs->status) == ST_CONVERGED;
    const double uf = uniform_d(rs->u_frame);
    const double meta = -S.eta;
    for (int64_t p = td.p0 + threadIdx.x; p < td.p1; p += blockDim.x) {
        double fp, ff, fi;
        np_forces(S, Uc, p, uf, fp, ff, fi);
        S.f_pot[p] = fp;
        S.f_frame[p] = ff;
        S.f_int[p] = fi;
        if (conv) {
            S.v[p] = 0.0;
            S.a[p] = 0.0;
            const double fd = meta * 0.0;
            S.f_damp[p] = fd;
            S.f[p] = ff + fp + fi + fd;
        }
        else {
            S.f[p] = fp + fi + ff;
        }
    }
}

} // namespace fq
