// Resident Velocity-Verlet kernel: one workgroup owns one system for many
// steps. Each thread keeps PPT consecutive particles (u, v, a, the two yield
// positions of the current well, its local well index) in registers for the
// whole launch; neighbours are exchanged through LDS. It implements
//   timeStep            detail.h:1539-1570
//   timeSteps/flowSteps detail.h:1577-1583, 1637-1645   (MODE_STEPS)
//   timeStepsUntilEvent detail.h:1595-1622              (MODE_UNTIL_EVENT)
//   minimise (inertial) detail.h:1754-1785              (MODE_MINIMISE)
//   minimise_truncate   detail.h:1858-1886              (MODE_TRUNCATE)
//   RandomNormalForcing::force detail.h:930-943         (EXT = 1)
// and, as k_quasi, the examples' quasistatic event loop around eventDrivenStep
// (detail.h:1933-1960) and minimise (examples/Line1d_Cuspy_Laplace.py:39-61).
//
// One __syncthreads per step. Step `it` is started speculatively (new u into
// registers and the LDS exchange buffer it&1) before the stop decision for
// step it-1 is taken from the reductions written at the end of step it-1;
// on a stop the speculative u is discarded. Everything with a side effect
// outside registers (well alignment, chunk refill, thermal draws) happens
// after the decision.
//
// Arithmetic: per particle, exactly the reference's expression order without
// FMA; the StopList is evaluated with trailing counters equivalent to
// GooseFEM's ring of `niter_tol` residuals initialised to +inf:
//   descending() && all_less(tol)  <=>  dec >= n-1 && below >= n
//   all_less(tol^2)                <=>  below2 >= n
// Norms are reduced in a fixed order (per thread, then a butterfly per
// wavefront, then wavefronts in order), so they are deterministic but not in
// the reference's sequential order.
#pragma once

#include "fq_common.hpp"
#include "fq_landscape.hpp"
#include "fq_physics.hpp"

namespace fq {

constexpr int RES_MAX_WAVES = 16;

template <int INT>
constexpr bool is_chain()
{
    return INT == INT_NONE || INT == INT_LAPLACE1D || INT == INT_QUARTIC1D || INT == INT_QUARTICGRADIENT1D;
}

template <int PPT>
constexpr int res_max_threads()
{
    return PPT <= 8 ? 1024 : 512;
}

// Values that are equal in every lane (read back from LDS after a barrier),
// made visibly uniform so that loop control and counters stay scalar.
__device__ __forceinline__ int uniform_i(int x)
{
    return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ long long uniform_ll(long long x)
{
    const uint64_t b = (uint64_t)x;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
    return (long long)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double uniform_d(double x)
{
    return __builtin_bit_cast(double, uniform_ll(__builtin_bit_cast(long long, x)));
}

// x of the lane that the DPP control CTRL names (both 32-bit halves moved);
// lanes of rows outside ROW_MASK receive 0.0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double x)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes of a wavefront, in a fixed order, with DPP lane
// moves (VALU only, no LDS round trips): pairs within each quad, the two
// quads of each half-row, the two halves of each row (every lane then holds
// its row's sum), rows 0 + 1 and 2 + 3 (row_bcast:15), and the two halves
// (row_bcast:31). Lane 63 ends with the total, which every lane receives.
__device__ __forceinline__ double wave_sum(double x)
{
    x += dpp_move<0xB1, 0xf>(x);  // quad_perm [1, 0, 3, 2]
    x += dpp_move<0x4E, 0xf>(x);  // quad_perm [2, 3, 0, 1]
    x += dpp_move<0x141, 0xf>(x); // row_half_mirror
    x += dpp_move<0x140, 0xf>(x); // row_mirror
    x += dpp_move<0x142, 0xa>(x); // row_bcast:15 into rows 1 and 3
    x += dpp_move<0x143, 0xc>(x); // row_bcast:31 into rows 2 and 3
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ long long wave_sum_i(long long x)
{
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
    }
    return x;
}

__device__ __forceinline__ long long wave_min_i(long long x)
{
    for (int off = 32; off > 0; off >>= 1) {
        long long y = __shfl_xor(x, off, 64);
        x = y < x ? y : x;
    }
    return x;
}

// Inclusive prefix sum within the wavefront.
__device__ __forceinline__ int wave_incl_scan(int x, int lane)
{
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) {
            x += y;
        }
    }
    return x;
}

// Block-wide exclusive prefix sum of `x` in thread order, plus the total.
// Uses s_w[RES_MAX_WAVES]; contains two barriers.
__device__ __forceinline__ int block_excl_scan(int x, int* s_w, int& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int incl = wave_incl_scan(x, lane);
    if (lane == 63) {
        s_w[wave] = incl;
    }
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < nw; ++w) {
        int v = s_w[w];
        if (w < wave) {
            before += v;
        }
        total += v;
    }
    __syncthreads();
    return before + incl - x;
}

// Block-wide minimum; uses s_w[RES_MAX_WAVES]; contains two barriers.
__device__ __forceinline__ long long block_min_i(long long x, long long* s_w)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    x = wave_min_i(x);
    if (lane == 0) {
        s_w[wave] = x;
    }
    __syncthreads();
    long long r = s_w[0];
    for (int w = 1; w < nw; ++w) {
        r = s_w[w] < r ? s_w[w] : r;
    }
    __syncthreads();
    return r;
}

// Is the well of slip x reachable from well j (yl, yr) by walking inside the
// safe zone [jmin, jmax] of the window `row`? Reads only.
__device__ __forceinline__ bool reachable(const double* row, int32_t j, double lo, double hi, double x, int32_t jmin,
                                          int32_t jmax)
{
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
    return lo <= x && x < hi;
}

// LDS needed by a resident launch (bytes of dynamic shared memory).
template <int INT>
FQ_HD size_t resident_lds_bytes(int64_t N, int threads)
{
    if constexpr (is_chain<INT>()) {
        return sizeof(double) * 4 * (size_t)threads;
    }
    else if constexpr (INT == INT_LONGRANGE1D) {
        return sizeof(double) * 3 * (size_t)N;
    }
    else {
        return sizeof(double) * 2 * (size_t)N;
    }
}

// Landscape refill inside a resident launch: particle base + i of a lane
// with bit i of `mask` set left the window's safe zone; it is aligned to the
// slip of the pending step, u + dt v + dt^2/2 a, or to u itself (at_u, a
// uniform displacement of the quasistatic driver), from the state in memory,
// exactly as k_refill does. The particles are taken one at a time by the
// whole wavefront (chunk_align_wave: the new yield positions are drawn in
// parallel). Called with the state written back (and a barrier after it), so
// that no registers are live across the call.
__device__ __noinline__ void refill_wave(const SysView* __restrict__ views, int64_t base, int at_u, uint32_t mask)
{
    const SysView& S = views[blockIdx.x];
    const Landscape& L = S.ls;
    const double dt = S.dt;
    const double c2 = 0.5 * dt * dt;
    // another lane's particle, written back by that lane: read through volatile
    // (vector) loads
    const volatile double* const vu = S.u;
    const volatile double* const vv = S.v;
    const volatile double* const va = S.a;
    const volatile int32_t* const vj = L.jloc;
    uint64_t lanes = __ballot(mask != 0);
    while (lanes) {
        const int src = __builtin_ctzll(lanes);
        lanes &= lanes - 1;
        uint32_t m = (uint32_t)__shfl((int)mask, src, 64);
        const int64_t b0 = __shfl(base, src, 64);
        while (m) {
            const int i = __builtin_ctz(m);
            m &= m - 1;
            const int64_t p = b0 + i;
            const double x = at_u ? vu[p] : vu[p] + dt * vv[p] + c2 * va[p];
            chunk_align_wave(L, p, x, vj[p]);
        }
    }
}

// The body of k_resident (QS = false) and of the quasistatic driver k_quasi
// (QS = true). The driver wraps the minimise loop in the event protocol of
// QsArgs: when the StopList is met it quenches, records the event, applies
// the next events' displacements (eventDrivenStep, detail.h:1933-1960, with
// advanceUniformly detail.h:2027-2050) and restarts the minimisation at the
// same iteration, all inside the workgroup; it leaves the launch only for the
// step budget, a NaN, max_iter or the end of the run.
// A step (or displacement) that needs a landscape refill writes the state
// back, refills (refill_wave) and re-enters with the state reloaded.
template <int POT, int INT, int EXT, int PPT, int MAXT, bool QS, bool FULL = false>
__device__ __forceinline__ void resident_body(const SysView* __restrict__ views, const RunArgs& ra, const QsArgs& qa)
{
    constexpr bool CHAIN = is_chain<INT>();
    const SysView& S = views[blockIdx.x];
    RunState* const rs = S.rs;
    if (rs->status != ST_RUNNING && rs->status != ST_BUDGET) {
        return;
    }

    __shared__ double s_red_d[2][RES_MAX_WAVES][2];
    __shared__ long long s_red_i[2][RES_MAX_WAVES][2];
    __shared__ int s_nan[2];
    __shared__ int s_refill[2];
    __shared__ int s_scan_w[RES_MAX_WAVES];
    __shared__ long long s_min_w[RES_MAX_WAVES];
    __shared__ long long s_qs_i[QS ? RES_MAX_WAVES : 1];
    __shared__ double s_qs_d[QS ? RES_MAX_WAVES : 1];
    extern __shared__ double s_dyn[];

    const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6, nw = T >> 6;
    const int64_t N = S.N;
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(N - base >= PPT ? PPT : (N - base > 0 ? N - base : 0));
    // FULL (every system of the launch has N divisible by PPT): a thread
    // holds PPT particles or none, so whether register i holds a particle is
    // one test per thread, and the last particle is register PPT - 1
    const int last = FULL ? PPT - 1 : cnt - 1;
    auto own = [&](int i) -> bool { return FULL ? cnt > 0 : i < cnt; };
    const int Tact = (int)((N + PPT - 1) / PPT);
    const Landscape& L = S.ls;
    const bool track = ra.mode != MODE_STEPS;
    // per-step counts of changed wells (events, activity, truncation)
    const bool counts = ra.mode == MODE_UNTIL_EVENT || ra.mode == MODE_TRUNCATE || ra.time_activity;

    // hot scalars
    double* const Y = L.y;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);

    // LDS exchange buffers
    double* h_first = s_dyn;        // [2][T]   chain: first slip of each thread
    double* h_last = s_dyn + 2 * T; // [2][T]   chain: last slip of each thread
    double* mir = s_dyn;            // [2][N]   lattice / long range: all slips
    double* pref = s_dyn + 2 * N;   // [N]      long range prefactor

    double u[PPT], v[PPT], a[PPT], yl[PPT], yr[PPT], fth[PPT];
    int32_t jl[PPT], dl[PPT];

    int64_t inc = rs->inc;
    double uf = rs->u_frame;
    int32_t below = rs->below, below2 = rs->below2, dec = rs->dec;
    double prev = rs->prev;
    int64_t s_n = rs->s_n;
    int32_t act_init = rs->act_init;
    int64_t qs_first = rs->qs_first, qs_last = rs->qs_last;
    uint64_t rng = rs->rng_state;
    const uint64_t rng_inc = rs->rng_inc;
    int64_t min_next = rs->min_next;

    // quasistatic driver: event, phase, sum of dl when the event began, steps
    // of the current minimisation before this entry, the iteration at which
    // it (re)started in this entry, and whether a refill aligns at u
    int64_t ev = 0, isum = 0, ev_done0 = 0;
    int32_t phase = QS_MINIMISE, refill_u = 0;
    int it0 = 1;
    if constexpr (QS) {
        ev = rs->qs_event;
        phase = rs->qs_phase;
        isum = rs->qs_isum;
        ev_done0 = rs->steps;
    }
    int64_t launch_done = 0; // steps of the earlier entries of this launch
    uint32_t rmask = 0;      // this thread's particles that need a refill (bit i: particle base + i)
    int32_t status;
    int it;

    for (;;) { // entries: the first, then one after each in-kernel refill
        // steps this entry may take: what is left of the launch budget, capped by
        // the run's target (the quasistatic driver: max_iter is per minimisation)
        const int64_t budget = ra.max_steps - launch_done;
        const int64_t room = QS ? budget : rs->target - rs->steps - launch_done;
        const int limit = (int)(room < budget ? (room > 0 ? room : 0) : budget);

#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            u[i] = v[i] = a[i] = yl[i] = yr[i] = fth[i] = 0.0;
            jl[i] = 0;
            dl[i] = 0;
            if (own(i)) {
                const int64_t p = base + i;
                u[i] = S.u[p];
                v[i] = S.v[p];
                a[i] = S.a[p];
                jl[i] = L.jloc[p];
                yl[i] = Y[p * nch + jl[i]];
                yr[i] = Y[p * nch + jl[i] + 1];
                if constexpr (EXT) {
                    fth[i] = S.f_th[p];
                }
                if (track) {
                    dl[i] = (int32_t)(L.start[p] + jl[i] - S.i_ref[p]);
                }
            }
        }

        const double dt = S.dt;
        const double c2 = 0.5 * dt * dt;
        const double hdt = 0.5 * dt;
        const double meta = -S.eta;
        const double inv_m = S.inv_m;
        const double kf = S.k_frame;
        const double dv = ra.v_frame * dt;

        // Publish the slips that other threads read: the current ones (`next`
        // false) or those after the position update of the coming step, computed
        // without touching the registers.
        auto publish = [&](int b, bool next) {
            if constexpr (CHAIN) {
                if (cnt > 0) {
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        if (i == 0 || i == last) {
                            const double x = next ? u[i] + dt * v[i] + c2 * a[i] : u[i];
                            if (i == 0) {
                                h_first[b * T + t] = x;
                            }
                            if (i == last) {
                                h_last[b * T + t] = x;
                            }
                        }
                    }
                }
            }
            else {
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    if (own(i)) {
                        mir[(int64_t)b * N + base + i] = next ? u[i] + dt * v[i] + c2 * a[i] : u[i];
                    }
                }
            }
        };

        if constexpr (INT == INT_LONGRANGE1D) {
            for (int64_t q = t; q < N; q += T) {
                pref[q] = S.pref[q];
            }
        }
        publish(0, false);
        if (t == 0) {
            s_nan[0] = 0;
            s_nan[1] = 0;
            s_refill[0] = 0;
            s_refill[1] = 0;
        }


        if constexpr (EXT) {
            long long mn = 0x7fffffffffffffffLL;
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (own(i)) {
                    long long nx = S.next[base + i];
                    mn = nx < mn ? nx : mn;
                }
            }
            min_next = uniform_ll(block_min_i(mn, s_min_w));
        }
        __syncthreads();

        auto left_of = [&](int b) -> double { return h_last[b * T + (t == 0 ? Tact - 1 : t - 1)]; };
        auto right_of = [&](int b) -> double { return h_first[b * T + (t == Tact - 1 ? 0 : t + 1)]; };

        // interaction force of local particle i, neighbours from exchange buffer b
        auto interaction = [&](int i, int b) -> double {
            if constexpr (INT == INT_NONE) {
                return 0.0;
            }
            else if constexpr (CHAIN) {
                double um = 0.0, up = 0.0;
#pragma unroll
                for (int q = 0; q < PPT; ++q) {
                    if (q == i) {
                        um = (q == 0) ? left_of(b) : u[q > 0 ? q - 1 : 0];
                        up = (q == last) ? right_of(b) : u[q + 1 < PPT ? q + 1 : q];
                    }
                }
                return chain_force<INT>(S, um, u[i], up);
            }
            else if constexpr (INT == INT_LONGRANGE1D) {
                const double* U = mir + (int64_t)b * N;
                return longrange_force(
                    S, [&](int64_t q) { return U[q]; }, [&](int64_t d) { return pref[d]; }, base + i);
            }
            else {
                const double* U = mir + (int64_t)b * N;
                const int64_t p = base + i;
                return lattice_force<INT>(S, [&](int64_t q) { return U[q]; }, p / S.cols, p % S.cols);
            }
        };

        // quasistatic driver ------------------------------------------------------
        // block sums of (xi, xd) in a fixed order; two barriers
        auto qs_sum = [&](long long xi, double xd, long long& ri, double& rd) __attribute__((always_inline)) {
            xi = wave_sum_i(xi);
            xd = wave_sum(xd);
            if (lane == 0) {
                s_qs_i[wave] = xi;
                s_qs_d[wave] = xd;
            }
            __syncthreads();
            long long ti = 0;
            double td = 0.0;
            for (int w = 0; w < nw; ++w) {
                ti += s_qs_i[w];
                td += s_qs_d[w];
            }
            __syncthreads();
            ri = uniform_ll(ti);
            rd = uniform_d(td);
        };

        // block minimum of xd and OR of xi; two barriers
        auto qs_min = [&](long long xi, double xd, long long& ri, double& rd) __attribute__((always_inline)) {
            for (int off = 32; off > 0; off >>= 1) {
                const double y = __shfl_xor(xd, off, 64);
                xd = y < xd ? y : xd;
                xi |= __shfl_xor(xi, off, 64);
            }
            if (lane == 0) {
                s_qs_i[wave] = xi;
                s_qs_d[wave] = xd;
            }
            __syncthreads();
            long long ti = s_qs_i[0];
            double td = s_qs_d[0];
            for (int w = 1; w < nw; ++w) {
                ti |= s_qs_i[w];
                td = s_qs_d[w] < td ? s_qs_d[w] : td;
            }
            __syncthreads();
            ri = uniform_ll(ti);
            rd = uniform_d(td);
        };

        // advanceUniformly(du, input_is_frame = false): u += du, u_frame += du
        // (k_frame + mu) / k_frame, and each particle's well found by a walk inside
        // its window; returns true if some particle left the window (the state is
        // then left for k_refill to align at u)
        auto qs_advance = [&](double du) __attribute__((always_inline)) -> bool {
            uint32_t need = 0;
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (own(i)) {
                    u[i] = u[i] + du;
                    const double x = u[i];
                    if (!(yl[i] <= x && x < yr[i]) && __builtin_isfinite(x)) {
                        const double* row = Y + (base + i) * nch;
                        int32_t j = jl[i];
                        double lo = yl[i], hi = yr[i];
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
                            dl[i] += j - jl[i];
                            jl[i] = j;
                            yl[i] = lo;
                            yr[i] = hi;
                        }
                        else {
                            need |= 1u << i;
                        }
                    }
                }
            }
            uf = uf + du * (kf + S.mu) / kf;
            rmask = need;
            long long any;
            double unused;
            qs_sum(need != 0, 0.0, any, unused);
            return any != 0;
        };

        // From equilibrium (phase QS_RECORD) or an event's start (QS_BEGIN): record
        // finished events and apply the displacements of the next ones until a
        // minimisation starts (ST_RUNNING, counters reset), the run is complete
        // (ST_DONE) or a displacement needs a refill (ST_REFILL).
        auto qs_next = [&]() __attribute__((always_inline)) -> int32_t {
            for (;;) {
                if (phase == QS_RECORD) {
                    long long sd = 0;
                    double sfr = 0.0;
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        if (own(i)) {
                            sd += dl[i];
                            sfr += kf * (uf - u[i]);
                        }
                    }
                    long long tot;
                    double tfr;
                    qs_sum(sd, sfr, tot, tfr);
                    if (t == 0) {
                        const int64_t o = (int64_t)blockIdx.x * qa.nevents + ev;
                        if (qa.out_S) {
                            qa.out_S[o] = tot - isum;
                        }
                        if (qa.out_uf) {
                            qa.out_uf[o] = uf;
                        }
                        if (qa.out_ff) {
                            qa.out_ff[o] = tfr / (double)N;
                        }
                        if (qa.out_inc) {
                            qa.out_inc[o] = inc;
                        }
                    }
                    ++ev;
                    phase = QS_BEGIN;
                }
                if (ev >= qa.nevents) {
                    return ST_DONE;
                }
                // QS_BEGIN
                {
                    long long sd = 0;
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        if (own(i)) {
                            sd += dl[i];
                        }
                    }
                    long long tot;
                    double unused;
                    qs_sum(sd, 0.0, tot, unused);
                    isum = tot;
                }
                const int64_t k = qa.first + ev;
                bool move = false;
                double du = 0.0;
                if (k == 0) { // initial quench: u_frame = 0, then minimise
                    uf = 0.0;
                    phase = QS_MINIMISE;
                }
                else if ((k & 1) && !qa.kicks) { // elastic: eventDrivenStep(eps, false)
                    // (Smooth has no maxUniformDisplacement: the host refuses
                    // elastic events for it, detail.h:414-422)
                    if constexpr (POT != POT_SMOOTH) {
                        double m = __builtin_inf();
                        int outside = 0;
#pragma unroll
                        for (int i = 0; i < PPT; ++i) {
                            if (own(i)) {
                                const double x =
                                    max_uniform_particle<POT>(S, u[i], yl[i], yr[i], qa.direction, outside);
                                m = x < m ? x : m;
                            }
                        }
                        long long out;
                        qs_min(outside, m, out, m);
                        if (out) {
                            m = 0.0;
                        }
                        if (!(m < 0.5 * qa.eps)) {
                            du = qa.direction > 0 ? m - 0.5 * qa.eps : 0.5 * qa.eps - m;
                            move = true;
                        }
                    }
                    phase = QS_RECORD;
                }
                else { // kick: eventDrivenStep(eps, true), then minimise
                    du = qa.direction > 0 ? qa.eps : -qa.eps;
                    move = true;
                    phase = QS_MINIMISE;
                }
                if (phase == QS_MINIMISE) {
                    below = 0;
                    below2 = 0;
                    dec = 0;
                    prev = __builtin_inf();
                    ev_done0 = 0;
                }
                if (move && qs_advance(du)) {
                    refill_u = 1;
                    return ST_REFILL;
                }
                if (phase == QS_MINIMISE) {
                    return ST_RUNNING;
                }
            }
        };

        status = ST_RUNNING;
        it = 1;
        it0 = 1;
        // quasistatic driver: an entry that starts outside a minimisation goes
        // through the transition at its first iteration
        bool qs_pending = QS && phase != QS_MINIMISE;
        for (;;) { // the quasistatic driver's minimisations (one pass otherwise)
            for (;; ++it) {
                const int b = it & 1;
                // A: publish the slips after the position update of step `it`, and
                // check (read only) that every particle's new well is reachable inside
                // its window; otherwise step `it` waits for an in-launch refill
                publish(b, true);
                {
                    uint32_t refill = 0; // particles whose next well is outside the safe zone
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        if (own(i)) {
                            const double x = u[i] + dt * v[i] + c2 * a[i];
                            if (!(yl[i] <= x && x < yr[i]) && __builtin_isfinite(x)) {
                                if (!reachable(Y + (base + i) * nch, jl[i], yl[i], yr[i], x, jmin, jmax)) {
                                    refill |= 1u << i;
                                }
                            }
                        }
                    }
                    rmask = refill;
                    if (refill) {
                        s_refill[b] = it;
                    }
                }
                __syncthreads();

                // B: decisions on step it-1
                if (it > it0 && uniform_i(s_nan[(it - 1) & 1]) == it - 1) {
                    status = ST_NAN;
                    break;
                }
                if (track && it > it0) {
                    const int pb = (it - 1) & 1;
                    double sf = 0.0, sff = 0.0;
                    long long nchg = 0, sabs = 0;
                    for (int w = 0; w < nw; ++w) {
                        sf += s_red_d[pb][w][0];
                        sff += s_red_d[pb][w][1];
                    }
                    if (counts) {
                        for (int w = 0; w < nw; ++w) {
                            nchg += s_red_i[pb][w][0];
                            sabs += s_red_i[pb][w][1];
                        }
                    }
                    sf = uniform_d(sf);
                    sff = uniform_d(sff);
                    nchg = uniform_ll(nchg);
                    sabs = uniform_ll(sabs);
                    if (ra.mode == MODE_UNTIL_EVENT && nchg > 0) {
                        status = ST_EVENT;
                    }
                    else {
                        const double rf = __builtin_sqrt(sf), rff = __builtin_sqrt(sff);
                        const double r = (rff != 0.0) ? rf / rff : rf;
                        below = !(r >= ra.tol) ? below + 1 : 0;
                        below2 = !(r >= ra.tol2) ? below2 + 1 : 0;
                        dec = !(r >= prev) ? dec + 1 : 0;
                        prev = r;
                        if (ra.mode == MODE_TRUNCATE || (ra.mode == MODE_MINIMISE && ra.time_activity)) {
                            if (sabs != s_n) {
                                if (act_init) {
                                    act_init = 0;
                                    qs_first = inc;
                                }
                                qs_last = inc;
                            }
                            s_n = sabs;
                        }
                        const int64_t nt = ra.niter_tol;
                        const bool conv = nt <= 0 || (dec >= nt - 1 && below >= nt) || below2 >= nt;
                        if (conv) {
                            status = ST_CONVERGED;
                        }
                        else if (ra.mode == MODE_TRUNCATE) {
                            if (ra.A_truncate > 0 && nchg >= ra.A_truncate) {
                                status = ST_TRUNCATED;
                            }
                            else if (ra.S_truncate > 0 && sabs >= ra.S_truncate) {
                                status = ST_TRUNCATED;
                            }
                        }
                        if constexpr (QS) {
                            if (status == ST_RUNNING && ev_done0 + (it - it0) >= qa.max_iter) {
                                status = ST_MAXITER;
                            }
                        }
                    }
                }
                if (status != ST_RUNNING || it > limit || qs_pending) {
                    break;
                }
                if (uniform_i(s_refill[b]) == it) {
                    status = ST_REFILL;
                    break;
                }

                // C: step `it`
                inc += 1;
                if (ra.flow) {
                    uf = uf + dv;
                }
                if constexpr (EXT) {
                    if (inc >= min_next) {
                        int due = 0;
#pragma unroll
                        for (int i = 0; i < PPT; ++i) {
                            if (own(i) && inc >= S.next[base + i]) {
                                ++due;
                            }
                        }
                        int total = 0;
                        int rank = block_excl_scan(due, s_scan_w, total);
                        if (due > 0) {
                            uint64_t st = pcg_advance(rng, rng_inc, 2 * (int64_t)rank);
#pragma unroll
                            for (int i = 0; i < PPT; ++i) {
                                if (own(i)) {
                                    const int64_t p = base + i;
                                    int64_t nx = S.next[p];
                                    if (inc >= nx) {
                                        fth[i] = draw_normal(st, rng_inc, S.mean, S.stddev);
                                        S.next[p] = nx + S.dinc[p];
                                    }
                                }
                            }
                        }
                        long long mn = 0x7fffffffffffffffLL;
#pragma unroll
                        for (int i = 0; i < PPT; ++i) {
                            if (own(i)) {
                                long long nx = S.next[base + i];
                                mn = nx < mn ? nx : mn;
                            }
                        }
                        rng = pcg_advance(rng, rng_inc, 2 * (int64_t)uniform_i(total));
                        min_next = uniform_ll(block_min_i(mn, s_min_w));
                    }
                }

                // position update and well search inside the window (phase A made
                // sure that the walk stays inside it; a non-finite slip keeps its well)
                int nanf = 0;
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    if (own(i)) {
                        u[i] = u[i] + dt * v[i] + c2 * a[i];
                        const double x = u[i];
                        nanf |= __builtin_isnan(x);
                        if (!(yl[i] <= x && x < yr[i])) {
                            const double* row = Y + (base + i) * nch;
                            int32_t j = jl[i];
                            double lo = yl[i], hi = yr[i];
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
                                dl[i] += j - jl[i];
                                jl[i] = j;
                                yl[i] = lo;
                                yr[i] = hi;
                            }
                        }
                    }
                }
                if (nanf) {
                    s_nan[b] = it;
                }

                double sf = 0.0, sff = 0.0;
                int nchg = 0, sabs = 0;
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    if (own(i)) {
                        const double fi = interaction(i, b);
                        const double fp = potential_force<POT>(S, u[i], yl[i], yr[i]);
                        const double ff = kf * (uf - u[i]);
                        const double sum3 = ff + fp + fi;
                        const double vn = v[i], an = a[i];
                        double vv = vn + dt * an;
                        double f = EXT ? sum3 + meta * vv + fth[i] : sum3 + meta * vv;
                        double aa = f * inv_m;
                        vv = vn + hdt * (an + aa);
                        f = EXT ? sum3 + meta * vv + fth[i] : sum3 + meta * vv;
                        aa = f * inv_m;
                        vv = vn + hdt * (an + aa);
                        f = EXT ? sum3 + meta * vv + fth[i] : sum3 + meta * vv;
                        aa = f * inv_m;
                        v[i] = vv;
                        a[i] = aa;
                        if (track) {
                            sf += f * f;
                            sff += ff * ff;
                        }
                        if (counts) {
                            nchg += dl[i] != 0;
                            sabs += dl[i] < 0 ? -dl[i] : dl[i];
                        }
                    }
                }
                if (track) {
                    sf = wave_sum(sf);
                    sff = wave_sum(sff);
                    if (lane == 0) {
                        s_red_d[b][wave][0] = sf;
                        s_red_d[b][wave][1] = sff;
                    }
                }
                if (counts) {
                    const long long n1 = wave_sum_i(nchg), n2 = wave_sum_i(sabs);
                    if (lane == 0) {
                        s_red_i[b][wave][0] = n1;
                        s_red_i[b][wave][1] = n2;
                    }
                }
            }
            if constexpr (QS) {
                if (status == ST_CONVERGED || qs_pending) {
                    if (!qs_pending) {
                        // quench (detail.h:1782), then record the event
#pragma unroll
                        for (int i = 0; i < PPT; ++i) {
                            v[i] = 0.0;
                            a[i] = 0.0;
                        }
                        phase = QS_RECORD;
                    }
                    qs_pending = false;
                    status = qs_next();
                    if (phase == QS_MINIMISE) {
                        it0 = it; // a new minimisation starts at iteration `it`
                    }
                    if (status == ST_RUNNING) {
                        // phase A's speculative slips of iteration `it` are stale
                        if (t == 0) {
                            s_refill[it & 1] = 0;
                        }
                        __syncthreads();
                        continue;
                    }
                }
            }
            break;
        }

        const int done = it - 1; // completed steps
        if (status == ST_CONVERGED && !QS) {
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                v[i] = 0.0;
                a[i] = 0.0;
            }
        }
        if (status == ST_RUNNING) {
            status = ST_BUDGET;
        }

        // write back the state and the forces of the final configuration; the
        // exchange buffer done&1 holds its slips (published at the start of the
        // iteration after the last completed step, or initially), except after a
        // displacement of the quasistatic driver: publish them again
        const int fb = done & 1;
        const bool forces = status != ST_NAN && status != ST_REFILL;
        if constexpr (QS) {
            if (forces) {
                publish(fb, false);
                __syncthreads();
            }
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (own(i)) {
                const int64_t p = base + i;
                S.u[p] = u[i];
                S.v[p] = v[i];
                S.a[p] = a[i];
                L.jloc[p] = jl[i];
                if (forces) {
                    const double fi = interaction(i, fb);
                    const double fp = potential_force<POT>(S, u[i], yl[i], yr[i]);
                    const double ff = kf * (uf - u[i]);
                    const double fd = meta * v[i];
                    S.f_int[p] = fi;
                    S.f_pot[p] = fp;
                    S.f_frame[p] = ff;
                    S.f_damp[p] = fd;
                    S.f[p] = EXT ? ff + fp + fi + fd + fth[i] : ff + fp + fi + fd;
                }
                if constexpr (EXT) {
                    S.f_th[p] = fth[i];
                }
            }
        }
        if constexpr (QS) {
            ev_done0 = phase == QS_MINIMISE ? ev_done0 + (it - it0) : 0;
        }
        launch_done += done;
        if (status != ST_REFILL) {
            break;
        }
        // refill from the written-back state, then reload it: the LDS of this
        // entry must no longer be read when the next one publishes
        __syncthreads(); // the state of every lane's particles is in memory
        refill_wave(views, base, refill_u, rmask);
        refill_u = 0;
        __syncthreads();
    }

    if (t == 0) {
        rs->inc = inc;
        rs->u_frame = uf;
        if constexpr (QS) {
            rs->steps = ev_done0;
            rs->qs_event = ev;
            rs->qs_phase = phase;
            rs->qs_isum = isum;
        }
        else {
            rs->steps += launch_done;
        }
        rs->status = status;
        rs->below = below;
        rs->below2 = below2;
        rs->dec = dec;
        rs->prev = prev;
        rs->s_n = s_n;
        rs->act_init = act_init;
        rs->qs_first = qs_first;
        rs->qs_last = qs_last;
        rs->rng_state = rng;
        rs->min_next = min_next;
    }
}

template <int POT, int INT, int EXT, int PPT, int MAXT>
__global__ void __launch_bounds__(MAXT) k_resident(const SysView* __restrict__ views, RunArgs ra)
{
    resident_body<POT, INT, EXT, PPT, MAXT, false>(views, ra, QsArgs{});
}

// Quasistatic driver: the events of QsArgs for every system, one workgroup
// per system, relaunched when its step budget runs out.
// FULL: every system of the launch has N divisible by PPT.
template <int POT, int INT, int PPT, int MAXT, bool FULL = false>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(MAXT <= 256 && PPT <= 8 ? 2 : 1)))
k_quasi(const SysView* __restrict__ views, RunArgs ra, QsArgs qa)
{
    resident_body<POT, INT, 0, PPT, MAXT, true, FULL>(views, ra, qa);
}

} // namespace fq
