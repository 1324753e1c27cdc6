// Resident kernel for timeSteps / flowSteps of chains (MODE_STEPS only):
//   timeStep            detail.h:1539-1570
//   timeSteps/flowSteps detail.h:1577-1583, 1637-1645
//   RandomNormalForcing::force detail.h:930-943         (EXT = 1)
// Same ownership, expressions and results as k_resident (fq_resident.hpp),
// bit for bit, with less work per particle-step:
//  * software-pipelined: the slips of step it+1, u + dt v + dt^2/2 a, are
//    computed once at the end of step it (into separate registers, so a stop
//    before step it+1 discards them), checked for reachability and published
//    to the LDS exchange buffer (it+1)&1 for the neighbours; step it+1 then
//    commits them. One __syncthreads per step, at the top of the step.
//  * the mid-point of the current well, 0.5 (y_l + y_r), is kept in a
//    register (recomputed when the well changes) for the Cuspy force;
//  * with m = 1 the multiplication by 1/m (exact) is skipped;
//  * no event / residual bookkeeping;
//  * FULL (every system of the launch has N divisible by PPT, so a thread
//    holds PPT particles or none): the thread's last particle is register
//    PPT - 1 at compile time, which removes a select per register and step
//    from the interaction and makes the publication of the last slip one
//    store.
#pragma once

#include "fq_resident.hpp"

#include <type_traits>

namespace fq {

template <int POT>
FQ_HD double potential_force_mid(const SysView& S, double u, double yl, double yr, double ym)
{
    if constexpr (POT == POT_CUSPY) {
        return (ym - u) * S.mu; // (0.5 * (yl + yr) - u) * mu with the mid-point cached
    }
    else {
        return potential_force<POT>(S, u, yl, yr);
    }
}

template <int POT, int INT, int EXT, int PPT, int MAXT, bool UNIT, bool FULL = false>
__global__ void __launch_bounds__(MAXT) k_flow(const SysView* __restrict__ views, RunArgs ra)
{
    static_assert(is_chain<INT>(), "k_flow handles chains");
    const SysView& S = views[blockIdx.x];
    // the system's parameters as a local copy: registers for the whole launch
    // instead of a reload of the ones the forces use after every barrier
    const SysView P = S;
    RunState* const rs = S.rs;
    if (rs->status != ST_RUNNING && rs->status != ST_BUDGET) {
        return;
    }
    const int64_t room = rs->target - rs->steps;
    const int limit = uniform_i((int)(room < ra.max_steps ? (room > 0 ? room : 0) : ra.max_steps));

    __shared__ int s_nan[2];
    __shared__ int s_refill[2];
    __shared__ int s_scan_w[RES_MAX_WAVES];
    __shared__ long long s_min_w[RES_MAX_WAVES];
    extern __shared__ double s_dyn[];

    const int t = threadIdx.x, T = blockDim.x;
    const int64_t N = S.N;
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(N - base >= PPT ? PPT : (N - base > 0 ? N - base : 0));
    const int Tact = (int)((N + PPT - 1) / PPT);
    // register of the thread's last particle (FULL: cnt is PPT or 0, and a
    // thread without particles neither publishes nor stores)
    const int last = FULL ? PPT - 1 : cnt - 1;
    const Landscape& L = S.ls;
    double* const Y = L.y;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);
    double* h_first = s_dyn;        // [2][T] first slip of each thread
    double* h_last = s_dyn + 2 * T; // [2][T] last slip of each thread

    // Slips live in two register arrays used in turn: one holds the state of
    // the current step, the other the slips prepared for the next step (and,
    // before they are committed, the state of the previous one). The step
    // loop is unrolled by two so that the arrays swap roles without copies.
    // Registers i >= cnt are padding: zero state (reset after every step) in
    // an unbounded well, so they never walk, never request a refill, never
    // set the NaN flag and never publish; the hot loops run over all PPT
    // registers without per-particle branches and only the stores are
    // guarded. The reset runs only in wavefronts that hold padding (a scalar
    // branch: a full wavefront skips it).
    double sa[PPT], sb[PPT], v[PPT], a[PPT], yl[PPT], yr[PPT], ym[PPT], fth[PPT];
    int32_t jl[PPT];
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        sa[i] = sb[i] = v[i] = a[i] = ym[i] = fth[i] = 0.0;
        yl[i] = -__builtin_inf();
        yr[i] = __builtin_inf();
        jl[i] = 0;
        if (i < cnt) {
            const int64_t p = base + i;
            sa[i] = S.u[p];
            v[i] = S.v[p];
            a[i] = S.a[p];
            jl[i] = L.jloc[p];
            yl[i] = Y[p * nch + jl[i]];
            yr[i] = Y[p * nch + jl[i] + 1];
            ym[i] = 0.5 * (yl[i] + yr[i]);
            if constexpr (EXT) {
                fth[i] = S.f_th[p];
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

    if (t == 0) {
        s_nan[0] = 0;
        s_nan[1] = 0;
        s_refill[0] = 0;
        s_refill[1] = 0;
    }
    int64_t inc = uniform_ll(rs->inc); // scalar: its increment is a scalar add
    double uf = rs->u_frame;
    const bool pad_wave = uniform_i(__any(cnt < PPT)) != 0;
    uint64_t rng = rs->rng_state;
    const uint64_t rng_inc = rs->rng_inc;
    int64_t min_next = rs->min_next;
    if constexpr (EXT) {
        long long mn = 0x7fffffffffffffffLL;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                long long nx = S.next[base + i];
                mn = nx < mn ? nx : mn;
            }
        }
        min_next = uniform_ll(block_min_i(mn, s_min_w));
    }
    __syncthreads(); // flags cleared before any thread sets them

    // publish the first and last value of x (registers) into exchange buffer
    // b; the last one is stored under a predicate per register (an indexed
    // read x[cnt - 1] would move the array to scratch memory)
    auto publish = [&](const double(&x)[PPT], int b) {
        if (cnt > 0) {
            h_first[b * T + t] = x[0];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (i == last) {
                    h_last[b * T + t] = x[i];
                }
            }
        }
    };

    // slips of step `it` from the state x into y, their reachability, and
    // publication into exchange buffer it&1; pend: some slip of this thread
    // leaves its well (or is NaN), so that the commit of the step walks and
    // checks only then
    int pend = 0;
    auto prepare = [&](int it, const double(&x)[PPT], double(&y)[PPT]) {
        const int b = it & 1;
        int out = 0;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            y[i] = x[i] + dt * v[i] + c2 * a[i];
            out |= !(yl[i] <= y[i] && y[i] < yr[i]);
        }
        pend = out;
        if (out) { // rare: some slip leaves its well
            int refill = 0;
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                // opaque copy: the comparisons of this rare path stay in it
                // instead of being evaluated for every step
                double z = y[i];
                __asm__ volatile("" : "+v"(z));
                if (!(yl[i] <= z && z < yr[i]) && __builtin_isfinite(z)) {
                    refill |= !reachable(Y + (base + i) * nch, jl[i], yl[i], yr[i], z, jmin, jmax);
                }
            }
            if (refill) {
                s_refill[b] = it;
            }
        }
        publish(y, b);
    };

    auto left_of = [&](int b) -> double { return h_last[b * T + (t == 0 ? Tact - 1 : t - 1)]; };
    auto right_of = [&](int b) -> double { return h_first[b * T + (t == Tact - 1 ? 0 : t + 1)]; };
    auto interaction = [&](int i, int b, const double(&x)[PPT]) -> double {
        if constexpr (INT == INT_NONE) {
            return 0.0;
        }
        else {
            double um = 0.0, up = 0.0;
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                if (q == i) {
                    um = (q == 0) ? left_of(b) : x[q > 0 ? q - 1 : 0];
                    up = (q == last) ? right_of(b) : x[q + 1 < PPT ? q + 1 : q];
                }
            }
            return chain_force<INT>(P, um, x[i], up);
        }
    };

    int32_t status = ST_RUNNING;

    // step `it` with the prepared slips in x (the previous state in y, which
    // receives the slips of step it+1); false if the run stops before it
    auto step = [&](int it, double(&x)[PPT], double(&y)[PPT]) -> bool {
        const int b = it & 1;
        __syncthreads(); // slips of step `it` published, flags of step it-1 set

        if (it > 1 && uniform_i(s_nan[(it - 1) & 1]) == it - 1) {
            status = ST_NAN;
            return false;
        }
        if (it > limit) {
            return false;
        }
        if (uniform_i(s_refill[b]) == it) {
            status = ST_REFILL;
            return false;
        }

        inc += 1;
        if (ra.flow) {
            uf = uf + dv;
        }
        if constexpr (EXT) {
            if (inc >= min_next) {
                int due = 0;
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    if (i < cnt && inc >= S.next[base + i]) {
                        ++due;
                    }
                }
                int total = 0;
                int rank = block_excl_scan(due, s_scan_w, total);
                if (due > 0) {
                    uint64_t st = pcg_advance(rng, rng_inc, 2 * (int64_t)rank);
#pragma unroll
                    for (int i = 0; i < PPT; ++i) {
                        if (i < cnt) {
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
                    if (i < cnt) {
                        long long nx = S.next[base + i];
                        mn = nx < mn ? nx : mn;
                    }
                }
                rng = pcg_advance(rng, rng_inc, 2 * (int64_t)uniform_i(total));
                min_next = uniform_ll(block_min_i(mn, s_min_w));
            }
        }

        // commit: the slips in x are the state now; walk to their wells
        // (reachable: checked when they were prepared; a non-finite slip
        // keeps its well)
        if (pend) { // rare
            int nanf = 0;
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const double z = x[i];
                nanf |= __builtin_isnan(z);
                if (!(yl[i] <= z && z < yr[i])) {
                    const double* row = Y + (base + i) * nch;
                    int32_t j = jl[i];
                    double lo = yl[i], hi = yr[i];
                    while (z >= hi && j < jmax) {
                        ++j;
                        lo = hi;
                        hi = row[j + 1];
                    }
                    while (z < lo && j > jmin) {
                        --j;
                        hi = lo;
                        lo = row[j];
                    }
                    if (lo <= z && z < hi) {
                        jl[i] = j;
                        yl[i] = lo;
                        yr[i] = hi;
                        ym[i] = 0.5 * (lo + hi);
                    }
                }
            }
            if (nanf) {
                s_nan[b] = it;
            }
        }

        // the reference's three velocity / acceleration passes (detail.h:1552-1565)
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const double fi = interaction(i, b, x);
            const double fp = potential_force_mid<POT>(P, x[i], yl[i], yr[i], ym[i]);
            const double ff = kf * (uf - x[i]);
            const double sum3 = ff + fp + fi;
            const double vn = v[i], an = a[i];
            double vv = vn + dt * an;
            double f = EXT ? sum3 + meta * vv + fth[i] : sum3 + meta * vv;
            double aa = UNIT ? f : f * inv_m;
            vv = vn + hdt * (an + aa);
            f = EXT ? sum3 + meta * vv + fth[i] : sum3 + meta * vv;
            aa = UNIT ? f : f * inv_m;
            vv = vn + hdt * (an + aa);
            f = EXT ? sum3 + meta * vv + fth[i] : sum3 + meta * vv;
            aa = UNIT ? f : f * inv_m;
            v[i] = vv;
            a[i] = aa;
        }
        if (pad_wave) { // padding registers stay at rest (FULL: all of them)
            // (the empty asm keeps this branch apart from the per-lane one
            // below, so that it stays a scalar branch)
            __asm__ volatile("");
            if (cnt < PPT) {
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    if (FULL || i >= cnt) {
                        v[i] = 0.0;
                        a[i] = 0.0;
                    }
                }
            }
        }
        prepare(it + 1, x, y);
        return true;
    };

    // state and forces of the final configuration (slips in x, `done` steps
    // taken); its slips are in exchange buffer done&1 (published before step
    // `done`, or initially)
    auto finish = [&](const double(&x)[PPT], int done) {
        if (status == ST_RUNNING) {
            status = ST_BUDGET;
        }
        const int fb = done & 1;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const int64_t p = base + i;
                S.u[p] = x[i];
                S.v[p] = v[i];
                S.a[p] = a[i];
                L.jloc[p] = jl[i];
                if (status != ST_NAN) {
                    const double fi = interaction(i, fb, x);
                    const double fp = potential_force_mid<POT>(S, x[i], yl[i], yr[i], ym[i]);
                    const double ff = kf * (uf - x[i]);
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
        if (t == 0) {
            rs->inc = inc;
            rs->u_frame = uf;
            rs->steps += done;
            rs->status = status;
            rs->rng_state = rng;
            rs->min_next = min_next;
        }
    };

    // the current slips into buffer 0 (forces of the final state when no step
    // is taken), the slips of step 1 into buffer 1
    publish(sa, 0);
    prepare(1, sa, sb);
    for (int it = 1;; it += 2) {
        if (!step(it, sb, sa)) {
            finish(sa, it - 1);
            return;
        }
        if (!step(it + 1, sa, sb)) {
            finish(sb, it);
            return;
        }
    }
}

} // namespace fq
