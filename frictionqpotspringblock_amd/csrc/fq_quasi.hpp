// Quasistatic driver for chains: the examples' event loop
// (examples/Line1d_Cuspy_Laplace.py:39-61: eventDrivenStep, detail.h:1933-1960,
// then minimise, detail.h:1676-1792) inside one workgroup per system, as
// k_quasi (fq_resident.hpp) runs it for every interaction, with the same
// ownership, expressions, sums and decisions, so the same results bit for
// bit, and the minimisation's steps in the form of k_flow (fq_flow.hpp):
//  * software-pipelined: the slips of step it+1 are computed once at the end
//    of step it (into the second register array), checked for reachability
//    and published for the neighbours; step it+1 commits them. The loop is
//    unrolled by two so that the arrays swap roles without copies (with 4 or
//    8 particles per thread one array: the committed state goes to LDS, read
//    back only when the entry stops, and the next slips replace it). One
//    __syncthreads per step, at its top, after which the StopList decision on
//    step it-1 is taken from the wavefronts' residual partial sums;
//  * the Cuspy mid-point of the current well kept in a register;
//  * with m = 1 the multiplication by 1/m (exact) is skipped;
//  * padding registers (i >= cnt) at rest in an unbounded well: no
//    per-particle branches in the step, the sums masked.
// The event protocol (records, elastic and kick displacements, quench), the
// in-launch refills and the step budget are those of k_quasi. Between
// minimisations the protocol works on the state in memory: the entry whose
// minimisation converged writes the state back, the events run on it, and a
// new entry reloads the registers, so that neither the step loop's registers
// nor the protocol's arguments are live across each other's code.
#pragma once

#include "fq_flow.hpp"

namespace fq {

template <int POT, int INT, int PPT, int MAXT, bool UNIT, bool FULL>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(MAXT <= 256 && PPT <= 8 ? 2 : 1)))
k_quasi_chain(const SysView* __restrict__ views, RunArgs ra, QsArgs qa)
{
    static_assert(is_chain<INT>(), "k_quasi_chain handles chains");
    const SysView& S = views[blockIdx.x];
    // the system's parameters as a local copy (no reload after every barrier)
    const SysView P = S;
    RunState* const rs = S.rs;
    if (rs->status != ST_RUNNING && rs->status != ST_BUDGET) {
        return;
    }

    __shared__ double s_red_d[2][RES_MAX_WAVES][2];
    __shared__ int s_nan[2];
    __shared__ int s_refill[2];
    __shared__ long long s_qs_i[RES_MAX_WAVES];
    __shared__ double s_qs_d[RES_MAX_WAVES];
    __shared__ QsArgs s_qa; // the protocol's arguments, read where the events use them
    extern __shared__ double s_dyn[];

    const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6, nw = T >> 6;
    const int64_t N = S.N;
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(N - base >= PPT ? PPT : (N - base > 0 ? N - base : 0));
    const int Tact = (int)((N + PPT - 1) / PPT);
    const int last = FULL ? PPT - 1 : cnt - 1;
    auto own = [&](int i) -> bool { return FULL ? cnt > 0 : i < cnt; };
    const Landscape& L = S.ls;
    double* const Y = L.y;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);
    double* h_first = s_dyn;        // [2][T] first slip of each thread
    double* h_last = s_dyn + 2 * T; // [2][T] last slip of each thread

    const double dt = S.dt;
    const double c2 = 0.5 * dt * dt;
    const double hdt = 0.5 * dt;
    const double meta = -S.eta;
    const double inv_m = S.inv_m;
    const double kf = S.k_frame;
    const bool pad_wave = uniform_i(__any(cnt < PPT)) != 0;
    if (t == 0) {
        s_qa = qa;
    }
    // The system's view as re-read by the paths outside the step loop
    // (entries, events, write-back): an opaque pointer, so that the array
    // pointers are loaded where these paths use them instead of being kept
    // in scalar registers through the step loop.
    auto V = [&]() -> const SysView& {
        const SysView* q = views + blockIdx.x;
        __asm__ volatile("" : "+s"(q));
        return *q;
    };

    int64_t inc = uniform_ll(rs->inc);
    double uf = rs->u_frame;
    int32_t below = rs->below, below2 = rs->below2, dec = rs->dec;
    double prev = rs->prev;
    // event, phase, sum of the well index changes when the event began, steps
    // of the current minimisation before this entry
    int64_t ev = rs->qs_event;
    int32_t phase = rs->qs_phase;
    int64_t isum = rs->qs_isum;
    int64_t ev_done0 = rs->steps;
    int32_t refill_u = 0;    // the pending refill aligns at u (a displacement)
    int64_t launch_done = 0; // steps of the earlier entries of this launch
    uint32_t rmask = 0;      // this thread's particles that need a refill (bit i)
    int32_t status = ST_RUNNING;
    int limit = 0;
    int64_t maxit = 0; // steps the entry's minimisation may take before max_iter

    // sa / sb: the state's slips and the slips prepared for the next step, in
    // turn; registers i >= cnt are padding
    // INPLACE (4 or 8 particles per thread, up to 512 threads, for the LDS):
    // one slip array; the state of the last committed step is kept in this
    // thread's slots of LDS (read back only when the entry stops) and the
    // next slips are prepared in place: no second register array (8
    // particles per thread: no spills) and one step per loop iteration
    constexpr bool INPLACE = quasi_inplace(PPT, MAXT);
    double sa[PPT], sb[INPLACE ? 1 : PPT], v[PPT], a[PPT], ym[PPT];
    int32_t jl[PPT];
    double* const s_keep = s_dyn + 4 * T; // INPLACE: [PPT][T] the state of the last committed step
    // the bounds (y_l, y_r) of each register's well and, for up to 4 particles
    // per thread (NEXT; 8 would spill), the right bound of the next well to
    // the right (yield position j + 2; +inf past the window's end), read when
    // the well changes: its load is not waited for until a later crossing
    // uses it, so that a move of one well to the right (most moves of an
    // avalanche under forward driving) reads no memory
    constexpr bool NEXT = PPT <= 4;
    double yl_r[PPT], yr_r[PPT], yn_r[NEXT ? PPT : 1];
    auto next_right = [&](const double* row, int32_t j) -> double {
        return j + 2 < nch ? row[j + 2] : __builtin_inf();
    };
    auto bounds = [&](int i, double& lo, double& hi) {
        lo = yl_r[i];
        hi = yr_r[i];
    };
    auto set_bounds = [&](int i, double lo, double hi) {
        yl_r[i] = lo;
        yr_r[i] = hi;
    };

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

    // slips of step `it` from the state x into y, their reachability (rmask,
    // s_refill) and publication into exchange buffer it&1
    int pend = 0;
    auto prepare = [&](int it, const double(&x)[PPT], double(&y)[PPT]) {
        const int b = it & 1;
        int out = 0;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if constexpr (INPLACE) {
                s_keep[i * T + t] = x[i];
            }
            y[i] = x[i] + dt * v[i] + c2 * a[i];
            double lo, hi;
            bounds(i, lo, hi);
            out |= !(lo <= y[i] && y[i] < hi);
        }
        pend = out;
        uint32_t refill = 0;
        if (out) { // rare: some slip leaves its well
            // opaque copy of the first index: the window addresses of this
            // rare path are computed in it, not kept live through the loop
            int64_t bq = base;
            __asm__ volatile("" : "+v"(bq));
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                double z = y[i];
                __asm__ volatile("" : "+v"(z));
                double lo, hi;
                bounds(i, lo, hi);
                if (!(lo <= z && z < hi) && __builtin_isfinite(z)) {
                    // one well to the right: reachable (the walk's first move)
                    const bool near = NEXT && z >= hi && z < yn_r[NEXT ? i : 0] && jl[i] < jmax;
                    if (!near && !reachable(Y + (bq + i) * nch, jl[i], lo, hi, z, jmin, jmax)) {
                        refill |= 1u << i;
                    }
                }
            }
            if (refill) {
                s_refill[b] = it;
            }
        }
        rmask = refill;
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

    // The event protocol between minimisations runs on the state in memory
    // (written back by the entry that ended; the next entry reloads it), each
    // thread on its own particles p = base + i, i < cnt, in register order,
    // so the registers of the step loop are not live across it.

    // advanceUniformly(du, input_is_frame = false): true if some particle left
    // its window (rmask; the state is then refilled at u)
    auto qs_advance = [&](double du) __attribute__((always_inline)) -> bool {
        const SysView& W = V();
        uint32_t need = 0;
#pragma unroll 1
        for (int i = 0; i < cnt; ++i) {
            const int64_t p = base + i;
            const double z = W.u[p] + du;
            W.u[p] = z;
            const double* row = Y + p * nch;
            int32_t j = W.ls.jloc[p];
            double lo = row[j], hi = row[j + 1];
            if (!(lo <= z && z < hi) && __builtin_isfinite(z)) {
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
                    W.ls.jloc[p] = j;
                }
                else {
                    need |= 1u << i;
                }
            }
        }
        uf = uf + du * (kf + P.mu) / kf;
        rmask = need;
        long long any;
        double unused;
        qs_sum(need != 0, 0.0, any, unused);
        return any != 0;
    };

    // sum over the thread's particles of (well index - i_ref)
    auto well_sum = [&]() -> long long {
        const SysView& W = V();
        long long sd = 0;
#pragma unroll 1
        for (int i = 0; i < cnt; ++i) {
            const int64_t p = base + i;
            sd += (int32_t)(W.ls.start[p] + W.ls.jloc[p] - W.i_ref[p]);
        }
        return sd;
    };

    // From equilibrium (QS_RECORD) or an event's start (QS_BEGIN): record
    // finished events and apply the displacements of the next ones until a
    // minimisation starts (ST_RUNNING, counters reset), the run is complete
    // (ST_DONE) or a displacement needs a refill (ST_REFILL).
    auto qs_next = [&]() -> int32_t {
        const SysView& W = V();
        for (;;) {
            if (phase == QS_RECORD) {
                double sfr = 0.0;
#pragma unroll 1
                for (int i = 0; i < cnt; ++i) {
                    sfr += kf * (uf - W.u[base + i]);
                }
                long long tot;
                double tfr;
                qs_sum(well_sum(), sfr, tot, tfr);
                if (t == 0) {
                    const int64_t o = (int64_t)blockIdx.x * s_qa.nevents + ev;
                    if (s_qa.out_S) {
                        s_qa.out_S[o] = tot - isum;
                    }
                    if (s_qa.out_uf) {
                        s_qa.out_uf[o] = uf;
                    }
                    if (s_qa.out_ff) {
                        s_qa.out_ff[o] = tfr / (double)N;
                    }
                    if (s_qa.out_inc) {
                        s_qa.out_inc[o] = inc;
                    }
                }
                ++ev;
                phase = QS_BEGIN;
            }
            if (ev >= s_qa.nevents) {
                return ST_DONE;
            }
            {
                long long tot;
                double unused;
                qs_sum(well_sum(), 0.0, tot, unused);
                isum = tot;
            }
            const int64_t k = s_qa.first + ev;
            bool move = false;
            double du = 0.0;
            if (k == 0) { // initial quench: u_frame = 0, then minimise
                uf = 0.0;
                phase = QS_MINIMISE;
            }
            else if ((k & 1) && !s_qa.kicks) { // elastic: eventDrivenStep(eps, false)
                if constexpr (POT != POT_SMOOTH) {
                    double m = __builtin_inf();
                    int outside = 0;
#pragma unroll 1
                    for (int i = 0; i < cnt; ++i) {
                        const int64_t p = base + i;
                        const int32_t j = W.ls.jloc[p];
                        const double z = max_uniform_particle<POT>(P, W.u[p], Y[p * nch + j], Y[p * nch + j + 1],
                                                                   s_qa.direction, outside);
                        m = z < m ? z : m;
                    }
                    long long out;
                    qs_min(outside, m, out, m);
                    if (out) {
                        m = 0.0;
                    }
                    if (!(m < 0.5 * s_qa.eps)) {
                        du = s_qa.direction > 0 ? m - 0.5 * s_qa.eps : 0.5 * s_qa.eps - m;
                        move = true;
                    }
                }
                phase = QS_RECORD;
            }
            else { // kick: eventDrivenStep(eps, true), then minimise
                du = s_qa.direction > 0 ? s_qa.eps : -s_qa.eps;
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

    // step `it` of the entry's minimisation with the prepared slips in x and
    // the state in y (y receives the slips of step it+1); false if the entry
    // stops before it, with the state in y: a NaN, the step budget, max_iter,
    // a refill, or convergence
    auto step = [&](int it, double(&x)[PPT], double(&y)[PPT]) -> bool {
        const int b = it & 1;
        __syncthreads(); // slips of step `it` published; flags and residual sums of step it-1 written

        // decisions on step it-1 (detail.h:1771-1787 with the StopList as counters)
        if (it > 1) {
            if (uniform_i(s_nan[(it - 1) & 1]) == it - 1) {
                status = ST_NAN;
                return false;
            }
            const int pb = (it - 1) & 1;
            double sf = 0.0, sff = 0.0;
            for (int w = 0; w < nw; ++w) {
                sf += s_red_d[pb][w][0];
                sff += s_red_d[pb][w][1];
            }
            sf = uniform_d(sf);
            sff = uniform_d(sff);
            const double rf = __builtin_sqrt(sf), rff = __builtin_sqrt(sff);
            const double r = (rff != 0.0) ? rf / rff : rf;
            below = !(r >= ra.tol) ? below + 1 : 0;
            below2 = !(r >= ra.tol2) ? below2 + 1 : 0;
            dec = !(r >= prev) ? dec + 1 : 0;
            prev = r;
            const int64_t nt = ra.niter_tol;
            if (nt <= 0 || (dec >= nt - 1 && below >= nt) || below2 >= nt) {
                status = ST_CONVERGED;
                return false;
            }
            if (it - 1 >= maxit) {
                status = ST_MAXITER;
                return false;
            }
        }
        if (it > limit) {
            return false;
        }
        if (uniform_i(s_refill[b]) == it) {
            status = ST_REFILL;
            return false;
        }

        inc += 1;
        // commit: walk to the wells of the slips in x (reachable: checked when
        // they were prepared; a non-finite slip keeps its well)
        if (pend) { // rare
            int nanf = 0;
            int64_t bq = base; // (opaque, as in prepare)
            __asm__ volatile("" : "+v"(bq));
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const double z = x[i];
                nanf |= __builtin_isnan(z);
                double lo, hi;
                bounds(i, lo, hi);
                if (!(lo <= z && z < hi)) {
                    const double* row = Y + (bq + i) * nch;
                    if (NEXT && z >= hi && z < yn_r[NEXT ? i : 0] && jl[i] < jmax) { // one well to the right
                        const int32_t j = jl[i] + 1;
                        const double hn = yn_r[NEXT ? i : 0];
                        jl[i] = j;
                        set_bounds(i, hi, hn);
                        ym[i] = 0.5 * (hi + hn);
                        yn_r[NEXT ? i : 0] = next_right(row, j);
                    }
                    else {
                        int32_t j = jl[i];
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
                            set_bounds(i, lo, hi);
                            ym[i] = 0.5 * (lo + hi);
                            if constexpr (NEXT) {
                                yn_r[i] = next_right(row, j);
                            }
                        }
                    }
                }
            }
            if (nanf) {
                s_nan[b] = it;
            }
        }

        // the reference's three velocity / acceleration passes (detail.h:1552-1565)
        // and the residual's two squared norms (own particles, in register order)
        double sf = 0.0, sff = 0.0;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const double fi = interaction(i, b, x);
            double lo, hi;
            bounds(i, lo, hi);
            const double fp = potential_force_mid<POT>(P, x[i], lo, hi, ym[i]);
            const double ff = kf * (uf - x[i]);
            const double sum3 = ff + fp + fi;
            const double vn = v[i], an = a[i];
            double vv = vn + dt * an;
            double f = sum3 + meta * vv;
            double aa = UNIT ? f : f * inv_m;
            vv = vn + hdt * (an + aa);
            f = sum3 + meta * vv;
            aa = UNIT ? f : f * inv_m;
            vv = vn + hdt * (an + aa);
            f = sum3 + meta * vv;
            aa = UNIT ? f : f * inv_m;
            v[i] = vv;
            a[i] = aa;
            if (own(i)) {
                sf += f * f;
                sff += ff * ff;
            }
        }
        sf = wave_sum(sf);
        sff = wave_sum(sff);
        if (lane == 0) {
            s_red_d[b][wave][0] = sf;
            s_red_d[b][wave][1] = sff;
        }
        if (pad_wave) { // padding registers stay at rest
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

    // write back the state x after `done` steps of this entry, with the forces
    // of that configuration unless a NaN or a refill stopped the entry
    auto finish = [&](double(&x)[PPT], int done) {
        const SysView& W = V();
        if (status == ST_RUNNING) {
            status = ST_BUDGET;
        }
        const int fb = done & 1;
        const bool forces = status != ST_NAN && status != ST_REFILL;
        if (forces) { // (the slips of x are in buffer fb unless no step was taken)
            publish(x, fb);
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (own(i)) {
                const int64_t p = base + i;
                W.u[p] = x[i];
                W.v[p] = v[i];
                W.a[p] = a[i];
                W.ls.jloc[p] = jl[i];
                if (forces) {
                    const double fi = interaction(i, fb, x);
                    double lo, hi;
                    bounds(i, lo, hi);
                    const double fp = potential_force_mid<POT>(P, x[i], lo, hi, ym[i]);
                    const double ff = kf * (uf - x[i]);
                    const double fd = meta * v[i];
                    W.f_int[p] = fi;
                    W.f_pot[p] = fp;
                    W.f_frame[p] = ff;
                    W.f_damp[p] = fd;
                    W.f[p] = ff + fp + fi + fd;
                }
            }
        }
    };

    // Entries: each loads the state, minimises (when the protocol is inside a
    // minimisation) and ends with the state written back: at convergence (or
    // at the start of an entry outside a minimisation) the protocol's next
    // events run on memory and a new entry follows; after a refill as well;
    // when the run is complete, a last entry writes the forces of the final
    // state; the step budget, max_iter and a NaN end the launch.
    bool complete = false;
    for (;;) {
        const int64_t budget = ra.max_steps - launch_done;
        limit = uniform_i((int)(budget > 0 ? budget : 0));
        const SysView& W = V();
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            sa[i] = v[i] = a[i] = ym[i] = 0.0;
            if constexpr (!INPLACE) {
                sb[i] = 0.0;
            }
            set_bounds(i, -__builtin_inf(), __builtin_inf());
            if constexpr (NEXT) {
                yn_r[i] = __builtin_inf();
            }
            jl[i] = 0;
            if (own(i)) {
                const int64_t p = base + i;
                sa[i] = W.u[p];
                v[i] = W.v[p];
                a[i] = W.a[p];
                jl[i] = W.ls.jloc[p];
                const double lo = Y[p * nch + jl[i]], hi = Y[p * nch + jl[i] + 1];
                set_bounds(i, lo, hi);
                ym[i] = 0.5 * (lo + hi);
                if constexpr (NEXT) {
                    yn_r[i] = next_right(Y + p * nch, jl[i]);
                }
            }
        }
        if (t == 0) {
            s_nan[0] = 0;
            s_nan[1] = 0;
            s_refill[0] = 0;
            s_refill[1] = 0;
        }
        status = complete ? ST_DONE : ST_RUNNING;
        __syncthreads(); // flags cleared before any thread sets them
        publish(sa, 0);
        int it = 1;
        const bool minimising = !complete && phase == QS_MINIMISE;
        maxit = s_qa.max_iter - ev_done0;
        if (minimising) {
            if constexpr (INPLACE) {
                prepare(1, sa, sa);
                for (;; ++it) {
                    if (!step(it, sa, sa)) {
                        break;
                    }
                }
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    sa[i] = s_keep[i * T + t];
                }
            }
            else {
                prepare(1, sa, sb);
                for (;; it += 2) { // steps in pairs, the register arrays in turn; the state ends in sa
                    if (!step(it, sb, sa)) {
                        break;
                    }
                    if (!step(it + 1, sa, sb)) {
                        ++it;
#pragma unroll
                        for (int i = 0; i < PPT; ++i) {
                            sa[i] = sb[i];
                        }
                        break;
                    }
                }
            }
        }
        const int done = it - 1;
        launch_done += done;
        if (status == ST_CONVERGED || (!complete && !minimising)) {
            if (status == ST_CONVERGED) { // quench (detail.h:1782), then record the event
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    v[i] = 0.0;
                    a[i] = 0.0;
                }
                phase = QS_RECORD;
            }
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (own(i)) {
                    const int64_t p = base + i;
                    W.u[p] = sa[i];
                    W.v[p] = v[i];
                    W.a[p] = a[i];
                    W.ls.jloc[p] = jl[i];
                }
            }
            status = qs_next();
            if (status == ST_DONE) {
                complete = true;
                continue;
            }
            if (status == ST_RUNNING) {
                continue;
            }
        }
        else {
            finish(sa, done);
            ev_done0 = phase == QS_MINIMISE ? ev_done0 + done : 0;
            if (status != ST_REFILL) {
                break;
            }
        }
        // refill from the written-back state, then reload it
        __syncthreads();
        refill_wave(views, base, refill_u, rmask);
        refill_u = 0;
        __syncthreads();
    }

    if (t == 0) {
        rs->inc = inc;
        rs->u_frame = uf;
        rs->steps = ev_done0;
        rs->qs_event = ev;
        rs->qs_phase = phase;
        rs->qs_isum = isum;
        rs->status = status;
        rs->below = below;
        rs->below2 = below2;
        rs->dec = dec;
        rs->prev = prev;
    }
}

} // namespace fq
