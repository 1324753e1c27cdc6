// Streaming Velocity-Verlet kernel: one workgroup per system for systems that
// do not fit a resident workgroup (any N, any interaction). The state is
// double-buffered in HBM (buffer k&1 holds the state after step k), so a step
// never has to be undone: a stop decision or a pending refill simply leaves the
// last complete buffer as the result. Same modes, StopList, thermal draws and
// refill-by-relaunch protocol as k_resident (fq_resident.hpp); the same
// per-particle expressions (fq_physics.hpp), so results are bit-identical.
//
// Per step, two block-stride passes separated by barriers:
//   1. u_new = u + dt v + dt^2/2 a, well walk inside the window (or a refill
//      request), written to the other buffer;
//   2. interactions from the new slips, potential, frame, the three
//      velocity/acceleration passes of timeStep (detail.h:1552-1565) and the
//      per-thread partial norms.
// Roofline: HBM / L2, ~104 B per particle-step (44 B of state read and
// written in pass 1, neighbours, v, a in pass 2).
#pragma once

#include "fq_resident.hpp"

namespace fq {

constexpr int STREAM_THREADS = 1024;

template <int POT, int INT, int EXT>
__global__ void __launch_bounds__(STREAM_THREADS) k_stream(const SysView* __restrict__ views, RunArgs ra)
{
    const SysView& S = views[blockIdx.x];
    RunState* const rs = S.rs;
    if (rs->status != ST_RUNNING && rs->status != ST_BUDGET) {
        return;
    }
    const int64_t room = rs->target - rs->steps;
    const int limit = (int)(room < ra.max_steps ? (room > 0 ? room : 0) : ra.max_steps);

    __shared__ double s_red_d[2][RES_MAX_WAVES][2];
    __shared__ long long s_red_i[2][RES_MAX_WAVES][2];
    __shared__ int s_nan[2];
    __shared__ int s_refill[2];
    __shared__ int s_scan_w[RES_MAX_WAVES];
    __shared__ long long s_min_w[RES_MAX_WAVES];

    const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6, nw = T >> 6;
    const int64_t N = S.N;
    const Landscape& L = S.ls;
    const bool track = ra.mode != MODE_STEPS;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);

    double* const U[2] = {S.u, S.u_b};
    double* const V[2] = {S.v, S.v_b};
    double* const A[2] = {S.a, S.a_b};
    double* const YL[2] = {S.yl_a, S.yl_b};
    double* const YR[2] = {S.yr_a, S.yr_b};
    int32_t* const J[2] = {L.jloc, S.jl_b};

    const double dt = S.dt;
    const double c2 = 0.5 * dt * dt;
    const double hdt = 0.5 * dt;
    const double meta = -S.eta;
    const double inv_m = S.inv_m;
    const double kf = S.k_frame;
    const double dv = ra.v_frame * dt;

    // the well of every particle from its window
    for (int64_t p = t; p < N; p += T) {
        const int32_t j = J[0][p];
        YL[0][p] = L.y[p * nch + j];
        YR[0][p] = L.y[p * nch + j + 1];
    }
    if (t == 0) {
        s_nan[0] = s_nan[1] = 0;
        s_refill[0] = s_refill[1] = 0;
    }

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
    if constexpr (EXT) {
        long long mn = 0x7fffffffffffffffLL;
        for (int64_t p = t; p < N; p += T) {
            long long nx = S.next[p];
            mn = nx < mn ? nx : mn;
        }
        min_next = uniform_ll(block_min_i(mn, s_min_w));
    }
    __syncthreads();

    auto interaction = [&](const double* Un, int64_t p) -> double {
        if constexpr (INT == INT_NONE) {
            return 0.0;
        }
        else if constexpr (is_chain<INT>()) {
            return chain_force<INT>(S, Un[p == 0 ? N - 1 : p - 1], Un[p], Un[p == N - 1 ? 0 : p + 1]);
        }
        else if constexpr (INT == INT_LONGRANGE1D) {
            return longrange_force(
                S, [&](int64_t q) { return Un[q]; }, [&](int64_t d) { return S.pref[d]; }, p);
        }
        else {
            return lattice_force<INT>(S, [&](int64_t q) { return Un[q]; }, p / S.cols, p % S.cols);
        }
    };

    int32_t status = ST_RUNNING;
    int completed = 0;
    for (int it = 1;; ++it) {
        const int cb = (it - 1) & 1, nb = it & 1;

        // decisions on step it-1
        if (it > 1) {
            if (uniform_i(s_nan[cb]) == it - 1) {
                status = ST_NAN;
            }
            else if (track) {
                double sf = 0.0, sff = 0.0;
                long long nchg = 0, sabs = 0;
                for (int w = 0; w < nw; ++w) {
                    sf += s_red_d[cb][w][0];
                    sff += s_red_d[cb][w][1];
                    nchg += s_red_i[cb][w][0];
                    sabs += s_red_i[cb][w][1];
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
                    if (nt <= 0 || (dec >= nt - 1 && below >= nt) || below2 >= nt) {
                        status = ST_CONVERGED;
                    }
                    else if (ra.mode == MODE_TRUNCATE) {
                        if ((ra.A_truncate > 0 && nchg >= ra.A_truncate) ||
                            (ra.S_truncate > 0 && sabs >= ra.S_truncate)) {
                            status = ST_TRUNCATED;
                        }
                    }
                }
            }
        }
        if (status != ST_RUNNING || it > limit) {
            completed = it - 1;
            break;
        }

        // pass 1: position update and well walk into buffer nb
        int refill = 0, nanf = 0;
        for (int64_t p = t; p < N; p += T) {
            const double x = U[cb][p] + dt * V[cb][p] + c2 * A[cb][p];
            nanf |= __builtin_isnan(x);
            int32_t j = J[cb][p];
            double lo = YL[cb][p], hi = YR[cb][p];
            if (!(lo <= x && x < hi) && __builtin_isfinite(x)) {
                const double* row = L.y + p * nch;
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
                refill |= !(lo <= x && x < hi);
            }
            U[nb][p] = x;
            J[nb][p] = j;
            YL[nb][p] = lo;
            YR[nb][p] = hi;
        }
        if (refill) {
            s_refill[nb] = it;
        }
        if (nanf) {
            s_nan[nb] = it;
        }
        __syncthreads();
        if (uniform_i(s_refill[nb]) == it) {
            status = ST_REFILL;
            completed = it - 1;
            break;
        }

        // pass 2: the forces of the new configuration and the velocity passes
        inc += 1;
        if (ra.flow) {
            uf = uf + dv;
        }
        if constexpr (EXT) {
            if (inc >= min_next) {
                int64_t before = 0;
                long long mn = 0x7fffffffffffffffLL;
                for (int64_t b0 = 0; b0 < N; b0 += T) {
                    const int64_t p = b0 + t;
                    const int due = (p < N && inc >= S.next[p]) ? 1 : 0;
                    int total = 0;
                    const int rank = block_excl_scan(due, s_scan_w, total);
                    if (due) {
                        uint64_t st = pcg_advance(rng, rng_inc, 2 * (before + rank));
                        S.f_th[p] = draw_normal(st, rng_inc, S.mean, S.stddev);
                        S.next[p] += S.dinc[p];
                    }
                    if (p < N) {
                        long long nx = S.next[p];
                        mn = nx < mn ? nx : mn;
                    }
                    before += uniform_i(total);
                }
                rng = pcg_advance(rng, rng_inc, 2 * before);
                min_next = uniform_ll(block_min_i(mn, s_min_w));
            }
        }
        double sf = 0.0, sff = 0.0;
        long long nchg = 0, sabs = 0;
        const double* Un = U[nb];
        for (int64_t p = t; p < N; p += T) {
            const double x = Un[p];
            const double fi = interaction(Un, p);
            const double fp = potential_force<POT>(S, x, YL[nb][p], YR[nb][p]);
            const double ff = kf * (uf - x);
            const double sum3 = ff + fp + fi;
            const double vn = V[cb][p], an = A[cb][p];
            const double fth = EXT ? S.f_th[p] : 0.0;
            double vv = vn + dt * an;
            double f = EXT ? sum3 + meta * vv + fth : sum3 + meta * vv;
            double aa = f * inv_m;
            vv = vn + hdt * (an + aa);
            f = EXT ? sum3 + meta * vv + fth : sum3 + meta * vv;
            aa = f * inv_m;
            vv = vn + hdt * (an + aa);
            f = EXT ? sum3 + meta * vv + fth : sum3 + meta * vv;
            aa = f * inv_m;
            V[nb][p] = vv;
            A[nb][p] = aa;
            if (track) {
                const int64_t dl = L.start[p] + J[nb][p] - S.i_ref[p];
                sf += f * f;
                sff += ff * ff;
                nchg += dl != 0;
                sabs += dl < 0 ? -dl : dl;
            }
        }
        if (track) {
            sf = wave_sum(sf);
            sff = wave_sum(sff);
            nchg = wave_sum_i(nchg);
            sabs = wave_sum_i(sabs);
            if (lane == 0) {
                s_red_d[nb][wave][0] = sf;
                s_red_d[nb][wave][1] = sff;
                s_red_i[nb][wave][0] = nchg;
                s_red_i[nb][wave][1] = sabs;
            }
        }
        __syncthreads();
    }

    // final state: buffer completed&1 (copied into the primary arrays) and
    // the forces of that configuration
    const int fb = completed & 1;
    if (status == ST_RUNNING) {
        status = ST_BUDGET;
    }
    __syncthreads();
    const double* Uf = U[fb];
    for (int64_t p = t; p < N; p += T) {
        const double x = Uf[p];
        double vv = V[fb][p], aa = A[fb][p];
        if (status == ST_CONVERGED) {
            vv = 0.0;
            aa = 0.0;
        }
        if (status != ST_NAN) {
            const double fi = interaction(Uf, p);
            const double fp = potential_force<POT>(S, x, YL[fb][p], YR[fb][p]);
            const double ff = kf * (uf - x);
            const double fd = meta * vv;
            S.f_int[p] = fi;
            S.f_pot[p] = fp;
            S.f_frame[p] = ff;
            S.f_damp[p] = fd;
            S.f[p] = EXT ? ff + fp + fi + fd + S.f_th[p] : ff + fp + fi + fd;
        }
        S.v[p] = vv;
        S.a[p] = aa;
        if (fb == 1) {
            L.jloc[p] = J[1][p];
        }
    }
    __syncthreads();
    if (fb == 1) {
        for (int64_t p = t; p < N; p += T) {
            S.u[p] = Uf[p];
        }
    }
    if (t == 0) {
        rs->inc = inc;
        rs->u_frame = uf;
        rs->steps += completed;
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

} // namespace fq
