// Discrete adjoint of System::timeStep (detail.h:1539-1570) for Cuspy +
// Laplace1d replicas (no reference counterpart). With the well index of every
// particle taken as a constant at every step (the indices the trajectory
// visited), one step is the map (u, v, a) -> (u1, v3, a3):
//   u1 = u + dt v + dt^2/2 a
//   g  = k_frame (u_frame - u1) + mu (ym - u1) + k L u1     ym = (y_l + y_r) / 2
//   v1 = v + dt a,               a1 = (g - eta v1) / m
//   v2 = v + dt/2 (a + a1),      a2 = (g - eta v2) / m
//   v3 = v + dt/2 (a + a2),      a3 = (g - eta v3) / m
// (L: periodic second difference). The adjoint runs the step backwards for the
// cotangents of (u1, v3, a3) and accumulates the derivatives with respect to
// mu, k, k_frame, eta and the yield positions y_l, y_r of the visited wells.
//
// A run of n steps is split into windows of K steps. For each window, last
// to first: k_adj_record re-runs the window from its checkpoint (the same
// expressions in the same order as the resident kernel, so the same
// trajectory and the same wells) and writes a tape of u1, v and a of every
// step and the well used by it; k_adj_reverse then walks the tape backwards,
// loading the tape of each step while it finishes the one after it. One
// workgroup per replica, PPT consecutive particles per thread, neighbours
// exchanged through LDS; the tape is stored so that each of a wavefront's
// tape loads and stores is contiguous (AdjBuffers).
#pragma once

#include "fq_physics.hpp"
#include "fq_resident.hpp"

namespace fq {


// Record windows w0 + blockIdx.y (each of min(K, n - w K) steps) from their
// start states, one workgroup per replica and window: every window is
// independent given its checkpoint, so all of a group run at once.
template <int PPT, int MAXT>
__global__ void __launch_bounds__(MAXT) k_adj_record(const SysView* __restrict__ views, AdjBuffers ab, int flow,
                                                     double v_frame)
{
    __shared__ double h_first[2][MAXT];
    __shared__ double h_last[2][MAXT];
    const int r = blockIdx.x;
    const int64_t w = ab.w0 + blockIdx.y;
    const int steps = (int)(ab.n - w * ab.K < ab.K ? ab.n - w * ab.K : ab.K);
    // this window's start state and tape
    ab.wu += blockIdx.y * ab.ws;
    ab.wv += blockIdx.y * ab.ws;
    ab.wa += blockIdx.y * ab.ws;
    ab.wj += blockIdx.y * ab.ws;
    ab.wuf += blockIdx.y * ab.R;
    ab.tu += blockIdx.y * ab.ts;
    ab.tv += blockIdx.y * ab.ts;
    ab.ta += blockIdx.y * ab.ts;
    ab.tj += blockIdx.y * ab.ts;
    ab.tuf += blockIdx.y * ab.K * ab.R;
    const SysView& S = views[r];
    const int t = threadIdx.x;
    const int64_t N = ab.N;
    const int64_t off = (int64_t)r * N;
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(N - base >= PPT ? PPT : (N - base > 0 ? N - base : 0));
    const int Tact = (int)((N + PPT - 1) / PPT);
    const Landscape& L = S.ls;
    const double* const Y = L.y;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);
    // tape entry of register 0 of this thread in step 0 (register i: + 64 i)
    const int64_t Rnp = ab.R * ab.np;
    const int64_t tq = (int64_t)r * ab.np + (int64_t)(t >> 6) * 64 * PPT + (t & 63);

    double u[PPT], v[PPT], a[PPT], yl[PPT], yr[PPT];
    int32_t jl[PPT];
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        u[i] = v[i] = a[i] = yl[i] = yr[i] = 0.0;
        jl[i] = 0;
        if (i < cnt) {
            const int64_t p = base + i;
            u[i] = ab.wu[off + p];
            v[i] = ab.wv[off + p];
            a[i] = ab.wa[off + p];
            jl[i] = ab.wj[off + p];
            yl[i] = Y[p * nch + jl[i]];
            yr[i] = Y[p * nch + jl[i] + 1];
        }
    }
    const double dt = S.dt, c2 = 0.5 * dt * dt, hdt = 0.5 * dt, meta = -S.eta, inv_m = S.inv_m, kf = S.k_frame;
    const double dv = v_frame * dt;
    double uf = ab.wuf[r];
    int bad = 0;

    for (int n = 0; n < steps; ++n) {
        const int b = n & 1;
        if (flow) {
            uf = uf + dv;
        }
        const int64_t tn = (int64_t)n * Rnp + tq;
        if (t == 0) {
            ab.tuf[(int64_t)n * ab.R + r] = uf;
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const int64_t p = base + i;
                const int64_t q = tn + 64 * i;
                ab.tv[q] = v[i];
                ab.ta[q] = a[i];
                u[i] = u[i] + dt * v[i] + c2 * a[i];
                ab.tu[q] = u[i];
                const double x = u[i];
                if (!(yl[i] <= x && x < yr[i])) {
                    const double* row = Y + p * nch;
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
                    bad |= !(lo <= x && x < hi);
                    jl[i] = j;
                    yl[i] = lo;
                    yr[i] = hi;
                }
                ab.tj[q] = jl[i];
            }
        }
        if (cnt > 0) {
            h_first[b][t] = u[0];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (i == cnt - 1) { // (an indexed read would move the array to scratch)
                    h_last[b][t] = u[i];
                }
            }
        }
        __syncthreads();
        const double left = h_last[b][t == 0 ? Tact - 1 : t - 1];
        const double right = h_first[b][t == Tact - 1 ? 0 : t + 1];
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const double um = i == 0 ? left : u[i > 0 ? i - 1 : 0];
                const double up = i == cnt - 1 ? right : u[i + 1 < PPT ? i + 1 : i];
                const double fi = chain_force<INT_LAPLACE1D>(S, um, u[i], up);
                const double fp = potential_force<POT_CUSPY>(S, u[i], yl[i], yr[i]);
                const double ff = kf * (uf - u[i]);
                const double sum3 = ff + fp + fi;
                const double vn = v[i], an = a[i];
                double vv = vn + dt * an;
                double f = sum3 + meta * vv;
                double aa = f * inv_m;
                vv = vn + hdt * (an + aa);
                f = sum3 + meta * vv;
                aa = f * inv_m;
                vv = vn + hdt * (an + aa);
                f = sum3 + meta * vv;
                aa = f * inv_m;
                v[i] = vv;
                a[i] = aa;
            }
        }
    }
    if (bad) {
        *ab.err = 1;
    }
}

// Walk the tape of one window backwards, updating the adjoint state and the
// parameter / yield-position derivatives. Since dg/dy_l = dg/dy_r = mu/2, one
// accumulator per particle collects the derivative for both yield positions
// of its current well; it is added to the output when the well changes.
template <int PPT, int MAXT>
__global__ void __launch_bounds__(MAXT) k_adj_reverse(const SysView* __restrict__ views, AdjBuffers ab, int steps)
{
    // neighbour exchange: slips u1 (h_*) and cotangents of g (g_*) of a step;
    // two barriers per step (a buffer is rewritten only after the barrier
    // that follows every read of it)
    __shared__ double h_first[MAXT];
    __shared__ double h_last[MAXT];
    __shared__ double g_first[MAXT];
    __shared__ double g_last[MAXT];
    __shared__ double s_red[4][MAXT / 64];
    const int r = blockIdx.x;
    const SysView& S = views[r];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = blockDim.x >> 6;
    const int64_t N = ab.N;
    const int64_t off = (int64_t)r * N;
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(N - base >= PPT ? PPT : (N - base > 0 ? N - base : 0));
    const int Tact = (int)((N + PPT - 1) / PPT);
    const double* const Y = S.ls.y;
    const int64_t nch = S.ls.nchunk;
    const int64_t Rnp = ab.R * ab.np;
    const int64_t tq = (int64_t)r * ab.np + (int64_t)(t >> 6) * 64 * PPT + (t & 63);

    const double dt = S.dt, c2 = 0.5 * dt * dt, hdt = 0.5 * dt, meta = -S.eta, w = S.inv_m, kf = S.k_frame;
    const double mu = S.mu, k = S.k;
    const double mweta = meta * w; // -eta / m
    const double hmu = 0.5 * mu;

    // U, V, A: cotangents of u, v, a after the step being reversed
    // jc, ymc: the well of the current tape step and its mid-point (read from
    // the chunk only when the well changes)
    double U[PPT], V[PPT], A[PPT], acc[PPT], ymc[PPT];
    int32_t jc[PPT];
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        U[i] = V[i] = A[i] = acc[i] = ymc[i] = 0.0;
        jc[i] = -1;
        if (i < cnt) {
            const int64_t p = base + i;
            U[i] = ab.au[off + p];
            V[i] = ab.av[off + p];
            A[i] = ab.aa[off + p];
        }
    }
    double gmu = 0.0, gk = 0.0, gkf = 0.0, geta = 0.0;

    // the tape of the step being reversed (u1, v, a, well, frame position),
    // loaded while the step before it in the walk is finished: u1 as soon as
    // it has been read (at the top of the step), the rest once the first
    // half of the step has used it
    double pu[PPT], pv[PPT], pa[PPT];
    int32_t pj[PPT];
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        pu[i] = pv[i] = pa[i] = 0.0;
        pj[i] = 0;
    }
    double puf = 0.0;
    auto load_u1 = [&](int n) {
        const int64_t tn = (int64_t)n * Rnp + tq;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                pu[i] = ab.tu[tn + 64 * i];
            }
        }
    };
    auto load_rest = [&](int n) {
        const int64_t tn = (int64_t)n * Rnp + tq;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const int64_t q = tn + 64 * i;
                pv[i] = ab.tv[q];
                pa[i] = ab.ta[q];
                pj[i] = ab.tj[q];
            }
        }
        puf = ab.tuf[(int64_t)n * ab.R + r];
    };
    if (steps > 0) {
        load_u1(steps - 1);
        load_rest(steps - 1);
    }
    // every load above complete: the loop's wait counts then concern only
    // the tape loads of the step before, which stay in flight until the top
    // of the next iteration
    __builtin_amdgcn_s_waitcnt(0);

    auto flush = [&](int i) {
        if (ab.gy && jc[i] >= 0) {
            double* row = ab.gy + (off + base + i) * nch;
            row[jc[i]] += acc[i];
            row[jc[i] + 1] += acc[i];
        }
        acc[i] = 0.0;
    };

    for (int n = steps - 1; n >= 0; --n) {
        const double uf = puf;
        double x[PPT]; // u1 of the step, then the cotangent of g
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            x[i] = pu[i];
        }
        if (cnt > 0) {
            h_first[t] = x[0];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (i == cnt - 1) { // (an indexed read would move the array to scratch)
                    h_last[t] = x[i];
                }
            }
        }
        __syncthreads();
        const double left = h_last[t == 0 ? Tact - 1 : t - 1];
        const double right = h_first[t == Tact - 1 ? 0 : t + 1];
        if (n > 0) {
            load_u1(n - 1);
        }
        double xm = left; // u1 of the particle before i (its original value)
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const int64_t p = base + i;
                const double u1 = x[i];
                const double up = i == cnt - 1 ? right : x[i + 1 < PPT ? i + 1 : i];
                const double lu = xm - 2 * u1 + up;
                xm = u1;
                const int32_t j = pj[i];
                if (j != jc[i]) {
                    flush(i);
                    jc[i] = j;
                    ymc[i] = 0.5 * (Y[p * nch + j] + Y[p * nch + j + 1]);
                }
                const double ym = ymc[i];
                const double v = pv[i], a = pa[i];
                // forward quantities of the step
                const double g = kf * (uf - u1) + (ym - u1) * mu + lu * k;
                const double v1 = v + dt * a;
                const double a1 = (g + meta * v1) * w;
                const double v2 = v + hdt * (a + a1);
                const double a2 = (g + meta * v2) * w;
                const double v3 = v + hdt * (a + a2);
                // a3 = w (g - eta v3)
                double G = w * A[i];
                const double V3 = V[i] + mweta * A[i];
                geta -= w * v3 * A[i];
                // v3 = v + hdt a + hdt a2
                double vb = V3, ab_ = hdt * V3;
                const double A2 = hdt * V3;
                G += w * A2;
                const double V2 = mweta * A2;
                geta -= w * v2 * A2;
                // v2 = v + hdt a + hdt a1
                vb += V2;
                ab_ += hdt * V2;
                const double A1 = hdt * V2;
                G += w * A1;
                const double V1 = mweta * A1;
                geta -= w * v1 * A1;
                // v1 = v + dt a
                vb += V1;
                ab_ += dt * V1;
                gkf += (uf - u1) * G;
                gmu += (ym - u1) * G;
                gk += lu * G;
                acc[i] += hmu * G;
                V[i] = vb;  // partial cotangents of v, a (the u1 part follows)
                A[i] = ab_;
                x[i] = G;
            }
        }
        if (n > 0) {
            load_rest(n - 1);
        }
        if (cnt > 0) {
            g_first[t] = x[0];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (i == cnt - 1) { // (an indexed read would move the array to scratch)
                    g_last[t] = x[i];
                }
            }
        }
        __syncthreads();
        const double gleft = g_last[t == 0 ? Tact - 1 : t - 1];
        const double gright = g_first[t == Tact - 1 ? 0 : t + 1];
        double gm = gleft;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const double G = x[i];
                const double gp = i == cnt - 1 ? gright : x[i + 1 < PPT ? i + 1 : i];
                // g depends on u1 through -(k_frame + mu) u1 + k L u1 (L symmetric)
                const double U1 = U[i] - (kf + mu) * G + k * (gm - 2 * G + gp);
                gm = G;
                // u1 = u + dt v + c2 a
                U[i] = U1;
                V[i] = V[i] + dt * U1;
                A[i] = A[i] + c2 * U1;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        if (i < cnt) {
            const int64_t p = base + i;
            ab.au[off + p] = U[i];
            ab.av[off + p] = V[i];
            ab.aa[off + p] = A[i];
            flush(i);
        }
    }
    // parameter derivatives: block sums in a fixed order, added to the totals
    gmu = wave_sum(gmu);
    gk = wave_sum(gk);
    gkf = wave_sum(gkf);
    geta = wave_sum(geta);
    if (lane == 0) {
        s_red[0][wave] = gmu;
        s_red[1][wave] = gk;
        s_red[2][wave] = gkf;
        s_red[3][wave] = geta;
    }
    __syncthreads();
    if (t < 4) {
        double s = 0.0;
        for (int q = 0; q < nw; ++q) {
            s += s_red[t][q];
        }
        ab.gpar[(int64_t)r * 4 + t] += s;
    }
}

} // namespace fq
