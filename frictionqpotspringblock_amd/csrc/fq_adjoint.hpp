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
// step and the well used by it; k_adj_reverse then walks the tape backwards.
// The tape is particle-major per step ([step][replica][particle]) and both
// kernels give a thread the particles t, t + T, t + 2T, ... of its segment,
// so that every tape load and store of a wavefront is contiguous; neighbours
// are exchanged through LDS.
//
// k_adj_reverse spreads each replica over tiles (many workgroups per
// replica): a tile walks its particles together with a halo of K + 1
// particles on each side, whose cotangents start from the adjoint state of
// their owners. The cotangent of a particle at the open end of the extended
// segment lacks one neighbour, and the error this makes moves inwards by one
// particle per reversed step (through the second difference of the
// cotangents of the forces), so after the K steps of a window the tile's own
// particles are exact. The adjoint state is double-buffered across windows
// (tiles read their halo from the buffer that no tile writes in the same
// launch); the parameter derivatives are summed per tile and then over the
// tiles in a fixed order (k_adj_params).
#pragma once

#include "fq_physics.hpp"
#include "fq_resident.hpp"

namespace fq {


// Record windows w0 + blockIdx.y (each of min(K, n - w K) steps) from their
// start states, one workgroup per replica, tile and window (blockIdx.x =
// replica * rnt + tile): every window is independent given its checkpoint,
// so all of a group run at once. With one tile per replica the workgroup
// holds the whole periodic chain; with several, a tile steps its particles
// together with a halo of rhalo >= K particles on each side (recomputed from
// their owners' checkpoint with the same expressions: the open ends' error
// moves inwards by one particle per step and never reaches the tile's own
// particles within a window) and writes the tape of its own particles only.
// Thread t holds local particles t + T k (k < SLOTS) of the segment; the
// slips of a step are exchanged through LDS, double-buffered by step parity
// (one barrier per step). RING: one tile per replica (local = global index,
// every particle its own: compile-time offsets of the tape entries).
template <int SLOTS, int MAXT, bool RING>
__global__ void __launch_bounds__(MAXT) k_adj_record(const SysView* __restrict__ views, AdjBuffers ab, int flow,
                                                     double v_frame)
{
    extern __shared__ double sx[]; // [2][E]
    const int r = blockIdx.x / ab.rnt, tile = blockIdx.x % ab.rnt;
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
    constexpr int T = MAXT; // launched with exactly MAXT threads
    const int t = threadIdx.x;
    const int N = (int)ab.N;
    const int64_t off = (int64_t)r * N;
    constexpr bool ring = RING;
    const int p0 = tile * (int)ab.rtp;
    const int p1 = p0 + (int)ab.rtp < N ? p0 + (int)ab.rtp : N;
    const int H = ring ? 0 : (int)ab.rhalo;
    const int E = ring ? N : (p1 - p0) + 2 * H; // particles of the segment
    const Landscape& L = S.ls;
    const double* const Y = L.y;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);
    const int64_t Rnp = ab.R * ab.np;

    double u[SLOTS], v[SLOTS], a[SLOTS], yl[SLOTS], yr[SLOTS];
    int32_t jl[SLOTS], pp[SLOTS];
    bool own[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; ++k) {
        u[k] = v[k] = a[k] = yl[k] = yr[k] = 0.0;
        jl[k] = 0;
        const int e = t + T * k;
        int p = e;
        if constexpr (!RING) {
            p = p0 - H + e;
            p = p < 0 ? p + N : (p >= N ? p - N : p);
        }
        pp[k] = p;
        own[k] = e < E && (RING || (e >= H && e < H + (p1 - p0)));
        if (e < E) {
            u[k] = ab.wu[off + p];
            v[k] = ab.wv[off + p];
            a[k] = ab.wa[off + p];
            jl[k] = ab.wj[off + p];
            yl[k] = Y[p * nch + jl[k]];
            yr[k] = Y[p * nch + jl[k] + 1];
        }
    }
    const double dt = S.dt, c2 = 0.5 * dt * dt, hdt = 0.5 * dt, meta = -S.eta, inv_m = S.inv_m, kf = S.k_frame;
    const double dv = v_frame * dt;
    double uf = ab.wuf[r];
    int bad = 0;

    for (int n = 0; n < steps; ++n) {
        double* const xb = sx + (n & 1) * E;
        if (flow) {
            uf = uf + dv;
        }
        const int64_t tn = (int64_t)n * Rnp + (int64_t)r * ab.np;
        if (t == 0 && tile == 0) {
            ab.tuf[(int64_t)n * ab.R + r] = uf;
        }
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            const int e = t + T * k;
            if (e < E) {
                const int64_t q = tn + (RING ? t + T * k : pp[k]);
                if (own[k]) {
                    ab.tv[q] = v[k];
                    ab.ta[q] = a[k];
                }
                u[k] = u[k] + dt * v[k] + c2 * a[k];
                const double x = u[k];
                if (own[k]) {
                    ab.tu[q] = x;
                }
                if (!(yl[k] <= x && x < yr[k])) {
                    const double* row = Y + (int64_t)pp[k] * nch;
                    int32_t j = jl[k];
                    double lo = yl[k], hi = yr[k];
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
                    bad |= own[k] && !(lo <= x && x < hi);
                    jl[k] = j;
                    yl[k] = lo;
                    yr[k] = hi;
                }
                if (own[k]) {
                    ab.tj[q] = jl[k];
                }
                xb[e] = x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            const int e = t + T * k;
            if (e < E) {
                // neighbours: periodic on the ring; at the open ends of a
                // tile's segment, any finite value (the particle's own slip)
                double um, up;
                if (ring) {
                    um = xb[e == 0 ? N - 1 : e - 1];
                    up = xb[e == N - 1 ? 0 : e + 1];
                }
                else {
                    um = e == 0 ? u[k] : xb[e - 1];
                    up = e == E - 1 ? u[k] : xb[e + 1];
                }
                const double fi = chain_force<INT_LAPLACE1D>(S, um, u[k], up);
                const double fp = potential_force<POT_CUSPY>(S, u[k], yl[k], yr[k]);
                const double ff = kf * (uf - u[k]);
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
            }
        }
    }
    if (bad) {
        *ab.err = 1;
    }
}

// Walk the tape of one window backwards on one tile of one replica
// (blockIdx.x = replica * nt + tile), updating the adjoint state of the
// tile's particles and its partial parameter derivatives. Since dg/dy_l =
// dg/dy_r = mu/2, one accumulator per particle collects the derivative for
// both yield positions of its current well; it is added to the output when
// the well changes.
template <int SLOTS>
__global__ void __launch_bounds__(ADJ_REV_THREADS) k_adj_reverse(const SysView* __restrict__ views, AdjBuffers ab,
                                                                   int steps)
{
    // neighbour exchange: slips u1 (su) and cotangents of g (sg) of a step;
    // two barriers per step (a buffer is rewritten only after the barrier
    // that follows every read of it)
    __shared__ double su[ADJ_REV_THREADS * SLOTS];
    __shared__ double sg[ADJ_REV_THREADS * SLOTS];
    __shared__ double s_red[4][ADJ_REV_THREADS / 64];
    const int r = blockIdx.x / ab.nt, tile = blockIdx.x % ab.nt;
    const SysView& S = views[r];
    constexpr int T = ADJ_REV_THREADS;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr int nw = T / 64;
    const int N = (int)ab.N;
    const int64_t off = (int64_t)r * N;
    const bool ring = ab.nt == 1; // one tile: the whole periodic chain, no halo
    const int p0 = tile * (int)ab.tp;
    const int p1 = p0 + (int)ab.tp < N ? p0 + (int)ab.tp : N;
    const int H = ring ? 0 : (int)ab.halo;
    const int E = ring ? N : (p1 - p0) + 2 * H; // particles of the extended segment
    const double* const Y = S.ls.y;
    const int64_t nch = S.ls.nchunk;
    const int64_t Rnp = ab.R * ab.np;

    const double dt = S.dt, c2 = 0.5 * dt * dt, hdt = 0.5 * dt, meta = -S.eta, w = S.inv_m, kf = S.k_frame;
    const double mu = S.mu, k = S.k;
    const double mweta = meta * w; // -eta / m
    const double hmu = 0.5 * mu;

    // slot i: local index e = t + T i of the extended segment, particle p
    int pp[SLOTS];
    bool own[SLOTS], live[SLOTS];
#pragma unroll
    for (int i = 0; i < SLOTS; ++i) {
        const int e = t + T * i;
        live[i] = e < E;
        own[i] = live[i] && (ring || (e >= H && e < H + (p1 - p0)));
        int q = ring ? e : p0 - H + e;
        q = q < 0 ? q + N : (q >= N ? q - N : q);
        pp[i] = q;
    }
    // U, V, A: cotangents of u, v, a after the step being reversed
    // jc, ymc: the well of the current tape step and its mid-point (read from
    // the chunk only when the well changes)
    double U[SLOTS], V[SLOTS], A[SLOTS], acc[SLOTS], ymc[SLOTS];
    int32_t jc[SLOTS];
#pragma unroll
    for (int i = 0; i < SLOTS; ++i) {
        U[i] = V[i] = A[i] = acc[i] = ymc[i] = 0.0;
        jc[i] = -1;
        if (live[i]) {
            U[i] = ab.au[off + pp[i]];
            V[i] = ab.av[off + pp[i]];
            A[i] = ab.aa[off + pp[i]];
        }
    }
    double gmu = 0.0, gk = 0.0, gkf = 0.0, geta = 0.0;

    auto flush = [&](int i) {
        if (ab.gy && own[i] && jc[i] >= 0) {
            double* row = ab.gy + (off + pp[i]) * nch;
            row[jc[i]] += acc[i];
            row[jc[i] + 1] += acc[i];
        }
        acc[i] = 0.0;
    };
    // neighbour of local index e at offset d (+1 or -1): periodic on the
    // ring; at the open ends of a tile's segment, any finite value (the
    // error zone that never reaches the tile's own particles)
    auto nb = [&](const double* s, int e, int d, double self) -> double {
        int q = e + d;
        if (ring) {
            q = q < 0 ? q + E : (q >= E ? q - E : q);
            return s[q];
        }
        return (q < 0 || q >= E) ? self : s[q];
    };

    // the tape of step n is loaded while step n + 1 is reversed (registers
    // qu, qv, qa, qj, quf), so its memory latency is hidden behind that work
    double qu[SLOTS], qv[SLOTS], qa[SLOTS], quf = 0.0;
    int32_t qj[SLOTS];
    auto load = [&](int n) {
        const int64_t tn = (int64_t)n * Rnp + (int64_t)r * ab.np;
#pragma unroll
        for (int i = 0; i < SLOTS; ++i) {
            qu[i] = qv[i] = qa[i] = 0.0;
            qj[i] = 0;
            if (live[i]) {
                const int64_t q = tn + pp[i];
                qu[i] = ab.tu[q];
                qv[i] = ab.tv[q];
                qa[i] = ab.ta[q];
                qj[i] = ab.tj[q];
            }
        }
        quf = ab.tuf[(int64_t)n * ab.R + r];
    };
    if (steps > 0) {
        load(steps - 1);
    }
    for (int n = steps - 1; n >= 0; --n) {
        const double uf = quf;
        double x[SLOTS], pv[SLOTS], pa[SLOTS];
        int32_t pj[SLOTS];
#pragma unroll
        for (int i = 0; i < SLOTS; ++i) {
            x[i] = qu[i];
            pv[i] = qv[i];
            pa[i] = qa[i];
            pj[i] = qj[i];
            if (live[i]) {
                su[t + T * i] = x[i];
            }
        }
        __syncthreads();
        if (n > 0) {
            load(n - 1);
        }
#pragma unroll
        for (int i = 0; i < SLOTS; ++i) {
            if (live[i]) {
                const int e = t + T * i;
                const int64_t p = pp[i];
                const double u1 = x[i];
                const double lu = nb(su, e, -1, u1) - 2 * u1 + nb(su, e, +1, u1);
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
                double ge = -(w * v3 * A[i]);
                // v3 = v + hdt a + hdt a2
                double vb = V3, ab_ = hdt * V3;
                const double A2 = hdt * V3;
                G += w * A2;
                const double V2 = mweta * A2;
                ge -= w * v2 * A2;
                // v2 = v + hdt a + hdt a1
                vb += V2;
                ab_ += hdt * V2;
                const double A1 = hdt * V2;
                G += w * A1;
                const double V1 = mweta * A1;
                ge -= w * v1 * A1;
                // v1 = v + dt a
                vb += V1;
                ab_ += dt * V1;
                if (own[i]) {
                    geta += ge;
                    gkf += (uf - u1) * G;
                    gmu += (ym - u1) * G;
                    gk += lu * G;
                    acc[i] += hmu * G;
                }
                V[i] = vb; // partial cotangents of v, a (the u1 part follows)
                A[i] = ab_;
                x[i] = G;
                sg[e] = G;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SLOTS; ++i) {
            if (live[i]) {
                const int e = t + T * i;
                const double G = x[i];
                // g depends on u1 through -(k_frame + mu) u1 + k L u1 (L symmetric)
                const double U1 = U[i] - (kf + mu) * G + k * (nb(sg, e, -1, G) - 2 * G + nb(sg, e, +1, G));
                // u1 = u + dt v + c2 a
                U[i] = U1;
                V[i] = V[i] + dt * U1;
                A[i] = A[i] + c2 * U1;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < SLOTS; ++i) {
        if (own[i]) {
            ab.bu[off + pp[i]] = U[i];
            ab.bv[off + pp[i]] = V[i];
            ab.ba[off + pp[i]] = A[i];
            flush(i);
        }
    }
    // parameter derivatives of the tile: block sums in a fixed order
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
        ab.gpart[(int64_t)blockIdx.x * 4 + t] = s;
    }
}

// The parameter derivatives of a window: the tiles' partial sums of each
// replica, in tile order, added to the totals (one thread per replica and
// parameter).
__global__ void __launch_bounds__(256) k_adj_params(AdjBuffers ab)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ab.R * 4) {
        return;
    }
    const int64_t r = i / 4, c = i % 4;
    double s = 0.0;
    for (int tile = 0; tile < ab.nt; ++tile) {
        s += ab.gpart[(r * ab.nt + tile) * 4 + c];
    }
    ab.gpar[i] += s;
}

} // namespace fq
