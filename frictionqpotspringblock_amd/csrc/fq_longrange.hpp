// Batched LongRange1d (detail.h:812-867) for R replicas of one chain: the
// interaction force of all replicas at once is a circulant matrix product,
//   f[r][p] = sum_{i != p} (u[r][i] - u[r][p]) pref[d(i, p)]
//           = (K (u[r] - ub[r]))[p] - c (u[r][p] - ub[r]),
// K[p][i] = pref[d(i, p)] (zero diagonal), c = sum_i K[p][i] (the same for
// every p), ub[r] = u[r][0] (a per-replica shift that keeps the operands small
// and the cancellation benign). The product runs on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64); K is never stored: each K-block's operand comes
// from a window of g[m] = pref[fold(m)], m = p - i, staged in LDS.
// This path accumulates in a different order (and with fused multiply-add)
// than the reference's sequential sum, so it is compared to the exact path
// within a relative tolerance, not bit for bit.
//
// One step of the batched run is two kernels:
//   k_lr_gemm : part[s] = K[:, I_s] (un - ub)[I_s] for the slips un of the
//               step, one partial product per range I_s of the summation
//               index (split-K, so that a launch fills the chip)
//   k_lr_elem : fg = sum_s part[s] (fixed order) - c (un - ub); well walk for
//               un (or a refill request), potential, frame, thermal draws,
//               the three velocity passes of timeStep, and the slips of the
//               next step un = u + dt v + dt^2/2 a.
#pragma once

#include "fq_resident.hpp"

namespace fq {

typedef double lr_double4 __attribute__((ext_vector_type(4)));

constexpr int LR_BM = 128; // rows (particles) per workgroup
constexpr int LR_BN = 64;  // columns (replicas) per workgroup
constexpr int LR_BK = 32;  // K-block
constexpr int LR_THREADS = 256;
constexpr int LR_GW = LR_BM + LR_BK;            // window of g per K-block
constexpr int LR_BPT = LR_BN * LR_BK / LR_THREADS; // B values per thread per K-block (8)
static_assert(LR_BPT * 4 == LR_BK && LR_THREADS / 4 == LR_BN, "B staging: 4 threads per replica column");
constexpr int LR_XCDS = 8;                      // XCDs of an MI355X

FQ_HD int64_t lr_fold(int64_t m, int64_t N, int64_t lr_m)
{
    int64_t a = m < 0 ? -m : m;
    return a > lr_m ? N - a : a;
}

// Launch geometry of k_lr_gemm: tm x tn output tiles, nsplit ranges of the
// summation index of klen (a multiple of LR_BK) each.
struct LrGeom {
    int tm, tn, nsplit;
    int64_t klen;
};

// part[s][r][p] for one BM x BN tile and one summation range. 4 wavefronts in
// 2 x 2; each computes a 64 x 32 sub-tile as 4 x 2 MFMA 16x16 tiles. Lane
// layout of v_mfma_f64_16x16x4f64: A[i = l%16][k = l/16], B[k = l/16][j = l%16],
// D[i = 4v + l/16][j = l%16]. The next K-block is loaded into registers while
// the current one is multiplied (two LDS buffers, one barrier per K-block).
// Workgroups are dispatched round-robin over the XCDs; the linear id is
// remapped so that each XCD works on contiguous row tiles of the same replica
// columns and summation range, whose B operand then stays in that XCD's L2.
__global__ void __launch_bounds__(LR_THREADS) k_lr_gemm(const double* __restrict__ pref, LrBuffers lb,
                                                         int64_t lr_m, LrGeom geo)
{
    __shared__ double s_g[2][LR_GW];
    __shared__ double s_b[2][LR_BN][LR_BK + 1];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t N = lb.N, R = lb.R;

    const int total = geo.tm * geo.tn * geo.nsplit;
    int id = (int)blockIdx.x;
    if (total % LR_XCDS == 0) {
        id = (id % LR_XCDS) * (total / LR_XCDS) + id / LR_XCDS;
    }
    const int rb = id % geo.tm;
    const int cb = (id / geo.tm) % geo.tn;
    const int sp = id / (geo.tm * geo.tn);

    const int64_t p0 = (int64_t)rb * LR_BM;
    const int64_t r0 = (int64_t)cb * LR_BN;
    const int64_t kbeg = (int64_t)sp * geo.klen;
    const int64_t kend = kbeg + geo.klen < N ? kbeg + geo.klen : N;
    const int nkb = (int)((kend - kbeg + LR_BK - 1) / LR_BK);
    const int wr = wave >> 1, wc = wave & 1; // wave position in the 2 x 2 grid
    const int li = lane & 15, lk = lane >> 4;

    // register staging of one K-block: the g window value (threads < LR_GW)
    // and LR_BPT values of B - ub: replica column bj = t / 4, summation
    // indices bk0 .. bk0 + 7 (bk0 = (t % 4) * 8); zero outside the matrix and
    // outside the summation range
    const int bj = t >> 2, bk0 = (t & 3) * LR_BPT;
    const int64_t br = r0 + bj;
    const bool bvalid = br < R;
    const double* const brow = lb.un + (bvalid ? br : 0) * N;
    const double ubt = bvalid ? brow[0] : 0.0;
    double rg = 0.0, rbv[LR_BPT];
    auto load = [&](int64_t i0) {
        if (t < LR_GW) {
            const int64_t m = p0 - i0 - (LR_BK - 1) + t;
            rg = (m > -N && m < N) ? pref[lr_fold(m, N, lr_m)] : 0.0;
        }
        const int64_t ib = i0 + bk0;
        if (bvalid && ib + LR_BPT <= kend) {
#pragma unroll
            for (int q = 0; q < LR_BPT; ++q) {
                rbv[q] = brow[ib + q] - ubt;
            }
        }
        else {
#pragma unroll
            for (int q = 0; q < LR_BPT; ++q) {
                rbv[q] = (bvalid && ib + q < kend) ? brow[ib + q] - ubt : 0.0;
            }
        }
    };
    auto store = [&](int buf) {
        if (t < LR_GW) {
            s_g[buf][t] = rg;
        }
#pragma unroll
        for (int q = 0; q < LR_BPT; ++q) {
            s_b[buf][bj][bk0 + q] = rbv[q];
        }
    };

    lr_double4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            acc[a][b] = lr_double4{0.0, 0.0, 0.0, 0.0};
        }
    }

    load(kbeg);
    store(0);
    __syncthreads();

    for (int kb = 0; kb < nkb; ++kb) {
        const int buf = kb & 1;
        const bool more = kb + 1 < nkb;
        if (more) {
            load(kbeg + (int64_t)(kb + 1) * LR_BK);
        }
#pragma unroll
        for (int ks = 0; ks < LR_BK; ks += 4) {
            double af[4], bf[2];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                // row p = p0 + wr*64 + a*16 + li, column i = i0 + ks + lk
                const int row = wr * 64 + a * 16 + li;
                af[a] = s_g[buf][row - (ks + lk) + (LR_BK - 1)];
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                bf[b] = s_b[buf][wc * 32 + b * 16 + li][ks + lk];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                }
            }
        }
        if (more) {
            store(buf ^ 1);
        }
        __syncthreads();
    }

    double* const part = lb.part + (size_t)sp * (size_t)(R * N);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t r = r0 + wc * 32 + b * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t p = p0 + wr * 64 + a * 16 + 4 * v + lk;
                if (r < R && p < N) {
                    part[r * N + p] = acc[a][b][v];
                }
            }
        }
    }
}

// The rest of one step for every replica (one workgroup per replica), after
// k_lr_gemm produced the partial products for the slips un of this step.
template <int EXT>
__global__ void __launch_bounds__(1024) k_lr_elem(const SysView* __restrict__ views, LrBuffers lb, RunArgs ra)
{
    const SysView& S = views[blockIdx.x];
    RunState* const rs = S.rs;
    // every thread reads the run state before the first barrier; thread 0
    // writes it back only after that barrier
    const int32_t status = uniform_i(rs->status);
    const int64_t steps = uniform_ll(rs->steps);
    const int64_t target = uniform_ll(rs->target);
    if ((status != ST_RUNNING && status != ST_BUDGET) || steps >= target) {
        return;
    }
    const int64_t inc = uniform_ll(rs->inc) + 1;
    const double uf0 = uniform_d(rs->u_frame);
    const int64_t min_next = EXT ? uniform_ll(rs->min_next) : 0;
    const uint64_t rng0 = EXT ? (uint64_t)uniform_ll((long long)rs->rng_state) : 0;
    const uint64_t rng_inc = EXT ? (uint64_t)uniform_ll((long long)rs->rng_inc) : 0;
    const double ub = lb.un[(int64_t)blockIdx.x * S.N]; // the shift of the product (un[0] is rewritten below)
    __shared__ int s_flag[2];
    __shared__ int s_scan_w[RES_MAX_WAVES];
    __shared__ long long s_min_w[RES_MAX_WAVES];

    const int t = threadIdx.x, T = blockDim.x;
    const int64_t N = S.N;
    const int64_t off = (int64_t)blockIdx.x * N;
    const Landscape& L = S.ls;
    const int64_t nch = L.nchunk;
    const int32_t jmin = (int32_t)L.buffer;
    const int32_t jmax = (int32_t)(L.nchunk - 2 - L.buffer);
    double* const un = lb.un + off;
    const double* const part = lb.part + off;
    const size_t pstride = (size_t)(lb.R * N);
    const int nsplit = lb.nsplit;
    const double c = lb.c;
    double* const yl = lb.yl + off;
    double* const yr = lb.yr + off;

    if (t == 0) {
        s_flag[0] = 0; // refill needed
        s_flag[1] = 0; // NaN
    }
    __syncthreads();

    // well walk for the slips of this step (read only): refill request?
    int refill = 0, nanf = 0;
    for (int64_t p = t; p < N; p += T) {
        const double x = un[p];
        nanf |= __builtin_isnan(x);
        double lo = yl[p], hi = yr[p];
        if (!(lo <= x && x < hi) && __builtin_isfinite(x)) {
            refill |= !reachable(L.y + p * nch, L.jloc[p], lo, hi, x, jmin, jmax);
        }
    }
    if (refill) {
        s_flag[0] = 1;
    }
    if (nanf) {
        s_flag[1] = 1;
    }
    __syncthreads();
    if (uniform_i(s_flag[0])) {
        if (t == 0) {
            rs->status = ST_REFILL;
        }
        return;
    }

    const double dt = S.dt, c2 = 0.5 * dt * dt, hdt = 0.5 * dt, meta = -S.eta, inv_m = S.inv_m, kf = S.k_frame;
    const double uf = ra.flow ? uf0 + ra.v_frame * dt : uf0;
    // minimise (detail.h:1754-1785): every step's forces are stored, since
    // any step can be the last one (the StopList decides after the step)
    const bool mini = ra.mode == MODE_MINIMISE;
    const bool last = mini || steps + 1 >= target;
    double sf = 0.0, sff = 0.0;

    if constexpr (EXT) {
        uint64_t rng = rng0;
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
            const long long m2 = uniform_ll(block_min_i(mn, s_min_w));
            if (t == 0) {
                rs->rng_state = rng;
                rs->min_next = m2;
            }
        }
    }

    for (int64_t p = t; p < N; p += T) {
        const double x = un[p];
        int32_t j = L.jloc[p];
        double lo = yl[p], hi = yr[p];
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
            L.jloc[p] = j;
            yl[p] = lo;
            yr[p] = hi;
        }
        double acc = part[p];
        for (int sp = 1; sp < nsplit; ++sp) {
            acc += part[(size_t)sp * pstride + p];
        }
        const double fi = acc - c * (x - ub);
        const double fp = potential_force<POT_CUSPY>(S, x, lo, hi);
        const double ff = kf * (uf - x);
        const double sum3 = ff + fp + fi;
        const double vn = S.v[p], an = S.a[p];
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
        S.u[p] = x;
        S.v[p] = vv;
        S.a[p] = aa;
        un[p] = x + dt * vv + c2 * aa;
        if (last) {
            S.f_int[p] = fi;
            S.f_pot[p] = fp;
            S.f_frame[p] = ff;
            S.f_damp[p] = meta * vv;
            S.f[p] = f;
        }
        sf += f * f;
        sff += ff * ff;
    }
    int32_t stop = 0;
    if (mini) {
        // residual (detail.h:1512-1520) with the norms in a fixed order, and
        // the StopList (GooseFEM Iterate::StopList, as in k_resident)
        __shared__ double s_red[2][RES_MAX_WAVES];
        const int lane = t & 63, wave = t >> 6, nw = T >> 6;
        sf = wave_sum(sf);
        sff = wave_sum(sff);
        if (lane == 0) {
            s_red[0][wave] = sf;
            s_red[1][wave] = sff;
        }
        __syncthreads();
        double a0 = 0.0, a1 = 0.0;
        for (int w = 0; w < nw; ++w) {
            a0 += s_red[0][w];
            a1 += s_red[1][w];
        }
        const double rf = __builtin_sqrt(a0), rff = __builtin_sqrt(a1);
        const double r = (rff != 0.0) ? rf / rff : rf;
        const int32_t below = !(r >= ra.tol) ? uniform_i(rs->below) + 1 : 0;
        const int32_t below2 = !(r >= ra.tol2) ? uniform_i(rs->below2) + 1 : 0;
        const int32_t dec = !(r >= uniform_d(rs->prev)) ? uniform_i(rs->dec) + 1 : 0;
        const int64_t nt = ra.niter_tol;
        stop = (nt <= 0 || (dec >= nt - 1 && below >= nt) || below2 >= nt) ? 1 : 0;
        __syncthreads(); // every thread read the counters before thread 0 updates them
        if (t == 0) {
            rs->below = below;
            rs->below2 = below2;
            rs->dec = dec;
            rs->prev = r;
        }
        if (stop && !uniform_i(s_flag[1])) {
            // quench (detail.h:1527-1532): v = a = 0 and the forces again
            for (int64_t p = t; p < N; p += T) {
                S.v[p] = 0.0;
                S.a[p] = 0.0;
                const double fd = meta * 0.0;
                S.f_damp[p] = fd;
                S.f[p] = S.f_frame[p] + S.f_pot[p] + S.f_int[p] + fd;
            }
        }
    }
    if (t == 0) {
        rs->inc = inc;
        rs->u_frame = uf;
        rs->steps = steps + 1;
        if (uniform_i(s_flag[1])) {
            rs->status = ST_NAN;
        }
        else if (stop) {
            rs->status = ST_CONVERGED;
        }
    }
}

// Prologue of a batched long-range run: the wells of the current state, the
// slips of the first step and (random forcing) the earliest pending draw.
__global__ void __launch_bounds__(1024) k_lr_prologue(const SysView* __restrict__ views, LrBuffers lb)
{
    __shared__ long long s_min_w[RES_MAX_WAVES];
    const SysView& S = views[blockIdx.x];
    const int64_t N = S.N;
    const int64_t off = (int64_t)blockIdx.x * N;
    const Landscape& L = S.ls;
    const double dt = S.dt, c2 = 0.5 * dt * dt;
    long long mn = 0x7fffffffffffffffLL;
    for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
        const int32_t j = L.jloc[p];
        lb.yl[off + p] = L.y[p * L.nchunk + j];
        lb.yr[off + p] = L.y[p * L.nchunk + j + 1];
        lb.un[off + p] = S.u[p] + dt * S.v[p] + c2 * S.a[p];
        if (S.next) {
            long long nx = S.next[p];
            mn = nx < mn ? nx : mn;
        }
    }
    if (S.next) {
        mn = block_min_i(mn, s_min_w);
        if (threadIdx.x == 0) {
            S.rs->min_next = mn;
        }
    }
}

} // namespace fq
