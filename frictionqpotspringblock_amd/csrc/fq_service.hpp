// Service kernels: everything that is not the many-step resident loop.
// One workgroup per system, block-stride loops over the particles, so they
// handle any system size. They implement the reference's force bookkeeping
// (updated_u / updated_v / updated_inc / computeForceFrame, detail.h:1321-1395),
// quench (detail.h:1527-1532), advanceUniformly (detail.h:2027-2050),
// maxUniformDisplacement (detail.h:176-187, 282-334), trigger (detail.h:193-205,
// 339-350, 427-438), construction of the landscape (Line1d.h:148-157) and
// chunk restore, and the no-passing minimiser (detail.h:1694-1753).
#pragma once

#include "fq_resident.hpp"

namespace fq {

constexpr int SVC_THREADS = 256;

// First window of every particle's landscape: cumsum of nchunk increments,
// then += offset; align to u (which is 0 at construction).
__global__ void __launch_bounds__(SVC_THREADS) k_landscape_init(const SysView* __restrict__ views)
{
    const SysView S = views[blockIdx.y];
    const Landscape& L = S.ls;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.N) {
        return;
    }
    const uint64_t seed = S.seed;
    const double offset = S.offset;
    uint64_t st, inc;
    pcg_seed(seed + (uint64_t)p, 0u, st, inc);
    double* y = L.y + p * L.nchunk;
    double cs = 0.0;
    for (int64_t g = 0; g < L.nchunk; ++g) {
        double d = draw_increment(L, st);
        cs = (g == 0) ? d : cs + d;
        y[g] = cs + offset;
        if (g == 0) {
            L.st_begin[p] = st;
        }
    }
    L.st_end[p] = st;
    L.start[p] = 0;
    L.jloc[p] = 0;
    double yl = y[0], yr = y[1];
    int32_t j = 0;
    if (!(yl <= S.u[p] && S.u[p] < yr)) {
        j = chunk_align_slow(L, p, S.u[p], 0);
    }
    L.jloc[p] = j;
}

// Rebuild every particle's window from (state, value, index) and re-align.
__global__ void __launch_bounds__(SVC_THREADS) k_restore(const SysView* __restrict__ views, const uint64_t* state,
                                                         const double* value, const int64_t* index)
{
    const SysView S = views[blockIdx.y];
    const Landscape& L = S.ls;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S.N) {
        return;
    }
    double* y = L.y + p * L.nchunk;
    uint64_t st = state[p];
    L.start[p] = index[p];
    L.st_begin[p] = st;
    y[0] = value[p];
    for (int64_t g = 1; g < L.nchunk; ++g) {
        y[g] = y[g - 1] + draw_increment(L, st);
    }
    L.st_end[p] = st;
    int32_t j = 0;
    if (!(y[0] <= S.u[p] && S.u[p] < y[1])) {
        j = chunk_align_slow(L, p, S.u[p], 0);
    }
    L.jloc[p] = j;
}

// RandomNormalForcing::force detail.h:930-943 for the whole system at the
// current increment: due particles draw in flat order from the one stream.
template <int T>
__device__ void thermal_update_block(const SysView& S, RunState* rs)
{
    __shared__ int s_w[RES_MAX_WAVES];
    __shared__ long long s_m[RES_MAX_WAVES];
    const int64_t inc = rs->inc;
    uint64_t rng = rs->rng_state;
    const uint64_t rng_inc = rs->rng_inc;
    long long mn = 0x7fffffffffffffffLL;
    for (int64_t b0 = 0; b0 < S.N; b0 += T) {
        const int64_t p = b0 + threadIdx.x;
        int due = (p < S.N && inc >= S.next[p]) ? 1 : 0;
        int total = 0;
        int rank = block_excl_scan(due, s_w, total);
        if (due) {
            uint64_t st = pcg_advance(rng, rng_inc, 2 * (int64_t)rank);
            S.f_th[p] = draw_normal(st, rng_inc, S.mean, S.stddev);
            S.next[p] += S.dinc[p];
        }
        if (p < S.N) {
            long long nx = S.next[p];
            mn = nx < mn ? nx : mn;
        }
        rng = pcg_advance(rng, rng_inc, 2 * (int64_t)total);
    }
    mn = block_min_i(mn, s_m);
    if (threadIdx.x == 0) {
        rs->rng_state = rng;
        rs->min_next = mn;
    }
}

// LongRange1d interaction forces of every system (blockIdx.y) from its slips
// u, each thread one particle's sum over the whole chain in the reference's
// order (detail.h:849-867), the chain spread over many workgroups: the O(N^2)
// part of k_forces for long-range systems.
__global__ void __launch_bounds__(SVC_THREADS) k_lr_fint(const SysView* __restrict__ views)
{
    const SysView& S = views[blockIdx.y];
    const int64_t N = S.N;
    const double* __restrict__ u = S.u;
    const double* __restrict__ pref = S.pref;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        S.f_int[p] = longrange_force(S, [&](int64_t q) { return u[q]; }, [&](int64_t d) { return pref[d]; }, p);
    }
}

template <int POT, int INT, int EXT>
__global__ void __launch_bounds__(SVC_THREADS) k_forces(const SysView* __restrict__ views, int32_t flags)
{
    const SysView S = views[blockIdx.x];
    RunState* rs = S.rs;
    const Landscape& L = S.ls;
    const int64_t N = S.N;
    if (flags & FF_ADD_U) {
        const double du = rs->du_add;
        if (du != 0.0) {
            for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
                S.u[p] += du;
            }
        }
        __syncthreads();
    }
    if (flags & FF_QUENCH) {
        for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
            S.v[p] = 0.0;
            S.a[p] = 0.0;
        }
    }
    if constexpr (EXT) {
        if (flags & FF_THERMAL) {
            thermal_update_block<SVC_THREADS>(S, rs);
        }
    }
    __syncthreads();
    const double uf = rs->u_frame;
    const double meta = -S.eta;
    for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
        const double up = S.u[p];
        if (flags & FF_POT) {
            int32_t j = L.jloc[p];
            double yl = L.y[p * L.nchunk + j], yr = L.y[p * L.nchunk + j + 1];
            align_particle(L, p, up, yl, yr, j);
            S.f_pot[p] = potential_force<POT>(S, up, yl, yr);
        }
        if (flags & FF_INT) {
            double fi = 0.0;
            if constexpr (is_chain<INT>() && INT != INT_NONE) {
                const double um = S.u[p == 0 ? N - 1 : p - 1];
                const double upp = S.u[p == N - 1 ? 0 : p + 1];
                fi = chain_force<INT>(S, um, up, upp);
            }
            else if constexpr (INT == INT_LONGRANGE1D) {
                fi = longrange_force(
                    S, [&](int64_t q) { return S.u[q]; }, [&](int64_t d) { return S.pref[d]; }, p);
            }
            else if constexpr (INT == INT_LAPLACE2D || INT == INT_QUARTICGRADIENT2D) {
                fi = lattice_force<INT>(S, [&](int64_t q) { return S.u[q]; }, p / S.cols, p % S.cols);
            }
            S.f_int[p] = fi;
        }
        if (flags & FF_FRAME) {
            S.f_frame[p] = S.k_frame * (uf - up);
        }
        if (flags & FF_DAMP) {
            S.f_damp[p] = meta * S.v[p];
        }
        if (flags & FF_TOTAL) {
            double f = S.f_frame[p] + S.f_pot[p] + S.f_int[p] + S.f_damp[p];
            if constexpr (EXT) {
                f = f + S.f_th[p];
            }
            S.f[p] = f;
        }
    }
}

// maxUniformDisplacement: rs->red0 = min over particles, rs->red_flag = 1 if
// a SemiSmooth particle is outside the parabolic part of its well.
template <int POT>
__global__ void __launch_bounds__(SVC_THREADS) k_max_uniform(const SysView* __restrict__ views, int32_t direction)
{
    __shared__ double s_d[RES_MAX_WAVES];
    __shared__ int s_f[RES_MAX_WAVES];
    const SysView S = views[blockIdx.x];
    const Landscape& L = S.ls;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double r = __builtin_inf();
    int outside = 0;
    for (int64_t p = threadIdx.x; p < S.N; p += blockDim.x) {
        const double up = S.u[p];
        int32_t j = L.jloc[p];
        double yl = L.y[p * L.nchunk + j], yr = L.y[p * L.nchunk + j + 1];
        align_particle(L, p, up, yl, yr, j);
        const double x = max_uniform_particle<POT>(S, up, yl, yr, direction, outside);
        r = x < r ? x : r;
    }
    for (int off = 32; off > 0; off >>= 1) {
        double y = __shfl_xor(r, off, 64);
        r = y < r ? y : r;
        outside |= __shfl_xor(outside, off, 64);
    }
    if (lane == 0) {
        s_d[wave] = r;
        s_f[wave] = outside;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = s_d[0];
        int o = s_f[0];
        for (int w = 1; w < nw; ++w) {
            m = s_d[w] < m ? s_d[w] : m;
            o |= s_f[w];
        }
        S.rs->red0 = m;
        S.rs->red_flag = o;
    }
}

// trigger: move particle p just past the edge of its well (forces untouched,
// detail.h:1972-1977; Cuspy aligns first, detail.h:193-205).
__device__ void trigger_one(const SysView& S, int64_t p, double eps, int32_t direction, int32_t do_align)
{
    const Landscape& L = S.ls;
    int32_t j = L.jloc[p];
    double yl = L.y[p * L.nchunk + j], yr = L.y[p * L.nchunk + j + 1];
    if (do_align) {
        align_particle(L, p, S.u[p], yl, yr, j);
    }
    S.u[p] = direction > 0 ? yr + 0.5 * eps : yl - 0.5 * eps;
}

__global__ void k_trigger(const SysView* __restrict__ views, int64_t p, double eps, int32_t direction,
                          int32_t do_align)
{
    trigger_one(views[blockIdx.x], p, eps, direction, do_align);
}

// trigger of particle ps[r] in every system r (one thread per system; a
// negative index leaves the system as it is).
__global__ void k_trigger_each(const SysView* __restrict__ views, const int64_t* __restrict__ ps, double eps,
                               int32_t direction, int32_t do_align)
{
    const int64_t p = ps[blockIdx.x];
    if (p >= 0) {
        trigger_one(views[blockIdx.x], p, eps, direction, do_align);
    }
}

// Landscape refill requested by a launch (status ST_REFILL): align every
// particle to its slip after the position update of the pending step,
// u + dt v + dt^2/2 a (bit-identical to the launching kernel), shifting and
// regenerating its window like the CPU align; then the run continues.
__global__ void __launch_bounds__(SVC_THREADS) k_refill(const SysView* __restrict__ views)
{
    const SysView& S = views[blockIdx.x];
    RunState* rs = S.rs;
    if (rs->status != ST_REFILL) {
        return;
    }
    const Landscape& L = S.ls;
    const double dt = S.dt;
    const double c2 = 0.5 * dt * dt;
    for (int64_t p = threadIdx.x; p < S.N; p += blockDim.x) {
        const double x = S.u[p] + dt * S.v[p] + c2 * S.a[p];
        const int32_t j = L.jloc[p];
        const double* row = L.y + p * L.nchunk;
        if (!(row[j] <= x && x < row[j + 1])) {
            chunk_align_slow(L, p, x, j);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rs->status = ST_RUNNING;
    }
}

// Reference index for event / activity counting: i_ref = start + jloc.
__global__ void __launch_bounds__(SVC_THREADS) k_set_iref(const SysView* __restrict__ views)
{
    const SysView S = views[blockIdx.y];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < S.N) {
        S.i_ref[p] = S.ls.start[p] + S.ls.jloc[p];
    }
}

// Quasistatic protocol driven from the host (systems without a k_quasi
// variant): begin = 1 stores rs->qs_isum = sum_p (start + jloc - i_ref);
// begin = 0 records event e, out_*[r * nevents + e] = (sum - qs_isum,
// u_frame, mean f_frame, inc). Fixed summation order.
__global__ void __launch_bounds__(SVC_THREADS) k_qs_record(const SysView* __restrict__ views, QsArgs qa, int64_t e,
                                                           int32_t begin)
{
    __shared__ long long s_i[RES_MAX_WAVES];
    __shared__ double s_d[RES_MAX_WAVES];
    const SysView& S = views[blockIdx.x];
    RunState* rs = S.rs;
    if (rs->status == ST_IDLE) {
        return;
    }
    const Landscape& L = S.ls;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    long long si = 0;
    double sd = 0.0;
    for (int64_t p = threadIdx.x; p < S.N; p += blockDim.x) {
        si += L.start[p] + L.jloc[p] - S.i_ref[p];
        sd += S.f_frame[p];
    }
    si = wave_sum_i(si);
    sd = wave_sum(sd);
    if (lane == 0) {
        s_i[wave] = si;
        s_d[wave] = sd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long ti = 0;
        double td = 0.0;
        for (int w = 0; w < nw; ++w) {
            ti += s_i[w];
            td += s_d[w];
        }
        if (begin) {
            rs->qs_isum = ti;
        }
        else {
            const int64_t o = (int64_t)blockIdx.x * qa.nevents + e;
            if (qa.out_S) {
                qa.out_S[o] = ti - rs->qs_isum;
            }
            if (qa.out_uf) {
                qa.out_uf[o] = rs->u_frame;
            }
            if (qa.out_ff) {
                qa.out_ff[o] = td / (double)S.N;
            }
            if (qa.out_inc) {
                qa.out_inc[o] = rs->inc;
            }
        }
    }
}

// No-passing minimiser (Jacobi sweep), Cuspy + Laplace1d only
// (detail.h:1694-1753): each particle jumps to the minimum of its quadratic
// well given the previous iterate of its neighbours, re-aligning until its
// well index is stable; then the residual of f = f_pot + f_int + f_frame is
// pushed through the StopList. Two barriers per sweep.
template <int PPT>
__global__ void __launch_bounds__(res_max_threads<PPT>()) k_nopassing(const SysView* __restrict__ views, RunArgs ra)
{
    const SysView S = views[blockIdx.x];
    RunState* rs = S.rs;
    if (rs->status != ST_RUNNING && rs->status != ST_BUDGET) {
        return;
    }
    __shared__ double s_red_d[2][RES_MAX_WAVES][2];
    extern __shared__ double s_dyn[];
    const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6, nw = T >> 6;
    const int64_t N = S.N;
    const int64_t base = (int64_t)t * PPT;
    const int cnt = (int)(N - base >= PPT ? PPT : (N - base > 0 ? N - base : 0));
    const int Tact = (int)((N + PPT - 1) / PPT);
    const Landscape& L = S.ls;
    double* h_first = s_dyn;
    double* h_last = s_dyn + 2 * T;

    double u[PPT], yl[PPT], yr[PPT], fp[PPT], ff[PPT], usnap[PPT];
    int32_t jl[PPT];
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        u[i] = yl[i] = yr[i] = fp[i] = ff[i] = usnap[i] = 0.0;
        jl[i] = 0;
        if (i < cnt) {
            const int64_t p = base + i;
            u[i] = S.u[p];
            jl[i] = L.jloc[p];
            yl[i] = L.y[p * L.nchunk + jl[i]];
            yr[i] = L.y[p * L.nchunk + jl[i] + 1];
        }
    }
    auto publish = [&](int b) {
        if (cnt > 0) {
            h_first[b * T + t] = u[0];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (i == cnt - 1) {
                    h_last[b * T + t] = u[i];
                }
            }
        }
    };
    auto left_of = [&](int b) -> double { return h_last[b * T + (t == 0 ? Tact - 1 : t - 1)]; };
    auto right_of = [&](int b) -> double { return h_first[b * T + (t == Tact - 1 ? 0 : t + 1)]; };

    const double k = S.k, kf = S.k_frame, mu = S.mu;
    const double uf = rs->u_frame;
    int32_t below = rs->below, below2 = rs->below2, dec = rs->dec;
    double prev = rs->prev;
    int32_t status = ST_RUNNING;
    publish(0);
    __syncthreads();
    int it;
    int b = 0;
    for (it = 1;; ++it) {
        if (it > ra.max_steps) {
            break;
        }
        // Jacobi sweep against the snapshot in buffer b
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            usnap[i] = u[i];
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const int64_t p = base + i;
                double um = 0.0, up = 0.0;
#pragma unroll
                for (int q = 0; q < PPT; ++q) {
                    if (q == i) {
                        um = (q == 0) ? left_of(b) : usnap[q - 1];
                        up = (q == cnt - 1) ? right_of(b) : (q + 1 < PPT ? usnap[q + 1] : 0.0);
                    }
                }
                const double uneigh = um + up;
                double umin, un;
                // the well sequence is monotone (the minimum moves with the
                // well), so this terminates; the cap only bounds the work
                for (int guard = 0; guard < (1 << 20); ++guard) {
                    umin = 0.5 * (yl[i] + yr[i]);
                    un = (k * uneigh + kf * uf + mu * umin) / (2 * k + kf + mu);
                    const int32_t jold = jl[i];
                    align_particle(L, p, un, yl[i], yr[i], jl[i]);
                    if (jl[i] == jold) {
                        break;
                    }
                }
                u[i] = un;
                ff[i] = kf * (uf - un);
                fp[i] = mu * (umin - un);
            }
        }
        b ^= 1;
        publish(b);
        __syncthreads();
        double sf = 0.0, sff = 0.0;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            if (i < cnt) {
                const double um = (i == 0) ? left_of(b) : u[i > 0 ? i - 1 : 0];
                const double up = (i == cnt - 1) ? right_of(b) : u[i + 1 < PPT ? i + 1 : i];
                const double fi = chain_force<INT_LAPLACE1D>(S, um, u[i], up);
                const double f = fp[i] + fi + ff[i];
                sf += f * f;
                sff += ff[i] * ff[i];
            }
        }
        sf = wave_sum(sf);
        sff = wave_sum(sff);
        if (lane == 0) {
            s_red_d[it & 1][wave][0] = sf;
            s_red_d[it & 1][wave][1] = sff;
        }
        __syncthreads();
        sf = 0.0;
        sff = 0.0;
        for (int w = 0; w < nw; ++w) {
            sf += s_red_d[it & 1][w][0];
            sff += s_red_d[it & 1][w][1];
        }
        const double rf = __builtin_sqrt(sf), rff = __builtin_sqrt(sff);
        const double r = (rff != 0.0) ? rf / rff : rf;
        below = !(r >= ra.tol) ? below + 1 : 0;
        below2 = !(r >= ra.tol2) ? below2 + 1 : 0;
        dec = !(r >= prev) ? dec + 1 : 0;
        prev = r;
        const int64_t nt = ra.niter_tol;
        if (nt <= 0 || (dec >= nt - 1 && below >= nt) || below2 >= nt) {
            status = ST_CONVERGED;
            ++it;
            break;
        }
    }
    const int done = it - 1;
    if (status == ST_RUNNING) {
        status = ST_BUDGET;
    }
    const double meta = -S.eta;
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        if (i < cnt) {
            const int64_t p = base + i;
            const double um = (i == 0) ? left_of(b) : u[i > 0 ? i - 1 : 0];
            const double up = (i == cnt - 1) ? right_of(b) : u[i + 1 < PPT ? i + 1 : i];
            const double fi = chain_force<INT_LAPLACE1D>(S, um, u[i], up);
            S.u[p] = u[i];
            L.jloc[p] = jl[i];
            if (done > 0) {
                // state after the last sweep (f without damping); quench on convergence
                S.f_pot[p] = fp[i];
                S.f_frame[p] = ff[i];
                S.f_int[p] = fi;
                if (status == ST_CONVERGED) {
                    S.v[p] = 0.0;
                    S.a[p] = 0.0;
                    const double fd = meta * 0.0;
                    S.f_damp[p] = fd;
                    S.f[p] = ff[i] + fp[i] + fi + fd;
                }
                else {
                    S.f[p] = fp[i] + fi + ff[i];
                }
            }
        }
    }
    if (t == 0) {
        rs->steps += done;
        rs->status = status;
        rs->below = below;
        rs->below2 = below2;
        rs->dec = dec;
        rs->prev = prev;
    }
}

} // namespace fq
