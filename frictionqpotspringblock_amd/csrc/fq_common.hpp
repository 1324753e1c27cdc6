// Shared definitions of the MI355X (gfx950) implementation of the
// FrictionQPotSpringBlock dynamics hot path: enums, the per-system device view,
// the per-system run state, pcg32 and the elementary functions used by the
// landscape transforms, the thermal draw and the Smooth potential.
//
// Every function here is evaluated without fused multiply-add (all sources are
// compiled with -ffp-contract=off) in the operation order of the reference's
// expressions, so that device results are bit-identical to a CPU evaluation of
// the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FQ_HD __host__ __device__ __forceinline__

namespace fq {

enum Potential : int32_t { POT_CUSPY = 0, POT_SEMISMOOTH = 1, POT_SMOOTH = 2 };

enum Interaction : int32_t {
    INT_NONE = 0,
    INT_LAPLACE1D = 1,
    INT_QUARTIC1D = 2,
    INT_QUARTICGRADIENT1D = 3,
    INT_LONGRANGE1D = 4,
    INT_LAPLACE2D = 5,
    INT_QUARTICGRADIENT2D = 6
};

enum Distribution : int32_t {
    DIST_RANDOM = 0,
    DIST_DELTA = 1,
    DIST_EXPONENTIAL = 2,
    DIST_POWER = 3,
    DIST_GAMMA = 4,
    DIST_PARETO = 5,
    DIST_WEIBULL = 6,
    DIST_NORMAL = 7
};

// What a resident launch does per step (see fq_resident.hpp).
enum Mode : int32_t {
    MODE_STEPS = 0,       // timeSteps / flowSteps
    MODE_MINIMISE = 1,    // minimise (inertial), optional time_activity
    MODE_UNTIL_EVENT = 2, // timeStepsUntilEvent
    MODE_TRUNCATE = 3     // minimise_truncate
};

// Why a run stopped.
enum Status : int32_t {
    ST_RUNNING = 0,
    ST_CONVERGED = 1, // StopList satisfied (state quenched)
    ST_EVENT = 2,     // a well index changed (timeStepsUntilEvent)
    ST_TRUNCATED = 3, // A_truncate / S_truncate reached
    ST_NAN = 4,       // NaN in u
    ST_BUDGET = 5,    // step budget of this launch used up
    ST_REFILL = 6,    // the next step needs a landscape refill (k_refill, then relaunch)
    ST_IDLE = 7,      // left out of this run by the caller
    ST_DONE = 8,      // quasistatic driver: every event of the run recorded
    ST_MAXITER = 9    // quasistatic driver: a minimisation used max_iter steps
};

// The quasistatic driver for chains (k_quasi_chain) keeps the state of the
// last committed step in LDS (8 bytes per particle, beyond the exchange
// buffers) instead of a second register array for these launch variants
// (particles per thread, max threads).
FQ_HD constexpr bool quasi_inplace(int ppt, int maxt)
{
    return (ppt == 8 || ppt == 4) && maxt <= 512;
}

// Phase of the quasistatic driver (k_quasi) within an event.
enum QsPhase : int32_t {
    QS_BEGIN = 0,    // the event's displacement is still to be applied
    QS_MINIMISE = 1, // minimising after the displacement
    QS_RECORD = 2    // equilibrium reached: record the event
};

// Components recomputed by k_forces (fq_service.hpp).
enum ForceFlags : int32_t {
    FF_POT = 1,      // align + potential force
    FF_INT = 2,      // interactions
    FF_FRAME = 4,    // frame force
    FF_DAMP = 8,     // damping force
    FF_TOTAL = 16,   // residual force
    FF_ADD_U = 32,   // u += du first (advanceUniformly)
    FF_QUENCH = 64,  // v = a = 0 first
    FF_THERMAL = 128 // draw the due thermal forces first (updated_inc)
};

// One yield-position landscape row per particle (the reference's prrng chunk):
// y[p * nchunk + j], global index of y[p*nchunk] is start[p]; jloc[p] is the
// local index of the current well, y[j] <= u < y[j+1].
struct Landscape {
    double* y;
    int64_t* start;
    int32_t* jloc;
    uint64_t* st_begin; // pcg32 state after producing y[start]
    uint64_t* st_end;   // pcg32 state after producing y[start + nchunk - 1]
    int64_t nchunk;
    int64_t buffer;
    int64_t margin;
    uint64_t inc; // stream increment ((initseq << 1) | 1 = 1)
    int32_t dist;
    int32_t pad;
    double p0, p1, p2; // transform parameters
};

// Per-system scalars that live in device memory between launches.
struct RunState {
    int64_t inc;
    double u_frame;
    int64_t steps;  // steps completed by the current run (all launches)
    int64_t target; // steps the current run may take in total
    int32_t status;
    int32_t act_init; // time_activity: still waiting for the first event
    // StopList (GooseFEM Iterate::StopList) as counters, see fq_resident.hpp
    int32_t below;  // trailing count of residuals < tol
    int32_t below2; // trailing count of residuals < tol^2
    int32_t dec;    // trailing count of strictly decreasing pairs
    int32_t cur;    // tiled path: the state buffer (0: u, v, a; 1: u_b, v_b, a_b) holding the state
    double prev;    // last residual (+inf initially)
    int64_t s_n;    // activity sum of the previous step
    int64_t qs_first, qs_last;
    // thermal (RandomNormalForcing)
    uint64_t rng_state;
    uint64_t rng_inc;
    int64_t min_next;
    // reductions written by service kernels
    double red0;
    int32_t red_flag;
    int32_t pad2;
    // uniform displacement of the particles applied by k_forces(FF_ADD_U)
    double du_add;
    // quasistatic driver: next event of the run (0-based), its phase, and the
    // sum of (well index - i_ref) over the particles when the event began
    int64_t qs_event;
    int32_t qs_phase;
    // tiled blocks (k_tile_block): steps the next block must redo (0: a new
    // block), and the failure (TILE_FLAG_*) to report after committing them
    int32_t tb_redo;
    int64_t qs_isum;
    int32_t tb_fail;
    int32_t tb_pad;
};

// Device view of one system (all pointers are device pointers).
struct SysView {
    int64_t N, rows, cols;
    int32_t pot, inter, ext, pad;
    double m, inv_m, eta, mu, kappa, k_frame, dt;
    double k, a1, a2, k2, k4, k4_4, alpha;
    int64_t lr_m;
    double* pref; // LongRange1d prefactor [N]
    double *u, *v, *a, *f, *f_pot, *f_int, *f_frame, *f_damp, *f_th;
    // streaming path (systems larger than a resident workgroup): second state
    // buffer and the yield positions of the current well of both buffers
    double *u_b, *v_b, *a_b, *yl_a, *yr_a, *yl_b, *yr_b;
    int32_t* jl_b;
    int64_t* i_ref; // reference well index for activity / truncate / events
    int64_t *next, *dinc;
    double mean, stddev; // thermal draw
    uint64_t seed;       // landscape: particle p uses pcg32 stream (seed + p, 0)
    double offset;       // landscape: added to the first window
    Landscape ls;
    RunState* rs;
};

// particles per tile of the blocked tiled path of chains (k_tile_block), and
// the longest block of steps its extended segment allows
constexpr int64_t TB_TILE_PARTICLES = 1024;
constexpr int TB_KMAX_STEPS = 128;

// blocked tiled path of lattices (k_tile_block2d): a tile of rows x columns
// plus K halo rows and columns on every side holds at most TL_SITES sites
// (TL_PPT per thread of TL_THREADS) and is at most TL_WIDTH sites wide;
// halos (and longest blocks) of TL_K_DEFAULT steps unless set otherwise
constexpr int TL_THREADS = 1024;
constexpr int TL_PPT = 4;
constexpr int TL_SITES = TL_THREADS * TL_PPT;
constexpr int TL_WIDTH = 64;
constexpr int TL_K_DEFAULT = 8;

// One tile of the tiled paths (fq_tile.hpp): particles [p0, p1) of system
// `sys` (whole rows for lattices); tiles of the lattice blocks also give
// their rows [r0, r0 + nr) and columns [c0, c0 + nc).
struct TileDesc {
    int32_t sys;
    int32_t pad;
    int64_t p0, p1;
    int32_t r0, nr, c0, nc;
};

// Extra device pointers of a batched long-range group (all [R, N], row r =
// replica r), used by the matrix-core form of LongRange1d.
struct LrBuffers {
    double* un;   // slips of the step being computed
    double* part; // [nsplit][R][N] partial products K[:, I_s] (un - ub)[I_s]
    double* yl;   // current well, left
    double* yr;   // current well, right
    int64_t N;
    int64_t R;
    double c;   // sum of one row of the circulant matrix
    int nsplit; // number of partial products
};

// adjoint reverse pass (k_adj_reverse): threads per workgroup and the most
// particles of a tile's extended segment per thread
constexpr int ADJ_REV_THREADS = 512;
constexpr int ADJ_REV_MAX_SLOTS = 8;

// Buffers of the adjoint of a batch (fq_adjoint.hpp); [R*N] arrays are
// replica after replica. The tape holds, per step, R blocks of np = N entries
// (particle-major).
struct AdjBuffers {
    double* tu;     // tape [K][R*np]: u1 = u + dt v + dt^2/2 a of step n
    double* tv;     // tape: v at the start of step n
    double* ta;     // tape: a at the start of step n
    int32_t* tj;    // tape: chunk-local well index used by step n
    double* tuf;    // tape [K][R]: frame position of step n
    double* wu;     // state at the window start [R*N] (its checkpoint)
    double* wv;
    double* wa;
    int32_t* wj;
    double* wuf;    // [R] frame position at the window start
    double* au;     // adjoint state [R*N] (cotangent of u, v, a) after the window
    double* av;
    double* aa;
    double* bu;     // adjoint state [R*N] before the window (written by k_adj_reverse)
    double* bv;
    double* ba;
    double* gpar;   // [R][4] accumulated d/d(mu, k_interactions, k_frame, eta)
    double* gpart;  // [R][nt][4] the window's partial sums per tile
    double* gy;     // [R][N][nchunk] accumulated d/dy, or null
    int32_t* err;   // set to 1 if a well walk left the landscape window
    int64_t N;
    int64_t R;
    // k_adj_record over several windows at once (workgroup row blockIdx.y =
    // window w0 + blockIdx.y): strides between consecutive windows of the
    // start state (elements of wu / wj; of wuf: R) and of the tape (elements
    // of tu / tj; of tuf: K R), the window length K and the run length n
    int64_t ws, ts;
    int64_t K, n;
    int64_t w0;
    int64_t np; // tape entries per replica and step
    // k_adj_reverse: tiles per replica, particles per tile, halo width
    int32_t nt;
    int32_t pad1;
    int64_t tp, halo;
    // k_adj_record: tiles per replica (1: the whole periodic chain), particles
    // per tile, halo width
    int32_t rnt;
    int32_t pad2;
    int64_t rtp, rhalo;
};

struct RunArgs {
    int32_t mode;
    int32_t flow;          // flowSteps: u_frame += v_frame * dt before each step
    int32_t time_activity; // minimise: track quasistaticActivityFirst/Last
    int32_t pad;
    int64_t max_steps;     // steps this launch may take
    double v_frame;
    double tol, tol2;
    int64_t niter_tol;
    int64_t A_truncate, S_truncate;
};

// Arguments of the quasistatic driver (k_quasi): events k = first + e,
// e < nevents, of the protocol
//   k == 0:       u_frame = 0, then minimise
//   k odd:        eventDrivenStep(eps, kick = false)
//   k even > 0:   eventDrivenStep(eps, kick = true), then minimise
// (kicks != 0: every k > 0 a kick followed by minimise)
// and, after event e of system r, out_*[r * nevents + e]: S (sum of the well
// index changes), u_frame, mean frame force, inc.
struct QsArgs {
    int64_t first;
    int64_t nevents;
    double eps;
    int32_t direction;
    int32_t kicks;
    int64_t max_iter; // steps per minimisation
    int64_t* out_S;
    double* out_uf;
    double* out_ff;
    int64_t* out_inc;
};

// ---------------------------------------------------------------------------
// pcg32 (PCG-XSH-RR 64/32, O'Neill 2014)
// ---------------------------------------------------------------------------

constexpr uint64_t PCG_MULT = 6364136223846793005ULL;
constexpr uint64_t PCG_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;

FQ_HD uint32_t pcg_output(uint64_t old)
{
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

FQ_HD uint32_t pcg_next(uint64_t& state, uint64_t inc)
{
    uint64_t old = state;
    state = old * PCG_MULT + inc;
    return pcg_output(old);
}

FQ_HD void pcg_seed(uint64_t initstate, uint64_t initseq, uint64_t& state, uint64_t& inc)
{
    state = 0u;
    inc = (initseq << 1u) | 1u;
    pcg_next(state, inc);
    state += initstate;
    pcg_next(state, inc);
}

// Jump ahead by delta (negative: backwards), F. Brown 1994.
FQ_HD uint64_t pcg_advance(uint64_t state, uint64_t inc, int64_t delta)
{
    uint64_t cur_mult = PCG_MULT, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
    uint64_t d = (uint64_t)delta;
    while (d > 0) {
        if (d & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        d >>= 1u;
    }
    return acc_mult * state + acc_plus;
}

// One step back: the LCG state before `state` (inverse multiplier mod 2^64).
FQ_HD uint64_t pcg_step_back(uint64_t state, uint64_t inc)
{
    return (state - inc) * 0xc097ef87329e28a5ULL;
}

FQ_HD double u32_to_double(uint32_t x)
{
    uint64_t b = ((uint64_t)x << 20) | 0x3ff0000000000000ULL;
    return __builtin_bit_cast(double, b) - 1.0;
}

// ---------------------------------------------------------------------------
// Elementary functions (fdlibm 5.3 methods, see the oracle for the pins).
// log, exp, sin, cos and the argument reduction below follow fdlibm 5.3
// (e_log.c, e_exp.c, k_sin.c, k_cos.c, e_rem_pio2.c), whose notice is:
//
// Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
// Developed at SunPro, a Sun Microsystems, Inc. business.
// Permission to use, copy, modify, and distribute this
// software is freely granted, provided that this notice
// is preserved.
//
// The reduction always runs the second Cody-Waite round (fdlibm skips it
// when the first one is already accurate), so fq_sin / fq_cos can differ
// from fdlibm in the last bit; the CPU oracle does the same.
// ---------------------------------------------------------------------------

FQ_HD uint32_t hiword(double x)
{
    return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32);
}

FQ_HD double with_hiword(double x, uint32_t hi)
{
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    b = ((uint64_t)hi << 32) | (b & 0xffffffffULL);
    return __builtin_bit_cast(double, b);
}

constexpr double LN2_HI = 6.93147180369123816490e-01;
constexpr double LN2_LO = 1.90821492927058770002e-10;

FQ_HD double fq_log(double x)
{
    constexpr double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                     Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                     Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                     Lg7 = 1.479819860511658591e-01;
    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int32_t hx = (int32_t)(bits >> 32);
    uint32_t lx = (uint32_t)bits;
    int32_t k = 0;
    if (hx < 0x00100000) {
        if (((hx & 0x7fffffff) | (int32_t)lx) == 0) {
            return -__builtin_inf();
        }
        if (hx < 0) {
            return __builtin_nan("");
        }
        k -= 54;
        x *= 18014398509481984.0;
        hx = (int32_t)hiword(x);
    }
    if (hx >= 0x7ff00000) {
        return x + x;
    }
    k += (hx >> 20) - 1023;
    hx &= 0x000fffff;
    int32_t i = (hx + 0x95f64) & 0x100000;
    x = with_hiword(x, (uint32_t)(hx | (i ^ 0x3ff00000)));
    k += (i >> 20);
    double f = x - 1.0;
    double dk;
    if ((0x000fffff & (2 + hx)) < 3) {
        if (f == 0.0) {
            if (k == 0) {
                return 0.0;
            }
            dk = (double)k;
            return dk * LN2_HI + dk * LN2_LO;
        }
        double R = f * f * (0.5 - 0.33333333333333333 * f);
        if (k == 0) {
            return f - R;
        }
        dk = (double)k;
        return dk * LN2_HI - ((R - dk * LN2_LO) - f);
    }
    double s = f / (2.0 + f);
    dk = (double)k;
    double z = s * s;
    i = hx - 0x6147a;
    double w = z * z;
    int32_t j = 0x6b851 - hx;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    i |= j;
    double R = t2 + t1;
    if (i > 0) {
        double hfsq = 0.5 * f * f;
        if (k == 0) {
            return f - (hfsq - s * (hfsq + R));
        }
        return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
    }
    if (k == 0) {
        return f - s * (f - R);
    }
    return dk * LN2_HI - ((s * (f - R) - dk * LN2_LO) - f);
}

FQ_HD double fq_exp(double x)
{
    constexpr double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                     P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                     P5 = 4.13813679705723846039e-08, INVLN2 = 1.44269504088896338700e+00;
    if (x != x) {
        return x;
    }
    if (x > 7.09782712893383973096e+02) {
        return __builtin_inf();
    }
    if (x < -7.45133219101941108420e+02) {
        return 0.0;
    }
    uint32_t hx = hiword(x) & 0x7fffffff;
    int xsb = (int)((hiword(x) >> 31) & 1u);
    double hi = 0.0, lo = 0.0;
    int k = 0;
    if (hx > 0x3fd62e42) {
        if (hx < 0x3FF0A2B2) {
            hi = xsb ? x + LN2_HI : x - LN2_HI;
            lo = xsb ? -LN2_LO : LN2_LO;
            k = 1 - xsb - xsb;
        }
        else {
            k = (int)(INVLN2 * x + (xsb ? -0.5 : 0.5));
            double t = (double)k;
            hi = x - t * LN2_HI;
            lo = t * LN2_LO;
        }
        x = hi - lo;
    }
    else if (hx < 0x3e300000) {
        return 1.0 + x;
    }
    double t = x * x;
    double c = x - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
    if (k == 0) {
        return 1.0 - ((x * c) / (c - 2.0) - x);
    }
    double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
    uint64_t b = __builtin_bit_cast(uint64_t, y);
    if (k >= -1021) {
        b += (uint64_t)((int64_t)k) << 52;
        return __builtin_bit_cast(double, b);
    }
    b += (uint64_t)((int64_t)(k + 1000)) << 52;
    return __builtin_bit_cast(double, b) * 9.33263618503218878990e-302;
}

FQ_HD double fq_ksin(double x, double y)
{
    constexpr double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                     S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                     S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = x * x;
    double v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

FQ_HD double fq_kcos(double x, double y)
{
    constexpr double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                     C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                     C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    uint32_t ix = hiword(x) & 0x7fffffff;
    if (ix < 0x3e400000) {
        if ((int)x == 0) {
            return 1.0;
        }
    }
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    if (ix < 0x3FD33333) {
        return 1.0 - (0.5 * z - (z * r - x * y));
    }
    double qx;
    if (ix > 0x3fe90000) {
        qx = 0.28125;
    }
    else {
        qx = __builtin_bit_cast(double, ((uint64_t)(ix - 0x00200000)) << 32);
    }
    double hz = 0.5 * z - qx;
    double a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}

FQ_HD int fq_rem_pio2(double x, double& y0, double& y1)
{
    constexpr double INVPIO2 = 6.36619772367581382433e-01, PIO2_1 = 1.57079632673412561417e+00,
                     PIO2_1T = 6.07710050650619224932e-11, PIO2_2 = 6.07710050630396597660e-11,
                     PIO2_2T = 2.02226624879595063154e-21;
    double t = __builtin_fabs(x);
    int n = (int)(t * INVPIO2 + 0.5);
    double fn = (double)n;
    double r = t - fn * PIO2_1;
    double w = fn * PIO2_1T;
    double tt = r;
    w = fn * PIO2_2;
    r = tt - w;
    w = fn * PIO2_2T - ((tt - r) - w);
    double a = r - w;
    double b = (r - a) - w;
    if (x < 0.0) {
        y0 = -a;
        y1 = -b;
        return -n;
    }
    y0 = a;
    y1 = b;
    return n;
}

// sin / cos for |x| < 823549.6875 (the hot path only produces |x| <= 2 pi).
FQ_HD double fq_sin(double x)
{
    if (__builtin_fabs(x) < 0.7853981633974483) {
        return fq_ksin(x, 0.0);
    }
    double y0, y1;
    int n = fq_rem_pio2(x, y0, y1);
    switch (n & 3) {
        case 0:
            return fq_ksin(y0, y1);
        case 1:
            return fq_kcos(y0, y1);
        case 2:
            return -fq_ksin(y0, y1);
        default:
            return -fq_kcos(y0, y1);
    }
}

FQ_HD double fq_cos(double x)
{
    if (__builtin_fabs(x) < 0.7853981633974483) {
        return fq_kcos(x, 0.0);
    }
    double y0, y1;
    int n = fq_rem_pio2(x, y0, y1);
    switch (n & 3) {
        case 0:
            return fq_kcos(y0, y1);
        case 1:
            return -fq_ksin(y0, y1);
        case 2:
            return -fq_kcos(y0, y1);
        default:
            return fq_ksin(y0, y1);
    }
}

FQ_HD double fq_pow(double x, double y)
{
    if (y == 1.0) {
        return x;
    }
    if (y == 0.5) {
        return __builtin_sqrt(x);
    }
    if (y == 2.0) {
        return x * x;
    }
    if (x == 0.0) {
        return y > 0.0 ? 0.0 : __builtin_inf();
    }
    return fq_exp(y * fq_log(x));
}

// One landscape increment from one uniform r in [0, 1).
FQ_HD double transform(const Landscape& L, double r)
{
    switch (L.dist) {
        case DIST_RANDOM:
            return L.p0 * r + L.p1;
        case DIST_DELTA:
            return L.p0 + L.p1;
        case DIST_EXPONENTIAL:
            return L.p0 * (-fq_log(1.0 - r)) + L.p1;
        case DIST_POWER:
            return L.p1 * fq_pow(r, 1.0 / L.p0) + L.p2;
        case DIST_PARETO:
            return L.p1 * fq_pow(1.0 - r, -1.0 / L.p0) + L.p2;
        case DIST_WEIBULL:
            return L.p1 * fq_pow(-fq_log(1.0 - r), 1.0 / L.p0) + L.p2;
        default:
            return __builtin_nan("");
    }
}

FQ_HD double draw_increment(const Landscape& L, uint64_t& state)
{
    return transform(L, u32_to_double(pcg_next(state, L.inc)));
}

// Box-Muller normal from two consecutive uniforms (thermal force draw).
FQ_HD double draw_normal(uint64_t& state, uint64_t inc, double mean, double stddev)
{
    double r1 = u32_to_double(pcg_next(state, inc));
    double r2 = u32_to_double(pcg_next(state, inc));
    double rad = __builtin_sqrt(-2.0 * fq_log(1.0 - r1));
    double z = rad * fq_cos(6.283185307179586 * r2);
    return mean + stddev * z;
}

} // namespace fq
