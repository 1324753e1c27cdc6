/*
 * fqsb_oracle.c -- CPU oracle of the FrictionQPotSpringBlock dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product. Built with -ffp-contract=off so that every expression is evaluated
 * exactly in the order written, which is the order of the reference's xtensor
 * expressions (no fused multiply-add).
 *
 * What this file restates (all citations relative to the reference tree):
 *   policies      Cuspy          detail.h:139-205
 *                 SemiSmooth     detail.h:236-350
 *                 Smooth         detail.h:379-438
 *                 Laplace1d      detail.h:474-487
 *                 Laplace2d      detail.h:545-583
 *                 QuarticGradient1d detail.h:639-652
 *                 QuarticGradient2d detail.h:682-750
 *                 Quartic1d      detail.h:778-804
 *                 LongRange1d    detail.h:829-867
 *                 RandomNormalForcing detail.h:904-943
 *   engine        initSystem/refresh/computeForce/updated_*  detail.h:1096-1395
 *                 residual/temperature/quench  detail.h:1500-1532
 *                 timeStep       detail.h:1539-1570
 *                 timeSteps/timeStepsUntilEvent/flowSteps detail.h:1577-1645
 *                 minimise (inertial + no-passing) detail.h:1676-1792
 *                 minimise_truncate detail.h:1833-1893
 *                 maxUniformDisplacement/eventDrivenStep/trigger/
 *                 advanceToFixedForce/advanceUniformly detail.h:1901-2050
 *   constructors  Line1d.h:134-161 (seed + arange(N), initseq 0, m_gen += offset)
 *                 Line2d.h:28-34, 88-115
 *
 * External dependencies restated from their published algorithms:
 *   pcg32         PCG-XSH-RR 64/32 (M.E. O'Neill, "PCG: A Family of Simple Fast
 *                 Space-Efficient Statistically Good Algorithms for Random Number
 *                 Generation", 2014), seeding as in O'Neill's pcg32_srandom_r,
 *                 uint32 -> double by the [1,2)-mantissa trick, jump-ahead by
 *                 F. Brown, "Random Number Generation with Arbitrary Stride" (1994).
 *   cumsum chunk  prrng::pcg32_tensor_cumsum (Line1d.h:51-52): yield positions are
 *                 a per-particle running sum of i.i.d. increments; the first chunk
 *                 is cumsum(draws) and then `+= offset` (Line1d.h:157); later
 *                 chunks continue the running sum from the last held value.
 *   StopList      GooseFEM::Iterate::StopList (detail.h:1601): ring of the last n
 *                 residuals initialised to +inf; descending() is a strictly
 *                 decreasing check, all_less(tol) is "no entry >= tol".
 *   log/exp/sin/cos  the fdlibm 5.3 methods (Sun Microsystems 1993, freely
 *                 redistributable): e_log.c, e_exp.c, k_sin.c, k_cos.c and the
 *                 medium-range Cody-Waite reduction of e_rem_pio2.c. Used for the
 *                 Weibull/exponential/power transforms, the Box-Muller normal draw
 *                 and the Smooth potential, so that the CPU and the device compute
 *                 bit-identical values.
 *
 * Pins (tests/): the example golden files (exact S sequences, which fix the
 * pcg32 seeding, the uint32->double map, the uniform transform, the cumsum and
 * offset handling and the forward refill), the reference tests' closed forms,
 * brute-force landscape checks, and numpy for the math functions. Parity
 * UNPINNED (no runnable source of truth in this container): prrng's Weibull,
 * exponential, power and normal transforms, backward chunk regeneration, the
 * exact chunk shift policy (alignment buffer/margin), and GooseFEM's StopList
 * tie semantics.
 */
#include "fqsb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* errors                                                                    */
/* ------------------------------------------------------------------------- */

static _Thread_local char g_err[512];

static int fail(const char* msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return 1;
}

const char* fqo_last_error(void)
{
    return g_err;
}

/* ------------------------------------------------------------------------- */
/* fdlibm-method elementary functions                                        */
/* ------------------------------------------------------------------------- */
/*
 * log, exp, sin, cos and the argument reduction below follow fdlibm 5.3
 * (e_log.c, e_exp.c, k_sin.c, k_cos.c, e_rem_pio2.c), whose notice is:
 *
 * Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
 *
 * Developed at SunPro, a Sun Microsystems, Inc. business.
 * Permission to use, copy, modify, and distribute this
 * software is freely granted, provided that this notice
 * is preserved.
 *
 * rem_pio2 always runs the second Cody-Waite round (fdlibm skips it when the
 * first is already accurate), so sin/cos may differ from fdlibm in the last
 * bit; the device code does the same.
 */

typedef union {
    double d;
    uint64_t u;
} dbits;

static inline uint32_t hiword(double x)
{
    dbits b;
    b.d = x;
    return (uint32_t)(b.u >> 32);
}

static inline double with_hiword(double x, uint32_t hi)
{
    dbits b;
    b.d = x;
    b.u = ((uint64_t)hi << 32) | (b.u & 0xffffffffULL);
    return b.d;
}

static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;

/* natural logarithm, fdlibm e_log.c method: x = 2^k (1+f), s = f/(2+f),
 * log(1+f) = f - s(f - R(s^2)) with a degree-14 minimax R. */
double fqo_log(double x)
{
    static const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                        Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                        Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                        Lg7 = 1.479819860511658591e-01;
    dbits b;
    b.d = x;
    int32_t hx = (int32_t)(b.u >> 32);
    uint32_t lx = (uint32_t)b.u;
    int32_t k = 0;
    if (hx < 0x00100000) {
        if (((hx & 0x7fffffff) | (int32_t)lx) == 0) {
            return -INFINITY;
        }
        if (hx < 0) {
            return NAN;
        }
        k -= 54;
        x *= 18014398509481984.0; /* 2^54 */
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

/* exponential, fdlibm e_exp.c method: x = k ln2 + r, exp(r) from a Remez
 * rational form in r^2, scaled by 2^k. */
double fqo_exp(double x)
{
    static const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                        P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                        P5 = 4.13813679705723846039e-08, INVLN2 = 1.44269504088896338700e+00;
    if (x != x) {
        return x;
    }
    if (x > 7.09782712893383973096e+02) {
        return INFINITY;
    }
    if (x < -7.45133219101941108420e+02) {
        return 0.0;
    }
    uint32_t hx = hiword(x) & 0x7fffffff;
    int xsb = (hiword(x) >> 31) & 1;
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
    if (k >= -1021) {
        dbits b;
        b.d = y;
        b.u += (uint64_t)((int64_t)k) << 52;
        return b.d;
    }
    dbits b;
    b.d = y;
    b.u += (uint64_t)((int64_t)(k + 1000)) << 52;
    return b.d * 9.33263618503218878990e-302; /* 2^-1000 */
}

/* kernel sin on [-pi/4, pi/4], x + y the reduced argument (fdlibm k_sin.c). */
static double ksin(double x, double y)
{
    static const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                        S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                        S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = x * x;
    double v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

/* kernel cos on [-pi/4, pi/4] (fdlibm k_cos.c). */
static double kcos(double x, double y)
{
    static const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
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
        dbits b;
        b.u = ((uint64_t)(ix - 0x00200000)) << 32;
        qx = b.d;
    }
    double hz = 0.5 * z - qx;
    double a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}

/* Reduce x (|x| < 2^19 pi/2) to n*pi/2 + (y0 + y1), two Cody-Waite rounds
 * (33+33+53 bits of pi/2, the medium-size branch of fdlibm e_rem_pio2.c). */
static int rem_pio2(double x, double* y0, double* y1)
{
    static const double INVPIO2 = 6.36619772367581382433e-01, PIO2_1 = 1.57079632673412561417e+00,
                        PIO2_1T = 6.07710050650619224932e-11, PIO2_2 = 6.07710050630396597660e-11,
                        PIO2_2T = 2.02226624879595063154e-21;
    double t = fabs(x);
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
        *y0 = -a;
        *y1 = -b;
        return -n;
    }
    *y0 = a;
    *y1 = b;
    return n;
}

double fqo_sin(double x)
{
    if (fabs(x) < 0.7853981633974483) {
        return ksin(x, 0.0);
    }
    if (!(fabs(x) < 823549.6875)) {
        return sin(x); /* outside the range ever produced on the hot path */
    }
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    switch (n & 3) {
        case 0:
            return ksin(y0, y1);
        case 1:
            return kcos(y0, y1);
        case 2:
            return -ksin(y0, y1);
        default:
            return -kcos(y0, y1);
    }
}

double fqo_cos(double x)
{
    if (fabs(x) < 0.7853981633974483) {
        return kcos(x, 0.0);
    }
    if (!(fabs(x) < 823549.6875)) {
        return cos(x);
    }
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    switch (n & 3) {
        case 0:
            return kcos(y0, y1);
        case 1:
            return -ksin(y0, y1);
        case 2:
            return -kcos(y0, y1);
        default:
            return ksin(y0, y1);
    }
}

/* x^y for x >= 0: exact special cases, sqrt for y = 1/2, else exp(y log x). */
double fqo_pow(double x, double y)
{
    if (y == 1.0) {
        return x;
    }
    if (y == 0.5) {
        return sqrt(x);
    }
    if (y == 2.0) {
        return x * x;
    }
    if (x == 0.0) {
        return y > 0.0 ? 0.0 : INFINITY;
    }
    return fqo_exp(y * fqo_log(x));
}

/* ------------------------------------------------------------------------- */
/* pcg32                                                                     */
/* ------------------------------------------------------------------------- */

#define PCG_MULT 6364136223846793005ULL
#define PCG_DEFAULT_STREAM 0xda3e39cb94b95bdbULL

static inline uint32_t pcg_output(uint64_t old)
{
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

uint32_t fqo_pcg32_next(uint64_t* state, uint64_t inc)
{
    uint64_t old = *state;
    *state = old * PCG_MULT + inc;
    return pcg_output(old);
}

void fqo_pcg32_seed(uint64_t initstate, uint64_t initseq, uint64_t* state, uint64_t* inc)
{
    *state = 0u;
    *inc = (initseq << 1u) | 1u;
    fqo_pcg32_next(state, *inc);
    *state += initstate;
    fqo_pcg32_next(state, *inc);
}

uint64_t fqo_pcg32_advance(uint64_t state, uint64_t inc, int64_t delta)
{
    uint64_t cur_mult = PCG_MULT, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
    uint64_t d = (uint64_t)delta; /* the period is 2^64: negative = wrap-around */
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

double fqo_uint32_to_double(uint32_t x)
{
    dbits b;
    b.u = ((uint64_t)x << 20) | 0x3ff0000000000000ULL;
    return b.d - 1.0;
}

/* One increment of the landscape from one uniform r in [0,1). Parameters
 * follow the reference's usage: random [scale, offset], delta [value, offset],
 * exponential [scale, offset], power [k, scale, offset], pareto [k, scale,
 * offset], weibull [k, scale, offset] (test_Line1d.py:289, 302). */
typedef struct {
    int dist;
    double p0, p1, p2;
} fq_transform;

static inline double transform(const fq_transform* t, double r)
{
    switch (t->dist) {
        case FQO_DIST_RANDOM:
            return t->p0 * r + t->p1;
        case FQO_DIST_DELTA:
            return t->p0 + t->p1;
        case FQO_DIST_EXPONENTIAL:
            return t->p0 * (-fqo_log(1.0 - r)) + t->p1;
        case FQO_DIST_POWER:
            return t->p1 * fqo_pow(r, 1.0 / t->p0) + t->p2;
        case FQO_DIST_PARETO:
            return t->p1 * fqo_pow(1.0 - r, -1.0 / t->p0) + t->p2;
        case FQO_DIST_WEIBULL:
            return t->p1 * fqo_pow(-fqo_log(1.0 - r), 1.0 / t->p0) + t->p2;
        default:
            return NAN;
    }
}

static int make_transform(int dist, const double* par, int npar, fq_transform* t)
{
    double p[3] = {0.0, 0.0, 0.0};
    switch (dist) {
        case FQO_DIST_RANDOM:
        case FQO_DIST_DELTA:
        case FQO_DIST_EXPONENTIAL:
            p[0] = 1.0;
            p[1] = 0.0;
            if (npar > 2) {
                return fail("Too many parameters for distribution");
            }
            break;
        case FQO_DIST_POWER:
        case FQO_DIST_PARETO:
        case FQO_DIST_WEIBULL:
            p[0] = 1.0;
            p[1] = 1.0;
            p[2] = 0.0;
            if (npar > 3) {
                return fail("Too many parameters for distribution");
            }
            break;
        default:
            return fail("Distribution not supported for the yield-position landscape");
    }
    for (int i = 0; i < npar; ++i) {
        p[i] = par[i];
    }
    t->dist = dist;
    t->p0 = p[0];
    t->p1 = p[1];
    t->p2 = p[2];
    return 0;
}

static inline double draw_one(const fq_transform* t, uint64_t* state, uint64_t inc)
{
    return transform(t, fqo_uint32_to_double(fqo_pcg32_next(state, inc)));
}

int fqo_draw(int distribution, const double* parameters, int nparameters, uint64_t* state, uint64_t inc, double* out,
             uint64_t n)
{
    fq_transform t;
    if (make_transform(distribution, parameters, nparameters, &t)) {
        return 1;
    }
    for (uint64_t i = 0; i < n; ++i) {
        out[i] = draw_one(&t, state, inc);
    }
    return 0;
}

/* Box-Muller normal from two consecutive uniforms (the thermal force draw). */
static inline double draw_normal(uint64_t* state, uint64_t inc, double mean, double stddev)
{
    double r1 = fqo_uint32_to_double(fqo_pcg32_next(state, inc));
    double r2 = fqo_uint32_to_double(fqo_pcg32_next(state, inc));
    double rad = sqrt(-2.0 * fqo_log(1.0 - r1));
    double z = rad * fqo_cos(6.283185307179586 * r2);
    return mean + stddev * z;
}

int fqo_draw_normal(uint64_t* state, uint64_t inc, double mean, double stddev, double* out, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) {
        out[i] = draw_normal(state, inc, mean, stddev);
    }
    return 0;
}

uint64_t fqo_pcg32_default_stream(void)
{
    return PCG_DEFAULT_STREAM;
}

/* ------------------------------------------------------------------------- */
/* chunked landscape                                                         */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t N, n;
    double* y;           /* [N][n] */
    int64_t* start;      /* global index of y[p][0] */
    int64_t* j;          /* local index: y[j] <= u < y[j+1] */
    uint64_t* st_begin;  /* pcg state after producing y[start] */
    uint64_t* st_end;    /* pcg state after producing y[start + n - 1] */
    uint64_t inc;        /* stream increment (initseq 0 -> 1) */
    fq_transform tr;
    int64_t buffer, margin;
} fq_chunk;

/* Move the window of particle p so that it starts at global index ns. Forward
 * values continue the running sum from the last held value; backward values
 * subtract the increment that produced the value to their right. */
static void chunk_shift_to(fq_chunk* c, uint64_t p, int64_t ns)
{
    const int64_t n = (int64_t)c->n;
    double* y = c->y + p * c->n;
    int64_t s = c->start[p];
    int64_t d = ns - s;
    if (d == 0) {
        return;
    }
    if (d > 0) {
        double val = y[n - 1];
        uint64_t st = c->st_end[p];
        if (d < n) {
            memmove(y, y + d, (size_t)(n - d) * sizeof(double));
        }
        for (int64_t g = s + n; g < ns + n; ++g) {
            val = val + draw_one(&c->tr, &st, c->inc);
            if (g >= ns) {
                y[g - ns] = val;
            }
        }
        c->st_end[p] = st;
        c->st_begin[p] = fqo_pcg32_advance(c->st_begin[p], c->inc, d);
    }
    else {
        int64_t e = -d;
        double val = y[0];
        uint64_t st = c->st_begin[p]; /* state after producing y[s] */
        if (e < n) {
            memmove(y + e, y, (size_t)(n - e) * sizeof(double));
        }
        for (int64_t g = s; g > ns; --g) {
            /* y[g-1] = y[g] - d_g, d_g drawn from the state that produced it */
            st = (st - c->inc) * 0xc097ef87329e28a5ULL; /* one step back */
            uint64_t tmp = st;
            double dg = draw_one(&c->tr, &tmp, c->inc);
            val = val - dg;
            if (g - 1 - ns < n) {
                y[g - 1 - ns] = val;
            }
        }
        c->st_begin[p] = fqo_pcg32_advance(c->st_begin[p], c->inc, d);
        c->st_end[p] = fqo_pcg32_advance(c->st_end[p], c->inc, d);
    }
    c->start[p] = ns;
}

/* Align particle p to slip u: afterwards y[j] <= u < y[j+1] with
 * buffer <= j <= n-2-buffer (own alignment rule standing in for
 * prrng::alignment(buffer=2, margin=30, min_margin=6, strict=false)). */
static void chunk_align_p(fq_chunk* c, uint64_t p, double u)
{
    const int64_t n = (int64_t)c->n;
    double* y = c->y + p * c->n;
    int64_t j = c->j[p];
    if (y[j] <= u && u < y[j + 1]) {
        return;
    }
    if (!isfinite(u)) {
        return;
    }
    const int64_t b = c->buffer, m = c->margin;
    /* bounded search: a slip further than 2^20 wells away is not followed */
    int64_t budget = (int64_t)1 << 20;
    while (!(u < y[n - 1])) {
        if (budget <= 0) {
            c->j[p] = n - 2;
            return;
        }
        budget -= n - 1 - m;
        chunk_shift_to(c, p, c->start[p] + (n - 1 - m));
        j = m;
    }
    while (u < y[0]) {
        if (budget <= 0) {
            c->j[p] = 0;
            return;
        }
        budget -= n - 1 - m;
        chunk_shift_to(c, p, c->start[p] - (n - 1 - m));
        j = n - 1 - m;
    }
    if (j < 0) {
        j = 0;
    }
    if (j > n - 2) {
        j = n - 2;
    }
    while (u >= y[j + 1]) {
        ++j;
    }
    while (u < y[j]) {
        --j;
    }
    if (j < b || j > n - 2 - b) {
        chunk_shift_to(c, p, c->start[p] + j - m);
        j = m;
    }
    c->j[p] = j;
}

static void chunk_align(fq_chunk* c, const double* u)
{
    for (uint64_t p = 0; p < c->N; ++p) {
        chunk_align_p(c, p, u[p]);
    }
}

static int chunk_init(fq_chunk* c, uint64_t N, uint64_t n, uint64_t seed, int dist, const double* par, int npar,
                      double offset, int64_t buffer, int64_t margin)
{
    if (n < 8) {
        return fail("nchunk must be at least 8");
    }
    if (make_transform(dist, par, npar, &c->tr)) {
        return 1;
    }
    c->N = N;
    c->n = n;
    c->y = (double*)malloc(sizeof(double) * N * n);
    c->start = (int64_t*)calloc(N, sizeof(int64_t));
    c->j = (int64_t*)calloc(N, sizeof(int64_t));
    c->st_begin = (uint64_t*)malloc(sizeof(uint64_t) * N);
    c->st_end = (uint64_t*)malloc(sizeof(uint64_t) * N);
    if (!c->y || !c->start || !c->j || !c->st_begin || !c->st_end) {
        return fail("out of memory");
    }
    int64_t b = buffer < 0 ? 0 : buffer;
    if (b > ((int64_t)n - 2) / 4) {
        b = ((int64_t)n - 2) / 4;
    }
    int64_t m = margin < b ? b : margin;
    if (m > (int64_t)n - 2 - b) {
        m = (int64_t)n - 2 - b;
    }
    c->buffer = b;
    c->margin = m;
    for (uint64_t p = 0; p < N; ++p) {
        uint64_t st, inc;
        fqo_pcg32_seed(seed + p, 0u, &st, &inc);
        c->inc = inc;
        double* y = c->y + p * n;
        double cs = 0.0;
        for (uint64_t g = 0; g < n; ++g) {
            double d = draw_one(&c->tr, &st, inc);
            cs = (g == 0) ? d : cs + d;
            y[g] = cs;
            if (g == 0) {
                c->st_begin[p] = st;
            }
        }
        c->st_end[p] = st;
        for (uint64_t g = 0; g < n; ++g) {
            y[g] = y[g] + offset;
        }
        c->start[p] = 0;
        c->j[p] = 0;
    }
    return 0;
}

static void chunk_free(fq_chunk* c)
{
    free(c->y);
    free(c->start);
    free(c->j);
    free(c->st_begin);
    free(c->st_end);
}

/* ------------------------------------------------------------------------- */
/* system                                                                    */
/* ------------------------------------------------------------------------- */

struct fqo_system {
    fqo_desc d;
    uint64_t N, rows, cols;
    double m, inv_m, eta, mu, kappa, k_frame, dt, u_frame;
    double k, a1, a2, k2, k4, k4_4, alpha;
    double* pref; /* LongRange1d prefactor */
    int64_t lr_m;
    double *u, *v, *a, *f, *f_pot, *f_int, *f_frame, *f_damp, *f_th, *v_n, *a_n;
    int64_t inc, qs_first, qs_last;
    fq_chunk ch;
    /* RandomNormalForcing */
    uint64_t rng_state, rng_inc;
    int64_t *next, *dinc;
    int has_ext;
    /* smallest relative change of one residual that would change the step at
     * which the last minimise-type call stopped (see stoplist_margin) */
    double stop_margin;
};

uint64_t fqo_size(const fqo_system* s)
{
    return s->N;
}

static inline double* field_ptr(const fqo_system* s, int field)
{
    switch (field) {
        case FQO_U:
            return s->u;
        case FQO_V:
            return s->v;
        case FQO_A:
            return s->a;
        case FQO_F:
            return s->f;
        case FQO_F_POTENTIAL:
            return s->f_pot;
        case FQO_F_INTERACTIONS:
            return s->f_int;
        case FQO_F_FRAME:
            return s->f_frame;
        case FQO_F_DAMPING:
            return s->f_damp;
        case FQO_F_THERMAL:
            return s->f_th;
        default:
            return NULL;
    }
}

/* --- policies ------------------------------------------------------------ */

/* Cuspy::force detail.h:139-170 */
static void force_cuspy(fqo_system* s)
{
    chunk_align(&s->ch, s->u);
    const uint64_t n = s->ch.n;
    for (uint64_t p = 0; p < s->N; ++p) {
        const double* l = s->ch.y + p * n + s->ch.j[p];
        s->f_pot[p] = 0.5 * (l[0] + l[1]) - s->u[p];
    }
    for (uint64_t p = 0; p < s->N; ++p) {
        s->f_pot[p] *= s->mu;
    }
}

/* SemiSmooth::force detail.h:236-277 */
static void force_semismooth(fqo_system* s)
{
    chunk_align(&s->ch, s->u);
    const uint64_t n = s->ch.n;
    const double mu = s->mu, kappa = s->kappa;
    for (uint64_t p = 0; p < s->N; ++p) {
        const double* y = s->ch.y + p * n + s->ch.j[p];
        double xi = 0.5 * (y[0] + y[1]);
        double u_r = (mu * xi + kappa * y[1]) / (mu + kappa);
        double u_l = (mu * xi + kappa * y[0]) / (mu + kappa);
        double up = s->u[p];
        if (up < u_l) {
            s->f_pot[p] = kappa * (up - y[0]);
        }
        else if (up <= u_r) {
            s->f_pot[p] = mu * (0.5 * (y[0] + y[1]) - up);
        }
        else {
            s->f_pot[p] = kappa * (up - y[1]);
        }
    }
}

/* Smooth::force detail.h:379-409 (sin evaluated with fqo_sin) */
static void force_smooth(fqo_system* s)
{
    chunk_align(&s->ch, s->u);
    const uint64_t n = s->ch.n;
    const double PI = 3.14159265358979323846;
    for (uint64_t p = 0; p < s->N; ++p) {
        const double* y = s->ch.y + p * n + s->ch.j[p];
        double up = s->u[p];
        double umin = 0.5 * (y[0] + y[1]);
        double dy = 0.5 * (y[1] - y[0]);
        s->f_pot[p] = -s->mu * dy / PI * fqo_sin(PI * (up - umin) / dy);
    }
}

static void compute_force_potential(fqo_system* s)
{
    switch (s->d.potential) {
        case FQO_POT_CUSPY:
            force_cuspy(s);
            break;
        case FQO_POT_SEMISMOOTH:
            force_semismooth(s);
            break;
        default:
            force_smooth(s);
            break;
    }
}

static inline uint64_t per(int64_t i, uint64_t n)
{
    int64_t r = i % (int64_t)n;
    return (uint64_t)(r < 0 ? r + (int64_t)n : r);
}

static void compute_force_interactions(fqo_system* s)
{
    const uint64_t N = s->N;
    const double* u = s->u;
    double* f = s->f_int;
    switch (s->d.interactions) {
        case FQO_INT_NONE:
            break;
        case FQO_INT_LAPLACE1D: /* detail.h:474-487 */
            for (uint64_t p = 1; p < N - 1; ++p) {
                f[p] = u[p - 1] - 2 * u[p] + u[p + 1];
            }
            f[0] = u[N - 1] - 2 * u[0] + u[1];
            f[N - 1] = u[N - 2] - 2 * u[N - 1] + u[0];
            for (uint64_t p = 0; p < N; ++p) {
                f[p] *= s->k;
            }
            break;
        case FQO_INT_QUARTIC1D: /* detail.h:778-804 */
            for (uint64_t p = 0; p < N; ++p) {
                uint64_t l = p == 0 ? N - 1 : p - 1;
                uint64_t r = p == N - 1 ? 0 : p + 1;
                double dup = u[r] - u[p];
                double dun = u[l] - u[p];
                f[p] = s->a1 * (u[l] - 2 * u[p] + u[r]) + s->a2 * (dup * dup * dup + dun * dun * dun);
            }
            break;
        case FQO_INT_QUARTICGRADIENT1D: /* detail.h:639-652 */
            for (uint64_t p = 0; p < N; ++p) {
                uint64_t l = p == 0 ? N - 1 : p - 1;
                uint64_t r = p == N - 1 ? 0 : p + 1;
                double du = u[r] - u[l];
                f[p] = (u[l] - 2 * u[p] + u[r]) * (s->k2 + s->k4_4 * du * du);
            }
            break;
        case FQO_INT_LONGRANGE1D: { /* detail.h:849-867 */
            const int64_t n = (int64_t)N;
            for (int64_t p = 0; p < n; ++p) {
                double fp = 0.0;
                double up = u[p];
                for (int64_t i = 0; i < n; ++i) {
                    if (i == p) {
                        continue;
                    }
                    int64_t dd = i > p ? i - p : p - i;
                    if (dd > s->lr_m) {
                        dd = n - dd;
                    }
                    fp += (u[i] - up) * s->pref[dd];
                }
                f[p] = fp;
            }
            break;
        }
        case FQO_INT_LAPLACE2D: { /* detail.h:545-583 */
            const uint64_t R = s->rows, C = s->cols;
            for (uint64_t i = 0; i < R; ++i) {
                for (uint64_t j = 0; j < C; ++j) {
                    double um = u[per((int64_t)i - 1, R) * C + j];
                    double upp = u[per((int64_t)i + 1, R) * C + j];
                    double ul = u[i * C + per((int64_t)j - 1, C)];
                    double ur = u[i * C + per((int64_t)j + 1, C)];
                    f[i * C + j] = um + upp + ul + ur - 4 * u[i * C + j];
                }
            }
            for (uint64_t p = 0; p < N; ++p) {
                f[p] *= s->k;
            }
            break;
        }
        case FQO_INT_QUARTICGRADIENT2D: { /* detail.h:682-750 */
            const uint64_t R = s->rows, C = s->cols;
            const double mk4_3 = s->k4 / 3.0;
            const double mk4_23 = 2.0 * mk4_3;
            for (uint64_t i = 0; i < R; ++i) {
                uint64_t ip = per((int64_t)i + 1, R), im = per((int64_t)i - 1, R);
                for (uint64_t j = 0; j < C; ++j) {
                    uint64_t jp = per((int64_t)j + 1, C), jm = per((int64_t)j - 1, C);
                    double c0 = u[i * C + j];
                    double l = u[ip * C + j] + u[im * C + j] + u[i * C + jp] + u[i * C + jm] - 4 * c0;
                    double dudx = 0.5 * (u[ip * C + j] - u[im * C + j]);
                    double dudy = 0.5 * (u[i * C + jp] - u[i * C + jm]);
                    double d2udxdy =
                        0.25 * (u[ip * C + jp] - u[ip * C + jm] - u[im * C + jp] + u[im * C + jm]);
                    double d2udx2 = u[ip * C + j] - 2 * c0 + u[im * C + j];
                    double d2udy2 = u[i * C + jp] - 2 * c0 + u[i * C + jm];
                    f[i * C + j] = l * (s->k2 + mk4_3) +
                                   mk4_23 * (dudx * dudx * d2udx2 + dudy * dudy * d2udy2 + 2.0 * dudx * dudy * d2udxdy);
                }
            }
            break;
        }
    }
}

/* computeForceFrame detail.h:1353-1356 */
static void compute_force_frame(fqo_system* s)
{
    for (uint64_t p = 0; p < s->N; ++p) {
        s->f_frame[p] = s->k_frame * (s->u_frame - s->u[p]);
    }
}

/* computeForceDamping detail.h:1361-1364 */
static void compute_force_damping(fqo_system* s)
{
    const double me = -s->eta;
    for (uint64_t p = 0; p < s->N; ++p) {
        s->f_damp[p] = me * s->v[p];
    }
}

/* computeForce detail.h:1321-1330 */
static void compute_force(fqo_system* s)
{
    if (s->has_ext) {
        for (uint64_t p = 0; p < s->N; ++p) {
            s->f[p] = s->f_frame[p] + s->f_pot[p] + s->f_int[p] + s->f_damp[p] + s->f_th[p];
        }
    }
    else {
        for (uint64_t p = 0; p < s->N; ++p) {
            s->f[p] = s->f_frame[p] + s->f_pot[p] + s->f_int[p] + s->f_damp[p];
        }
    }
}

/* RandomNormalForcing::force detail.h:930-943 */
static void external_force(fqo_system* s)
{
    for (uint64_t p = 0; p < s->N; ++p) {
        if (s->inc >= s->next[p]) {
            s->f_th[p] = draw_normal(&s->rng_state, s->rng_inc, s->d.mean, s->d.stddev);
            s->next[p] += s->dinc[p];
        }
    }
}

/* updated_inc detail.h:1369-1375 */
static void updated_inc(fqo_system* s)
{
    if (s->has_ext) {
        external_force(s);
    }
    compute_force(s);
}

/* updated_u detail.h:1380-1386 */
static void updated_u(fqo_system* s)
{
    compute_force_potential(s);
    compute_force_interactions(s);
    compute_force_frame(s);
    compute_force(s);
}

/* updated_v detail.h:1391-1395 */
static void updated_v(fqo_system* s)
{
    compute_force_damping(s);
    compute_force(s);
}

int fqo_refresh(fqo_system* s)
{
    updated_u(s);
    updated_v(s);
    updated_inc(s);
    return 0;
}

int fqo_create(const fqo_desc* d, fqo_system** out)
{
    *out = NULL;
    if (d->rank != 1 && d->rank != 2) {
        return fail("rank must be 1 or 2");
    }
    uint64_t rows = d->rows, cols = d->rank == 1 ? 1 : d->cols;
    uint64_t N = rows * cols;
    if (N < 1) {
        return fail("empty system");
    }
    if (d->interactions == FQO_INT_LAPLACE1D || d->interactions == FQO_INT_QUARTIC1D ||
        d->interactions == FQO_INT_QUARTICGRADIENT1D) {
        if (N < 3) {
            return fail("a chain needs at least 3 particles");
        }
    }
    if ((d->interactions == FQO_INT_LAPLACE2D || d->interactions == FQO_INT_QUARTICGRADIENT2D) && d->rank != 2) {
        return fail("2d interactions need rank 2");
    }
    fqo_system* s = (fqo_system*)calloc(1, sizeof(fqo_system));
    if (!s) {
        return fail("out of memory");
    }
    s->stop_margin = INFINITY;
    s->d = *d;
    s->N = N;
    s->rows = rows;
    s->cols = cols;
    /* initSystem detail.h:1096-1139 */
    s->m = d->m;
    s->inv_m = 1.0 / d->m;
    s->eta = d->eta;
    s->mu = d->mu;
    s->kappa = d->kappa;
    s->k_frame = d->k_frame;
    s->dt = d->dt;
    s->u_frame = 0.0;
    s->k = d->k_interactions;
    s->a1 = d->a1;
    s->a2 = d->a2;
    s->k2 = d->k2;
    s->k4 = d->k4;
    s->k4_4 = 0.25 * d->k4;
    s->alpha = d->alpha;
    s->has_ext = d->external == FQO_EXT_RANDOMFORCING;
    double** arrs[] = {&s->u, &s->v, &s->a, &s->f, &s->f_pot, &s->f_int, &s->f_frame, &s->f_damp, &s->f_th, &s->v_n,
                       &s->a_n};
    for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); ++i) {
        *arrs[i] = (double*)calloc(N, sizeof(double));
        if (!*arrs[i]) {
            fqo_destroy(s);
            return fail("out of memory");
        }
    }
    if (d->interactions == FQO_INT_LONGRANGE1D) { /* detail.h:829-844 */
        const int64_t n = (int64_t)N;
        s->lr_m = (n - n % 2) / 2;
        s->pref = (double*)calloc(N, sizeof(double));
        for (int64_t dd = 1; dd < n; ++dd) {
            s->pref[dd] = d->k_interactions / pow((double)dd, d->alpha + 1.0);
        }
    }
    if (s->has_ext) { /* RandomNormalForcing ctor detail.h:904-922 */
        fqo_pcg32_seed(d->seed_forcing, PCG_DEFAULT_STREAM, &s->rng_state, &s->rng_inc);
        s->next = (int64_t*)malloc(sizeof(int64_t) * N);
        s->dinc = (int64_t*)malloc(sizeof(int64_t) * N);
        memcpy(s->next, d->dinc_init, sizeof(int64_t) * N);
        memcpy(s->dinc, d->dinc, sizeof(int64_t) * N);
    }
    /* Line1d.h:148-157: seeds seed + arange(N), initseq 0, then += offset */
    if (chunk_init(&s->ch, N, d->nchunk, d->seed, d->distribution, d->parameters, d->nparameters, d->offset,
                   d->align_buffer, d->align_margin)) {
        fqo_destroy(s);
        return 1;
    }
    fqo_refresh(s);
    *out = s;
    return 0;
}

void fqo_destroy(fqo_system* s)
{
    if (!s) {
        return;
    }
    free(s->u);
    free(s->v);
    free(s->a);
    free(s->f);
    free(s->f_pot);
    free(s->f_int);
    free(s->f_frame);
    free(s->f_damp);
    free(s->f_th);
    free(s->v_n);
    free(s->a_n);
    free(s->pref);
    free(s->next);
    free(s->dinc);
    chunk_free(&s->ch);
    free(s);
}

/* --- accessors (detail.h:1231-1305, 1402-1503) ----------------------------- */

int fqo_get(const fqo_system* s, int field, double* out)
{
    double* p = field_ptr(s, field);
    if (!p) {
        return fail("unknown field");
    }
    memcpy(out, p, sizeof(double) * s->N);
    return 0;
}

int fqo_set(fqo_system* s, int field, const double* in)
{
    switch (field) {
        case FQO_U: /* set_u detail.h:1276-1281 */
            memcpy(s->u, in, sizeof(double) * s->N);
            updated_u(s);
            return 0;
        case FQO_V: /* set_v detail.h:1290-1295 */
            memcpy(s->v, in, sizeof(double) * s->N);
            updated_v(s);
            return 0;
        case FQO_A: /* set_a detail.h:1301-1305 */
            memcpy(s->a, in, sizeof(double) * s->N);
            return 0;
        case FQO_F_THERMAL: /* RandomNormalForcing::set_f_thermal detail.h:976-980 */
            memcpy(s->f_th, in, sizeof(double) * s->N);
            return 0;
        default:
            return fail("field is read-only");
    }
}

int64_t fqo_get_inc(const fqo_system* s)
{
    return s->inc;
}

int fqo_set_inc(fqo_system* s, int64_t inc)
{
    s->inc = inc;
    s->qs_first = inc;
    s->qs_last = inc;
    updated_inc(s);
    return 0;
}

int fqo_set_t(fqo_system* s, double t)
{
    s->inc = (int64_t)round(t / s->dt);
    return 0;
}

double fqo_get_u_frame(const fqo_system* s)
{
    return s->u_frame;
}

int fqo_set_u_frame(fqo_system* s, double u_frame)
{
    s->u_frame = u_frame;
    compute_force_frame(s);
    compute_force(s);
    return 0;
}

int64_t fqo_qs_first(const fqo_system* s)
{
    return s->qs_first;
}

int64_t fqo_qs_last(const fqo_system* s)
{
    return s->qs_last;
}

static double norm_sq(const double* x, uint64_t n)
{
    double r = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        r += x[i] * x[i];
    }
    return r;
}

/* residual detail.h:1512-1520 */
double fqo_residual(const fqo_system* s)
{
    double r_fres = sqrt(norm_sq(s->f, s->N));
    double r_fext = sqrt(norm_sq(s->f_frame, s->N));
    if (r_fext != 0.0) {
        return r_fres / r_fext;
    }
    return r_fres;
}

/* temperature detail.h:1500-1503 */
double fqo_temperature(const fqo_system* s)
{
    return 0.5 * s->m * norm_sq(s->v, s->N) / (double)s->N;
}

/* quench detail.h:1527-1532 */
int fqo_quench(fqo_system* s)
{
    memset(s->v, 0, sizeof(double) * s->N);
    memset(s->a, 0, sizeof(double) * s->N);
    updated_v(s);
    return 0;
}

/* --- dynamics -------------------------------------------------------------- */

static int check_nan(const fqo_system* s)
{
    for (uint64_t p = 0; p < s->N; ++p) {
        if (isnan(s->u[p])) {
            return fail("NaN entries found");
        }
    }
    return 0;
}

/* timeStep detail.h:1539-1570 */
int fqo_time_step(fqo_system* s)
{
    if (s->d.minimisation == FQO_MIN_OVERDAMPED) {
        return fail("timeStep is not available for no-passing systems");
    }
    const uint64_t N = s->N;
    const double dt = s->dt;
    const double half_dt2 = 0.5 * dt * dt;
    const double half_dt = 0.5 * dt;
    s->inc++;
    if (s->has_ext) {
        updated_inc(s);
    }
    memcpy(s->v_n, s->v, sizeof(double) * N);
    memcpy(s->a_n, s->a, sizeof(double) * N);

    for (uint64_t p = 0; p < N; ++p) {
        s->u[p] = s->u[p] + dt * s->v[p] + half_dt2 * s->a[p];
    }
    updated_u(s);

    for (uint64_t p = 0; p < N; ++p) {
        s->v[p] = s->v_n[p] + dt * s->a_n[p];
    }
    updated_v(s);

    for (uint64_t p = 0; p < N; ++p) {
        s->a[p] = s->f[p] * s->inv_m;
    }

    for (int pass = 0; pass < 2; ++pass) {
        for (uint64_t p = 0; p < N; ++p) {
            s->v[p] = s->v_n[p] + half_dt * (s->a_n[p] + s->a[p]);
        }
        updated_v(s);
        for (uint64_t p = 0; p < N; ++p) {
            s->a[p] = s->f[p] * s->inv_m;
        }
    }
    return check_nan(s);
}

/* timeSteps detail.h:1577-1583 */
int fqo_time_steps(fqo_system* s, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) {
        if (fqo_time_step(s)) {
            return 1;
        }
    }
    return 0;
}

/* flowSteps detail.h:1637-1645 */
int fqo_flow_steps(fqo_system* s, uint64_t n, double v_frame)
{
    for (uint64_t i = 0; i < n; ++i) {
        s->u_frame += v_frame * s->dt;
        if (fqo_time_step(s)) {
            return 1;
        }
    }
    return 0;
}

/* GooseFEM::Iterate::StopList (used at detail.h:1601, 1689, 1851) */
typedef struct {
    double* r;
    uint64_t n;
} stoplist;

static int stoplist_init(stoplist* l, uint64_t n)
{
    l->n = n;
    l->r = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (!l->r) {
        return fail("out of memory");
    }
    for (uint64_t i = 0; i < n; ++i) {
        l->r[i] = INFINITY;
    }
    return 0;
}

static void stoplist_roll_insert(stoplist* l, double res)
{
    if (l->n == 0) {
        return;
    }
    memmove(l->r, l->r + 1, sizeof(double) * (l->n - 1));
    l->r[l->n - 1] = res;
}

static int stoplist_descending(const stoplist* l)
{
    for (uint64_t i = 1; i < l->n; ++i) {
        if (l->r[i] >= l->r[i - 1]) {
            return 0;
        }
    }
    return 1;
}

static int stoplist_all_less(const stoplist* l, double tol)
{
    for (uint64_t i = 0; i < l->n; ++i) {
        if (l->r[i] >= tol) {
            return 0;
        }
    }
    return 1;
}

static int stoplist_stop(const stoplist* l, double tol, double tol2)
{
    return (stoplist_descending(l) && stoplist_all_less(l, tol)) || stoplist_all_less(l, tol2);
}

/* Relative distance of a from the threshold b (infinite when either is not
 * finite: a comparison with the ring's initial +inf cannot flip). */
static inline double rel_gap(double a, double b)
{
    if (!isfinite(a) || !isfinite(b) || b == 0.0) {
        return INFINITY;
    }
    return fabs(a - b) / fabs(b);
}

/* How robust the StopList's decision at this step is: the smallest relative
 * perturbation of the residuals in the ring that would flip it. Not stopped:
 * each branch ((descending && all_less(tol)) or all_less(tol^2)) becomes true
 * only if every failing comparison of it flips, so the branch needs the
 * largest gap of its failing comparisons; the decision flips at the smaller of
 * the two branches. Stopped: every true branch must become false, each by its
 * smallest gap. Not a check of the reference: it measures how far the
 * residuals were from a tie, so that a residual computed in another summation
 * order (a few ulp away) is known to give the same stop step. */
static double stoplist_margin(const stoplist* l, double tol, double tol2, int stopped)
{
    const uint64_t n = l->n;
    if (n == 0) {
        return INFINITY;
    }
    double fail_ad = 0.0, fail_b = 0.0; /* largest gap of a failing comparison */
    double min_ad = INFINITY, min_b = INFINITY; /* smallest gap of any comparison */
    int ad = 1, b = 1;
    for (uint64_t i = 0; i < n; ++i) {
        const double x = l->r[i];
        double g = rel_gap(x, tol);
        min_ad = g < min_ad ? g : min_ad;
        if (!(x < tol)) {
            ad = 0;
            fail_ad = g > fail_ad ? g : fail_ad;
        }
        g = rel_gap(x, tol2);
        min_b = g < min_b ? g : min_b;
        if (!(x < tol2)) {
            b = 0;
            fail_b = g > fail_b ? g : fail_b;
        }
        if (i > 0) {
            g = rel_gap(x, l->r[i - 1]);
            min_ad = g < min_ad ? g : min_ad;
            if (!(x < l->r[i - 1])) {
                ad = 0;
                fail_ad = g > fail_ad ? g : fail_ad;
            }
        }
    }
    if (!stopped) {
        const double make_ad = ad ? 0.0 : fail_ad, make_b = b ? 0.0 : fail_b;
        return make_ad < make_b ? make_ad : make_b;
    }
    const double break_ad = ad ? min_ad : 0.0, break_b = b ? min_b : 0.0;
    return break_ad > break_b ? break_ad : break_b;
}

static inline void track_margin(fqo_system* s, const stoplist* l, double tol, double tol2, int stopped)
{
    const double m = stoplist_margin(l, tol, tol2, stopped);
    if (m < s->stop_margin) {
        s->stop_margin = m;
    }
}

double fqo_last_stop_margin(const fqo_system* s)
{
    return s->stop_margin;
}

int fqo_stoplist_margin(const double* residuals, uint64_t n, uint64_t niter_tol, double tol, double* margin)
{
    stoplist l;
    if (stoplist_init(&l, niter_tol)) {
        return 1;
    }
    double m = INFINITY;
    for (uint64_t i = 0; i < n; ++i) {
        stoplist_roll_insert(&l, residuals[i]);
        const int stop = stoplist_stop(&l, tol, tol * tol);
        const double g = stoplist_margin(&l, tol, tol * tol, stop);
        m = g < m ? g : m;
        if (stop) {
            break;
        }
    }
    *margin = m;
    free(l.r);
    return 0;
}

int fqo_stoplist_run(const double* residuals, uint64_t n, uint64_t niter_tol, double tol, int64_t* stop_index)
{
    stoplist l;
    if (stoplist_init(&l, niter_tol)) {
        return 1;
    }
    *stop_index = -1;
    for (uint64_t i = 0; i < n; ++i) {
        stoplist_roll_insert(&l, residuals[i]);
        if (stoplist_stop(&l, tol, tol * tol)) {
            *stop_index = (int64_t)i;
            break;
        }
    }
    free(l.r);
    return 0;
}

static inline int64_t gidx(const fqo_system* s, uint64_t p)
{
    return s->ch.start[p] + s->ch.j[p];
}

/* timeStepsUntilEvent detail.h:1595-1622 */
int fqo_time_steps_until_event(fqo_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, uint64_t* ret)
{
    if (s->d.minimisation == FQO_MIN_OVERDAMPED) {
        return fail("timeStepsUntilEvent is not available for no-passing systems");
    }
    double tol2 = tol * tol;
    stoplist res;
    if (stoplist_init(&res, niter_tol)) {
        return 1;
    }
    s->stop_margin = INFINITY;
    int64_t* i_n = (int64_t*)malloc(sizeof(int64_t) * s->N);
    for (uint64_t p = 0; p < s->N; ++p) {
        i_n[p] = gidx(s, p);
    }
    uint64_t step;
    int rc = 0;
    for (step = 1; step < max_iter + 1; ++step) {
        if (fqo_time_step(s)) {
            rc = 1;
            break;
        }
        int changed = 0;
        for (uint64_t p = 0; p < s->N; ++p) {
            if (gidx(s, p) != i_n[p]) {
                changed = 1;
                break;
            }
        }
        if (changed) {
            break;
        }
        stoplist_roll_insert(&res, fqo_residual(s));
        track_margin(s, &res, tol, tol2, stoplist_stop(&res, tol, tol2));
        if (stoplist_stop(&res, tol, tol2)) {
            fqo_quench(s);
            step = 0;
            break;
        }
    }
    *ret = step;
    free(i_n);
    free(res.r);
    return rc;
}

/* no-passing minimise, detail.h:1694-1753 */
static int minimise_nopassing(fqo_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                              uint64_t* ret)
{
    const uint64_t N = s->N;
    const uint64_t n = s->ch.n;
    double tol2 = tol * tol;
    stoplist res;
    if (stoplist_init(&res, niter_tol)) {
        return 1;
    }
    s->stop_margin = INFINITY;
    s->qs_first = s->inc;
    s->qs_last = s->inc;
    const double k = s->k;
    uint64_t step;
    for (step = 1; step < max_iter + 1; ++step) {
        memcpy(s->v_n, s->u, sizeof(double) * N);
        for (uint64_t p = 0; p < N; ++p) {
            double uneigh;
            if (p == 0) {
                uneigh = s->v_n[N - 1] + s->v_n[1];
            }
            else if (p == N - 1) {
                uneigh = s->v_n[N - 2] + s->v_n[0];
            }
            else {
                uneigh = s->v_n[p - 1] + s->v_n[p + 1];
            }
            int64_t i = s->ch.j[p];
            const double* y = s->ch.y + p * n;
            double u, umin;
            while (1) {
                umin = 0.5 * (y[i] + y[i + 1]);
                u = (k * uneigh + s->k_frame * s->u_frame + s->mu * umin) / (2 * k + s->k_frame + s->mu);
                chunk_align_p(&s->ch, p, u);
                int64_t j = s->ch.j[p];
                if (j == i) {
                    break;
                }
                i = j;
            }
            s->u[p] = u;
            s->f_frame[p] = s->k_frame * (s->u_frame - u);
            s->f_pot[p] = s->mu * (umin - u);
        }
        compute_force_interactions(s);
        for (uint64_t p = 0; p < N; ++p) {
            s->f[p] = s->f_pot[p] + s->f_int[p] + s->f_frame[p];
        }
        stoplist_roll_insert(&res, fqo_residual(s));
        track_margin(s, &res, tol, tol2, stoplist_stop(&res, tol, tol2));
        if (stoplist_stop(&res, tol, tol2)) {
            fqo_quench(s);
            free(res.r);
            *ret = 0;
            return 0;
        }
    }
    free(res.r);
    if (max_iter_is_error) {
        return fail("No convergence found");
    }
    *ret = step;
    return 0;
}

/* minimise detail.h:1676-1792 */
int fqo_minimise(fqo_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, int time_activity,
                 int max_iter_is_error, uint64_t* ret)
{
    if (s->d.minimisation == FQO_MIN_NONE) {
        return fail("Minimisation not implementated");
    }
    if (s->d.minimisation == FQO_MIN_OVERDAMPED) {
        if (s->d.interactions != FQO_INT_LAPLACE1D || s->d.potential != FQO_POT_CUSPY) {
            return fail("no-passing minimisation needs Cuspy + Laplace1d");
        }
        return minimise_nopassing(s, tol, niter_tol, max_iter, max_iter_is_error, ret);
    }
    double tol2 = tol * tol;
    stoplist res;
    if (stoplist_init(&res, niter_tol)) {
        return 1;
    }
    s->stop_margin = INFINITY;
    int64_t* i_n = NULL;
    int64_t sa = 0, s_n = 0;
    int init = 1;
    if (time_activity) {
        i_n = (int64_t*)malloc(sizeof(int64_t) * s->N);
        for (uint64_t p = 0; p < s->N; ++p) {
            i_n[p] = gidx(s, p);
        }
    }
    uint64_t step;
    for (step = 1; step < max_iter + 1; ++step) {
        if (fqo_time_step(s)) {
            free(i_n);
            free(res.r);
            return 1;
        }
        stoplist_roll_insert(&res, fqo_residual(s));
        if (time_activity) {
            sa = 0;
            for (uint64_t p = 0; p < s->N; ++p) {
                int64_t dd = gidx(s, p) - i_n[p];
                sa += dd < 0 ? -dd : dd;
            }
            if (sa != s_n) {
                if (init) {
                    init = 0;
                    s->qs_first = s->inc;
                }
                s->qs_last = s->inc;
            }
            s_n = sa;
        }
        track_margin(s, &res, tol, tol2, stoplist_stop(&res, tol, tol2));
        if (stoplist_stop(&res, tol, tol2)) {
            fqo_quench(s);
            free(i_n);
            free(res.r);
            *ret = 0;
            return 0;
        }
    }
    free(i_n);
    free(res.r);
    if (max_iter_is_error) {
        return fail("No convergence found");
    }
    *ret = step;
    return 0;
}

/* minimise_truncate detail.h:1833-1893 */
int fqo_minimise_truncate(fqo_system* s, const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate, double tol,
                          uint64_t niter_tol, uint64_t max_iter, int time_activity, int max_iter_is_error,
                          uint64_t* ret)
{
    (void)time_activity;
    if (s->d.minimisation != FQO_MIN_INERTIAL) {
        return fail("minimise_truncate is not available for this system");
    }
    double tol2 = tol * tol;
    stoplist res;
    if (stoplist_init(&res, niter_tol)) {
        return 1;
    }
    s->stop_margin = INFINITY;
    int64_t sa = 0, s_n = 0;
    int init = 1;
    uint64_t step;
    for (step = 1; step < max_iter + 1; ++step) {
        if (fqo_time_step(s)) {
            free(res.r);
            return 1;
        }
        stoplist_roll_insert(&res, fqo_residual(s));
        uint64_t a = 0;
        sa = 0;
        for (uint64_t p = 0; p < s->N; ++p) {
            int64_t dd = gidx(s, p) - i_n[p];
            a += dd != 0;
            sa += dd < 0 ? -dd : dd;
        }
        if (sa != s_n) {
            if (init) {
                init = 0;
                s->qs_first = s->inc;
            }
            s->qs_last = s->inc;
        }
        s_n = sa;
        track_margin(s, &res, tol, tol2, stoplist_stop(&res, tol, tol2));
        if (stoplist_stop(&res, tol, tol2)) {
            fqo_quench(s);
            free(res.r);
            *ret = 0;
            return 0;
        }
        if (A_truncate > 0 && a >= A_truncate) {
            free(res.r);
            *ret = step;
            return 0;
        }
        if (S_truncate > 0 && sa >= (int64_t)S_truncate) {
            free(res.r);
            *ret = step;
            return 0;
        }
    }
    free(res.r);
    if (max_iter_is_error) {
        return fail("No convergence found");
    }
    *ret = step;
    return 0;
}

/* maxUniformDisplacement detail.h:1901-1905 -> Cuspy detail.h:176-187,
 * SemiSmooth detail.h:282-334, Smooth detail.h:414-422 */
int fqo_max_uniform_displacement(fqo_system* s, int direction, double* ret)
{
    const uint64_t n = s->ch.n;
    if (s->d.potential == FQO_POT_SMOOTH) {
        return fail("Operation not possible.");
    }
    chunk_align(&s->ch, s->u);
    if (s->d.potential == FQO_POT_CUSPY) {
        double r = INFINITY;
        for (uint64_t p = 0; p < s->N; ++p) {
            const double* y = s->ch.y + p * n + s->ch.j[p];
            double v = direction > 0 ? y[1] - s->u[p] : s->u[p] - y[0];
            if (v < r) {
                r = v;
            }
        }
        *ret = r;
        return 0;
    }
    const double mu = s->mu, kappa = s->kappa;
    double r = INFINITY;
    for (uint64_t p = 0; p < s->N; ++p) {
        const double* y = s->ch.y + p * n + s->ch.j[p];
        double xi = 0.5 * (y[0] + y[1]);
        double u_r = (mu * xi + kappa * y[1]) / (mu + kappa);
        double u_l = (mu * xi + kappa * y[0]) / (mu + kappa);
        double up = s->u[p];
        if (up < u_l) {
            *ret = 0.0;
            return 0;
        }
        else if (up <= u_r) {
            double v = direction > 0 ? u_r - up : up - u_l;
            if (v < r) {
                r = v;
            }
        }
        else {
            *ret = 0.0;
            return 0;
        }
    }
    *ret = r;
    return 0;
}

/* advanceUniformly detail.h:2027-2050 (input_is_frame = false) */
static double advance_uniformly(fqo_system* s, double du)
{
    double du_particles = du;
    double du_frame = du * (s->k_frame + s->mu) / s->k_frame;
    for (uint64_t p = 0; p < s->N; ++p) {
        s->u[p] += du_particles;
    }
    s->u_frame += du_frame;
    updated_u(s);
    return du_frame;
}

/* eventDrivenStep detail.h:1933-1960 */
int fqo_event_driven_step(fqo_system* s, double eps, int kick, int direction, double* ret)
{
    if (s->has_ext) {
        return fail("eventDrivenStep is not available for this system");
    }
    if (direction > 0 && !kick) {
        double du;
        if (fqo_max_uniform_displacement(s, direction, &du)) {
            return 1;
        }
        if (du < 0.5 * eps) {
            *ret = 0.0;
            return 0;
        }
        *ret = advance_uniformly(s, du - 0.5 * eps);
        return 0;
    }
    if (direction > 0 && kick) {
        *ret = advance_uniformly(s, eps);
        return 0;
    }
    if (!kick) {
        double du;
        if (fqo_max_uniform_displacement(s, direction, &du)) {
            return 1;
        }
        if (du < 0.5 * eps) {
            *ret = 0.0;
            return 0;
        }
        *ret = advance_uniformly(s, 0.5 * eps - du);
        return 0;
    }
    *ret = advance_uniformly(s, -eps);
    return 0;
}

/* trigger detail.h:1972-1977 -> Cuspy detail.h:193-205 (aligns first),
 * SemiSmooth/Smooth detail.h:339-350, 427-438 (no align). Forces are not
 * updated (the reference returns before updated_u). */
int fqo_trigger(fqo_system* s, uint64_t p, double eps, int direction)
{
    if (p >= s->N) {
        return fail("particle index out of range");
    }
    if (s->d.potential == FQO_POT_CUSPY) {
        chunk_align(&s->ch, s->u);
    }
    const double* y = s->ch.y + p * s->ch.n + s->ch.j[p];
    if (direction > 0) {
        s->u[p] = y[1] + 0.5 * eps;
    }
    else {
        s->u[p] = y[0] - 0.5 * eps;
    }
    return 0;
}

/* advanceToFixedForce detail.h:1988-1995 */
int fqo_advance_to_fixed_force(fqo_system* s, double f_frame, int allow_plastic)
{
    int64_t* i_n = (int64_t*)malloc(sizeof(int64_t) * s->N);
    for (uint64_t p = 0; p < s->N; ++p) {
        i_n[p] = gidx(s, p);
    }
    double mean = 0.0;
    for (uint64_t p = 0; p < s->N; ++p) {
        mean += s->f_frame[p];
    }
    mean /= (double)s->N;
    advance_uniformly(s, (f_frame - mean) / s->mu);
    int ok = 1;
    for (uint64_t p = 0; p < s->N; ++p) {
        if (gidx(s, p) != i_n[p]) {
            ok = 0;
        }
    }
    free(i_n);
    if (!allow_plastic && !ok) {
        return fail("assertion failed (allow_plastic || xt::all(xt::equal(m_chunk->index_at_align(), i_n)))");
    }
    return 0;
}

/* --- chunk API ------------------------------------------------------------- */

uint64_t fqo_chunk_size(const fqo_system* s)
{
    return s->ch.n;
}

int fqo_chunk_data(const fqo_system* s, double* out)
{
    memcpy(out, s->ch.y, sizeof(double) * s->N * s->ch.n);
    return 0;
}

int fqo_chunk_index_at_align(const fqo_system* s, int64_t* out)
{
    for (uint64_t p = 0; p < s->N; ++p) {
        out[p] = gidx(s, p);
    }
    return 0;
}

int fqo_chunk_chunk_index_at_align(const fqo_system* s, int64_t* out)
{
    memcpy(out, s->ch.j, sizeof(int64_t) * s->N);
    return 0;
}

int fqo_chunk_start(const fqo_system* s, int64_t* out)
{
    memcpy(out, s->ch.start, sizeof(int64_t) * s->N);
    return 0;
}

int fqo_chunk_state_at(const fqo_system* s, const int64_t* index, uint64_t* out)
{
    for (uint64_t p = 0; p < s->N; ++p) {
        out[p] = fqo_pcg32_advance(s->ch.st_begin[p], s->ch.inc, index[p] - s->ch.start[p]);
    }
    return 0;
}

int fqo_chunk_restore(fqo_system* s, const uint64_t* state, const double* value, const int64_t* index)
{
    fq_chunk* c = &s->ch;
    for (uint64_t p = 0; p < s->N; ++p) {
        double* y = c->y + p * c->n;
        uint64_t st = state[p];
        c->start[p] = index[p];
        c->st_begin[p] = st;
        y[0] = value[p];
        for (uint64_t g = 1; g < c->n; ++g) {
            y[g] = y[g - 1] + draw_one(&c->tr, &st, c->inc);
        }
        c->st_end[p] = st;
        c->j[p] = 0;
        chunk_align_p(c, p, s->u[p]);
    }
    return 0;
}

/* --- thermal --------------------------------------------------------------- */

uint64_t fqo_thermal_state(const fqo_system* s)
{
    return s->rng_state;
}

int fqo_thermal_set_state(fqo_system* s, uint64_t state)
{
    s->rng_state = state;
    return 0;
}

int fqo_thermal_next(const fqo_system* s, int64_t* out)
{
    if (!s->has_ext) {
        return fail("system has no external forcing");
    }
    memcpy(out, s->next, sizeof(int64_t) * s->N);
    return 0;
}

int fqo_thermal_set_next(fqo_system* s, const int64_t* in)
{
    if (!s->has_ext) {
        return fail("system has no external forcing");
    }
    memcpy(s->next, in, sizeof(int64_t) * s->N);
    return 0;
}
