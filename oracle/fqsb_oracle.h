/*
 * fqsb_oracle.h -- CPU oracle of the FrictionQPotSpringBlock dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, not the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (frictionqpotspringblock_amd, HIP on gfx950) never links or calls it.
 *
 * The oracle restates, operation for operation, the CPU algorithm of the
 * reference header include/FrictionQPotSpringBlock/detail.h (policies at
 * detail.h:106-1000, engine detail::System at detail.h:1046-2051) and the
 * constructors of Line1d.h:112-672 / Line2d.h:77-162, plus its two external
 * dependencies that are not available here:
 *   - prrng (tdegeus/prrng >= 1.11.1, environment.yaml:13): pcg32 streams
 *     (O'Neill 2014, PCG-XSH-RR 64/32) and the chunked cumulative sum of
 *     yield positions (pcg32_tensor_cumsum, used at Line1d.h:51-52, 148-157);
 *   - GooseFEM Iterate::StopList (detail.h:1601, 1689, 1851).
 * See fqsb_oracle.c for the per-function citations and what pins each part.
 */
#ifndef FQSB_ORACLE_H
#define FQSB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FQO_POT_CUSPY = 0, FQO_POT_SEMISMOOTH = 1, FQO_POT_SMOOTH = 2 };
enum {
    FQO_INT_NONE = 0,
    FQO_INT_LAPLACE1D = 1,
    FQO_INT_QUARTIC1D = 2,
    FQO_INT_QUARTICGRADIENT1D = 3,
    FQO_INT_LONGRANGE1D = 4,
    FQO_INT_LAPLACE2D = 5,
    FQO_INT_QUARTICGRADIENT2D = 6
};
enum { FQO_EXT_NONE = 0, FQO_EXT_RANDOMFORCING = 1 };
enum { FQO_MIN_INERTIAL = 0, FQO_MIN_OVERDAMPED = 1, FQO_MIN_NONE = 2 };
enum {
    FQO_DIST_RANDOM = 0,
    FQO_DIST_DELTA = 1,
    FQO_DIST_EXPONENTIAL = 2,
    FQO_DIST_POWER = 3,
    FQO_DIST_GAMMA = 4,
    FQO_DIST_PARETO = 5,
    FQO_DIST_WEIBULL = 6,
    FQO_DIST_NORMAL = 7
};

/* Array fields of a system. */
enum {
    FQO_U = 0,
    FQO_V = 1,
    FQO_A = 2,
    FQO_F = 3,
    FQO_F_POTENTIAL = 4,
    FQO_F_INTERACTIONS = 5,
    FQO_F_FRAME = 6,
    FQO_F_DAMPING = 7,
    FQO_F_THERMAL = 8
};

typedef struct fqo_desc {
    int32_t rank;         /* 1 (Line1d) or 2 (Line2d) */
    int32_t potential;    /* FQO_POT_* */
    int32_t interactions; /* FQO_INT_* */
    int32_t external;     /* FQO_EXT_* */
    int32_t minimisation; /* FQO_MIN_* */
    int32_t distribution; /* FQO_DIST_* */
    uint64_t rows;        /* rank 1: number of particles */
    uint64_t cols;        /* rank 1: 1 */
    uint64_t seed;
    uint64_t nchunk;
    double m, eta, mu, kappa, k_frame, dt;
    double k_interactions; /* Laplace1d/2d, LongRange1d */
    double a1, a2;         /* Quartic1d */
    double k2, k4;         /* QuarticGradient1d/2d */
    double alpha;          /* LongRange1d */
    double offset;
    double parameters[4];
    int32_t nparameters;
    int32_t pad0;
    /* RandomNormalForcing (detail.h:904-922) */
    double mean, stddev;
    uint64_t seed_forcing;
    const int64_t* dinc_init;
    const int64_t* dinc;
    /* chunk alignment, Line1d.h:155: prrng::alignment(2, 30, 6, false) */
    int64_t align_buffer, align_margin;
} fqo_desc;

typedef struct fqo_system fqo_system;

const char* fqo_last_error(void);

int fqo_create(const fqo_desc* d, fqo_system** out);
void fqo_destroy(fqo_system* s);
uint64_t fqo_size(const fqo_system* s);

int fqo_get(const fqo_system* s, int field, double* out);
int fqo_set(fqo_system* s, int field, const double* in);
int64_t fqo_get_inc(const fqo_system* s);
int fqo_set_inc(fqo_system* s, int64_t inc);
int fqo_set_t(fqo_system* s, double t);
double fqo_get_u_frame(const fqo_system* s);
int fqo_set_u_frame(fqo_system* s, double u_frame);
int64_t fqo_qs_first(const fqo_system* s);
int64_t fqo_qs_last(const fqo_system* s);
double fqo_residual(const fqo_system* s);
double fqo_temperature(const fqo_system* s);

int fqo_refresh(fqo_system* s);
int fqo_quench(fqo_system* s);
int fqo_time_step(fqo_system* s);
int fqo_time_steps(fqo_system* s, uint64_t n);
int fqo_flow_steps(fqo_system* s, uint64_t n, double v_frame);
int fqo_time_steps_until_event(fqo_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, uint64_t* ret);
int fqo_minimise(fqo_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, int time_activity,
                 int max_iter_is_error, uint64_t* ret);
int fqo_minimise_truncate(fqo_system* s, const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate, double tol,
                          uint64_t niter_tol, uint64_t max_iter, int time_activity, int max_iter_is_error,
                          uint64_t* ret);
int fqo_max_uniform_displacement(fqo_system* s, int direction, double* ret);
int fqo_event_driven_step(fqo_system* s, double eps, int kick, int direction, double* ret);
int fqo_trigger(fqo_system* s, uint64_t p, double eps, int direction);
int fqo_advance_to_fixed_force(fqo_system* s, double f_frame, int allow_plastic);

/* chunk (yield-position landscape) */
uint64_t fqo_chunk_size(const fqo_system* s);
int fqo_chunk_data(const fqo_system* s, double* out);           /* [N, nchunk] */
int fqo_chunk_index_at_align(const fqo_system* s, int64_t* out); /* global */
int fqo_chunk_chunk_index_at_align(const fqo_system* s, int64_t* out);
int fqo_chunk_start(const fqo_system* s, int64_t* out);
int fqo_chunk_state_at(const fqo_system* s, const int64_t* index, uint64_t* out);
int fqo_chunk_restore(fqo_system* s, const uint64_t* state, const double* value, const int64_t* index);

/* RandomNormalForcing accessors (detail.h:949-999) */
uint64_t fqo_thermal_state(const fqo_system* s);
int fqo_thermal_set_state(fqo_system* s, uint64_t state);
int fqo_thermal_next(const fqo_system* s, int64_t* out);
int fqo_thermal_set_next(fqo_system* s, const int64_t* in);

/* Standalone generator helpers (landscape and math pins). */
void fqo_pcg32_seed(uint64_t initstate, uint64_t initseq, uint64_t* state, uint64_t* inc);
uint32_t fqo_pcg32_next(uint64_t* state, uint64_t inc);
uint64_t fqo_pcg32_advance(uint64_t state, uint64_t inc, int64_t delta);
double fqo_uint32_to_double(uint32_t x);
int fqo_draw(int distribution, const double* parameters, int nparameters, uint64_t* state, uint64_t inc, double* out,
             uint64_t n);
/* The thermal draw of RandomNormalForcing: n consecutive normals of one
 * pcg32 stream (two uniforms each); the stream a forcing seed s starts from
 * is fqo_pcg32_seed(s, fqo_pcg32_default_stream()). */
int fqo_draw_normal(uint64_t* state, uint64_t inc, double mean, double stddev, double* out, uint64_t n);
uint64_t fqo_pcg32_default_stream(void);
double fqo_log(double x);
double fqo_exp(double x);
double fqo_sin(double x);
double fqo_cos(double x);
double fqo_pow(double x, double y);

/* StopList restatement (GooseFEM Iterate::StopList) exposed for pinning. */
int fqo_stoplist_run(const double* residuals, uint64_t n, uint64_t niter_tol, double tol, int64_t* stop_index);
/* Smallest relative change of a single residual that would change the step at
 * which the StopList stops on this sequence (the decision margin), and the
 * same for the last minimise / minimise_truncate / timeStepsUntilEvent call of
 * a system (+inf if no step came near a decision). */
int fqo_stoplist_margin(const double* residuals, uint64_t n, uint64_t niter_tol, double tol, double* margin);
double fqo_last_stop_margin(const fqo_system* s);

#ifdef __cplusplus
}
#endif
#endif
