/*
 * fqsb.h -- C ABI of the MI355X (gfx950) implementation of the
 * FrictionQPotSpringBlock dynamics hot path (library libfqsb.so).
 *
 * Plain pointers and sizes only. Every array argument is host memory of the
 * system's size N = prod(shape) (flat, row-major for Line2d), copied in or out
 * by the call; the library owns all device memory. Every function returns 0 on
 * success and non-zero on failure, with the message in fqsb_last_error()
 * (thread-local). Messages of errors that the reference raises are the
 * reference's own texts ("NaN entries found", "No convergence found",
 * "Operation not possible.", "Minimisation not implementated", "Unknown
 * distribution: X").
 *
 * The reference binds these operations through pybind11 (python/main.cpp)
 * onto the C++ class template detail::System
 * (include/FrictionQPotSpringBlock/detail.h:1046-2051); each declaration
 * below cites the member it replaces. Not thread-safe per handle.
 */
#ifndef FQSB_H
#define FQSB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* potential policy: detail::Cuspy detail.h:113, SemiSmooth detail.h:212, Smooth detail.h:356 */
enum { FQSB_POT_CUSPY = 0, FQSB_POT_SEMISMOOTH = 1, FQSB_POT_SMOOTH = 2 };

/* interaction policy: Laplace1d detail.h:446, Quartic1d detail.h:757,
 * QuarticGradient1d detail.h:616, LongRange1d detail.h:812, Laplace2d
 * detail.h:524, QuarticGradient2d detail.h:658; NONE = no interactions */
enum {
    FQSB_INT_NONE = 0,
    FQSB_INT_LAPLACE1D = 1,
    FQSB_INT_QUARTIC1D = 2,
    FQSB_INT_QUARTICGRADIENT1D = 3,
    FQSB_INT_LONGRANGE1D = 4,
    FQSB_INT_LAPLACE2D = 5,
    FQSB_INT_QUARTICGRADIENT2D = 6
};

/* external policy: RandomNormalForcing detail.h:881 */
enum { FQSB_EXT_NONE = 0, FQSB_EXT_RANDOMFORCING = 1 };

/* minimisation tag: inertial (default), Overdamped (no-passing, detail.h:1005),
 * None (detail.h:1015: minimise raises) */
enum { FQSB_MIN_INERTIAL = 0, FQSB_MIN_OVERDAMPED = 1, FQSB_MIN_NONE = 2 };

/* landscape increment distribution: string_to_distribution detail.h:31-66 */
enum {
    FQSB_DIST_RANDOM = 0,
    FQSB_DIST_DELTA = 1,
    FQSB_DIST_EXPONENTIAL = 2,
    FQSB_DIST_POWER = 3,
    FQSB_DIST_GAMMA = 4,
    FQSB_DIST_PARETO = 5,
    FQSB_DIST_WEIBULL = 6,
    FQSB_DIST_NORMAL = 7
};

/* array fields: System::u/v/a/f/f_potential/f_interactions/f_frame/f_damping
 * detail.h:1402-1468, RandomNormalForcing::f_thermal detail.h:967 */
enum {
    FQSB_U = 0,
    FQSB_V = 1,
    FQSB_A = 2,
    FQSB_F = 3,
    FQSB_F_POTENTIAL = 4,
    FQSB_F_INTERACTIONS = 5,
    FQSB_F_FRAME = 6,
    FQSB_F_DAMPING = 7,
    FQSB_F_THERMAL = 8
};

/* Description of a system: the arguments of the reference's constructors
 * (Line1d.h:134-161, 199-225, 283-318, 348-376, 394-421, 451-479, 508-555,
 * 585-613, 643-671; Line2d.h:88-115, 133-161). */
typedef struct fqsb_desc {
    int32_t rank;         /* 1: Line1d (shape [rows]); 2: Line2d (shape [rows, cols]) */
    int32_t potential;    /* FQSB_POT_* */
    int32_t interactions; /* FQSB_INT_* */
    int32_t external;     /* FQSB_EXT_* */
    int32_t minimisation; /* FQSB_MIN_* */
    int32_t distribution; /* FQSB_DIST_* */
    uint64_t rows;
    uint64_t cols; /* 1 for rank 1 */
    uint64_t seed; /* particle p uses pcg32 stream (seed + p, 0), Line1d.h:150 */
    uint64_t nchunk;
    double m, eta, mu, kappa, k_frame, dt;
    double k_interactions; /* Laplace1d/2d, LongRange1d */
    double a1, a2;         /* Quartic1d */
    double k2, k4;         /* QuarticGradient1d/2d */
    double alpha;          /* LongRange1d */
    double offset;         /* m_gen += offset, Line1d.h:157 */
    double parameters[4];
    int32_t nparameters;
    int32_t pad0;
    double mean, stddev; /* RandomNormalForcing */
    uint64_t seed_forcing;
    const int64_t* dinc_init; /* [N] */
    const int64_t* dinc;      /* [N] */
    int64_t align_buffer;     /* prrng::alignment(buffer=2, margin=30, ...), Line1d.h:155 */
    int64_t align_margin;
} fqsb_desc;

typedef struct fqsb_system fqsb_system;

/* Last error message of the calling thread. */
const char* fqsb_last_error(void);

/* Library version string. */
const char* fqsb_version(void);

/* Construct (Line1d.h:134-161 etc.: landscape, policies, initSystem
 * detail.h:1096-1139). */
int fqsb_create(const fqsb_desc* desc, fqsb_system** out);
int fqsb_destroy(fqsb_system* s);

/* System::size detail.h:1155 */
uint64_t fqsb_size(const fqsb_system* s);

/* Getters detail.h:1402-1468 (copy out) and setters set_u/set_v/set_a
 * detail.h:1276-1305, RandomNormalForcing::set_f_thermal detail.h:976. */
int fqsb_get(fqsb_system* s, int field, double* out);
int fqsb_set(fqsb_system* s, int field, const double* in);

/* inc / set_inc detail.h:1486, 1241; set_t detail.h:1231 */
int fqsb_get_inc(const fqsb_system* s, int64_t* out);
int fqsb_set_inc(fqsb_system* s, int64_t inc);
int fqsb_set_t(fqsb_system* s, double t);

/* u_frame / set_u_frame detail.h:1264, 1253 */
int fqsb_get_u_frame(const fqsb_system* s, double* out);
int fqsb_set_u_frame(fqsb_system* s, double u_frame);

/* quasistaticActivityFirst/Last detail.h:1802-1818 */
int fqsb_quasistatic_activity(const fqsb_system* s, int64_t* first, int64_t* last);

/* residual detail.h:1512, temperature detail.h:1500 */
int fqsb_residual(fqsb_system* s, double* out);
int fqsb_temperature(fqsb_system* s, double* out);

/* refresh detail.h:1310, quench detail.h:1527 */
int fqsb_refresh(fqsb_system* s);
int fqsb_quench(fqsb_system* s);

/* timeStep detail.h:1539, timeSteps detail.h:1577, flowSteps detail.h:1637 */
int fqsb_time_step(fqsb_system* s);
int fqsb_time_steps(fqsb_system* s, uint64_t n);
int fqsb_flow_steps(fqsb_system* s, uint64_t n, double v_frame);

/* timeStepsUntilEvent detail.h:1595 */
int fqsb_time_steps_until_event(fqsb_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, uint64_t* ret);

/* minimise detail.h:1676 (inertial, or no-passing for FQSB_MIN_OVERDAMPED) */
int fqsb_minimise(fqsb_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, int time_activity,
                  int max_iter_is_error, uint64_t* ret);

/* minimise_truncate detail.h:1833 */
int fqsb_minimise_truncate(fqsb_system* s, const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate, double tol,
                           uint64_t niter_tol, uint64_t max_iter, int time_activity, int max_iter_is_error,
                           uint64_t* ret);

/* maxUniformDisplacement detail.h:1901, eventDrivenStep detail.h:1933,
 * trigger detail.h:1972, advanceToFixedForce detail.h:1988 */
int fqsb_max_uniform_displacement(fqsb_system* s, int direction, double* ret);
int fqsb_event_driven_step(fqsb_system* s, double eps, int kick, int direction, double* ret);
int fqsb_trigger(fqsb_system* s, uint64_t p, double eps, int direction);
int fqsb_advance_to_fixed_force(fqsb_system* s, double f_frame, int allow_plastic);

/* The quasistatic protocol of the reference's examples
 * (examples/Line1d_Cuspy_Laplace.py:39-61, the loop a Python caller writes
 * around eventDrivenStep and minimise), run as one call: events
 * k = first_event .. first_event + nevents - 1, where
 *   k == 0:      u_frame = 0 (detail.h:1253-1258), minimise
 *   k odd:       eventDrivenStep(eps, false, direction)
 *   k even > 0:  eventDrivenStep(eps, true, direction), minimise
 * (kicks != 0: every k > 0 is eventDrivenStep(eps, true, direction) and
 * minimise, the form available to the Smooth potential) with minimise(tol, niter_tol, max_iter) raising "No convergence found" at
 * max_iter as the example's `assert ret == 0`. After event e the host arrays
 * (each of nevents values, any may be NULL) receive S[e] = the sum over the
 * particles of the change of chunk.index_at_align during the event, the
 * frame position u_frame[e], the mean of f_frame[e] and inc[e]. The systems
 * that fit one workgroup run the whole loop on the device (one workgroup per
 * system, no host round trip per event); the others, and no-passing systems,
 * are driven event by event from the host. No counterpart function in the
 * reference (it replaces the Python loop). */
int fqsb_quasistatic(fqsb_system* s, int64_t first_event, int64_t nevents, double eps, int direction, int kicks,
                     double tol, uint64_t niter_tol, uint64_t max_iter, int64_t* S, double* u_frame, double* f_frame,
                     int64_t* inc);

/* The chunk of yield positions (System::chunk detail.h:1146; the prrng
 * pcg32_tensor_cumsum members used by the reference and its tests:
 * data(), index_at_align(), chunk_index_at_align(), start(), state_at(),
 * restore(), test_Line1d.py:309-341). data is [N, nchunk]. */
uint64_t fqsb_chunk_size(const fqsb_system* s);
int fqsb_chunk_data(fqsb_system* s, double* out);
int fqsb_chunk_index_at_align(fqsb_system* s, int64_t* out);
int fqsb_chunk_chunk_index_at_align(fqsb_system* s, int64_t* out);
int fqsb_chunk_start(fqsb_system* s, int64_t* out);
int fqsb_chunk_state_at(fqsb_system* s, const int64_t* index, uint64_t* out);
int fqsb_chunk_restore(fqsb_system* s, const uint64_t* state, const double* value, const int64_t* index);

/* RandomNormalForcing state/next detail.h:949-999 */
int fqsb_thermal_state(const fqsb_system* s, uint64_t* out);
int fqsb_thermal_set_state(fqsb_system* s, uint64_t state);
int fqsb_thermal_next(fqsb_system* s, int64_t* out);
int fqsb_thermal_set_next(fqsb_system* s, const int64_t* in);

/* ------------------------------------------------------------------------
 * Batched form (no reference counterpart: the reference runs one system per
 * object). R replicas of one description; replica r uses landscape seed
 * desc->seed + r * seed_stride (and, with random forcing, forcing seed
 * desc->seed_forcing + r * seed_stride), so with seed_stride = 1 replica r
 * equals the single system constructed with seed = desc->seed + r (and
 * seed_forcing = desc->seed_forcing + r). All replicas advance together in
 * one resident launch (one workgroup per replica).
 * ---------------------------------------------------------------------- */
typedef struct fqsb_batch fqsb_batch;

int fqsb_batch_create(const fqsb_desc* desc, uint64_t replicas, uint64_t seed_stride, fqsb_batch** out);
int fqsb_batch_destroy(fqsb_batch* b);
uint64_t fqsb_batch_replicas(const fqsb_batch* b);
/* Choose the resident launch shape (particles per thread, threads per
 * workgroup); 0, 0 restores the automatic choice (chains: 4 particles per
 * thread while the replicas' wavefronts fit two per SIMD, else 8, where a
 * kernel variant holds a replica so; otherwise the fewest particles per
 * thread that hold it). Results do not depend on the shape. */
int fqsb_batch_set_launch(fqsb_batch* b, int ppt, int threads);
int fqsb_batch_get_launch(const fqsb_batch* b, int* ppt, int* threads);
/* How a batch evaluates LongRange1d (detail.h:846-867):
 * FQSB_LR_EXACT  -- the reference's sequential sum per particle, in the
 *                   resident or streaming kernels; bit-identical to a single
 *                   system of the same seed.
 * FQSB_LR_MATRIX -- (default for Cuspy batches) the forces of all replicas as
 *                   one product with the circulant matrix on the FP64 matrix
 *                   cores; a different summation order, so each force differs
 *                   from the sequential sum by a few units in the last place
 *                   of the largest term. Used by timeSteps / flowSteps;
 *                   minimise and the event-driven steps always use the exact
 *                   sum. Fails on a batch without long-range interactions. */
enum { FQSB_LR_EXACT = 0, FQSB_LR_MATRIX = 1 };
int fqsb_batch_set_longrange_method(fqsb_batch* b, int method);
int fqsb_batch_get_longrange_method(const fqsb_batch* b, int* method);
/* Average duration (ms) of one matrix-product launch in the last timed
 * timeSteps / flowSteps call of the matrix form (HIP events around the first
 * product of every block, on the launch stream); 0 if none was timed. */
int fqsb_batch_longrange_product_ms(const fqsb_batch* b, float* ms);
/* flowSteps / timeSteps on all replicas, launched as resident blocks of
 * `block` steps (0: automatic; with the long-range matrix form, the number of
 * steps between checks of the run states). If ms is non-NULL it receives the
 * time between HIP events recorded on the launch stream around the launches. */
int fqsb_batch_flow_steps(fqsb_batch* b, uint64_t n, double v_frame, uint64_t block, float* ms);
int fqsb_batch_time_steps(fqsb_batch* b, uint64_t n, uint64_t block, float* ms);
/* minimise on all replicas; ret[r] as System::minimise. */
int fqsb_batch_minimise(fqsb_batch* b, double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                        uint64_t* ret);
/* eventDrivenStep on all replicas (each with its own displacement); ret[r]. */
int fqsb_batch_event_driven_step(fqsb_batch* b, double eps, int kick, int direction, double* ret);
/* fqsb_quasistatic for every replica, each at its own pace; outputs [R, nevents]. */
int fqsb_batch_quasistatic(fqsb_batch* b, int64_t first_event, int64_t nevents, double eps, int direction, int kicks,
                           double tol, uint64_t niter_tol, uint64_t max_iter, int64_t* S, double* u_frame,
                           double* f_frame, int64_t* inc);
/* fqsb_batch_quasistatic without any device allocation in the call. No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and errors as fqsb_batch_quasistatic;
 * the per-event records are staged in a fixed 8 MiB device array of the
 * library (one per GPU, held until the process ends) instead of buffers
 * allocated per call, so replicas x nevents is at most 262144 (a larger call
 * fails with an error and changes nothing). Runs only where the whole
 * protocol runs on the device (systems of one resident workgroup, inertial
 * minimisation); other batches fail. Calls that overlap on one GPU share the
 * array and race. */
int fqsb_batch_quasistatic_noalloc(fqsb_batch* b, int64_t first_event, int64_t nevents, double eps, int direction,
                                   int kicks, double tol, uint64_t niter_tol, uint64_t max_iter, int64_t* S,
                                   double* u_frame, double* f_frame, int64_t* inc);
/* The last fqsb_batch_quasistatic: time between HIP events on the launch
 * stream around its device work (ms), and the number of driver launches (0
 * if it was driven event by event from the host). */
int fqsb_batch_quasistatic_stats(const fqsb_batch* b, float* ms, int64_t* launches);
/* minimise_truncate (detail.h:1833-1893) on all replicas: i_n [R, N]; ret[r]. */
int fqsb_batch_minimise_truncate(fqsb_batch* b, const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate,
                                 double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                                 uint64_t* ret);
/* trigger (detail.h:1972-1977) of particle p[r] in replica r; forces are not
 * recomputed, as in the reference. */
int fqsb_batch_trigger(fqsb_batch* b, const uint64_t* p, double eps, int direction);
/* advanceToFixedForce (detail.h:1988-1995) with target mean frame force
 * f_frame[r] for replica r. */
int fqsb_batch_advance_to_fixed_force(fqsb_batch* b, const double* f_frame, int allow_plastic);
/* Checkpoint / restart of a batch (the batched form of the reference's
 * restart through chunk.restore and the state setters, test_Line1d.py:309-341):
 * every item that, with u, v and a (fqsb_batch_get / fqsb_batch_set), fixes
 * the continuation bit for bit. Per particle [R, N] (the landscape
 * [R, N, nchunk]) or per replica [R]; element sizes from
 * fqsb_batch_state_size. Setting an item does not recompute forces; call
 * fqsb_batch_refresh after the last one. */
enum {
    FQSB_STATE_Y = 0,         /* f64 [R, N, nchunk]: the landscape windows */
    FQSB_STATE_START = 1,     /* i64 [R, N]: global index of each window's first value */
    FQSB_STATE_JLOC = 2,      /* i32 [R, N]: local index of the current well */
    FQSB_STATE_ST_BEGIN = 3,  /* u64 [R, N]: pcg32 state after each window's first value */
    FQSB_STATE_ST_END = 4,    /* u64 [R, N]: pcg32 state after each window's last value */
    FQSB_STATE_F_THERMAL = 5, /* f64 [R, N]: random forcing (systems with external forcing) */
    FQSB_STATE_NEXT = 6,      /* i64 [R, N]: increment of each particle's next draw */
    FQSB_STATE_INC = 7,       /* i64 [R] */
    FQSB_STATE_U_FRAME = 8,   /* f64 [R] */
    FQSB_STATE_RNG = 9,       /* u64 [R]: pcg32 state of the random forcing */
    FQSB_STATE_QS_FIRST = 10, /* i64 [R]: quasistaticActivityFirst */
    FQSB_STATE_QS_LAST = 11   /* i64 [R]: quasistaticActivityLast */
};
int fqsb_batch_state_size(fqsb_batch* b, int what, uint64_t* count, int* bytes);
int fqsb_batch_get_state(fqsb_batch* b, int what, void* out);
int fqsb_batch_set_state(fqsb_batch* b, int what, const void* in);
/* refresh (detail.h:1310-1315) of every replica: all forces from the state. */
int fqsb_batch_refresh(fqsb_batch* b);
/* Arrays of all replicas, [R, N]; scalars [R]. */
int fqsb_batch_get(fqsb_batch* b, int field, double* out);
int fqsb_batch_set(fqsb_batch* b, int field, const double* in);
int fqsb_batch_get_u_frame(fqsb_batch* b, double* out);
int fqsb_batch_get_inc(fqsb_batch* b, int64_t* out);
int fqsb_batch_chunk_index_at_align(fqsb_batch* b, int64_t* out);
/* The HIP stream (hipStream_t) the batch launches on. */
void* fqsb_batch_stream(fqsb_batch* b);

/* Per-replica parameters [R][4] = (mu, k_interactions, k_frame, eta), host
 * arrays; setting them recomputes the forces. */
int fqsb_batch_set_parameters(fqsb_batch* b, const double* params);
int fqsb_batch_get_parameters(const fqsb_batch* b, double* params);
uint64_t fqsb_batch_nchunk(const fqsb_batch* b);
/* Device-side access (pointers on the batch's device, e.g. from torch):
 * fields [R, N] (setting u or v recomputes forces as fqsb_batch_set), the
 * yield-position chunks [R, N, nchunk] (setting them realigns and recomputes
 * the forces). */
int fqsb_batch_get_device(fqsb_batch* b, int field, double* out);
int fqsb_batch_set_device(fqsb_batch* b, int field, const double* in);
int fqsb_batch_chunk_data_device(fqsb_batch* b, double* out);
int fqsb_batch_set_chunk_data_device(fqsb_batch* b, const double* in);

/* Adjoint (no reference counterpart) of timeSteps / flowSteps (detail.h:
 * 1539-1583, 1637-1645) for batches of Cuspy + Laplace1d replicas with
 * N <= 4096 and no external forcing. fqsb_batch_adjoint_forward runs n steps
 * (flow != 0: flowSteps at v_frame) from the current state, keeping a
 * checkpoint every `window` steps; it fails if the trajectory needs a
 * landscape refill. fqsb_batch_adjoint_backward_device maps the cotangents
 * (gu, gv, ga) of the final u, v, a to those of the initial state (gu0, gv0,
 * ga0), the derivatives d/d(mu, k_interactions, k_frame, eta) per replica
 * (gparams [R][4]) and d/dy of the yield-position chunks (gy [R][N][nchunk],
 * NULL to skip), differentiating the exact discrete step with every well
 * index held at the value the trajectory used. All pointers are device
 * pointers. */
int fqsb_batch_adjoint_forward(fqsb_batch* b, uint64_t n, uint64_t window, int flow, double v_frame);
int fqsb_batch_adjoint_backward_device(fqsb_batch* b, const double* gu, const double* gv, const double* ga,
                                       double* gu0, double* gv0, double* ga0, double* gparams, double* gy);
/* Which tape the batch holds: a number that increases with every
 * fqsb_batch_adjoint_forward, or 0 once any later call has changed the
 * batch's state, landscape or parameters (a backward pass then has no valid
 * tape). A caller that keeps the number of its forward pass can tell that its
 * tape was replaced by a later forward pass on the same batch. */
int fqsb_batch_adjoint_generation(const fqsb_batch* b, uint64_t* gen);
/* Average duration (ms) of one window's recompute (record) and reverse
 * kernel in the last backward pass (HIP events on the batch's stream). */
int fqsb_batch_adjoint_kernel_ms(const fqsb_batch* b, float* record_ms, float* reverse_ms);

/* ------------------------------------------------------------------------
 * Ragged set (no reference counterpart): n systems of any classes, sizes and
 * parameters (descs[i] as for fqsb_create), advanced together. Systems that
 * share a kernel class and a size class (sizes within a factor 2) run in one
 * launch; the groups run concurrently. Every operation applies to each
 * system exactly as the corresponding single-system call would, so results
 * are bit-identical to n separate systems. Arrays are the systems' fields
 * concatenated in the order of descs (sum of the sizes); scalars [n].
 * `active` (NULL: all) selects the systems an operation applies to; the
 * others are left unchanged and get ret 0.
 * ---------------------------------------------------------------------- */
typedef struct fqsb_ragged fqsb_ragged;

int fqsb_ragged_create(const fqsb_desc* descs, uint64_t n, fqsb_ragged** out);
int fqsb_ragged_destroy(fqsb_ragged* r);
uint64_t fqsb_ragged_systems(const fqsb_ragged* r);
uint64_t fqsb_ragged_groups(const fqsb_ragged* r);
int fqsb_ragged_sizes(const fqsb_ragged* r, uint64_t* out);
/* timeSteps / flowSteps (detail.h:1577-1583, 1637-1645) on every system. */
int fqsb_ragged_time_steps(fqsb_ragged* r, uint64_t n);
int fqsb_ragged_flow_steps(fqsb_ragged* r, uint64_t n, double v_frame);
/* minimise (detail.h:1676-1792; no-passing systems: detail.h:1694-1753);
 * ret[i] as System::minimise. Fails ("Minimisation not implementated") if an
 * active system has no minimisation. */
int fqsb_ragged_minimise(fqsb_ragged* r, double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                         const uint8_t* active, uint64_t* ret);
/* eventDrivenStep (detail.h:1933-1960); fails with "Operation not possible."
 * if kick == 0 and an active system has the Smooth potential (its
 * maxUniformDisplacement raises, detail.h:414-422). */
int fqsb_ragged_event_driven_step(fqsb_ragged* r, double eps, int kick, int direction, const uint8_t* active,
                                  double* ret);
/* fqsb_quasistatic for every active system (active: NULL for all);
 * outputs [systems, nevents], rows of inactive systems zero. */
int fqsb_ragged_quasistatic(fqsb_ragged* r, int64_t first_event, int64_t nevents, double eps, int direction,
                            int kicks, double tol, uint64_t niter_tol, uint64_t max_iter, const uint8_t* active,
                            int64_t* S, double* u_frame, double* f_frame, int64_t* inc);
int fqsb_ragged_get(fqsb_ragged* r, int field, double* out);
int fqsb_ragged_set(fqsb_ragged* r, int field, const double* in);
int fqsb_ragged_get_u_frame(fqsb_ragged* r, double* out);
int fqsb_ragged_get_inc(fqsb_ragged* r, int64_t* out);
/* Steps (time steps, or no-passing sweeps) each system took in the last
 * minimise / timeSteps / flowSteps call ([n]). */
int fqsb_ragged_last_steps(fqsb_ragged* r, uint64_t* out);
int fqsb_ragged_chunk_index_at_align(fqsb_ragged* r, int64_t* out);

/* Launch groups of a ragged set, for tuning (no reference counterpart). The
 * path a group runs on: */
enum {
    FQSB_PATH_RESIDENT = 0,            /* one workgroup per system, many steps per launch */
    FQSB_PATH_TILES = 1,               /* tiles of a system over many workgroups (blocks of steps) */
    FQSB_PATH_STREAMING = 2,           /* one workgroup per system, state in HBM */
    FQSB_PATH_NOPASSING_RESIDENT = 3,  /* no-passing sweeps, one workgroup per system */
    FQSB_PATH_NOPASSING_TILES = 4      /* no-passing sweeps over tiles */
};
/* info[13] = systems, largest N, path, particles per thread, threads, block
 * length of the tiled path (0: not blocked), rank, potential, interactions,
 * external, minimisation, particles, longest block the tiles' halos allow
 * (chains 128; lattices: K halo rows, 8 unless FQSB_LATTICE_K is set). */
int fqsb_ragged_group_info(const fqsb_ragged* r, uint64_t group, int64_t* info);
/* Time (ms) of the group's part of every operation since the last reset,
 * between HIP events on the group's stream, and the particle-steps (no-passing
 * groups: particle-sweeps) it took. */
int fqsb_ragged_group_time(const fqsb_ragged* r, uint64_t group, double* ms, double* particle_steps);
int fqsb_ragged_reset_group_time(fqsb_ragged* r);
/* serial != 0: every operation runs the groups one after the other from the
 * calling thread instead of concurrently on their streams (same results; each
 * group's time is then its own, not shared with the groups running beside
 * it). */
int fqsb_ragged_set_serial(fqsb_ragged* r, int serial);
/* Indices (in the order of descs) of the group's systems: info[0] values. */
int fqsb_ragged_group_systems(const fqsb_ragged* r, uint64_t group, uint64_t* idx);
/* Resident launch shapes the group can use (at most cap; *count: all). */
int fqsb_ragged_group_shapes(const fqsb_ragged* r, uint64_t group, int* ppt, int* threads, int cap, int* count);
/* Set a resident group's launch shape, or a tiled group's block length
 * (1 .. 128 steps per launch; longer than info[12] runs blocks of info[12]).
 * Results do not depend on either. */
int fqsb_ragged_set_group_launch(fqsb_ragged* r, uint64_t group, int ppt, int threads);
int fqsb_ragged_set_group_block(fqsb_ragged* r, uint64_t group, int steps);

#ifdef __cplusplus
}
#endif

#endif
