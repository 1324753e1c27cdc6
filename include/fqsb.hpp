/*
 * fqsb.hpp -- C++ host API of the MI355X implementation of the
 * FrictionQPotSpringBlock dynamics hot path, on top of the C ABI (fqsb.h).
 *
 * Same namespaces, class names, constructor arguments (and defaults) and member
 * functions as the reference's headers Line1d.h / Line2d.h and
 * detail::System (detail.h:1046-2051), so that C++ code written against the
 * reference compiles against this header. Arrays are std::vector<double> (flat,
 * row-major for Line2d) instead of xtensor containers; getters return copies.
 * Errors are std::runtime_error with the reference's messages.
 */
#ifndef FQSB_HPP
#define FQSB_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fqsb.h"

namespace FrictionQPotSpringBlock {

inline std::string version()
{
    return fqsb_version();
}

namespace detail {

inline void check(int rc)
{
    if (rc != 0) {
        throw std::runtime_error(fqsb_last_error());
    }
}

/* string_to_distribution detail.h:31-66 */
inline int string_to_distribution(const std::string& str)
{
    if (str == "random") {
        return FQSB_DIST_RANDOM;
    }
    if (str == "delta") {
        return FQSB_DIST_DELTA;
    }
    if (str == "exponential") {
        return FQSB_DIST_EXPONENTIAL;
    }
    if (str == "power") {
        return FQSB_DIST_POWER;
    }
    if (str == "gamma") {
        return FQSB_DIST_GAMMA;
    }
    if (str == "pareto") {
        return FQSB_DIST_PARETO;
    }
    if (str == "weibull") {
        return FQSB_DIST_WEIBULL;
    }
    if (str == "normal") {
        return FQSB_DIST_NORMAL;
    }
    throw std::runtime_error("Unknown distribution: " + str);
}

inline fqsb_desc make_desc(int rank, const std::vector<size_t>& shape, uint64_t seed, const std::string& distribution,
                           const std::vector<double>& parameters, double offset, size_t nchunk)
{
    fqsb_desc d{};
    d.rank = rank;
    d.rows = shape.at(0);
    d.cols = rank == 2 ? shape.at(1) : 1;
    d.seed = seed;
    d.distribution = string_to_distribution(distribution);
    if (parameters.size() > 4) {
        throw std::runtime_error("Too many parameters for distribution");
    }
    for (size_t i = 0; i < parameters.size(); ++i) {
        d.parameters[i] = parameters[i];
    }
    d.nparameters = (int32_t)parameters.size();
    d.offset = offset;
    d.nchunk = nchunk;
    d.align_buffer = 2;
    d.align_margin = 30;
    return d;
}

struct Handle {
    fqsb_system* h = nullptr;
    explicit Handle(const fqsb_desc& d)
    {
        check(fqsb_create(&d, &h));
    }
    ~Handle()
    {
        fqsb_destroy(h);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

/* The chunk of yield positions (the reference's prrng::pcg32_tensor_cumsum,
 * Line1d.h:51-52): the members the reference and its tests use. */
class Chunk {
    std::shared_ptr<Handle> m_h;
    size_t m_N;

public:
    Chunk(std::shared_ptr<Handle> h, size_t N) : m_h(std::move(h)), m_N(N)
    {
    }
    size_t chunk_size() const
    {
        return fqsb_chunk_size(m_h->h);
    }
    std::vector<double> data() const
    {
        std::vector<double> r(m_N * chunk_size());
        check(fqsb_chunk_data(m_h->h, r.data()));
        return r;
    }
    std::vector<int64_t> index_at_align() const
    {
        std::vector<int64_t> r(m_N);
        check(fqsb_chunk_index_at_align(m_h->h, r.data()));
        return r;
    }
    std::vector<int64_t> chunk_index_at_align() const
    {
        std::vector<int64_t> r(m_N);
        check(fqsb_chunk_chunk_index_at_align(m_h->h, r.data()));
        return r;
    }
    std::vector<int64_t> start() const
    {
        std::vector<int64_t> r(m_N);
        check(fqsb_chunk_start(m_h->h, r.data()));
        return r;
    }
    std::vector<double> left_of_align() const
    {
        return at_offset(0);
    }
    std::vector<double> right_of_align() const
    {
        return at_offset(1);
    }
    std::vector<uint64_t> state_at(const std::vector<int64_t>& index) const
    {
        std::vector<uint64_t> r(m_N);
        check(fqsb_chunk_state_at(m_h->h, index.data(), r.data()));
        return r;
    }
    void restore(const std::vector<uint64_t>& state, const std::vector<double>& value,
                 const std::vector<int64_t>& index)
    {
        check(fqsb_chunk_restore(m_h->h, state.data(), value.data(), index.data()));
    }

private:
    std::vector<double> at_offset(int o) const
    {
        std::vector<double> y = data();
        std::vector<int64_t> j = chunk_index_at_align();
        std::vector<double> r(m_N);
        const size_t n = chunk_size();
        for (size_t p = 0; p < m_N; ++p) {
            r[p] = y[p * n + (size_t)j[p] + (size_t)o];
        }
        return r;
    }
};

/* RandomNormalForcing accessors (detail.h:949-999). */
class RandomNormalForcing {
    std::shared_ptr<Handle> m_h;
    size_t m_N;

public:
    RandomNormalForcing(std::shared_ptr<Handle> h, size_t N) : m_h(std::move(h)), m_N(N)
    {
    }
    uint64_t state() const
    {
        uint64_t s;
        check(fqsb_thermal_state(m_h->h, &s));
        return s;
    }
    void set_state(uint64_t s)
    {
        check(fqsb_thermal_set_state(m_h->h, s));
    }
    std::vector<double> f_thermal() const
    {
        std::vector<double> r(m_N);
        check(fqsb_get(m_h->h, FQSB_F_THERMAL, r.data()));
        return r;
    }
    void set_f_thermal(const std::vector<double>& f)
    {
        check(fqsb_set(m_h->h, FQSB_F_THERMAL, f.data()));
    }
    std::vector<int64_t> next() const
    {
        std::vector<int64_t> r(m_N);
        check(fqsb_thermal_next(m_h->h, r.data()));
        return r;
    }
    void set_next(const std::vector<int64_t>& n)
    {
        check(fqsb_thermal_set_next(m_h->h, n.data()));
    }
};

/* detail::System (detail.h:1053-2051). */
class System {
protected:
    std::shared_ptr<Handle> m_h;
    std::vector<size_t> m_shape;
    size_t m_N = 0;
    double m_dt = 0, m_mu = 0, m_eta = 0, m_m = 0, m_k_frame = 0;

    void init(fqsb_desc d, const std::vector<size_t>& shape)
    {
        m_shape = shape;
        m_N = 1;
        for (size_t s : shape) {
            m_N *= s;
        }
        m_dt = d.dt;
        m_mu = d.mu;
        m_eta = d.eta;
        m_m = d.m;
        m_k_frame = d.k_frame;
        m_h = std::make_shared<Handle>(d);
    }

    std::vector<double> get(int field) const
    {
        std::vector<double> r(m_N);
        check(fqsb_get(m_h->h, field, r.data()));
        return r;
    }

    void set(int field, const std::vector<double>& x)
    {
        if (x.size() != m_N) {
            throw std::runtime_error("array has the wrong size");
        }
        check(fqsb_set(m_h->h, field, x.data()));
    }

public:
    Chunk chunk() const
    {
        return Chunk(m_h, m_N);
    }
    size_t size() const
    {
        return m_N;
    }
    const std::vector<size_t>& shape() const
    {
        return m_shape;
    }
    double dt() const
    {
        return m_dt;
    }
    double mu() const
    {
        return m_mu;
    }
    double eta() const
    {
        return m_eta;
    }
    double m() const
    {
        return m_m;
    }
    double k_frame() const
    {
        return m_k_frame;
    }
    RandomNormalForcing external() const
    {
        return RandomNormalForcing(m_h, m_N);
    }

    void set_t(double arg)
    {
        check(fqsb_set_t(m_h->h, arg));
    }
    void set_inc(int64_t arg)
    {
        check(fqsb_set_inc(m_h->h, arg));
    }
    void set_u_frame(double arg)
    {
        check(fqsb_set_u_frame(m_h->h, arg));
    }
    double u_frame() const
    {
        double r;
        check(fqsb_get_u_frame(m_h->h, &r));
        return r;
    }
    void set_u(const std::vector<double>& x)
    {
        set(FQSB_U, x);
    }
    void set_v(const std::vector<double>& x)
    {
        set(FQSB_V, x);
    }
    void set_a(const std::vector<double>& x)
    {
        set(FQSB_A, x);
    }
    void refresh()
    {
        check(fqsb_refresh(m_h->h));
    }

    std::vector<double> u() const
    {
        return get(FQSB_U);
    }
    std::vector<double> v() const
    {
        return get(FQSB_V);
    }
    std::vector<double> a() const
    {
        return get(FQSB_A);
    }
    std::vector<double> f() const
    {
        return get(FQSB_F);
    }
    std::vector<double> f_potential() const
    {
        return get(FQSB_F_POTENTIAL);
    }
    std::vector<double> f_frame() const
    {
        return get(FQSB_F_FRAME);
    }
    std::vector<double> f_interactions() const
    {
        return get(FQSB_F_INTERACTIONS);
    }
    std::vector<double> f_damping() const
    {
        return get(FQSB_F_DAMPING);
    }
    double t() const
    {
        return static_cast<double>(inc()) * m_dt;
    }
    int64_t inc() const
    {
        int64_t r;
        check(fqsb_get_inc(m_h->h, &r));
        return r;
    }
    double temperature() const
    {
        double r;
        check(fqsb_temperature(m_h->h, &r));
        return r;
    }
    double residual() const
    {
        double r;
        check(fqsb_residual(m_h->h, &r));
        return r;
    }
    void quench()
    {
        check(fqsb_quench(m_h->h));
    }

    void timeStep()
    {
        check(fqsb_time_step(m_h->h));
    }
    void timeSteps(size_t n)
    {
        check(fqsb_time_steps(m_h->h, n));
    }
    size_t timeStepsUntilEvent(double tol = 1e-5, size_t niter_tol = 10, size_t max_iter = 1e9)
    {
        uint64_t r;
        check(fqsb_time_steps_until_event(m_h->h, tol, niter_tol, max_iter, &r));
        return r;
    }
    void flowSteps(size_t n, double v_frame)
    {
        check(fqsb_flow_steps(m_h->h, n, v_frame));
    }
    size_t minimise(double tol = 1e-5, size_t niter_tol = 10, size_t max_iter = 1e9, bool time_activity = false,
                    bool max_iter_is_error = true)
    {
        uint64_t r;
        check(fqsb_minimise(m_h->h, tol, niter_tol, max_iter, time_activity, max_iter_is_error, &r));
        return r;
    }
    size_t quasistaticActivityFirst() const
    {
        int64_t f, l;
        check(fqsb_quasistatic_activity(m_h->h, &f, &l));
        return (size_t)f;
    }
    size_t quasistaticActivityLast() const
    {
        int64_t f, l;
        check(fqsb_quasistatic_activity(m_h->h, &f, &l));
        return (size_t)l;
    }
    size_t minimise_truncate(const std::vector<int64_t>& i_n, size_t A_truncate = 0, size_t S_truncate = 0,
                             double tol = 1e-5, size_t niter_tol = 10, size_t max_iter = 1e9,
                             bool time_activity = true, bool max_iter_is_error = true)
    {
        if (i_n.size() != m_N) {
            throw std::runtime_error("array has the wrong size");
        }
        uint64_t r;
        check(fqsb_minimise_truncate(m_h->h, i_n.data(), A_truncate, S_truncate, tol, niter_tol, max_iter,
                                     time_activity, max_iter_is_error, &r));
        return r;
    }
    double maxUniformDisplacement(int direction)
    {
        double r;
        check(fqsb_max_uniform_displacement(m_h->h, direction, &r));
        return r;
    }
    double eventDrivenStep(double eps, bool kick, int direction = 1)
    {
        double r;
        check(fqsb_event_driven_step(m_h->h, eps, kick, direction, &r));
        return r;
    }
    void trigger(size_t p, double eps, int direction = 1)
    {
        check(fqsb_trigger(m_h->h, p, eps, direction));
    }
    void advanceToFixedForce(double f_frame, bool allow_plastic = false)
    {
        check(fqsb_advance_to_fixed_force(m_h->h, f_frame, allow_plastic));
    }

    // Per-event output of quasistatic(): S, u_frame, mean f_frame, inc.
    struct Quasistatic {
        std::vector<int64_t> S;
        std::vector<double> u_frame;
        std::vector<double> f_frame;
        std::vector<int64_t> inc;
    };

    // The examples' loop around eventDrivenStep and minimise in one call
    // (fqsb_quasistatic).
    Quasistatic quasistatic(double eps, int64_t nevents, int64_t first_event = 0, int direction = 1,
                            bool kicks_only = false, double tol = 1e-5, size_t niter_tol = 10,
                            size_t max_iter = size_t(1e9))
    {
        Quasistatic q;
        const size_t n = nevents > 0 ? (size_t)nevents : 0;
        q.S.resize(n);
        q.u_frame.resize(n);
        q.f_frame.resize(n);
        q.inc.resize(n);
        check(fqsb_quasistatic(m_h->h, first_event, nevents, eps, direction, kicks_only, tol, niter_tol, max_iter,
                               q.S.data(), q.u_frame.data(), q.f_frame.data(), q.inc.data()));
        return q;
    }
};

inline fqsb_desc physics(fqsb_desc d, double m, double eta, double mu, double k_frame, double dt)
{
    d.m = m;
    d.eta = eta;
    d.mu = mu;
    d.k_frame = k_frame;
    d.dt = dt;
    return d;
}

} // namespace detail

namespace Line1d {

using Shape = std::array<size_t, 1>;

inline std::vector<std::string> version_dependencies()
{
    return {"frictionqpotspringblock_amd=" + std::string(fqsb_version()), "hip", "pcg32"};
}

inline std::vector<std::string> version_compiler()
{
    return {"hipcc gfx950"};
}

#define FQSB_LINE1D_DESC                                                                                            \
    detail::make_desc(1, {shape[0]}, seed, distribution, parameters, offset, nchunk)

/* Line1d.h:112-161 */
class System_Cuspy_Laplace : public detail::System {
public:
    System_Cuspy_Laplace(double m, double eta, double mu, double k_interactions, double k_frame, double dt,
                         const Shape& shape, uint64_t seed, const std::string& distribution,
                         const std::vector<double>& parameters, double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_LAPLACE1D;
        d.k_interactions = k_interactions;
        init(d, {shape[0]});
    }
};

/* Line1d.h:173-238 */
class System_Cuspy_Laplace_Nopassing : public detail::System {
public:
    System_Cuspy_Laplace_Nopassing(double mu, double k_interactions, double k_frame, const Shape& shape,
                                   uint64_t seed, const std::string& distribution,
                                   const std::vector<double>& parameters, double offset = -100.0,
                                   size_t nchunk = 5000, double eta = 0.0, double dt = 0.0)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, 1.0, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_LAPLACE1D;
        d.minimisation = FQSB_MIN_OVERDAMPED;
        d.k_interactions = k_interactions;
        init(d, {shape[0]});
    }
};

/* Line1d.h:261-326 */
class System_Cuspy_Laplace_RandomForcing : public detail::System {
public:
    System_Cuspy_Laplace_RandomForcing(double m, double eta, double mu, double k_interactions, double k_frame,
                                       double dt, double mean, double stddev, uint64_t seed_forcing,
                                       const std::vector<int64_t>& dinc_init, const std::vector<int64_t>& dinc,
                                       const Shape& shape, uint64_t seed, const std::string& distribution,
                                       const std::vector<double>& parameters, double offset = -100.0,
                                       size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_LAPLACE1D;
        d.external = FQSB_EXT_RANDOMFORCING;
        d.minimisation = FQSB_MIN_NONE;
        d.k_interactions = k_interactions;
        d.mean = mean;
        d.stddev = stddev;
        d.seed_forcing = seed_forcing;
        if (dinc_init.size() != shape[0] || dinc.size() != shape[0]) {
            throw std::runtime_error("dinc_init and dinc must have the shape of the system");
        }
        d.dinc_init = dinc_init.data();
        d.dinc = dinc.data();
        init(d, {shape[0]});
    }
};

/* Line1d.h:486-559 */
class System_Cuspy_Quartic_RandomForcing : public detail::System {
public:
    System_Cuspy_Quartic_RandomForcing(double m, double eta, double mu, double a1, double a2, double k_frame,
                                       double dt, double mean, double stddev, uint64_t seed_forcing,
                                       const std::vector<int64_t>& dinc_init, const std::vector<int64_t>& dinc,
                                       const Shape& shape, uint64_t seed, const std::string& distribution,
                                       const std::vector<double>& parameters, double offset = -100.0,
                                       size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_QUARTIC1D;
        d.external = FQSB_EXT_RANDOMFORCING;
        d.minimisation = FQSB_MIN_NONE;
        d.a1 = a1;
        d.a2 = a2;
        d.mean = mean;
        d.stddev = stddev;
        d.seed_forcing = seed_forcing;
        if (dinc_init.size() != shape[0] || dinc.size() != shape[0]) {
            throw std::runtime_error("dinc_init and dinc must have the shape of the system");
        }
        d.dinc_init = dinc_init.data();
        d.dinc = dinc.data();
        init(d, {shape[0]});
    }
};

/* Line1d.h:336-376 */
class System_SemiSmooth_Laplace : public detail::System {
public:
    System_SemiSmooth_Laplace(double m, double eta, double mu, double kappa, double k_interactions, double k_frame,
                              double dt, const Shape& shape, uint64_t seed, const std::string& distribution,
                              const std::vector<double>& parameters, double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_SEMISMOOTH;
        d.interactions = FQSB_INT_LAPLACE1D;
        d.kappa = kappa;
        d.k_interactions = k_interactions;
        init(d, {shape[0]});
    }
};

/* Line1d.h:383-421 */
class System_Smooth_Laplace : public detail::System {
public:
    System_Smooth_Laplace(double m, double eta, double mu, double k_interactions, double k_frame, double dt,
                          const Shape& shape, uint64_t seed, const std::string& distribution,
                          const std::vector<double>& parameters, double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_SMOOTH;
        d.interactions = FQSB_INT_LAPLACE1D;
        d.k_interactions = k_interactions;
        init(d, {shape[0]});
    }
};

/* Line1d.h:428-479 */
class System_Cuspy_Quartic : public detail::System {
public:
    System_Cuspy_Quartic(double m, double eta, double mu, double a1, double a2, double k_frame, double dt,
                         const Shape& shape, uint64_t seed, const std::string& distribution,
                         const std::vector<double>& parameters, double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_QUARTIC1D;
        d.a1 = a1;
        d.a2 = a2;
        init(d, {shape[0]});
    }
};

/* Line1d.h:562-613 */
class System_Cuspy_QuarticGradient : public detail::System {
public:
    System_Cuspy_QuarticGradient(double m, double eta, double mu, double k2, double k4, double k_frame, double dt,
                                 const Shape& shape, uint64_t seed, const std::string& distribution,
                                 const std::vector<double>& parameters, double offset = -100.0,
                                 size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_QUARTICGRADIENT1D;
        d.k2 = k2;
        d.k4 = k4;
        init(d, {shape[0]});
    }
};

/* Line1d.h:620-671 */
class System_Cuspy_LongRange : public detail::System {
public:
    System_Cuspy_LongRange(double m, double eta, double mu, double k_interactions, double alpha, double k_frame,
                           double dt, const Shape& shape, uint64_t seed, const std::string& distribution,
                           const std::vector<double>& parameters, double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_LINE1D_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_LONGRANGE1D;
        d.k_interactions = k_interactions;
        d.alpha = alpha;
        init(d, {shape[0]});
    }
};

#undef FQSB_LINE1D_DESC

} // namespace Line1d

namespace Particles {

using Shape = std::array<size_t, 1>;

inline std::vector<std::string> version_dependencies()
{
    return Line1d::version_dependencies();
}

inline std::vector<std::string> version_compiler()
{
    return Line1d::version_compiler();
}

// Independent particles, each in its own landscape (no interactions).
#define FQSB_PARTICLES_DESC                                                                                         \
    detail::make_desc(1, {shape[0]}, seed, distribution, parameters, offset, nchunk)

/* Particles.h:93-151 */
class System_Cuspy : public detail::System {
public:
    System_Cuspy(double m, double eta, double mu, double k_frame, double dt, const Shape& shape, uint64_t seed,
                 const std::string& distribution, const std::vector<double>& parameters, double offset = -100.0,
                 size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_PARTICLES_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_NONE;
        init(d, {shape[0]});
    }
};

/* Particles.h:161-231 */
class System_Cuspy_RandomForcing : public detail::System {
public:
    System_Cuspy_RandomForcing(double m, double eta, double mu, double k_frame, double dt, double mean,
                               double stddev, uint64_t seed_forcing, const std::vector<int64_t>& dinc_init,
                               const std::vector<int64_t>& dinc, const Shape& shape, uint64_t seed,
                               const std::string& distribution, const std::vector<double>& parameters,
                               double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_PARTICLES_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_NONE;
        d.external = FQSB_EXT_RANDOMFORCING;
        d.minimisation = FQSB_MIN_NONE;
        d.mean = mean;
        d.stddev = stddev;
        d.seed_forcing = seed_forcing;
        if (dinc_init.size() != shape[0] || dinc.size() != shape[0]) {
            throw std::runtime_error("dinc_init and dinc must have the shape of the system");
        }
        d.dinc_init = dinc_init.data();
        d.dinc = dinc.data();
        init(d, {shape[0]});
    }
};

/* Particles.h:233-274 */
class System_SemiSmooth : public detail::System {
public:
    System_SemiSmooth(double m, double eta, double mu, double kappa, double k_frame, double dt, const Shape& shape,
                      uint64_t seed, const std::string& distribution, const std::vector<double>& parameters,
                      double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_PARTICLES_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_SEMISMOOTH;
        d.interactions = FQSB_INT_NONE;
        d.kappa = kappa;
        init(d, {shape[0]});
    }
};

/* Particles.h:276-311 */
class System_Smooth : public detail::System {
public:
    System_Smooth(double m, double eta, double mu, double k_frame, double dt, const Shape& shape, uint64_t seed,
                  const std::string& distribution, const std::vector<double>& parameters, double offset = -100.0,
                  size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(FQSB_PARTICLES_DESC, m, eta, mu, k_frame, dt);
        d.potential = FQSB_POT_SMOOTH;
        d.interactions = FQSB_INT_NONE;
        init(d, {shape[0]});
    }
};

#undef FQSB_PARTICLES_DESC

} // namespace Particles

namespace Line2d {

using Shape = std::array<size_t, 2>;

inline std::vector<std::string> version_dependencies()
{
    return Line1d::version_dependencies();
}

inline std::vector<std::string> version_compiler()
{
    return Line1d::version_compiler();
}

/* Line2d.h:77-115 */
class System_Cuspy_Laplace : public detail::System {
public:
    System_Cuspy_Laplace(double m, double eta, double mu, double k_interactions, double k_frame, double dt,
                         const Shape& shape, uint64_t seed, const std::string& distribution,
                         const std::vector<double>& parameters, double offset = -100.0, size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(
            detail::make_desc(2, {shape[0], shape[1]}, seed, distribution, parameters, offset, nchunk), m, eta, mu,
            k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_LAPLACE2D;
        d.k_interactions = k_interactions;
        init(d, {shape[0], shape[1]});
    }
};

/* Line2d.h:122-161 */
class System_Cuspy_QuarticGradient : public detail::System {
public:
    System_Cuspy_QuarticGradient(double m, double eta, double mu, double k2, double k4, double k_frame, double dt,
                                 const Shape& shape, uint64_t seed, const std::string& distribution,
                                 const std::vector<double>& parameters, double offset = -100.0,
                                 size_t nchunk = 5000)
    {
        fqsb_desc d = detail::physics(
            detail::make_desc(2, {shape[0], shape[1]}, seed, distribution, parameters, offset, nchunk), m, eta, mu,
            k_frame, dt);
        d.potential = FQSB_POT_CUSPY;
        d.interactions = FQSB_INT_QUARTICGRADIENT2D;
        d.k2 = k2;
        d.k4 = k4;
        init(d, {shape[0], shape[1]});
    }
};

} // namespace Line2d

} // namespace FrictionQPotSpringBlock

#endif
