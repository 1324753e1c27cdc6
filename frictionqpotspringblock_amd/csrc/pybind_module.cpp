// Python binding of the C++ host API (include/fqsb.hpp): the same module
// layout, class names, property/method split, argument names and defaults as
// the reference's binding (python/main.cpp:46-226 for the shared members,
// python/main.cpp:475-1006 for the constructors). Arrays are returned as numpy
// copies with the system's shape; assigning a property copies it in.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "fqsb.hpp"

namespace py = pybind11;
namespace M = FrictionQPotSpringBlock;

namespace {

template <class T>
py::array_t<T> to_array(const std::vector<T>& x, const std::vector<size_t>& shape)
{
    std::vector<py::ssize_t> sh(shape.begin(), shape.end());
    py::array_t<T> r(sh);
    std::copy(x.begin(), x.end(), r.mutable_data());
    return r;
}

template <class T>
std::vector<T> from_array(const py::object& obj, size_t n)
{
    py::array_t<T, py::array::c_style | py::array::forcecast> a(obj);
    if ((size_t)a.size() == 1 && n != 1) {
        return std::vector<T>(n, *a.data());
    }
    if ((size_t)a.size() != n) {
        throw std::runtime_error("array has the wrong size");
    }
    return std::vector<T>(a.data(), a.data() + n);
}

std::vector<size_t> with_chunk(std::vector<size_t> shape, size_t n)
{
    shape.push_back(n);
    return shape;
}

struct PyChunk {
    M::detail::Chunk c;
    std::vector<size_t> shape;
};

struct PyExternal {
    M::detail::RandomNormalForcing e;
    std::vector<size_t> shape;
};

void bind_base(py::module& m)
{
    using S = M::detail::System;
    py::class_<S> cls(m, "System");

    cls.def_property(
        "x", [](const S&) -> py::object { throw std::runtime_error("Deprecated, use 'u'"); },
        [](S&, const py::object&) { throw std::runtime_error("Deprecated, use 'u'"); });
    cls.def_property(
        "x_frame", [](const S&) -> py::object { throw std::runtime_error("Deprecated, use 'u_frame'"); },
        [](S&, double) { throw std::runtime_error("Deprecated, use 'u_frame'"); });
    cls.def_property_readonly("f_neighbours", [](const S&) -> py::object {
        throw std::runtime_error("Deprecated, use 'f_interactions'");
    });

    cls.def_property_readonly(
        "chunk", [](const S& s) { return PyChunk{s.chunk(), s.shape()}; }, "Chunk of random numbers");
    cls.def_property_readonly("size", &S::size, "Number of particles");
    cls.def_property_readonly("shape", &S::shape, "Shape of the system");
    cls.def_property_readonly("dt", &S::dt, "Time step (parameter)");
    cls.def_property_readonly("mu", &S::mu, "Curvature of each well (parameter)");
    cls.def_property_readonly("eta", &S::eta, "Damping coefficient (parameter)");
    cls.def_property_readonly("m", &S::m, "Mass of each particle (parameter)");
    cls.def_property_readonly("k_frame", &S::k_frame, "Loading frame stiffness (parameter)");

#define FQSB_ARRAY_RW(name, getter, setter, doc)                                                                   \
    cls.def_property(                                                                                              \
        name, [](const S& s) { return to_array(s.getter(), s.shape()); },                                         \
        [](S& s, const py::object& x) { s.setter(from_array<double>(x, s.size())); }, doc)
#define FQSB_ARRAY_RO(name, getter, doc)                                                                           \
    cls.def_property_readonly(name, [](const S& s) { return to_array(s.getter(), s.shape()); }, doc)

    FQSB_ARRAY_RW("u", u, set_u, "Particle slip ('positions').");
    FQSB_ARRAY_RW("v", v, set_v, "Particle velocities.");
    FQSB_ARRAY_RW("a", a, set_a, "Particle accelerations.");
    cls.def_property("inc", &S::inc, &S::set_inc, "Increment");
    cls.def_property("t", &S::t, &S::set_t, "Time");
    cls.def_property("u_frame", &S::u_frame, &S::set_u_frame, "Frame position");
    FQSB_ARRAY_RO("f", f, "Residual forces");
    FQSB_ARRAY_RO("f_potential", f_potential, "Elastic forces");
    FQSB_ARRAY_RO("f_frame", f_frame, "Frame forces");
    FQSB_ARRAY_RO("f_interactions", f_interactions, "Interaction forces");
    FQSB_ARRAY_RO("f_damping", f_damping, "Particle damping forces");
#undef FQSB_ARRAY_RW
#undef FQSB_ARRAY_RO

    cls.def_property_readonly("temperature", &S::temperature, "Temperature");
    cls.def_property_readonly("residual", &S::residual, "Residual");
    cls.def("refresh", &S::refresh, "refresh");
    cls.def("quench", &S::quench, "quench");
    cls.def("maxUniformDisplacement", &S::maxUniformDisplacement, "maxUniformDisplacement",
            py::arg("direction") = 1);
    cls.def("trigger", &S::trigger, "trigger", py::arg("p"), py::arg("eps"), py::arg("direction") = 1);
    cls.def("advanceToFixedForce", &S::advanceToFixedForce, "advanceToFixedForce", py::arg("f_frame"),
            py::arg("allow_plastic") = false);
}

// python/main.cpp:158-202
template <class C>
void athermal(C& cls)
{
    using S = M::detail::System;
    cls.def("minimise", &S::minimise, "minimise", py::arg("tol") = 1e-5, py::arg("niter_tol") = 10,
            py::arg("max_iter") = size_t(1e9), py::arg("time_activity") = false,
            py::arg("max_iter_is_error") = true);
    cls.def(
        "minimise_truncate",
        [](S& s, const py::object& i_n, size_t A_truncate, size_t S_truncate, double tol, size_t niter_tol,
           size_t max_iter, bool time_activity, bool max_iter_is_error) {
            return s.minimise_truncate(from_array<int64_t>(i_n, s.size()), A_truncate, S_truncate, tol, niter_tol,
                                       max_iter, time_activity, max_iter_is_error);
        },
        "minimise_truncate", py::arg("i_n"), py::arg("A_truncate") = 0, py::arg("S_truncate") = 0,
        py::arg("tol") = 1e-5, py::arg("niter_tol") = 10, py::arg("max_iter") = size_t(1e9),
        py::arg("time_activity") = true, py::arg("max_iter_is_error") = true);
    cls.def("eventDrivenStep", &S::eventDrivenStep, "eventDrivenStep", py::arg("eps"), py::arg("kick"),
            py::arg("direction") = 1);
    cls.def(
        "quasistatic",
        [](S& s, double eps, int64_t nevents, int64_t first_event, int direction, bool kicks_only, double tol,
           size_t niter_tol, size_t max_iter) {
            auto q = s.quasistatic(eps, nevents, first_event, direction, kicks_only, tol, niter_tol, max_iter);
            const std::vector<size_t> sh{q.S.size()};
            py::dict d;
            d["S"] = to_array(q.S, sh);
            d["u_frame"] = to_array(q.u_frame, sh);
            d["f_frame"] = to_array(q.f_frame, sh);
            d["inc"] = to_array(q.inc, sh);
            return d;
        },
        "The examples' quasistatic loop (eventDrivenStep + minimise per event) in one call", py::arg("eps"),
        py::arg("nevents"), py::arg("first_event") = 0, py::arg("direction") = 1, py::arg("kicks_only") = false,
        py::arg("tol") = 1e-5,
        py::arg("niter_tol") = 10, py::arg("max_iter") = size_t(1e9));
    cls.def_property_readonly("quasistaticActivityFirst", &S::quasistaticActivityFirst,
                              "quasistaticActivityFirst");
    cls.def_property_readonly("quasistaticActivityLast", &S::quasistaticActivityLast, "quasistaticActivityLast");
}

// python/main.cpp:204-208
template <class C>
void external(C& cls)
{
    using S = M::detail::System;
    cls.def_property_readonly(
        "external", [](const S& s) { return PyExternal{s.external(), s.shape()}; }, "Class adding external force");
}

// python/main.cpp:210-226
template <class C>
void dynamics(C& cls)
{
    using S = M::detail::System;
    cls.def("timeStep", &S::timeStep, "timeStep");
    cls.def("timeSteps", &S::timeSteps, "timeSteps", py::arg("n"));
    cls.def("timeStepsUntilEvent", &S::timeStepsUntilEvent, "timeStepsUntilEvent", py::arg("tol") = 1e-5,
            py::arg("niter_tol") = 10, py::arg("max_iter") = size_t(1e9));
    cls.def("flowSteps", &S::flowSteps, "flowSteps", py::arg("n"), py::arg("v_frame"));
}

std::vector<int64_t> as_i64(const py::object& x, size_t n)
{
    return from_array<int64_t>(x, n);
}

} // namespace

PYBIND11_MODULE(_fqsb, m)
{
    m.doc() = "Spring-block friction model, dynamics hot path on AMD MI355X (gfx950)";
    m.def("version", &M::version, "Return version string.");

    bind_base(m);

    // chunk
    py::class_<PyChunk> ch(m, "Chunk");
    ch.def_property_readonly("data", [](const PyChunk& c) {
        return to_array(c.c.data(), with_chunk(c.shape, c.c.chunk_size()));
    });
    ch.def_property_readonly("index_at_align",
                             [](const PyChunk& c) { return to_array(c.c.index_at_align(), c.shape); });
    ch.def_property_readonly("chunk_index_at_align",
                             [](const PyChunk& c) { return to_array(c.c.chunk_index_at_align(), c.shape); });
    ch.def_property_readonly("start", [](const PyChunk& c) { return to_array(c.c.start(), c.shape); });
    ch.def_property_readonly("left_of_align",
                             [](const PyChunk& c) { return to_array(c.c.left_of_align(), c.shape); });
    ch.def_property_readonly("right_of_align",
                             [](const PyChunk& c) { return to_array(c.c.right_of_align(), c.shape); });
    ch.def_property_readonly("chunk_size", [](const PyChunk& c) { return c.c.chunk_size(); });
    ch.def(
        "state_at",
        [](const PyChunk& c, const py::object& index) {
            size_t n = 1;
            for (size_t s : c.shape) {
                n *= s;
            }
            return to_array(c.c.state_at(as_i64(index, n)), c.shape);
        },
        py::arg("index"));
    ch.def(
        "restore",
        [](PyChunk& c, const py::object& state, const py::object& value, const py::object& index) {
            size_t n = 1;
            for (size_t s : c.shape) {
                n *= s;
            }
            c.c.restore(from_array<uint64_t>(state, n), from_array<double>(value, n), as_i64(index, n));
        },
        py::arg("state"), py::arg("value"), py::arg("index"));

    // FrictionQPotSpringBlock.detail
    {
        py::module sm = m.def_submodule("detail", "detail");
        py::class_<PyExternal> cls(sm, "RandomNormalForcing_1");
        cls.def_property(
            "state", [](const PyExternal& e) { return e.e.state(); },
            [](PyExternal& e, uint64_t s) { e.e.set_state(s); }, "State of RNG");
        cls.def_property(
            "f_thermal", [](const PyExternal& e) { return to_array(e.e.f_thermal(), e.shape); },
            [](PyExternal& e, const py::object& x) {
                size_t n = 1;
                for (size_t s : e.shape) {
                    n *= s;
                }
                e.e.set_f_thermal(from_array<double>(x, n));
            },
            "Random force");
        cls.def_property(
            "next", [](const PyExternal& e) { return to_array(e.e.next(), e.shape); },
            [](PyExternal& e, const py::object& x) {
                size_t n = 1;
                for (size_t s : e.shape) {
                    n *= s;
                }
                e.e.set_next(as_i64(x, n));
            },
            "Next draw increment");
        cls.def("__repr__",
                [](const PyExternal&) { return "<FrictionQPotSpringBlock.detail.RandomNormalForcing_1>"; });
    }

    // FrictionQPotSpringBlock.Line1d
    {
        py::module sm = m.def_submodule("Line1d", "Line1d");
        namespace SM = M::Line1d;
        sm.def("version_dependencies", &SM::version_dependencies,
               "Return version information of library and its dependencies.");
        sm.def("version_compiler", &SM::version_compiler, "Return compiler information.");
        using Sh = SM::Shape;
        using Par = std::vector<double>;

        {
            using S = SM::System_Cuspy_Laplace;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_Laplace");
            cls.def(py::init<double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_interactions"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_Laplace>"; });
        }
        {
            using S = SM::System_Cuspy_Laplace_Nopassing;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_Laplace_Nopassing");
            cls.def(py::init<double, double, double, const Sh&, uint64_t, const std::string&, const Par&, double,
                             size_t, double, double>(),
                    "Constructor.", py::arg("mu"), py::arg("k_interactions"), py::arg("k_frame"), py::arg("shape"),
                    py::arg("seed"), py::arg("distribution"), py::arg("parameters"), py::arg("offset") = -100.0,
                    py::arg("nchunk") = 5000, py::arg("eta") = 0.0, py::arg("dt") = 0.0);
            athermal(cls);
            cls.def("__repr__",
                    [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_Laplace_Nopassing>"; });
        }
        {
            using S = SM::System_Cuspy_Laplace_RandomForcing;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_Laplace_RandomForcing");
            cls.def(py::init([](double m_, double eta, double mu, double k_interactions, double k_frame, double dt,
                                double mean, double stddev, uint64_t seed_forcing, const py::object& dinc_init,
                                const py::object& dinc, const Sh& shape, uint64_t seed,
                                const std::string& distribution, const Par& parameters, double offset,
                                size_t nchunk) {
                        return new S(m_, eta, mu, k_interactions, k_frame, dt, mean, stddev, seed_forcing,
                                     as_i64(dinc_init, shape[0]), as_i64(dinc, shape[0]), shape, seed,
                                     distribution, parameters, offset, nchunk);
                    }),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_interactions"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("mean"), py::arg("stddev"),
                    py::arg("seed_forcing"), py::arg("dinc_init"), py::arg("dinc"), py::arg("shape"),
                    py::arg("seed"), py::arg("distribution"), py::arg("parameters"), py::arg("offset") = -100.0,
                    py::arg("nchunk") = 5000);
            dynamics(cls);
            external(cls);
            cls.def("__repr__", [](const S&) {
                return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_Laplace_RandomForcing>";
            });
        }
        {
            using S = SM::System_Cuspy_Quartic_RandomForcing;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_Quartic_RandomForcing");
            cls.def(py::init([](double m_, double eta, double mu, double a1, double a2, double k_frame, double dt,
                                double mean, double stddev, uint64_t seed_forcing, const py::object& dinc_init,
                                const py::object& dinc, const Sh& shape, uint64_t seed,
                                const std::string& distribution, const Par& parameters, double offset,
                                size_t nchunk) {
                        return new S(m_, eta, mu, a1, a2, k_frame, dt, mean, stddev, seed_forcing,
                                     as_i64(dinc_init, shape[0]), as_i64(dinc, shape[0]), shape, seed,
                                     distribution, parameters, offset, nchunk);
                    }),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("a1"), py::arg("a2"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("mean"), py::arg("stddev"),
                    py::arg("seed_forcing"), py::arg("dinc_init"), py::arg("dinc"), py::arg("shape"),
                    py::arg("seed"), py::arg("distribution"), py::arg("parameters"), py::arg("offset") = -100.0,
                    py::arg("nchunk") = 5000);
            dynamics(cls);
            external(cls);
            cls.def("__repr__", [](const S&) {
                return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_Quartic_RandomForcing>";
            });
        }
        {
            using S = SM::System_SemiSmooth_Laplace;
            py::class_<S, M::detail::System> cls(sm, "System_SemiSmooth_Laplace");
            cls.def(py::init<double, double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("kappa"),
                    py::arg("k_interactions"), py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"),
                    py::arg("distribution"), py::arg("parameters"), py::arg("offset") = -100.0,
                    py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__",
                    [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_SemiSmooth_Laplace>"; });
        }
        {
            using S = SM::System_Smooth_Laplace;
            py::class_<S, M::detail::System> cls(sm, "System_Smooth_Laplace");
            cls.def(py::init<double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_interactions"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_Smooth_Laplace>"; });
        }
        {
            using S = SM::System_Cuspy_Quartic;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_Quartic");
            cls.def(py::init<double, double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("a1"), py::arg("a2"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_Quartic>"; });
        }
        {
            using S = SM::System_Cuspy_QuarticGradient;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_QuarticGradient");
            cls.def(py::init<double, double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k2"), py::arg("k4"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__",
                    [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_QuarticGradient>"; });
        }
        {
            using S = SM::System_Cuspy_LongRange;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_LongRange");
            cls.def(py::init<double, double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_interactions"),
                    py::arg("alpha"), py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"),
                    py::arg("distribution"), py::arg("parameters"), py::arg("offset") = -100.0,
                    py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Line1d.System_Cuspy_LongRange>"; });
        }
    }

    // FrictionQPotSpringBlock.Particles (python/main.cpp:278-470)
    {
        py::module sm = m.def_submodule("Particles", "Particles");
        namespace SM = M::Particles;
        sm.def("version_dependencies", &SM::version_dependencies,
               "Return version information of library and its dependencies.");
        sm.def("version_compiler", &SM::version_compiler, "Return compiler information.");
        using Sh = SM::Shape;
        using Par = std::vector<double>;
        {
            using S = SM::System_Cuspy;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy");
            cls.def(py::init<double, double, double, double, double, const Sh&, uint64_t, const std::string&,
                             const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_frame"), py::arg("dt"),
                    py::arg("shape"), py::arg("seed"), py::arg("distribution"), py::arg("parameters"),
                    py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            // the reference's text (python/main.cpp:338)
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Particle.System_Cuspy>"; });
        }
        {
            using S = SM::System_Cuspy_RandomForcing;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_RandomForcing");
            cls.def(py::init([](double m_, double eta, double mu, double k_frame, double dt, double mean,
                                double stddev, uint64_t seed_forcing, const py::object& dinc_init,
                                const py::object& dinc, const Sh& shape, uint64_t seed,
                                const std::string& distribution, const Par& parameters, double offset,
                                size_t nchunk) {
                        return new S(m_, eta, mu, k_frame, dt, mean, stddev, seed_forcing,
                                     as_i64(dinc_init, shape[0]), as_i64(dinc, shape[0]), shape, seed,
                                     distribution, parameters, offset, nchunk);
                    }),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_frame"), py::arg("dt"),
                    py::arg("mean"), py::arg("stddev"), py::arg("seed_forcing"), py::arg("dinc_init"),
                    py::arg("dinc"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            dynamics(cls);
            external(cls);
            cls.def("__repr__",
                    [](const S&) { return "<FrictionQPotSpringBlock.Particles.System_Cuspy_RandomForcing>"; });
        }
        {
            using S = SM::System_SemiSmooth;
            py::class_<S, M::detail::System> cls(sm, "System_SemiSmooth");
            cls.def(py::init<double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("kappa"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Particles.System_SemiSmooth>"; });
        }
        {
            using S = SM::System_Smooth;
            py::class_<S, M::detail::System> cls(sm, "System_Smooth");
            cls.def(py::init<double, double, double, double, double, const Sh&, uint64_t, const std::string&,
                             const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_frame"), py::arg("dt"),
                    py::arg("shape"), py::arg("seed"), py::arg("distribution"), py::arg("parameters"),
                    py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Particles.System_Smooth>"; });
        }
    }

    // FrictionQPotSpringBlock.Line2d
    {
        py::module sm = m.def_submodule("Line2d", "Line2d");
        namespace SM = M::Line2d;
        sm.def("version_dependencies", &SM::version_dependencies,
               "Return version information of library and its dependencies.");
        sm.def("version_compiler", &SM::version_compiler, "Return compiler information.");
        using Sh = SM::Shape;
        using Par = std::vector<double>;
        {
            using S = SM::System_Cuspy_Laplace;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_Laplace");
            cls.def(py::init<double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k_interactions"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__", [](const S&) { return "<FrictionQPotSpringBlock.Line2d.System_Cuspy_Laplace>"; });
        }
        {
            using S = SM::System_Cuspy_QuarticGradient;
            py::class_<S, M::detail::System> cls(sm, "System_Cuspy_QuarticGradient");
            cls.def(py::init<double, double, double, double, double, double, double, const Sh&, uint64_t,
                             const std::string&, const Par&, double, size_t>(),
                    "Constructor.", py::arg("m"), py::arg("eta"), py::arg("mu"), py::arg("k2"), py::arg("k4"),
                    py::arg("k_frame"), py::arg("dt"), py::arg("shape"), py::arg("seed"), py::arg("distribution"),
                    py::arg("parameters"), py::arg("offset") = -100.0, py::arg("nchunk") = 5000);
            athermal(cls);
            dynamics(cls);
            cls.def("__repr__",
                    [](const S&) { return "<FrictionQPotSpringBlock.Line2d.System_Cuspy_QuarticGradient>"; });
        }
    }
}
