// bindings.cpp -- pybind11 module libaccint._core over the C++ host layer.
// Names and argument meaning follow the reference bindings
// (python/src/{basis,core_types,operator,screening,buffer,consumer,engine}_bindings.cpp).
#include <libaccint/data.hpp>
#include <libaccint/df.hpp>
#include <libaccint/engine.hpp>

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <sstream>

namespace py = pybind11;
using namespace libaccint;
using consumers::DFFockBuilder;
using consumers::DFFockBuilderConfig;
using consumers::DFMetricType;
using consumers::FockBuilder;
using consumers::FockThreadingStrategy;
using consumers::GpuFockBuilder;
using consumers::MixedPrecisionFockBuilder;
using screening::ScreeningOptions;
using screening::SchwarzBounds;

namespace {

py::array_t<Real> to_matrix(const std::vector<Real>& v, Size n) {
    py::array_t<Real> a({n, n});
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

// C-contiguous float64 n x n view (copied when needed is rejected, as the reference does)
py::array_t<Real, py::array::c_style> square(py::array D, Size n, const char* who) {
    if (!py::isinstance<py::array_t<Real>>(D)) throw std::invalid_argument(std::string(who) + ": density must have dtype float64");
    auto buf = D.request();
    if (buf.ndim != 2 || buf.shape[0] != py::ssize_t(n) || buf.shape[1] != py::ssize_t(n))
        throw std::invalid_argument(std::string(who) + ": density shape must be (" + std::to_string(n) + ", " +
                                    std::to_string(n) + ")");
    if (!(D.flags() & py::array::c_style)) throw std::invalid_argument(std::string(who) + ": density must be C-contiguous");
    return py::array_t<Real, py::array::c_style>::ensure(D);
}

PointChargeParams charges_of(const std::vector<data::Atom>& atoms) {
    PointChargeParams c;
    for (const auto& a : atoms) {
        c.x.push_back(a.position.x);
        c.y.push_back(a.position.y);
        c.z.push_back(a.position.z);
        c.charge.push_back(static_cast<Real>(a.atomic_number));
    }
    return c;
}

Point3D to_point(py::handle h) {
    if (py::isinstance<Point3D>(h)) return h.cast<Point3D>();
    auto v = h.cast<std::vector<Real>>();
    if (v.size() != 3) throw std::invalid_argument("position must have 3 components");
    return {v[0], v[1], v[2]};
}

// A Python callable as a two-electron consumer: accumulate(block, fa, fb, fc,
// fd) with block the (na, nb, nc, nd) integrals (the TwoElectronConsumer
// concept; Engine routes it like any consumer that is not a Fock builder).
struct PyConsumer {
    py::function fn;
    void accumulate(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb, int nc,
                    int nd) {
        py::array_t<Real> a({na, nb, nc, nd});
        Real* p = a.mutable_data();
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j)
                for (int k = 0; k < nc; ++k)
                    for (int l = 0; l < nd; ++l) *p++ = buf(i, j, k, l);
        fn(a, fa, fb, fc, fd);
    }
};

// The ShellSetQuartets named by `which` for f: None = the basis's canonical
// list (BasisSet::shell_set_quartets), else a sequence of (a, b, c, d)
// indices into BasisSet::shell_sets() (bra pair (a, b), ket pair (c, d)).
template <typename F>
void with_shell_set_quartets(Engine& e, py::object which, F&& f) {
    const BasisSet& bs = e.basis();
    if (which.is_none()) {
        const auto& q = bs.shell_set_quartets();
        f(std::span<const ShellSetQuartet>(q.data(), q.size()));
        return;
    }
    const auto sets = bs.shell_sets();
    auto idx = which.cast<std::vector<std::array<int, 4>>>();
    std::vector<ShellSetPair> pairs;
    pairs.reserve(2 * idx.size());  // stable addresses for the quartets
    std::vector<ShellSetQuartet> quartets;
    quartets.reserve(idx.size());
    for (const auto& t : idx) {
        for (int k = 0; k < 4; ++k)
            if (t[k] < 0 || static_cast<Size>(t[k]) >= sets.size())
                throw InvalidArgumentException("shell set index out of range");
        pairs.emplace_back(*sets[t[0]], *sets[t[1]]);
        pairs.emplace_back(*sets[t[2]], *sets[t[3]]);
        quartets.emplace_back(pairs[pairs.size() - 2], pairs.back());
    }
    f(std::span<const ShellSetQuartet>(quartets.data(), quartets.size()));
}

py::dict info_dict(const WorklistInfo& w) {
    py::dict d;
    d["n_quartets"] = w.n_quartets;
    d["n_fp32"] = w.n_fp32;
    d["checksum"] = w.checksum;
    d["flops_fp64"] = w.flops_fp64;
    d["flops_fp32"] = w.flops_fp32;
    d["n_eri"] = w.n_eri;
    d["class_counts"] = std::vector<Index>(w.class_counts.begin(), w.class_counts.end());
    return d;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "LibAccInt on MI355X: Rys ERI and Fock (J/K) build on gfx950";
    m.attr("__version__") = "0.1.0-alpha.2";

    py::register_exception<InvalidArgumentException>(m, "InvalidArgumentException", PyExc_ValueError);
    py::register_exception<InvalidStateException>(m, "InvalidStateException", PyExc_RuntimeError);
    py::register_exception<MemoryException>(m, "MemoryException", PyExc_MemoryError);
    py::register_exception<BackendError>(m, "BackendError", PyExc_RuntimeError);

    // ---- core types -------------------------------------------------------------
    py::enum_<AngularMomentum>(m, "AngularMomentum")
        .value("S", AngularMomentum::S).value("P", AngularMomentum::P)
        .value("D", AngularMomentum::D).value("F", AngularMomentum::F);
    py::enum_<BackendType>(m, "BackendType").value("CPU", BackendType::CPU).value("HIP", BackendType::HIP);
    py::enum_<BackendHint>(m, "BackendHint")
        .value("Auto", BackendHint::Auto).value("ForceCPU", BackendHint::ForceCPU)
        .value("ForceGPU", BackendHint::ForceGPU).value("PreferCPU", BackendHint::PreferCPU)
        .value("PreferGPU", BackendHint::PreferGPU);
    py::class_<Point3D>(m, "Point3D")
        .def(py::init<>())
        .def(py::init<Real, Real, Real>(), py::arg("x"), py::arg("y"), py::arg("z"))
        .def_readwrite("x", &Point3D::x).def_readwrite("y", &Point3D::y).def_readwrite("z", &Point3D::z)
        .def("distance_squared", &Point3D::distance_squared)
        .def("__repr__", [](const Point3D& p) {
            std::ostringstream o;
            o << "Point3D(" << p.x << ", " << p.y << ", " << p.z << ")";
            return o.str();
        });
    m.def("n_cartesian", &n_cartesian, py::arg("l"));
    m.def("n_spherical", &n_spherical, py::arg("l"));
    m.def("is_backend_available", [](BackendType t) { return t == BackendType::CPU || device::gpu_visible(); });
    m.def("gpu_available", &device::gpu_visible);

    // ---- basis ------------------------------------------------------------------
    py::class_<data::Atom>(m, "Atom")
        .def(py::init([](int z, py::handle pos) { return data::Atom{z, to_point(pos)}; }), py::arg("atomic_number"),
             py::arg("position"))
        .def_readwrite("atomic_number", &data::Atom::atomic_number)
        .def_readwrite("position", &data::Atom::position);

    py::class_<Shell>(m, "Shell")
        .def(py::init([](int am, py::handle c, std::vector<Real> e, std::vector<Real> k) {
                 return Shell(am, to_point(c), std::move(e), std::move(k));
             }),
             py::arg("am"), py::arg("center"), py::arg("exponents"), py::arg("coefficients"))
        .def("angular_momentum", &Shell::angular_momentum)
        .def("n_functions", &Shell::n_functions)
        .def("n_primitives", &Shell::n_primitives)
        .def("center", &Shell::center)
        .def("exponents", [](const Shell& s) { return std::vector<Real>(s.exponents().begin(), s.exponents().end()); })
        .def("coefficients",
             [](const Shell& s) { return std::vector<Real>(s.coefficients().begin(), s.coefficients().end()); })
        .def("shell_index", &Shell::shell_index)
        .def("function_index", &Shell::function_index)
        .def("atom_index", &Shell::atom_index);

    py::class_<ShellSet>(m, "ShellSet")
        .def("angular_momentum", &ShellSet::angular_momentum)
        .def("n_primitives_per_shell", &ShellSet::n_primitives_per_shell)
        .def("n_shells", &ShellSet::n_shells)
        .def("n_functions_per_shell", &ShellSet::n_functions_per_shell)
        .def("shell", &ShellSet::shell, py::return_value_policy::reference_internal);

    py::class_<BasisSet, std::shared_ptr<BasisSet>>(m, "BasisSet")
        .def(py::init([](std::vector<Shell> shells) { return std::make_shared<BasisSet>(std::move(shells)); }),
             py::arg("shells"))
        .def("n_shells", &BasisSet::n_shells)
        .def("n_basis_functions", &BasisSet::n_basis_functions)
        .def("n_basis_functions_spherical", &BasisSet::n_basis_functions_spherical)
        .def("max_angular_momentum", &BasisSet::max_angular_momentum)
        .def("max_n_primitives", &BasisSet::max_n_primitives)
        .def("n_shell_sets", &BasisSet::n_shell_sets)
        .def("shell", &BasisSet::shell, py::return_value_policy::reference_internal)
        .def("shells", [](const BasisSet& b) { return std::vector<Shell>(b.shells().begin(), b.shells().end()); })
        .def("shell_sets", &BasisSet::shell_sets, py::return_value_policy::reference_internal)
        .def("n_shell_set_pairs", [](const BasisSet& b) { return b.shell_set_pairs().size(); })
        .def("n_shell_set_quartets", [](const BasisSet& b) { return b.shell_set_quartets().size(); })
        .def("shell_set_quartet_indices",
             [](const BasisSet& b) {  // (a, b, c, d) indices into shell_sets() of shell_set_quartets()
                 const auto sets = b.shell_sets();
                 auto index = [&](const ShellSet* p) {
                     return int(std::find(sets.begin(), sets.end(), p) - sets.begin());
                 };
                 std::vector<std::array<int, 4>> out;
                 for (const auto& q : b.shell_set_quartets())
                     out.push_back({index(&q.bra_pair().shell_set_a()), index(&q.bra_pair().shell_set_b()),
                                    index(&q.ket_pair().shell_set_a()), index(&q.ket_pair().shell_set_b())});
                 return out;
             })
        .def("__repr__", [](const BasisSet& b) {
            return "BasisSet(n_shells=" + std::to_string(b.n_shells()) +
                   ", n_basis_functions=" + std::to_string(b.n_basis_functions()) + ")";
        });

    m.def("create_builtin_basis",
          [](const std::string& name, const std::vector<data::Atom>& atoms) {
              return std::make_shared<BasisSet>(data::create_builtin_basis(name, atoms));
          },
          py::arg("name"), py::arg("atoms"));
    m.def("load_basis_set",
          [](const std::string& name, const std::vector<data::Atom>& atoms) {
              return std::make_shared<BasisSet>(data::load_basis_set(name, atoms));
          },
          py::arg("name"), py::arg("atoms"));
    m.def("list_available_basis_sets", &data::list_available_basis_sets);
    m.def("basis_data_dir", &data::basis_data_dir);

    // ---- operators -------------------------------------------------------------
    py::enum_<OperatorKind>(m, "OperatorKind")
        .value("Overlap", OperatorKind::Overlap).value("Kinetic", OperatorKind::Kinetic)
        .value("Nuclear", OperatorKind::Nuclear).value("PointCharge", OperatorKind::PointCharge)
        .value("Coulomb", OperatorKind::Coulomb);
    py::class_<PointChargeParams>(m, "PointChargeParams")
        .def(py::init<>())
        .def_readwrite("x", &PointChargeParams::x).def_readwrite("y", &PointChargeParams::y)
        .def_readwrite("z", &PointChargeParams::z).def_readwrite("charge", &PointChargeParams::charge)
        .def("n_centers", &PointChargeParams::n_centers);
    py::class_<Operator>(m, "Operator")
        .def_static("coulomb", &Operator::coulomb)
        .def_static("overlap", &Operator::overlap)
        .def_static("kinetic", &Operator::kinetic)
        .def_static("nuclear", &Operator::nuclear, py::arg("charges"))
        .def("kind", &Operator::kind)
        .def("is_one_electron", &Operator::is_one_electron)
        .def("is_two_electron", &Operator::is_two_electron);
    py::class_<OneElectronOperator>(m, "OneElectronOperator")
        .def(py::init<const Operator&, Real>(), py::arg("op"), py::arg("scale") = 1.0)
        .def("add", &OneElectronOperator::add, py::arg("op"), py::arg("scale") = 1.0,
             py::return_value_policy::reference_internal);

    // ---- buffers -----------------------------------------------------------------
    py::class_<TwoElectronBuffer<0>>(m, "TwoElectronBuffer")
        .def(py::init<>())
        .def(py::init<int, int, int, int>(), py::arg("na"), py::arg("nb"), py::arg("nc"), py::arg("nd"))
        .def("__call__", [](const TwoElectronBuffer<0>& b, int a, int c, int d, int e) { return b(a, c, d, e); })
        .def("n_integrals", &TwoElectronBuffer<0>::n_integrals)
        .def("to_numpy", [](const TwoElectronBuffer<0>& b) {
            py::array_t<Real> a({b.na(), b.nb(), b.nc(), b.nd()});
            std::copy(b.data().begin(), b.data().end(), a.mutable_data());
            return a;
        })
        .def("from_numpy", [](TwoElectronBuffer<0>& b, py::array_t<Real, py::array::c_style | py::array::forcecast> x) {
            if (x.ndim() != 4) throw InvalidArgumentException("TwoElectronBuffer.from_numpy: expected 4 dimensions");
            b.resize(int(x.shape(0)), int(x.shape(1)), int(x.shape(2)), int(x.shape(3)));
            std::copy(x.data(), x.data() + x.size(), b.data().begin());
        });
    py::class_<TwoElectronBuffer<0, float>>(m, "TwoElectronBufferF32")
        .def(py::init<>())
        .def("__call__", [](const TwoElectronBuffer<0, float>& b, int a, int c, int d, int e) { return b(a, c, d, e); })
        .def("from_numpy",
             [](TwoElectronBuffer<0, float>& b, py::array_t<float, py::array::c_style | py::array::forcecast> x) {
                 if (x.ndim() != 4) throw InvalidArgumentException("TwoElectronBufferF32.from_numpy: expected 4 dimensions");
                 b.resize(int(x.shape(0)), int(x.shape(1)), int(x.shape(2)), int(x.shape(3)));
                 std::copy(x.data(), x.data() + x.size(), b.data().begin());
             })
        .def("to_numpy", [](const TwoElectronBuffer<0, float>& b) {
            py::array_t<float> a({b.na(), b.nb(), b.nc(), b.nd()});
            std::copy(b.data().begin(), b.data().end(), a.mutable_data());
            return a;
        });
    py::class_<TwoElectronBuffer<1>>(m, "TwoElectronGradientBuffer")
        .def(py::init<int, int, int, int>(), py::arg("na"), py::arg("nb"), py::arg("nc"), py::arg("nd"))
        .def_property_readonly_static("N_DERIV", [](py::object) { return TwoElectronBuffer<1>::N_DERIV; })
        .def("__call__", [](const TwoElectronBuffer<1>& b, int a, int c, int d, int e, int k) { return b(a, c, d, e, k); })
        .def("set", [](TwoElectronBuffer<1>& b, int a, int c, int d, int e, int k, Real v) { b(a, c, d, e, k) = v; })
        .def("size", &TwoElectronBuffer<1>::size)
        .def("to_numpy", [](const TwoElectronBuffer<1>& b) {
            py::array_t<Real> a({TwoElectronBuffer<1>::N_DERIV, b.na(), b.nb(), b.nc(), b.nd()});
            std::copy(b.data().begin(), b.data().end(), a.mutable_data());
            return a;
        });

    // ---- screening ---------------------------------------------------------------
    py::enum_<screening::ScreeningPreset>(m, "ScreeningPreset")
        .value("None_", screening::ScreeningPreset::None).value("Loose", screening::ScreeningPreset::Loose)
        .value("Normal", screening::ScreeningPreset::Normal).value("Tight", screening::ScreeningPreset::Tight)
        .value("Custom", screening::ScreeningPreset::Custom);
    py::class_<ScreeningOptions>(m, "ScreeningOptions")
        .def(py::init<>())
        .def(py::init([](Real thr, bool enabled, bool dw, bool sym) {
                 return ScreeningOptions{thr, enabled, dw, sym, false};
             }),
             py::arg("threshold") = 1e-12, py::arg("enabled") = true, py::arg("density_weighted") = false,
             py::arg("use_permutation_symmetry") = false)
        .def_readwrite("threshold", &ScreeningOptions::threshold)
        .def_readwrite("enabled", &ScreeningOptions::enabled)
        .def_readwrite("density_weighted", &ScreeningOptions::density_weighted)
        .def_readwrite("use_permutation_symmetry", &ScreeningOptions::use_permutation_symmetry)
        .def_static("none", &ScreeningOptions::none)
        .def_static("loose", &ScreeningOptions::loose)
        .def_static("normal", &ScreeningOptions::normal)
        .def_static("tight", &ScreeningOptions::tight)
        .def("effective_threshold", &ScreeningOptions::effective_threshold)
        .def("passes_screening", &ScreeningOptions::passes_screening)
        .def("preset_name", &ScreeningOptions::preset_name);
    py::class_<SchwarzBounds>(m, "SchwarzBounds")
        .def("__call__", &SchwarzBounds::operator())
        .def("n_shells", &SchwarzBounds::n_shells)
        .def("max_bound", &SchwarzBounds::max_bound)
        .def("passes_screening", &SchwarzBounds::passes_screening)
        .def("quartet_bound", &SchwarzBounds::quartet_bound)
        .def("count_passing_quartets", &SchwarzBounds::count_passing_quartets)
        .def("estimate_pass_fraction", &SchwarzBounds::estimate_pass_fraction)
        .def("packed", [](const SchwarzBounds& s) {
            py::array_t<Real> a(s.packed().size());
            std::copy(s.packed().begin(), s.packed().end(), a.mutable_data());
            return a;
        });

    // ---- consumers ---------------------------------------------------------------
    py::enum_<FockThreadingStrategy>(m, "FockThreadingStrategy")
        .value("Sequential", FockThreadingStrategy::Sequential)
        .value("Atomic", FockThreadingStrategy::Atomic)
        .value("ThreadLocal", FockThreadingStrategy::ThreadLocal);
    // FockBuilder keeps the caller's density pointer (reference semantics); the
    // Python wrapper keeps the array alive on the builder object.
    py::class_<FockBuilder>(m, "FockBuilder", py::dynamic_attr())
        .def(py::init<Size>(), py::arg("nbf"))
        .def("set_density",
             [](py::object self, py::array D) {
                 FockBuilder& fb = self.cast<FockBuilder&>();
                 auto a = square(D, fb.nbf(), "FockBuilder.set_density");
                 self.attr("_density_ref") = a;
                 fb.set_density(a.data(), fb.nbf());
             },
             py::arg("D"))
        .def("get_coulomb_matrix", [](const FockBuilder& f) {
            return to_matrix(std::vector<Real>(f.get_coulomb_matrix().begin(), f.get_coulomb_matrix().end()), f.nbf());
        })
        .def("get_exchange_matrix", [](const FockBuilder& f) {
            return to_matrix(std::vector<Real>(f.get_exchange_matrix().begin(), f.get_exchange_matrix().end()),
                             f.nbf());
        })
        .def("get_fock_matrix",
             [](const FockBuilder& f, py::array_t<Real, py::array::c_style | py::array::forcecast> H, Real x) {
                 return to_matrix(f.get_fock_matrix(std::span<const Real>(H.data(), H.size()), x), f.nbf());
             },
             py::arg("H_core"), py::arg("exchange_fraction") = 1.0)
        .def("reset", &FockBuilder::reset)
        .def("nbf", &FockBuilder::nbf)
        .def("set_threading_strategy", &FockBuilder::set_threading_strategy)
        .def("threading_strategy", &FockBuilder::threading_strategy)
        .def("prepare_parallel", &FockBuilder::prepare_parallel, py::arg("n_threads") = 0)
        .def("finalize_parallel", &FockBuilder::finalize_parallel)
        .def("__repr__", [](const FockBuilder& f) { return "FockBuilder(nbf=" + std::to_string(f.nbf()) + ")"; });

    py::enum_<MixedPrecisionMode>(m, "MixedPrecisionMode")
        .value("Pure32", MixedPrecisionMode::Pure32)
        .value("Pure64", MixedPrecisionMode::Pure64)
        .value("Compute32Accumulate64", MixedPrecisionMode::Compute32Accumulate64)
        .value("Adaptive", MixedPrecisionMode::Adaptive);
    py::class_<MixedPrecisionFockBuilder>(m, "MixedPrecisionFockBuilder", py::dynamic_attr())
        .def(py::init<Size>(), py::arg("nbf"))
        .def("set_density",
             [](py::object self, py::array D) {
                 MixedPrecisionFockBuilder& fb = self.cast<MixedPrecisionFockBuilder&>();
                 auto a = square(D, fb.nbf(), "MixedPrecisionFockBuilder.set_density");
                 self.attr("_density_ref") = a;
                 fb.set_density(a.data(), fb.nbf());
             },
             py::arg("D"))
        .def("accumulate",
             [](MixedPrecisionFockBuilder& f, const TwoElectronBuffer<0, float>& b, Index fa, Index fb, Index fc,
                Index fd) { f.accumulate(b, fa, fb, fc, fd, b.na(), b.nb(), b.nc(), b.nd()); },
             py::arg("buffer"), py::arg("fa"), py::arg("fb"), py::arg("fc"), py::arg("fd"))
        .def("accumulate",
             [](MixedPrecisionFockBuilder& f, const TwoElectronBuffer<0>& b, Index fa, Index fb, Index fc, Index fd) {
                 f.accumulate(b, fa, fb, fc, fd, b.na(), b.nb(), b.nc(), b.nd());
             },
             py::arg("buffer"), py::arg("fa"), py::arg("fb"), py::arg("fc"), py::arg("fd"))
        .def("get_coulomb_matrix", [](const MixedPrecisionFockBuilder& f) {
            return to_matrix(std::vector<Real>(f.get_coulomb_matrix().begin(), f.get_coulomb_matrix().end()), f.nbf());
        })
        .def("get_exchange_matrix", [](const MixedPrecisionFockBuilder& f) {
            return to_matrix(std::vector<Real>(f.get_exchange_matrix().begin(), f.get_exchange_matrix().end()),
                             f.nbf());
        })
        .def("get_fock_matrix",
             [](const MixedPrecisionFockBuilder& f, py::array_t<Real, py::array::c_style | py::array::forcecast> H,
                Real x) { return to_matrix(f.get_fock_matrix(std::span<const Real>(H.data(), H.size()), x), f.nbf()); },
             py::arg("H_core"), py::arg("exchange_fraction") = 1.0)
        .def("reset", &MixedPrecisionFockBuilder::reset)
        .def("nbf", &MixedPrecisionFockBuilder::nbf)
        .def("mode", &MixedPrecisionFockBuilder::mode)
        .def("n_float32_accumulations", &MixedPrecisionFockBuilder::n_float32_accumulations)
        .def("n_float64_accumulations", &MixedPrecisionFockBuilder::n_float64_accumulations)
        .def("set_fp32_threshold", &MixedPrecisionFockBuilder::set_fp32_threshold, py::arg("tau"))
        .def("fp32_threshold", &MixedPrecisionFockBuilder::fp32_threshold);

    py::class_<DeviceEriBatch>(m, "DeviceEriBatch")
        .def("size", &DeviceEriBatch::size)
        .def("n_blocks", &DeviceEriBatch::n_blocks)
        .def("empty", &DeviceEriBatch::empty)
        .def("data_ptr", [](const DeviceEriBatch& b) { return reinterpret_cast<std::uintptr_t>(b.data()); })
        .def("blocks_ptr", [](const DeviceEriBatch& b) { return reinterpret_cast<std::uintptr_t>(b.blocks()); });

    py::class_<GpuFockBuilder>(m, "GpuFockBuilder")
        .def(py::init([](Size nbf, std::uintptr_t stream) {
                 return std::make_unique<GpuFockBuilder>(nbf, reinterpret_cast<hipStream_t>(stream));
             }),
             py::arg("nbf"), py::arg("stream") = 0)
        .def("set_density",
             [](GpuFockBuilder& f, py::array D) {
                 auto a = square(D, f.nbf(), "GpuFockBuilder.set_density");
                 f.set_density(a.data(), f.nbf());
             },
             py::arg("D"))
        .def("get_coulomb_matrix", [](const GpuFockBuilder& f) { return to_matrix(f.get_coulomb_matrix(), f.nbf()); })
        .def("get_exchange_matrix", [](const GpuFockBuilder& f) { return to_matrix(f.get_exchange_matrix(), f.nbf()); })
        .def("get_fock_matrix",
             [](const GpuFockBuilder& f, py::array_t<Real, py::array::c_style | py::array::forcecast> H, Real x) {
                 return to_matrix(f.get_fock_matrix(std::span<const Real>(H.data(), H.size()), x), f.nbf());
             },
             py::arg("H_core"), py::arg("exchange_fraction") = 1.0)
        .def("accumulate",
             [](GpuFockBuilder& f, const TwoElectronBuffer<0>& b, Index fa, Index fb, Index fc, Index fd) {
                 f.accumulate(b, fa, fb, fc, fd, b.na(), b.nb(), b.nc(), b.nd());
             },
             py::arg("buffer"), py::arg("fa"), py::arg("fb"), py::arg("fc"), py::arg("fd"))
        .def("accumulate_device_eri_batch",
             [](GpuFockBuilder& f, const DeviceEriBatch& batch) { f.accumulate_device_eri_batch(batch, f.nbf()); },
             py::arg("batch"))
        .def("reset", &GpuFockBuilder::reset)
        .def("synchronize", &GpuFockBuilder::synchronize)
        .def("nbf", &GpuFockBuilder::nbf)
        .def("d_J", [](GpuFockBuilder& f) { return reinterpret_cast<std::uintptr_t>(f.d_J()); })
        .def("d_K", [](GpuFockBuilder& f) { return reinterpret_cast<std::uintptr_t>(f.d_K()); })
        .def("d_D", [](const GpuFockBuilder& f) { return reinterpret_cast<std::uintptr_t>(f.d_D()); })
        .def("__repr__", [](const GpuFockBuilder& f) { return "GpuFockBuilder(nbf=" + std::to_string(f.nbf()) + ")"; });

    // ---- engine ----------------------------------------------------------------------
    py::class_<DispatchConfig>(m, "DispatchConfig")
        .def(py::init<>())
        .def_readwrite("min_gpu_batch_size", &DispatchConfig::min_gpu_batch_size)
        .def_readwrite("min_gpu_primitives", &DispatchConfig::min_gpu_primitives)
        .def_readwrite("min_gpu_shells", &DispatchConfig::min_gpu_shells)
        .def_readwrite("n_gpu_slots", &DispatchConfig::n_gpu_slots)
        .def_readwrite("enable_auto_tuning", &DispatchConfig::enable_auto_tuning)
        .def_readwrite("auto_tune_min_batch", &DispatchConfig::auto_tune_min_batch);

    py::enum_<WorkUnitType>(m, "WorkUnitType")
        .value("SingleShellPair", WorkUnitType::SingleShellPair)
        .value("SingleShellQuartet", WorkUnitType::SingleShellQuartet)
        .value("ShellSetPair", WorkUnitType::ShellSetPair)
        .value("ShellSetQuartet", WorkUnitType::ShellSetQuartet)
        .value("FullBasis", WorkUnitType::FullBasis);

    py::class_<DispatchPolicy>(m, "DispatchPolicy")
        .def(py::init<>())
        .def(py::init<DispatchConfig>(), py::arg("config"))
        .def("select_backend", &DispatchPolicy::select_backend, py::arg("work_type"), py::arg("batch_size"),
             py::arg("total_am"), py::arg("n_primitives"), py::arg("hint"), py::arg("gpu_available"))
        .def("config", &DispatchPolicy::config)
        .def("set_config", &DispatchPolicy::set_config, py::arg("config"))
        .def("set_auto_tuning", &DispatchPolicy::set_auto_tuning, py::arg("enable"))
        .def("auto_tuning_enabled", &DispatchPolicy::auto_tuning_enabled);

    py::enum_<GpuExecutionStrategy>(m, "GpuExecutionStrategy")
        .value("ThreadPerQuartet", GpuExecutionStrategy::ThreadPerQuartet)
        .value("WarpPerQuartet", GpuExecutionStrategy::WarpPerQuartet)
        .value("BlockPerQuartet", GpuExecutionStrategy::BlockPerQuartet);
    m.def("gpu_execution_strategy_from_string", [](const std::string& s) { return gpu_execution_strategy_from_string(s); });
    m.def("strategy_name", [](GpuExecutionStrategy s) { return std::string(to_string(s)); });
    m.def("quartet_class_index", [](std::array<int, 4> am) { return quartet_class_index(am); }, py::arg("am"));
    m.def("quartet_class_am", &quartet_class_am, py::arg("cls"));

    py::class_<DispatchEntry>(m, "DispatchEntry")
        .def(py::init([](int cls, GpuExecutionStrategy st, int chunk, int bpc) {
                 DispatchEntry e;
                 e.cls = cls;
                 e.strategy = st;
                 e.chunk = chunk;
                 e.blocks_per_cu = bpc;
                 return e;
             }),
             py::arg("cls"), py::arg("strategy"), py::arg("chunk") = 0, py::arg("blocks_per_cu") = 0)
        .def_readwrite("cls", &DispatchEntry::cls)
        .def_property_readonly("am", [](const DispatchEntry& e) { return quartet_class_am(e.cls); })
        .def_readwrite("strategy", &DispatchEntry::strategy)
        .def_readwrite("chunk", &DispatchEntry::chunk)
        .def_readwrite("blocks_per_cu", &DispatchEntry::blocks_per_cu)
        .def_readonly("ms", &DispatchEntry::ms)
        .def_readonly("ms_default", &DispatchEntry::ms_default)
        .def("__repr__", [](const DispatchEntry& e) {
            return "DispatchEntry(cls=" + std::to_string(e.cls) + ", strategy=" + std::string(to_string(e.strategy)) +
                   ", chunk=" + std::to_string(e.chunk) + ", blocks_per_cu=" + std::to_string(e.blocks_per_cu) + ")";
        });

    py::class_<DispatchTable>(m, "DispatchTable")
        .def(py::init<>())
        .def(py::init([](std::vector<DispatchEntry> v) { return DispatchTable{std::move(v)}; }), py::arg("entries"))
        .def_readwrite("entries", &DispatchTable::entries)
        .def_static("from_json", &DispatchTable::from_json, py::arg("path"))
        .def("to_json", &DispatchTable::to_json, py::arg("path"))
        .def_static("parse_json", &DispatchTable::parse_json, py::arg("text"))
        .def("dump_json", &DispatchTable::dump_json)
        .def("__len__", [](const DispatchTable& t) { return t.entries.size(); });

    py::class_<JKOptions>(m, "JKOptions")
        .def(py::init([](Real thr, bool dw, bool mixed, Real tau, Real pcut, bool persistent) {
                 return JKOptions{thr, dw, mixed, tau, pcut, persistent};
             }),
             py::arg("schwarz_threshold") = 0.0, py::arg("density_weighted") = false,
             py::arg("mixed_precision") = false, py::arg("fp32_threshold") = 1e-7, py::arg("fp32_prim_cut") = 1e-7,
             py::arg("persistent") = false)
        .def_readwrite("schwarz_threshold", &JKOptions::schwarz_threshold)
        .def_readwrite("density_weighted", &JKOptions::density_weighted)
        .def_readwrite("mixed_precision", &JKOptions::mixed_precision)
        .def_readwrite("fp32_threshold", &JKOptions::fp32_threshold)
        .def_readwrite("fp32_prim_cut", &JKOptions::fp32_prim_cut)
        .def_readwrite("persistent", &JKOptions::persistent);

    py::class_<Engine>(m, "Engine")
        .def(py::init([](std::shared_ptr<BasisSet> b, DispatchConfig c) { return std::make_unique<Engine>(*b, c); }),
             py::arg("basis"), py::arg("config") = DispatchConfig{}, py::keep_alive<1, 2>())
        .def("max_angular_momentum", &Engine::max_angular_momentum)
        .def("gpu_available", &Engine::gpu_available)
        .def("compute_overlap_matrix",
             [](Engine& e, BackendHint h) {
                 std::vector<Real> r;
                 e.compute_overlap_matrix(r, h);
                 return to_matrix(r, e.basis().n_basis_functions());
             },
             py::arg("hint") = BackendHint::Auto)
        .def("compute_kinetic_matrix",
             [](Engine& e, BackendHint h) {
                 std::vector<Real> r;
                 e.compute_kinetic_matrix(r, h);
                 return to_matrix(r, e.basis().n_basis_functions());
             },
             py::arg("hint") = BackendHint::Auto)
        .def("compute_nuclear_matrix",
             [](Engine& e, const std::vector<data::Atom>& atoms, BackendHint h) {
                 std::vector<Real> r;
                 e.compute_nuclear_matrix(charges_of(atoms), r, h);
                 return to_matrix(r, e.basis().n_basis_functions());
             },
             py::arg("atoms"), py::arg("hint") = BackendHint::Auto)
        .def("compute_core_hamiltonian",
             [](Engine& e, const std::vector<data::Atom>& atoms, BackendHint h) {
                 std::vector<Real> r;
                 e.compute_core_hamiltonian(charges_of(atoms), r, h);
                 return to_matrix(r, e.basis().n_basis_functions());
             },
             py::arg("atoms"), py::arg("hint") = BackendHint::Auto)
        .def("compute_1e",
             [](Engine& e, const OneElectronOperator& op, BackendHint h) {
                 std::vector<Real> r;
                 e.compute_1e(op, r, h);
                 return to_matrix(r, e.basis().n_basis_functions());
             },
             py::arg("op"), py::arg("hint") = BackendHint::Auto)
        .def("compute_and_consume",
             [](Engine& e, const Operator& op, FockBuilder& fb, BackendHint h) { e.compute_and_consume(op, fb, h); },
             py::arg("op"), py::arg("consumer"), py::arg("hint") = BackendHint::Auto)
        .def("compute_and_consume",
             [](Engine& e, const Operator& op, GpuFockBuilder& fb, BackendHint h) { e.compute_and_consume(op, fb, h); },
             py::arg("op"), py::arg("consumer"), py::arg("hint") = BackendHint::Auto)
        .def("compute_and_consume",
             [](Engine& e, const Operator& op, MixedPrecisionFockBuilder& fb, BackendHint h) {
                 e.compute_and_consume(op, fb, h);
             },
             py::arg("op"), py::arg("consumer"), py::arg("hint") = BackendHint::Auto)
        .def("compute_and_consume_screened",
             [](Engine& e, const Operator& op, MixedPrecisionFockBuilder& fb, const ScreeningOptions& o) {
                 e.compute_and_consume(op, fb, o);
             },
             py::arg("op"), py::arg("consumer"), py::arg("options"))
        .def("compute_and_consume_parallel",
             [](Engine& e, const Operator& op, FockBuilder& fb, int n, BackendHint h) {
                 e.compute_and_consume_parallel(op, fb, n, h);
             },
             py::arg("op"), py::arg("consumer"), py::arg("n_threads") = 0, py::arg("hint") = BackendHint::Auto)
        .def("compute_and_consume_screened",
             [](Engine& e, const Operator& op, FockBuilder& fb, const ScreeningOptions& o) {
                 e.compute_and_consume(op, fb, o);
             },
             py::arg("op"), py::arg("consumer"), py::arg("options"))
        .def("compute_and_consume_screened",
             [](Engine& e, const Operator& op, GpuFockBuilder& fb, const ScreeningOptions& o) {
                 e.compute_and_consume(op, fb, o);
             },
             py::arg("op"), py::arg("consumer"), py::arg("options"))
        .def("compute_and_consume_screened_parallel",
             [](Engine& e, const Operator& op, FockBuilder& fb, const ScreeningOptions& o, int n) {
                 e.compute_and_consume_screened_parallel(op, fb, o, n);
             },
             py::arg("op"), py::arg("consumer"), py::arg("options"), py::arg("n_threads") = 0)
        .def("compute_eri_batch_device",
             [](Engine& e, py::array_t<int, py::array::c_style | py::array::forcecast> q, std::uintptr_t stream) {
                 if (q.ndim() != 2 || q.shape(1) != 4) throw std::invalid_argument("quartets must be (n, 4)");
                 std::vector<std::array<int, 4>> v(q.shape(0));
                 std::memcpy(v.data(), q.data(), sizeof(int) * 4 * v.size());
                 return e.compute_eri_batch_device(v, reinterpret_cast<hipStream_t>(stream));
             },
             py::arg("quartets"), py::arg("stream") = std::uintptr_t(0))
        .def("compute_shell_set_quartets",
             [](Engine& e, const Operator& op, FockBuilder& fb, py::object which) {
                 with_shell_set_quartets(e, which, [&](std::span<const ShellSetQuartet> q) { e.compute(op, q, fb); });
             },
             py::arg("op"), py::arg("consumer"), py::arg("quartets") = py::none())
        .def("compute_shell_set_quartets",
             [](Engine& e, const Operator& op, GpuFockBuilder& fb, py::object which) {
                 with_shell_set_quartets(e, which, [&](std::span<const ShellSetQuartet> q) { e.compute(op, q, fb); });
             },
             py::arg("op"), py::arg("consumer"), py::arg("quartets") = py::none())
        .def("compute_shell_set_quartets",
             [](Engine& e, const Operator& op, py::function fn, py::object which) {
                 PyConsumer c{fn};
                 with_shell_set_quartets(e, which, [&](std::span<const ShellSetQuartet> q) {
                     e.compute_and_consume(op, q, c);
                 });
             },
             py::arg("op"), py::arg("consumer"), py::arg("quartets") = py::none())
        .def("compute_shell_set_quartet",
             [](Engine& e, const Operator& op, GpuFockBuilder& fb, int a, int b, int c, int d) {
                 with_shell_set_quartets(e, py::make_tuple(py::make_tuple(a, b, c, d)),
                                         [&](std::span<const ShellSetQuartet> q) { e.compute(op, q[0], fb); });
             },
             py::arg("op"), py::arg("consumer"), py::arg("a"), py::arg("b"), py::arg("c"), py::arg("d"))
        .def("compute_shell_set_quartet",
             [](Engine& e, const Operator& op, FockBuilder& fb, int a, int b, int c, int d) {
                 with_shell_set_quartets(e, py::make_tuple(py::make_tuple(a, b, c, d)),
                                         [&](std::span<const ShellSetQuartet> q) { e.compute(op, q[0], fb); });
             },
             py::arg("op"), py::arg("consumer"), py::arg("a"), py::arg("b"), py::arg("c"), py::arg("d"))
        .def("compute_and_consume",
             [](Engine& e, const Operator& op, py::function fn, py::object options) {
                 PyConsumer c{fn};
                 if (options.is_none()) e.compute_and_consume(op, c);
                 else e.compute_and_consume(op, c, options.cast<ScreeningOptions>());
             },
             py::arg("op"), py::arg("consumer"), py::arg("options") = py::none())
        .def("compute_quartet",
             [](Engine& e, const Operator& op, int i, int j, int k, int l) {
                 TwoElectronBuffer<0> buf;
                 e.compute(op, e.basis().shell(i), e.basis().shell(j), e.basis().shell(k), e.basis().shell(l), buf);
                 py::array_t<Real> a({buf.na(), buf.nb(), buf.nc(), buf.nd()});
                 std::copy(buf.data().begin(), buf.data().end(), a.mutable_data());
                 return a;
             },
             py::arg("op"), py::arg("i"), py::arg("j"), py::arg("k"), py::arg("l"))
        .def("compute_eri_blocks",
             [](Engine& e, py::array_t<int, py::array::c_style | py::array::forcecast> q) {
                 if (q.ndim() != 2 || q.shape(1) != 4) throw std::invalid_argument("quartets must be (n, 4)");
                 std::vector<std::array<int, 4>> v(q.shape(0));
                 std::memcpy(v.data(), q.data(), sizeof(int) * 4 * v.size());
                 std::vector<Real> out;
                 std::vector<Index> offs;
                 e.compute_eri_blocks(v, out, offs);
                 return py::make_tuple(py::array_t<Real>(out.size(), out.data()),
                                       py::array_t<Index>(offs.size(), offs.data()));
             },
             py::arg("quartets"))
        .def("compute_eri_tensor",
             [](Engine& e) {
                 auto t = e.compute_eri_tensor();
                 const Size n = e.basis().n_basis_functions();
                 py::array_t<Real> a({n, n, n, n});
                 std::copy(t.begin(), t.end(), a.mutable_data());
                 return a;
             })
        .def("precompute_schwarz_bounds", &Engine::precompute_schwarz_bounds)
        .def("schwarz_bounds_precomputed", &Engine::schwarz_bounds_precomputed)
        .def("schwarz_bounds", &Engine::schwarz_bounds, py::return_value_policy::reference_internal)
        .def("set_schwarz_bounds",
             [](Engine& e, py::array_t<Real, py::array::c_style | py::array::forcecast> q) {
                 e.set_schwarz_bounds(std::vector<Real>(q.data(), q.data() + q.size()));
             })
        .def("set_density_matrix",
             [](Engine& e, py::array D, Size nbf) {
                 auto a = square(D, nbf, "set_density_matrix");
                 e.set_density_matrix(a.data(), nbf);
             },
             py::arg("D"), py::arg("nbf"))
        .def("density_matrix_set", &Engine::density_matrix_set)
        .def("prepare_worklist",
             [](Engine& e, const JKOptions& o, py::object D) {
                 if (D.is_none()) return info_dict(e.prepare_worklist(o, nullptr));
                 auto a = square(D.cast<py::array>(), e.basis().n_basis_functions(), "prepare_worklist");
                 return info_dict(e.prepare_worklist(o, a.data()));
             },
             py::arg("options") = JKOptions{}, py::arg("D") = py::none())
        .def("build_jk",
             [](Engine& e, py::array D, const JKOptions& o) {
                 const Size n = e.basis().n_basis_functions();
                 auto a = square(D, n, "build_jk");
                 std::vector<Real> J(n * n), K(n * n);
                 e.build_jk_host(o, a.data(), J.data(), K.data());
                 return py::make_tuple(to_matrix(J, n), to_matrix(K, n));
             },
             py::arg("D"), py::arg("options") = JKOptions{})
        .def("build_jk_device",
             [](Engine& e, std::uintptr_t dD, std::uintptr_t dJ, std::uintptr_t dK, const JKOptions& o,
                std::uintptr_t stream, py::object Dh) {
                 const Real* dh = nullptr;
                 py::array_t<Real, py::array::c_style> keep;
                 if (!Dh.is_none()) {
                     keep = square(Dh.cast<py::array>(), e.basis().n_basis_functions(), "build_jk_device");
                     dh = keep.data();
                 }
                 e.build_jk(o, reinterpret_cast<const Real*>(dD), reinterpret_cast<Real*>(dJ),
                            reinterpret_cast<Real*>(dK), reinterpret_cast<hipStream_t>(stream), dh);
             },
             py::arg("d_D"), py::arg("d_J"), py::arg("d_K"), py::arg("options") = JKOptions{},
             py::arg("stream") = 0, py::arg("D_host") = py::none())
        .def("capture_jk_graph",
             [](Engine& e, std::uintptr_t dD, std::uintptr_t dJ, std::uintptr_t dK, const JKOptions& o,
                std::uintptr_t stream) {
                 e.capture_jk_graph(o, reinterpret_cast<const Real*>(dD), reinterpret_cast<Real*>(dJ),
                                    reinterpret_cast<Real*>(dK), reinterpret_cast<hipStream_t>(stream));
             },
             py::arg("d_D"), py::arg("d_J"), py::arg("d_K"), py::arg("options"), py::arg("stream"))
        .def("launch_jk_graph", [](Engine& e, std::uintptr_t s) { e.launch_jk_graph(reinterpret_cast<hipStream_t>(s)); },
             py::arg("stream"))
        .def("compute_jk_gradient",
             [](Engine& e, py::array D, Real x, int n_atoms) {
                 auto a = square(D, e.basis().n_basis_functions(), "compute_jk_gradient");
                 auto g = e.compute_jk_gradient(a.data(), x, n_atoms);
                 py::array_t<Real> r({g.size() / 3, Size(3)});
                 std::copy(g.begin(), g.end(), r.mutable_data());
                 return r;
             },
             py::arg("D"), py::arg("exchange_scale") = 1.0, py::arg("n_atoms") = -1)
        .def("autotune",
             [](Engine& e, std::uintptr_t dD, const JKOptions& o, std::uintptr_t stream, int reps, py::object Dh) {
                 const Real* dh = nullptr;
                 py::array_t<Real, py::array::c_style> keep;
                 if (!Dh.is_none()) {
                     keep = square(Dh.cast<py::array>(), e.basis().n_basis_functions(), "autotune");
                     dh = keep.data();
                 }
                 return e.autotune(o, reinterpret_cast<const Real*>(dD), reinterpret_cast<hipStream_t>(stream), reps, dh);
             },
             py::arg("d_D"), py::arg("options") = JKOptions{}, py::arg("stream") = std::uintptr_t(0),
             py::arg("reps") = 3, py::arg("D_host") = py::none())
        .def("dispatch_table", &Engine::dispatch_table)
        .def("set_dispatch_table", &Engine::set_dispatch_table, py::arg("table"))
        .def("worklist_info", [](const Engine& e) { return info_dict(e.worklist_info()); })
        .def("device_basis_handle", [](Engine& e) { return reinterpret_cast<std::uintptr_t>(e.device_basis()); })
        .def("__repr__", [](const Engine& e) {
            return "Engine(n_shells=" + std::to_string(e.basis().n_shells()) +
                   ", gpu_available=" + (e.gpu_available() ? std::string("True") : std::string("False")) + ")";
        });

    py::class_<BatchEngine>(m, "BatchEngine")
        .def(py::init([](const std::vector<std::shared_ptr<BasisSet>>& bases) {
                 std::vector<std::shared_ptr<const BasisSet>> v(bases.begin(), bases.end());
                 return std::make_unique<BatchEngine>(std::move(v));
             }),
             py::arg("bases"))
        .def("n_molecules", &BatchEngine::n_molecules)
        .def("n_basis_functions", [](const BatchEngine& e) { return e.n_basis_functions(); })
        .def("matrix_offsets", [](const BatchEngine& e) { return e.matrix_offsets(); })
        .def("matrix_size", &BatchEngine::matrix_size)
        .def("precompute_schwarz_bounds", &BatchEngine::precompute_schwarz_bounds)
        .def("prepare_worklist",
             [](BatchEngine& e, const JKOptions& o, py::object D) {
                 if (D.is_none()) return info_dict(e.prepare_worklist(o, nullptr));
                 auto a = py::array_t<Real, py::array::c_style | py::array::forcecast>(D.cast<py::array>());
                 if (Size(a.size()) != e.matrix_size())
                     throw InvalidArgumentException("BatchEngine.prepare_worklist: D has the wrong length");
                 return info_dict(e.prepare_worklist(o, a.data()));
             },
             py::arg("options") = JKOptions{}, py::arg("D") = py::none())
        .def("worklist_info", [](const BatchEngine& e) { return info_dict(e.worklist_info()); })
        .def("build_jk",
             [](BatchEngine& e, py::array D, const JKOptions& o) {
                 auto a = py::array_t<Real, py::array::c_style | py::array::forcecast>(D);
                 if (Size(a.size()) != e.matrix_size())
                     throw InvalidArgumentException("BatchEngine.build_jk: D has the wrong length");
                 std::vector<Real> J(e.matrix_size()), K(e.matrix_size());
                 e.build_jk_host(o, a.data(), J.data(), K.data());
                 return py::make_tuple(py::array_t<Real>(J.size(), J.data()), py::array_t<Real>(K.size(), K.data()));
             },
             py::arg("D"), py::arg("options") = JKOptions{})
        .def("build_jk_device",
             [](BatchEngine& e, std::uintptr_t dD, std::uintptr_t dJ, std::uintptr_t dK, const JKOptions& o,
                std::uintptr_t stream, py::object Dh) {
                 const Real* dh = nullptr;
                 py::array_t<Real, py::array::c_style | py::array::forcecast> keep;
                 if (!Dh.is_none()) {
                     keep = py::array_t<Real, py::array::c_style | py::array::forcecast>(Dh.cast<py::array>());
                     if (Size(keep.size()) != e.matrix_size())
                         throw InvalidArgumentException("BatchEngine.build_jk_device: D has the wrong length");
                     dh = keep.data();
                 }
                 e.build_jk(o, reinterpret_cast<const Real*>(dD), reinterpret_cast<Real*>(dJ),
                            reinterpret_cast<Real*>(dK), reinterpret_cast<hipStream_t>(stream), dh);
             },
             py::arg("d_D"), py::arg("d_J"), py::arg("d_K"), py::arg("options") = JKOptions{},
             py::arg("stream") = std::uintptr_t(0), py::arg("D_host") = py::none())
        .def("capture_jk_graph",
             [](BatchEngine& e, std::uintptr_t dD, std::uintptr_t dJ, std::uintptr_t dK, const JKOptions& o,
                std::uintptr_t stream) {
                 e.capture_jk_graph(o, reinterpret_cast<const Real*>(dD), reinterpret_cast<Real*>(dJ),
                                    reinterpret_cast<Real*>(dK), reinterpret_cast<hipStream_t>(stream));
             },
             py::arg("d_D"), py::arg("d_J"), py::arg("d_K"), py::arg("options"), py::arg("stream"))
        .def("launch_jk_graph",
             [](BatchEngine& e, std::uintptr_t s) { e.launch_jk_graph(reinterpret_cast<hipStream_t>(s)); },
             py::arg("stream"))
        .def("autotune",
             [](BatchEngine& e, std::uintptr_t dD, const JKOptions& o, std::uintptr_t stream, int reps, py::object Dh) {
                 const Real* dh = nullptr;
                 py::array_t<Real, py::array::c_style | py::array::forcecast> keep;
                 if (!Dh.is_none()) {
                     keep = py::array_t<Real, py::array::c_style | py::array::forcecast>(Dh.cast<py::array>());
                     if (Size(keep.size()) != e.matrix_size())
                         throw InvalidArgumentException("BatchEngine.autotune: D has the wrong length");
                     dh = keep.data();
                 }
                 return e.autotune(o, reinterpret_cast<const Real*>(dD), reinterpret_cast<hipStream_t>(stream), reps, dh);
             },
             py::arg("d_D"), py::arg("options") = JKOptions{}, py::arg("stream") = std::uintptr_t(0),
             py::arg("reps") = 3, py::arg("D_host") = py::none())
        .def("dispatch_table", &BatchEngine::dispatch_table)
        .def("set_dispatch_table", &BatchEngine::set_dispatch_table, py::arg("table"))
        .def("device_basis_handle", [](BatchEngine& e) { return reinterpret_cast<std::uintptr_t>(e.device_handle()); })
        .def("__repr__", [](const BatchEngine& e) {
            return "BatchEngine(n_molecules=" + std::to_string(e.n_molecules()) + ")";
        });

    // ---- density fitting ---------------------------------------------------------------
    py::enum_<FittingType>(m, "FittingType")
        .value("RI", FittingType::RI).value("JKFIT", FittingType::JKFIT)
        .value("JFIT", FittingType::JFIT).value("RIFIT", FittingType::RIFIT);
    m.def("fitting_type_to_string", [](FittingType t) { return std::string(fitting_type_to_string(t)); });

    py::class_<AuxiliaryBasisSet, std::shared_ptr<AuxiliaryBasisSet>>(m, "AuxiliaryBasisSet")
        .def(py::init([](std::vector<Shell> shells, FittingType t, const std::string& name) {
                 return std::make_shared<AuxiliaryBasisSet>(std::move(shells), t, name);
             }),
             py::arg("shells"), py::arg("fitting_type") = FittingType::JKFIT, py::arg("name") = "")
        .def("n_shells", &AuxiliaryBasisSet::n_shells)
        .def("n_functions", &AuxiliaryBasisSet::n_functions)
        .def("n_primitives", &AuxiliaryBasisSet::n_primitives)
        .def("max_angular_momentum", &AuxiliaryBasisSet::max_angular_momentum)
        .def("shell", &AuxiliaryBasisSet::shell, py::return_value_policy::reference_internal)
        .def("shells", [](const AuxiliaryBasisSet& b) { return std::vector<Shell>(b.shells().begin(), b.shells().end()); })
        .def("shell_to_function", &AuxiliaryBasisSet::shell_to_function)
        .def("function_to_shell", &AuxiliaryBasisSet::function_to_shell)
        .def("fitting_type", &AuxiliaryBasisSet::fitting_type)
        .def("name", &AuxiliaryBasisSet::name)
        .def("__repr__", [](const AuxiliaryBasisSet& b) {
            return "AuxiliaryBasisSet(name='" + b.name() + "', n_shells=" + std::to_string(b.n_shells()) +
                   ", n_functions=" + std::to_string(b.n_functions()) + ")";
        });

    auto atom_lists = [](const std::vector<int>& Z, const std::vector<std::vector<Real>>& xyz) {
        std::vector<std::array<Real, 3>> c;
        for (const auto& v : xyz) {
            if (v.size() != 3) throw std::invalid_argument("centers must have 3 components each");
            c.push_back({v[0], v[1], v[2]});
        }
        return c;
    };
    m.def("load_auxiliary_basis",
          [atom_lists](const std::string& name, const std::vector<int>& Z, const std::vector<std::vector<Real>>& xyz) {
              const auto c = atom_lists(Z, xyz);
              return std::make_shared<AuxiliaryBasisSet>(load_auxiliary_basis(name, Z, c));
          },
          py::arg("basis_name"), py::arg("atomic_numbers"), py::arg("centers"));
    m.def("available_auxiliary_bases", &available_auxiliary_bases);
    m.def("is_auxiliary_basis_available",
          [](const std::string& name, const std::vector<int>& Z) { return is_auxiliary_basis_available(name, Z); },
          py::arg("basis_name"), py::arg("atomic_numbers"));

    py::enum_<DFMetricType>(m, "DFMetricType")
        .value("Coulomb", DFMetricType::Coulomb)
        .value("Overlap", DFMetricType::Overlap)
        .value("AttenuatedCoulomb", DFMetricType::AttenuatedCoulomb);

    py::class_<DFFockBuilderConfig>(m, "DFFockBuilderConfig")
        .def(py::init<>())
        .def_readwrite("exchange_fraction", &DFFockBuilderConfig::exchange_fraction)
        .def_readwrite("compute_coulomb", &DFFockBuilderConfig::compute_coulomb)
        .def_readwrite("compute_exchange", &DFFockBuilderConfig::compute_exchange)
        .def_readwrite("use_symmetry", &DFFockBuilderConfig::use_symmetry)
        .def_readwrite("incremental", &DFFockBuilderConfig::incremental)
        .def_readwrite("memory_limit_mb", &DFFockBuilderConfig::memory_limit_mb)
        .def_readwrite("n_threads", &DFFockBuilderConfig::n_threads)
        .def_readwrite("conditioning_threshold", &DFFockBuilderConfig::conditioning_threshold)
        .def_readwrite("conditioning_hard_limit", &DFFockBuilderConfig::conditioning_hard_limit)
        .def_readwrite("metric_type", &DFFockBuilderConfig::metric_type)
        .def_readwrite("range_separation_omega", &DFFockBuilderConfig::range_separation_omega);

    py::class_<DFFockBuilder>(m, "DFFockBuilder")
        .def(py::init([](std::shared_ptr<BasisSet> orb, std::shared_ptr<AuxiliaryBasisSet> aux, DFFockBuilderConfig c) {
                 return std::make_unique<DFFockBuilder>(*orb, *aux, c);
             }),
             py::arg("orbital"), py::arg("auxiliary"), py::arg("config") = DFFockBuilderConfig{}, py::keep_alive<1, 2>(),
             py::keep_alive<1, 3>())
        .def("set_density",
             [](DFFockBuilder& f, py::array D) {
                 auto a = square(D, f.n_orb(), "DFFockBuilder.set_density");
                 f.set_density(std::span<const Real>(a.data(), a.size()));
             },
             py::arg("D"))
        .def("set_density_unrestricted",
             [](DFFockBuilder& f, py::array Da, py::array Db) {
                 auto a = square(Da, f.n_orb(), "DFFockBuilder.set_density_unrestricted");
                 auto b = square(Db, f.n_orb(), "DFFockBuilder.set_density_unrestricted");
                 f.set_density_unrestricted(std::span<const Real>(a.data(), a.size()),
                                            std::span<const Real>(b.data(), b.size()));
             },
             py::arg("D_alpha"), py::arg("D_beta"))
        .def("initialize", &DFFockBuilder::initialize)
        .def("is_initialized", &DFFockBuilder::is_initialized)
        .def("compute", [](DFFockBuilder& f) { return to_matrix(f.compute(), f.n_orb()); })
        .def("compute_coulomb", [](DFFockBuilder& f) { return to_matrix(f.compute_coulomb(), f.n_orb()); })
        .def("compute_exchange", [](DFFockBuilder& f) { return to_matrix(f.compute_exchange(), f.n_orb()); })
        .def("compute_accumulate",
             [](DFFockBuilder& f, py::array_t<Real, py::array::c_style> F) {
                 if (F.size() != py::ssize_t(f.n_orb() * f.n_orb()))
                     throw InvalidArgumentException("Fock matrix size mismatch");
                 f.compute_accumulate(std::span<Real>(F.mutable_data(), F.size()));
             },
             py::arg("F"))
        .def("coulomb_matrix", [](const DFFockBuilder& f) {
            return to_matrix(std::vector<Real>(f.coulomb_matrix().begin(), f.coulomb_matrix().end()), f.n_orb());
        })
        .def("exchange_matrix", [](const DFFockBuilder& f) {
            return to_matrix(std::vector<Real>(f.exchange_matrix().begin(), f.exchange_matrix().end()), f.n_orb());
        })
        .def("exchange_matrix_alpha", [](const DFFockBuilder& f) {
            return to_matrix(std::vector<Real>(f.exchange_matrix_alpha().begin(), f.exchange_matrix_alpha().end()),
                             f.is_uhf() ? f.n_orb() : 0);
        })
        .def("exchange_matrix_beta", [](const DFFockBuilder& f) {
            return to_matrix(std::vector<Real>(f.exchange_matrix_beta().begin(), f.exchange_matrix_beta().end()),
                             f.is_uhf() ? f.n_orb() : 0);
        })
        .def("is_uhf", &DFFockBuilder::is_uhf)
        .def("fock_matrix",
             [](const DFFockBuilder& f, py::object H) {
                 if (H.is_none()) return to_matrix(f.fock_matrix(), f.n_orb());
                 auto h = py::array_t<Real, py::array::c_style | py::array::forcecast>::ensure(H);
                 if (h.size() != py::ssize_t(f.n_orb() * f.n_orb()))
                     throw InvalidArgumentException("H_core size mismatch");
                 return to_matrix(f.fock_matrix(std::span<const Real>(h.data(), h.size())), f.n_orb());
             },
             py::arg("H_core") = py::none())
        .def("n_orb", &DFFockBuilder::n_orb)
        .def("n_aux", &DFFockBuilder::n_aux)
        .def("b_tensor_memory", &DFFockBuilder::b_tensor_memory)
        .def("config", &DFFockBuilder::config)
        .def("metric_condition_estimate", &DFFockBuilder::metric_condition_estimate)
        .def("metric_ill_conditioned", &DFFockBuilder::metric_ill_conditioned)
        .def("device_handle", [](const DFFockBuilder& f) { return reinterpret_cast<std::uintptr_t>(f.device_handle()); });

    m.def("make_df_fock_builder",
          [](std::shared_ptr<BasisSet> orb, const std::string& name) {
              return consumers::make_df_fock_builder(*orb, name);
          },
          py::arg("orbital"), py::arg("aux_basis_name") = "", py::keep_alive<0, 1>());
    m.def("make_df_fock_builder",
          [atom_lists](std::shared_ptr<BasisSet> orb, const std::vector<int>& Z, const std::vector<std::vector<Real>>& xyz,
                       const std::string& name, DFFockBuilderConfig c) {
              const auto cc = atom_lists(Z, xyz);
              return consumers::make_df_fock_builder(*orb, Z, cc, name, c);
          },
          py::arg("orbital"), py::arg("atomic_numbers"), py::arg("centers"), py::arg("aux_basis_name") = "",
          py::arg("config") = DFFockBuilderConfig{}, py::keep_alive<0, 1>());
}
