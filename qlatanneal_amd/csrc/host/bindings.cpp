// pybind11 module _qanneal: the Python surface of the reference's python/bindings.cpp (same class,
// method and argument names), over this repository's C++ classes. Annealers default to
// backend="hip"; "hip_fp64" selects the FP64 validation replay; "cpu" constructs but cannot run
// a sweep (there is no CPU engine in this build).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <stdexcept>

#include "qanneal/core.hpp"

namespace py = pybind11;
using namespace qanneal;

namespace {

using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;

std::vector<double> vec1(const darr &a) {
    auto b = a.request();
    if (b.ndim != 1) throw std::invalid_argument("Expected 1D array.");
    const auto *p = static_cast<const double *>(b.ptr);
    return std::vector<double>(p, p + b.shape[0]);
}

std::vector<double> mat2(const darr &a, std::size_t &n) {
    auto b = a.request();
    if (b.ndim != 2) throw std::invalid_argument("Expected 2D array.");
    if (b.shape[0] != b.shape[1]) throw std::invalid_argument("Expected square matrix.");
    n = static_cast<std::size_t>(b.shape[0]);
    const auto *p = static_cast<const double *>(b.ptr);
    return std::vector<double>(p, p + n * n);
}

std::vector<int8_t> spins_of(const py::sequence &seq) {
    std::vector<int8_t> s;
    s.reserve(seq.size());
    for (auto v : seq) s.push_back(static_cast<int8_t>(py::cast<int>(v)));
    return s;
}

std::vector<std::size_t> term_vars(const py::handle &key) {
    std::vector<std::size_t> v;
    if (py::isinstance<py::int_>(key)) {
        v.push_back(py::cast<std::size_t>(key));
        return v;
    }
    if (py::isinstance<py::iterable>(key)) {
        for (auto x : py::reinterpret_borrow<py::iterable>(key)) v.push_back(py::cast<std::size_t>(x));
        return v;
    }
    throw std::invalid_argument("HigherOrderIsing term key must be an int or an iterable of ints.");
}

std::vector<HigherOrderIsing::RawTerm> raw_terms(const py::dict &d, std::size_t &n, bool infer) {
    std::vector<HigherOrderIsing::RawTerm> raw;
    std::size_t mx = 0;
    bool any = false;
    for (auto item : d) {
        auto v = term_vars(item.first);
        for (std::size_t x : v) {
            mx = std::max(mx, x);
            any = true;
        }
        raw.push_back({std::move(v), py::cast<double>(item.second)});
    }
    if (infer) n = any ? mx + 1 : 0;
    return raw;
}

std::shared_ptr<Backend> backend_for(std::shared_ptr<Hamiltonian> ham, const std::string &name) {
    return make_backend(parse_backend(name), std::move(ham));
}

}  // namespace

PYBIND11_MODULE(_qanneal, m) {
    m.doc() = "qanneal annealing-sweep API on the gfx950 HIP backend";
    m.def("version_string", &version_string);
    m.def("hip_context_count", &hip_context_count,
          "HIP contexts (one stream each) that device engines are spread over, round-robin");
    m.attr("version_major") = version_major;
    m.attr("version_minor") = version_minor;
    m.attr("version_patch") = version_patch;

    py::class_<State>(m, "State")
        .def(py::init<std::size_t>())
        .def_property("spins", [](const State &s) { return s.spins; },
                      [](State &s, const std::vector<int8_t> &v) { s.spins = v; })
        .def("size", &State::size);

    py::class_<Hamiltonian, std::shared_ptr<Hamiltonian>>(m, "Hamiltonian");

    py::class_<DenseIsing, Hamiltonian, std::shared_ptr<DenseIsing>>(m, "DenseIsing")
        .def(py::init([](darr h, darr J, double c) {
                 std::vector<double> hv = vec1(h);
                 std::size_t n = 0;
                 std::vector<double> Jv = mat2(J, n);
                 if (hv.size() != n) throw std::invalid_argument("h vector length mismatch.");
                 return DenseIsing(std::move(hv), std::move(Jv), n, c);
             }),
             py::arg("h"), py::arg("J"), py::arg("c") = 0.0)
        .def("size", &DenseIsing::size)
        .def("couplings", [](const DenseIsing &d) {
            const std::size_t n = d.size();
            py::array_t<double> out({n, n});
            std::memcpy(out.mutable_data(), d.J().data(), n * n * sizeof(double));
            return out;
        })
        .def("fields", [](const DenseIsing &d) {
            py::array_t<double> out(static_cast<py::ssize_t>(d.size()));
            std::memcpy(out.mutable_data(), d.h().data(), d.size() * sizeof(double));
            return out;
        })
        .def("constant", &DenseIsing::constant)
        .def("energy", [](const DenseIsing &d, const py::sequence &s) {
            auto v = spins_of(s);
            return d.energy(v.data(), v.size());
        })
        .def("delta_energy", [](const DenseIsing &d, const py::sequence &s, std::size_t f) {
            auto v = spins_of(s);
            return d.delta_energy(v.data(), v.size(), f);
        });

    py::class_<SparseEdge>(m, "SparseEdge")
        .def(py::init<std::size_t, std::size_t, double>())
        .def_readwrite("i", &SparseEdge::i)
        .def_readwrite("j", &SparseEdge::j)
        .def_readwrite("value", &SparseEdge::value);

    py::class_<SparseIsing, Hamiltonian, std::shared_ptr<SparseIsing>>(m, "SparseIsing")
        .def(py::init([](darr h, const std::vector<SparseEdge> &edges, std::size_t n, double c) {
                 std::vector<double> hv = vec1(h);
                 if (hv.size() != n) throw std::invalid_argument("h vector length mismatch.");
                 return SparseIsing(std::move(hv), edges, n, c);
             }),
             py::arg("h"), py::arg("edges"), py::arg("n"), py::arg("c") = 0.0)
        .def("size", &SparseIsing::size)
        .def("energy", [](const SparseIsing &d, const py::sequence &s) {
            auto v = spins_of(s);
            return d.energy(v.data(), v.size());
        })
        .def("delta_energy", [](const SparseIsing &d, const py::sequence &s, std::size_t f) {
            auto v = spins_of(s);
            return d.delta_energy(v.data(), v.size(), f);
        });

    py::class_<QUBO>(m, "QUBO")
        .def(py::init([](darr Q) {
            std::size_t n = 0;
            auto q = mat2(Q, n);
            return QUBO(std::move(q), n);
        }))
        .def(py::init([](py::object bqm) {
            // dimod BinaryQuadraticModel-like objects (bindings.cpp:195-247 of the reference)
            if (!py::hasattr(bqm, "variables") || !py::hasattr(bqm, "linear") || !py::hasattr(bqm, "quadratic"))
                throw std::invalid_argument("QUBO expects a dimod BinaryQuadraticModel-like object.");
            if (py::hasattr(bqm, "vartype")) {
                const std::string vt = py::str(bqm.attr("vartype"));
                if (vt.find("BINARY") == std::string::npos) {
                    if (!py::hasattr(bqm, "to_binary"))
                        throw std::invalid_argument("QUBO expects a BINARY dimod BQM or a BQM with to_binary().");
                    py::object conv = bqm.attr("to_binary")();
                    bqm = py::isinstance<py::tuple>(conv) ? conv.cast<py::tuple>()[0] : conv;
                }
            }
            py::list vars = py::list(bqm.attr("variables"));
            const std::size_t n = vars.size();
            if (n == 0) throw std::invalid_argument("QUBO BQM must have at least one variable.");
            py::dict index;
            for (std::size_t i = 0; i < n; ++i) index[vars[i]] = py::int_(i);
            std::vector<double> q(n * n, 0.0);
            for (auto item : bqm.attr("linear").attr("items")()) {
                auto pr = item.cast<py::tuple>();
                const std::size_t i = py::cast<std::size_t>(index[pr[0]]);
                q[i * n + i] += py::cast<double>(pr[1]);
            }
            for (auto item : bqm.attr("quadratic").attr("items")()) {
                auto pr = item.cast<py::tuple>();
                auto key = pr[0].cast<py::tuple>();
                const std::size_t i = py::cast<std::size_t>(index[key[0]]);
                const std::size_t j = py::cast<std::size_t>(index[key[1]]);
                const double b = py::cast<double>(pr[1]);
                q[i * n + j] += 0.5 * b;
                q[j * n + i] += 0.5 * b;
            }
            return QUBO(std::move(q), n);
        }))
        .def(py::init([](const std::vector<std::tuple<std::size_t, std::size_t, double>> &entries, std::size_t n) {
                 std::vector<std::pair<std::pair<std::size_t, std::size_t>, double>> e;
                 for (const auto &t : entries) e.push_back({{std::get<0>(t), std::get<1>(t)}, std::get<2>(t)});
                 return QUBO(e, n);
             }),
             py::arg("entries"), py::arg("n"))
        .def(py::init([](py::dict entries, std::size_t n) {
                 std::vector<std::pair<std::pair<std::size_t, std::size_t>, double>> e;
                 for (auto item : entries) {
                     auto key = item.first.cast<py::tuple>();
                     if (key.size() != 2) throw std::invalid_argument("QUBO dict keys must be (i, j) tuples.");
                     e.push_back({{py::cast<std::size_t>(key[0]), py::cast<std::size_t>(key[1])},
                                  py::cast<double>(item.second)});
                 }
                 return QUBO(e, n);
             }),
             py::arg("entries"), py::arg("n"))
        .def("size", &QUBO::size)
        .def("to_ising", &QUBO::to_ising);

    py::class_<HigherOrderIsing, Hamiltonian, std::shared_ptr<HigherOrderIsing>>(m, "HigherOrderIsing")
        .def(py::init([](py::dict terms, std::size_t n) {
                 std::size_t ignored = n;
                 return HigherOrderIsing(n, raw_terms(terms, ignored, false));
             }),
             py::arg("terms"), py::arg("n"))
        .def(py::init([](py::dict terms) {
                 std::size_t n = 0;
                 auto raw = raw_terms(terms, n, true);
                 if (n == 0) throw std::invalid_argument("HigherOrderIsing needs at least one variable index.");
                 return HigherOrderIsing(n, raw);
             }),
             py::arg("terms"))
        .def(py::init([](const std::vector<std::pair<std::vector<std::size_t>, double>> &entries, std::size_t n) {
                 std::vector<HigherOrderIsing::RawTerm> raw(entries.begin(), entries.end());
                 return HigherOrderIsing(n, raw);
             }),
             py::arg("entries"), py::arg("n"))
        .def("size", &HigherOrderIsing::size)
        .def("max_degree", &HigherOrderIsing::max_degree)
        .def("num_terms", &HigherOrderIsing::num_terms)
        .def("constant", &HigherOrderIsing::constant)
        .def("energy", [](const HigherOrderIsing &d, const py::sequence &s) {
            auto v = spins_of(s);
            return d.energy(v.data(), v.size());
        })
        .def("delta_energy", [](const HigherOrderIsing &d, const py::sequence &s, std::size_t f) {
            auto v = spins_of(s);
            return d.delta_energy(v.data(), v.size(), f);
        });

    py::class_<AnnealSchedule>(m, "AnnealSchedule")
        .def(py::init<>())
        .def_readwrite("betas", &AnnealSchedule::betas)
        .def_static("linear", &AnnealSchedule::linear, py::arg("beta_start"), py::arg("beta_end"), py::arg("steps"))
        .def_static("from_betas", [](const std::vector<double> &betas) {
            if (betas.empty()) throw std::invalid_argument("Schedule must contain betas.");
            AnnealSchedule s;
            s.betas = betas;
            return s;
        });

    py::class_<Observer, std::shared_ptr<Observer>>(m, "Observer");
    py::class_<MetricsObserver, Observer, std::shared_ptr<MetricsObserver>>(m, "MetricsObserver")
        .def(py::init<>())
        .def_readonly("energy_trace", &MetricsObserver::energy_trace)
        .def_readonly("magnetization_trace", &MetricsObserver::magnetization_trace)
        .def("clear", &MetricsObserver::clear);
    py::class_<StateTraceObserver, Observer, std::shared_ptr<StateTraceObserver>>(m, "StateTraceObserver")
        .def(py::init<>())
        .def_readwrite("stride", &StateTraceObserver::stride)
        .def_readonly("step_trace", &StateTraceObserver::step_trace)
        .def_readonly("sweep_trace", &StateTraceObserver::sweep_trace)
        .def_readonly("beta_trace", &StateTraceObserver::beta_trace)
        .def_readonly("energy_trace", &StateTraceObserver::energy_trace)
        .def_readonly("state_trace", &StateTraceObserver::state_trace)
        .def("clear", &StateTraceObserver::clear);
    py::class_<SQAObserver, std::shared_ptr<SQAObserver>>(m, "SQAObserver");
    py::enum_<SQASweepPhase>(m, "SQASweepPhase")
        .value("SLICE", SQASweepPhase::Slice)
        .value("WORLDLINE", SQASweepPhase::Worldline)
        .value("CLUSTER", SQASweepPhase::Cluster);
    py::class_<SQAMetricsObserver, SQAObserver, std::shared_ptr<SQAMetricsObserver>>(m, "SQAMetricsObserver")
        .def(py::init<>())
        .def_readonly("energy_trace", &SQAMetricsObserver::energy_trace)
        .def_readonly("magnetization_trace", &SQAMetricsObserver::magnetization_trace)
        .def("clear", &SQAMetricsObserver::clear);
    py::class_<SQAStateTraceObserver, SQAObserver, std::shared_ptr<SQAStateTraceObserver>>(m, "SQAStateTraceObserver")
        .def(py::init<>())
        .def_readwrite("stride", &SQAStateTraceObserver::stride)
        .def_readonly("replicas", &SQAStateTraceObserver::replicas)
        .def_readonly("slices", &SQAStateTraceObserver::slices)
        .def_readonly("spins", &SQAStateTraceObserver::spins)
        .def_readonly("step_trace", &SQAStateTraceObserver::step_trace)
        .def_readonly("replica_trace", &SQAStateTraceObserver::replica_trace)
        .def_readonly("sweep_trace", &SQAStateTraceObserver::sweep_trace)
        .def_readonly("phase_trace", &SQAStateTraceObserver::phase_trace)
        .def_readonly("beta_trace", &SQAStateTraceObserver::beta_trace)
        .def_readonly("gamma_trace", &SQAStateTraceObserver::gamma_trace)
        .def_readonly("avg_energy_trace", &SQAStateTraceObserver::avg_energy_trace)
        .def_readonly("replica_energy_trace", &SQAStateTraceObserver::replica_energy_trace)
        .def_readonly("state_trace", &SQAStateTraceObserver::state_trace)
        .def("clear", &SQAStateTraceObserver::clear);

    py::class_<AnnealResult>(m, "AnnealResult")
        .def_readonly("best_state", &AnnealResult::best_state)
        .def_readonly("best_energy", &AnnealResult::best_energy)
        .def_readonly("energy_trace", &AnnealResult::energy_trace);

    py::class_<Annealer>(m, "Annealer")
        .def(py::init([](std::shared_ptr<Hamiltonian> ham, AnnealSchedule schedule, const std::string &backend) {
                 return Annealer(backend_for(std::move(ham), backend), std::move(schedule));
             }),
             py::arg("hamiltonian"), py::arg("schedule"), py::arg("backend") = "hip", py::keep_alive<1, 2>())
        .def("set_seed", &Annealer::set_seed)
        .def("run", [](Annealer &self, std::size_t spb, std::shared_ptr<Observer> obs) { return self.run(spb, obs.get()); },
             py::arg("sweeps_per_beta"), py::arg("observer") = nullptr, py::call_guard<py::gil_scoped_release>())
        .def("run_reads", &Annealer::run_reads, py::arg("sweeps_per_beta"), py::arg("reads"), py::arg("seed"),
             py::call_guard<py::gil_scoped_release>());

    py::class_<ReplicaResult>(m, "ReplicaResult")
        .def_readonly("best_state", &ReplicaResult::best_state)
        .def_readonly("best_energy", &ReplicaResult::best_energy)
        .def_readonly("energy_trace", &ReplicaResult::energy_trace)
        .def_readonly("magnetization_trace", &ReplicaResult::magnetization_trace);

    py::class_<MultiAnnealResult>(m, "MultiAnnealResult")
        .def_readonly("replicas", &MultiAnnealResult::replicas)
        .def_readonly("global_best_state", &MultiAnnealResult::global_best_state)
        .def_readonly("global_best_energy", &MultiAnnealResult::global_best_energy)
        .def_readonly("average_energy_trace", &MultiAnnealResult::average_energy_trace)
        .def_readonly("average_magnetization_trace", &MultiAnnealResult::average_magnetization_trace);

    py::class_<ReplicaAnnealer>(m, "ReplicaAnnealer")
        .def(py::init([](std::shared_ptr<Hamiltonian> ham, AnnealSchedule schedule, std::size_t replicas,
                         const std::string &backend) {
                 return ReplicaAnnealer(backend_for(std::move(ham), backend), std::move(schedule), replicas);
             }),
             py::arg("hamiltonian"), py::arg("schedule"), py::arg("replicas"), py::arg("backend") = "hip",
             py::keep_alive<1, 2>())
        .def("set_seed", &ReplicaAnnealer::set_seed)
        .def("run", &ReplicaAnnealer::run, py::arg("sweeps_per_beta"), py::call_guard<py::gil_scoped_release>());

    py::class_<ParallelTemperingResult>(m, "ParallelTemperingResult")
        .def_readonly("final_states", &ParallelTemperingResult::final_states)
        .def_readonly("final_energies", &ParallelTemperingResult::final_energies)
        .def_readonly("best_state", &ParallelTemperingResult::best_state)
        .def_readonly("best_energy", &ParallelTemperingResult::best_energy)
        .def_readonly("average_energy_trace", &ParallelTemperingResult::average_energy_trace)
        .def_readonly("swap_acceptance_trace", &ParallelTemperingResult::swap_acceptance_trace);

    py::class_<ParallelTemperingAnnealer>(m, "ParallelTemperingAnnealer")
        .def(py::init([](std::shared_ptr<Hamiltonian> ham, const std::vector<double> &betas, const std::string &backend) {
                 return ParallelTemperingAnnealer(backend_for(std::move(ham), backend), betas);
             }),
             py::arg("hamiltonian"), py::arg("betas"), py::arg("backend") = "hip", py::keep_alive<1, 2>())
        .def("set_seed", &ParallelTemperingAnnealer::set_seed)
        .def("run", &ParallelTemperingAnnealer::run, py::arg("sweeps_per_step"), py::arg("steps"),
             py::arg("swap_interval") = 1, py::call_guard<py::gil_scoped_release>());

    py::class_<SQAParallelTemperingResult>(m, "SQAParallelTemperingResult")
        .def_readonly("final_states", &SQAParallelTemperingResult::final_states)
        .def_readonly("final_energies", &SQAParallelTemperingResult::final_energies)
        .def_readonly("best_state", &SQAParallelTemperingResult::best_state)
        .def_readonly("best_energy", &SQAParallelTemperingResult::best_energy)
        .def_readonly("average_energy_trace", &SQAParallelTemperingResult::average_energy_trace)
        .def_readonly("swap_acceptance_trace", &SQAParallelTemperingResult::swap_acceptance_trace)
        .def_readonly("j_perp_trace", &SQAParallelTemperingResult::j_perp_trace)
        .def_readonly("chi_B_trace", &SQAParallelTemperingResult::chi_B_trace)
        .def_readonly("calibrated_eps_tilde", &SQAParallelTemperingResult::calibrated_eps_tilde)
        .def_readonly("j_perp_start", &SQAParallelTemperingResult::j_perp_start)
        .def_readonly("resolved_j_perp_end", &SQAParallelTemperingResult::resolved_j_perp_end)
        .def_readonly("final_j_perp", &SQAParallelTemperingResult::final_j_perp);

    py::class_<SQAParallelTemperingAnnealer>(m, "SQAParallelTemperingAnnealer")
        .def(py::init([](std::shared_ptr<Hamiltonian> ham, const std::vector<double> &betas,
                         const std::vector<double> &gammas, std::size_t trotter_slices, const std::string &backend) {
                 return SQAParallelTemperingAnnealer(backend_for(std::move(ham), backend), betas, gammas, trotter_slices);
             }),
             py::arg("hamiltonian"), py::arg("betas"), py::arg("gammas"), py::arg("trotter_slices"),
             py::arg("backend") = "hip", py::keep_alive<1, 2>())
        .def("set_seed", &SQAParallelTemperingAnnealer::set_seed)
        .def("run", &SQAParallelTemperingAnnealer::run, py::arg("sweeps_per_step"), py::arg("worldline_sweeps"),
             py::arg("steps"), py::arg("swap_interval") = 1, py::arg("cluster_sweeps") = 0,
             py::arg("continuous_time_slices") = 0, py::call_guard<py::gil_scoped_release>())
        .def("run_ladders", &SQAParallelTemperingAnnealer::run_ladders, py::arg("ladders"),
             py::arg("sweeps_per_step"), py::arg("worldline_sweeps"), py::arg("steps"), py::arg("swap_interval") = 1,
             py::arg("cluster_sweeps") = 0, py::arg("step_betas") = std::vector<std::vector<double>>{},
             py::arg("step_gammas") = std::vector<std::vector<double>>{}, py::call_guard<py::gil_scoped_release>())
        .def("run_optimal", &SQAParallelTemperingAnnealer::run_optimal, py::arg("num_steps"),
             py::arg("sweeps_per_step"), py::arg("worldline_sweeps") = 0, py::arg("eps_tilde") = 0.0,
             py::arg("alpha") = 15.0 / 14.0, py::arg("j_perp_end") = 0.0, py::arg("cluster_sweeps") = 0,
             py::arg("swap_interval") = 1, py::arg("continuous_time_slices") = 0, py::arg("calib_probes") = 12,
             py::arg("calib_sweeps") = 10, py::arg("debug_csv_path") = "", py::call_guard<py::gil_scoped_release>());

    py::class_<SQASchedule>(m, "SQASchedule")
        .def(py::init<>())
        .def_readwrite("betas", &SQASchedule::betas)
        .def_readwrite("gammas", &SQASchedule::gammas)
        .def_static("from_vectors", &SQASchedule::from_vectors, py::arg("betas"), py::arg("gammas"));

    py::class_<SQAResult>(m, "SQAResult")
        .def_readonly("best_state", &SQAResult::best_state)
        .def_readonly("best_energy", &SQAResult::best_energy)
        .def_readonly("energy_trace", &SQAResult::energy_trace)
        .def_readonly("j_perp_trace", &SQAResult::j_perp_trace)
        .def_readonly("chi_B_trace", &SQAResult::chi_B_trace)
        .def_readonly("calibrated_eps_tilde", &SQAResult::calibrated_eps_tilde)
        .def_readonly("j_perp_start", &SQAResult::j_perp_start)
        .def_readonly("resolved_j_perp_end", &SQAResult::resolved_j_perp_end)
        .def_readonly("final_j_perp", &SQAResult::final_j_perp);

    py::class_<SQAAnnealer>(m, "SQAAnnealer")
        .def(py::init([](std::shared_ptr<Hamiltonian> ham, SQASchedule schedule, std::size_t trotter_slices,
                         std::size_t replicas, const std::string &backend) {
                 return SQAAnnealer(backend_for(std::move(ham), backend), std::move(schedule), trotter_slices, replicas);
             }),
             py::arg("hamiltonian"), py::arg("schedule"), py::arg("trotter_slices"), py::arg("replicas") = 1,
             py::arg("backend") = "hip", py::keep_alive<1, 2>())
        .def("set_seed", &SQAAnnealer::set_seed)
        .def("run",
             [](SQAAnnealer &self, std::size_t spb, std::size_t wl, std::size_t cl, std::size_t cts,
                std::shared_ptr<SQAObserver> obs) { return self.run(spb, wl, cl, cts, obs.get()); },
             py::arg("sweeps_per_beta"), py::arg("worldline_sweeps"), py::arg("cluster_sweeps") = 0,
             py::arg("continuous_time_slices") = 0, py::arg("observer") = nullptr,
             py::call_guard<py::gil_scoped_release>())
        .def("run_optimal", &SQAAnnealer::run_optimal, py::arg("beta"), py::arg("j_perp_start"),
             py::arg("j_perp_end"), py::arg("eps_tilde"), py::arg("alpha") = 15.0 / 14.0, py::arg("num_steps") = 100,
             py::arg("sweeps_per_step") = 20, py::arg("worldline_sweeps") = 0, py::arg("cluster_sweeps") = 0,
             py::arg("calib_probes") = 12, py::arg("calib_sweeps") = 10, py::arg("debug_csv_path") = "",
             py::arg("beta_ramp_fraction") = 0.3, py::arg("beta_ramp_start") = 0.1,
             py::call_guard<py::gil_scoped_release>())
        .def("last_kernel_ms", &SQAAnnealer::last_kernel_ms)
        .def("last_accept_rate", &SQAAnnealer::last_accept_rate);

    py::class_<SQAChiResult>(m, "SQAChiResult")
        .def_readonly("best_state", &SQAChiResult::best_state)
        .def_readonly("best_energy", &SQAChiResult::best_energy)
        .def_readonly("energy_trace", &SQAChiResult::energy_trace)
        .def_readonly("beta_schedule", &SQAChiResult::beta_schedule)
        .def_readonly("gamma_schedule", &SQAChiResult::gamma_schedule)
        .def_readonly("s_schedule", &SQAChiResult::s_schedule)
        .def_readonly("scan_s", &SQAChiResult::scan_s)
        .def_readonly("scan_gamma", &SQAChiResult::scan_gamma)
        .def_readonly("scan_chi_B", &SQAChiResult::scan_chi_B)
        .def_readonly("s_star", &SQAChiResult::s_star)
        .def_readonly("gamma_star", &SQAChiResult::gamma_star)
        .def_readonly("j_perp_star", &SQAChiResult::j_perp_star)
        .def_readonly("chi_floor", &SQAChiResult::chi_floor)
        .def_readonly("driver_A0", &SQAChiResult::driver_A0);

    py::class_<SQAChiAnnealer>(m, "SQAChiAnnealer")
        .def(py::init([](std::shared_ptr<Hamiltonian> ham, std::size_t trotter_slices, std::size_t replicas,
                         const std::string &backend) {
                 return SQAChiAnnealer(backend_for(std::move(ham), backend), trotter_slices, replicas);
             }),
             py::arg("hamiltonian"), py::arg("trotter_slices"), py::arg("replicas") = 1, py::arg("backend") = "hip",
             py::keep_alive<1, 2>())
        .def("set_seed", &SQAChiAnnealer::set_seed)
        .def("run_chi", &SQAChiAnnealer::run_chi, py::arg("beta"), py::arg("gamma_start"), py::arg("gamma_end"),
             py::arg("num_steps") = 100, py::arg("sweeps_per_step") = 20, py::arg("scan_points") = 16,
             py::arg("scan_sweeps") = 30, py::arg("scan_burn") = 10, py::arg("chi_floor_fraction") = 1e-6,
             py::arg("driver_A0") = 0.0, py::arg("beta_ramp_fraction") = 0.3, py::arg("beta_ramp_start") = 0.1,
             py::arg("debug_csv_path") = "", py::call_guard<py::gil_scoped_release>())
        .def("run_chi_with_profile", &SQAChiAnnealer::run_chi_with_profile, py::arg("chi_B"), py::arg("beta"),
             py::arg("gamma_start"), py::arg("gamma_end"), py::arg("num_steps") = 100,
             py::arg("sweeps_per_step") = 20, py::arg("chi_floor_fraction") = 1e-6, py::arg("driver_A0") = 0.0,
             py::arg("beta_ramp_fraction") = 0.3, py::arg("beta_ramp_start") = 0.1,
             py::call_guard<py::gil_scoped_release>());

    m.def("magnetization", [](const py::sequence &s) {
        auto v = spins_of(s);
        return magnetization(v.data(), v.size());
    });
    m.def("overlap", [](const py::sequence &a, const py::sequence &b) {
        auto va = spins_of(a), vb = spins_of(b);
        if (va.size() != vb.size()) throw std::invalid_argument("Spin vectors must be same length.");
        return overlap(va.data(), vb.data(), va.size());
    });
}
