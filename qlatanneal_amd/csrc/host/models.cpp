// Host FP64 methods of the problem models, QUBO conversion, and the HIP backend's device plumbing.
// Semantics follow dense_ising.cpp, sparse_ising.cpp, higher_order_ising.cpp and qubo.cpp.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qanneal/backend.hpp"
#include "qanneal/dense_ising.hpp"
#include "qanneal/higher_order_ising.hpp"
#include "qanneal/qubo.hpp"
#include "qanneal/sparse_ising.hpp"
#include "qanneal_hip.h"

namespace qanneal {

namespace {
void need_size(std::size_t n, std::size_t expect) {
    if (n != expect) throw std::invalid_argument("State size mismatch.");
}
}  // namespace

// ---------------------------------------------------------------------------------- DenseIsing
DenseIsing::DenseIsing(std::vector<double> h, std::vector<double> J, std::size_t n, double c)
    : h_(std::move(h)), J_(std::move(J)), n_(n), c_(c) {
    if (n_ == 0) throw std::invalid_argument("DenseIsing size must be > 0.");
    if (h_.size() != n_) throw std::invalid_argument("DenseIsing h size mismatch.");
    if (J_.size() != n_ * n_) throw std::invalid_argument("DenseIsing J size mismatch.");
}

double DenseIsing::energy(const int8_t *s, std::size_t n) const {
    need_size(n, n_);
    validate_spins(s, n_);
    double e = c_;
    for (std::size_t i = 0; i < n_; ++i) e += h_[i] * static_cast<double>(s[i]);
    for (std::size_t i = 0; i < n_; ++i) {
        const double si = static_cast<double>(s[i]);
        for (std::size_t j = i + 1; j < n_; ++j) e += J_[i * n_ + j] * si * static_cast<double>(s[j]);
    }
    return e;
}

double DenseIsing::delta_energy(const int8_t *s, std::size_t n, std::size_t flip) const {
    need_size(n, n_);
    if (flip >= n_) throw std::invalid_argument("Flip index out of range.");
    double local = h_[flip];
    const double *row = &J_[flip * n_];
    for (std::size_t j = 0; j < n_; ++j)
        if (j != flip) local += row[j] * static_cast<double>(s[j]);
    return -2.0 * static_cast<double>(s[flip]) * local;
}

void DenseIsing::compute_local_fields(const int8_t *s, std::size_t n, double *f) const {
    need_size(n, n_);
    for (std::size_t i = 0; i < n_; ++i) {
        const double *row = &J_[i * n_];
        double acc = h_[i];
        for (std::size_t j = 0; j < n_; ++j) acc += row[j] * static_cast<double>(s[j]);
        f[i] = acc - row[i] * static_cast<double>(s[i]);
    }
}

void DenseIsing::update_local_fields_after_flip(double *f, const int8_t *, std::size_t n, std::size_t flip,
                                                int8_t old_spin) const {
    need_size(n, n_);
    const double ds = -2.0 * static_cast<double>(old_spin);
    const double *row = &J_[flip * n_];
    for (std::size_t j = 0; j < n_; ++j) f[j] += row[j] * ds;
    f[flip] -= row[flip] * ds;
}

// --------------------------------------------------------------------------------- SparseIsing
SparseIsing::SparseIsing(std::vector<double> h, std::vector<SparseEdge> edges, std::size_t n, double c)
    : h_(std::move(h)), edges_(std::move(edges)), n_(n), c_(c) {
    if (n_ == 0) throw std::invalid_argument("SparseIsing size must be > 0.");
    if (h_.size() != n_) throw std::invalid_argument("SparseIsing h size mismatch.");
    for (const auto &e : edges_) {
        if (e.i >= n_ || e.j >= n_) throw std::invalid_argument("SparseIsing edge index out of range.");
        if (e.i == e.j) throw std::invalid_argument("SparseIsing self-edge not allowed.");
    }
    adj_.assign(n_, {});
    for (const auto &e : edges_) {
        adj_[e.i].push_back({e.j, e.value});
        adj_[e.j].push_back({e.i, e.value});
    }
}

double SparseIsing::energy(const int8_t *s, std::size_t n) const {
    need_size(n, n_);
    validate_spins(s, n_);
    double e = c_;
    for (std::size_t i = 0; i < n_; ++i) e += h_[i] * static_cast<double>(s[i]);
    for (const auto &ed : edges_) e += ed.value * static_cast<double>(s[ed.i]) * static_cast<double>(s[ed.j]);
    return e;
}

double SparseIsing::delta_energy(const int8_t *s, std::size_t n, std::size_t flip) const {
    need_size(n, n_);
    if (flip >= n_) throw std::invalid_argument("Flip index out of range.");
    double local = h_[flip];
    for (const auto &nb : adj_[flip]) local += nb.second * static_cast<double>(s[nb.first]);
    return -2.0 * static_cast<double>(s[flip]) * local;
}

void SparseIsing::compute_local_fields(const int8_t *s, std::size_t n, double *f) const {
    need_size(n, n_);
    for (std::size_t i = 0; i < n_; ++i) f[i] = h_[i];
    for (const auto &ed : edges_) {
        f[ed.i] += ed.value * static_cast<double>(s[ed.j]);
        f[ed.j] += ed.value * static_cast<double>(s[ed.i]);
    }
}

void SparseIsing::update_local_fields_after_flip(double *f, const int8_t *, std::size_t n, std::size_t flip,
                                                 int8_t old_spin) const {
    need_size(n, n_);
    const double ds = -2.0 * static_cast<double>(old_spin);
    for (const auto &nb : adj_[flip]) f[nb.first] += nb.second * ds;
}

// ---------------------------------------------------------------------------- HigherOrderIsing
HigherOrderIsing::HigherOrderIsing(std::size_t n, const std::vector<RawTerm> &terms, double c)
    : h_(n, 0.0), n_(n), c_(c) {
    if (n_ == 0) throw std::invalid_argument("HigherOrderIsing size must be > 0.");
    std::map<std::vector<std::uint32_t>, double> support;  // ordered by sorted variable list
    for (const auto &t : terms) {
        if (t.second == 0.0) continue;
        std::vector<std::size_t> v = t.first;
        for (std::size_t x : v)
            if (x >= n_) throw std::invalid_argument("HigherOrderIsing term variable index out of range.");
        std::sort(v.begin(), v.end());
        std::vector<std::uint32_t> odd;
        for (std::size_t a = 0; a < v.size();) {
            std::size_t b = a;
            while (b < v.size() && v[b] == v[a]) ++b;
            if ((b - a) & 1u) odd.push_back(static_cast<std::uint32_t>(v[a]));
            a = b;
        }
        if (odd.empty()) c_ += t.second;
        else if (odd.size() == 1) h_[odd[0]] += t.second;
        else support[odd] += t.second;
    }
    offset_.push_back(0);
    incident_.assign(n_, {});
    for (const auto &kv : support) {
        if (kv.second == 0.0) continue;
        const auto id = static_cast<std::uint32_t>(coeff_.size());
        coeff_.push_back(kv.second);
        for (std::uint32_t x : kv.first) {
            vars_.push_back(x);
            incident_[x].push_back(id);
        }
        offset_.push_back(vars_.size());
        max_degree_ = std::max(max_degree_, kv.first.size());
    }
    if (max_degree_ == 0 && std::any_of(h_.begin(), h_.end(), [](double x) { return x != 0.0; })) max_degree_ = 1;
}

double HigherOrderIsing::energy(const int8_t *s, std::size_t n) const {
    need_size(n, n_);
    validate_spins(s, n_);
    // one running sum: c, then the linear terms, then every coupling term (the order the
    // reference accumulates in, so the FP64 result is identical)
    double e = c_;
    for (std::size_t i = 0; i < n_; ++i) e += h_[i] * static_cast<double>(s[i]);
    for (std::size_t t = 0; t < coeff_.size(); ++t) {
        double prod = coeff_[t];
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k) prod *= static_cast<double>(s[vars_[k]]);
        e += prod;
    }
    return e;
}

double HigherOrderIsing::coupling_energy(const int8_t *s, std::size_t n) const {
    need_size(n, n_);
    double e = 0.0;
    for (std::size_t t = 0; t < coeff_.size(); ++t) {
        double prod = coeff_[t];
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k) prod *= static_cast<double>(s[vars_[k]]);
        e += prod;
    }
    return e;
}

double HigherOrderIsing::delta_energy(const int8_t *s, std::size_t n, std::size_t flip) const {
    need_size(n, n_);
    if (flip >= n_) throw std::invalid_argument("Flip index out of range.");
    double local = h_[flip];
    for (std::uint32_t t : incident_[flip]) {
        double prod = coeff_[t];
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k)
            if (vars_[k] != flip) prod *= static_cast<double>(s[vars_[k]]);
        local += prod;
    }
    return -2.0 * static_cast<double>(s[flip]) * local;
}

void HigherOrderIsing::compute_local_fields(const int8_t *s, std::size_t n, double *f) const {
    need_size(n, n_);
    for (std::size_t i = 0; i < n_; ++i) f[i] = h_[i];
    for (std::size_t t = 0; t < coeff_.size(); ++t) {
        double prod = coeff_[t];
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k) prod *= static_cast<double>(s[vars_[k]]);
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k) f[vars_[k]] += prod * static_cast<double>(s[vars_[k]]);
    }
}

void HigherOrderIsing::update_local_fields_after_flip(double *f, const int8_t *ns, std::size_t n,
                                                      std::size_t flip, int8_t) const {
    need_size(n, n_);
    for (std::uint32_t t : incident_[flip]) {
        double prod = coeff_[t];
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k) prod *= static_cast<double>(ns[vars_[k]]);
        for (std::size_t k = offset_[t]; k < offset_[t + 1]; ++k)
            if (vars_[k] != flip) f[vars_[k]] += 2.0 * prod * static_cast<double>(ns[vars_[k]]);
    }
}

// --------------------------------------------------------------------------------------- QUBO
QUBO::QUBO(std::vector<double> q, std::size_t n) : q_(std::move(q)), n_(n) {
    if (n_ == 0) throw std::invalid_argument("QUBO size must be > 0.");
    if (q_.size() != n_ * n_) throw std::invalid_argument("QUBO matrix size mismatch.");
}

QUBO::QUBO(const std::vector<std::pair<std::pair<std::size_t, std::size_t>, double>> &entries, std::size_t n)
    : q_(n * n, 0.0), n_(n) {
    if (n_ == 0) throw std::invalid_argument("QUBO size must be > 0.");
    for (const auto &e : entries) {
        if (e.first.first >= n_ || e.first.second >= n_) throw std::invalid_argument("QUBO entry index out of range.");
        q_[e.first.first * n_ + e.first.second] += e.second;
    }
}

// With W = (Q + Q^T)/2 and x = (1 + s)/2: x^T Q x = 1/4 s^T W s + 1/2 s^T W 1 + 1/4 1^T W 1, so
// h = 1/2 W 1, J = 1/2 W (full matrix; energy() reads i<j only) and c = 1/4 (tr W + 1^T W 1).
DenseIsing QUBO::to_ising() const {
    const std::size_t n = n_;
    std::vector<double> W(n * n), h(n, 0.0), J(n * n, 0.0);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) W[i * n + j] = 0.5 * (q_[i * n + j] + q_[j * n + i]);
    for (std::size_t i = 0; i < n; ++i) {
        double rs = 0.0;
        for (std::size_t j = 0; j < n; ++j) rs += W[i * n + j];
        h[i] = 0.5 * rs;
    }
    double c = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        c += 0.25 * W[i * n + i];
        for (std::size_t j = 0; j < n; ++j) {
            J[i * n + j] = 0.5 * W[i * n + j];
            c += 0.25 * W[i * n + j];
        }
    }
    return DenseIsing(std::move(h), std::move(J), n, c);
}

// --------------------------------------------------------------------------------- backends
BackendSpec parse_backend(std::string_view name) {
    std::string s(name);
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return std::tolower(ch); });
    if (s == "cpu") return {BackendKind::CPU, DevicePrecision::FP32};
    if (s == "cuda") return {BackendKind::CUDA, DevicePrecision::FP32};
    if (s == "hip" || s == "hip_fp32" || s == "rocm") return {BackendKind::HIP, DevicePrecision::FP32};
    if (s == "hip_fp64" || s == "hip-fp64" || s == "hip_validate")
        return {BackendKind::HIP, DevicePrecision::FP64Validation};
    throw std::invalid_argument("Unknown backend: " + std::string(name));
}

void throw_on_status(int status) {
    if (status == QA_OK) return;
    const std::string msg = qa_last_error();
    if (status == QA_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

namespace {
struct DeviceContext {
    qa_ctx *ctx = nullptr;
    std::mutex run_mu;
};

// Contexts of the selected GPU, one HIP stream each. A backend is bound to one slot for its
// lifetime (its device problem lives in that slot's context); backends on different slots run
// concurrently when driven from different host threads. The pool size defaults to the GPU's
// hardware queues per process (4) and can be set with QANNEAL_HIP_STREAMS (1..16).
struct ContextPool {
    int device = 0;
    std::vector<std::unique_ptr<DeviceContext>> slots;
    std::once_flag once;
};

ContextPool &context_pool() {
    static ContextPool pool;
    return pool;
}

int pool_size() {
    static const int k = [] {
        const char *env = std::getenv("QANNEAL_HIP_STREAMS");
        const int v = env ? std::atoi(env) : 4;
        return std::clamp(v, 1, 16);
    }();
    return k;
}

int next_slot() {
    static std::atomic<unsigned> counter{0};
    return static_cast<int>(counter.fetch_add(1, std::memory_order_relaxed) % static_cast<unsigned>(pool_size()));
}

DeviceContext &device_context(int slot) {
    ContextPool &pool = context_pool();
    std::call_once(pool.once, [&pool] {
        const char *env = std::getenv("QANNEAL_HIP_DEVICE");
        pool.device = env ? std::atoi(env) : 0;
        std::vector<std::unique_ptr<DeviceContext>> slots;
        for (int k = 0; k < pool_size(); ++k) {
            auto dc = std::make_unique<DeviceContext>();
            throw_on_status(qa_ctx_create(pool.device, &dc->ctx));
            slots.push_back(std::move(dc));
        }
        pool.slots = std::move(slots);
    });
    return *pool.slots[static_cast<std::size_t>(slot)];
}
}  // namespace

int hip_context_count() { return pool_size(); }

namespace {
// Approximate device bytes of a problem's copy (dense: J in FP64 and FP32 plus the bf16 split;
// sparse and HUBO: their coupling lists and incidence arrays), for the choice of whether runs may
// copy it into another context.
int64_t problem_bytes(const Hamiltonian *h) {
    const auto n = static_cast<int64_t>(h->size());
    if (dynamic_cast<const DenseIsing *>(h)) return 16 * n * n + 32 * n;
    if (auto *sp = dynamic_cast<const SparseIsing *>(h)) return 40 * static_cast<int64_t>(sp->edges().size()) + 32 * n;
    if (auto *hb = dynamic_cast<const HigherOrderIsing *>(h))
        return 24 * static_cast<int64_t>(hb->term_variables().size()) + 24 * static_cast<int64_t>(hb->num_terms()) + 32 * n;
    return 32 * n;
}

constexpr int64_t kMaxMovable = int64_t(256) << 20;    // larger problems stay in their own context
constexpr int64_t kExtraCopyBudget = int64_t(16) << 30;  // all extra copies of all backends together
std::atomic<int64_t> g_extra_copy_bytes{0};

qa_problem *upload_problem(qa_ctx *ctx, const Hamiltonian *h) {
    qa_problem *p = nullptr;
    if (auto *d = dynamic_cast<const DenseIsing *>(h)) {
        throw_on_status(qa_problem_dense(ctx, static_cast<int64_t>(d->size()), d->h().data(), d->J().data(),
                                         d->constant(), &p));
    } else if (auto *sp = dynamic_cast<const SparseIsing *>(h)) {
        std::vector<int64_t> ei, ej;
        std::vector<double> ev;
        for (const auto &e : sp->edges()) {
            ei.push_back(static_cast<int64_t>(e.i));
            ej.push_back(static_cast<int64_t>(e.j));
            ev.push_back(e.value);
        }
        throw_on_status(qa_problem_sparse(ctx, static_cast<int64_t>(sp->size()), sp->h().data(),
                                          static_cast<int64_t>(ev.size()), ei.data(), ej.data(), ev.data(),
                                          sp->constant(), &p));
    } else if (auto *hb = dynamic_cast<const HigherOrderIsing *>(h)) {
        std::vector<int64_t> off(hb->term_offsets().begin(), hb->term_offsets().end());
        throw_on_status(qa_problem_hubo(ctx, static_cast<int64_t>(hb->size()), hb->h().data(),
                                        static_cast<int64_t>(hb->num_terms()), off.data(),
                                        hb->term_variables().data(), hb->term_coefficients().data(),
                                        hb->constant(), &p));
    } else {
        throw std::invalid_argument("the HIP backend runs DenseIsing, SparseIsing and HigherOrderIsing models");
    }
    return p;
}
}  // namespace

HIPBackend::HIPBackend(std::shared_ptr<const Hamiltonian> ham, DevicePrecision precision)
    : CPUBackend(std::move(ham)), precision_(precision), slot_(next_slot()), bytes_(problem_bytes(ham_.get())) {}

HIPBackend::~HIPBackend() {
    for (qa_problem *p : problems_)
        if (p) qa_problem_destroy(p);
    g_extra_copy_bytes.fetch_sub(extra_bytes_);
}

bool HIPBackend::has_problem(int slot) const {
    std::lock_guard<std::mutex> lock(upload_mu_);
    return !problems_.empty() && problems_[static_cast<std::size_t>(slot)] != nullptr;
}

qa_problem *HIPBackend::problem_on(int slot) const {
    qa_ctx *ctx = device_context(slot).ctx;
    std::lock_guard<std::mutex> lock(upload_mu_);
    if (problems_.empty()) problems_.assign(static_cast<std::size_t>(pool_size()), nullptr);
    qa_problem *&p = problems_[static_cast<std::size_t>(slot)];
    if (!p) p = upload_problem(ctx, ham_.get());
    return p;
}

// Reserves room for one more copy of this problem in another context; false when the problem is
// too large to copy or the copies of all backends together would exceed their budget.
bool HIPBackend::reserve_extra_copy() const {
    if (bytes_ > kMaxMovable) return false;
    if (g_extra_copy_bytes.fetch_add(bytes_) + bytes_ > kExtraCopyBudget) {
        g_extra_copy_bytes.fetch_sub(bytes_);
        return false;
    }
    std::lock_guard<std::mutex> lock(upload_mu_);
    extra_bytes_ += bytes_;
    return true;
}

HIPBackend::Lease HIPBackend::lease() const {
    const int k = pool_size();
    for (int d = 0; d < k; ++d) {
        const int s = (slot_ + d) % k;
        if (d > 0 && bytes_ > kMaxMovable && !has_problem(s)) continue;
        std::unique_lock<std::mutex> lk(device_context(s).run_mu, std::try_to_lock);
        if (!lk.owns_lock()) continue;
        if (d > 0 && !has_problem(s) && !reserve_extra_copy()) continue;
        qa_problem *p = problem_on(s);
        return Lease{std::move(lk), device_context(s).ctx, p};
    }
    std::unique_lock<std::mutex> lk(device_context(slot_).run_mu);
    qa_problem *p = problem_on(slot_);
    return Lease{std::move(lk), device_context(slot_).ctx, p};
}

std::shared_ptr<Backend> make_backend(BackendSpec spec, std::shared_ptr<const Hamiltonian> ham) {
    switch (spec.kind) {
        case BackendKind::CPU: return std::make_shared<CPUBackend>(std::move(ham));
        case BackendKind::CUDA: throw std::runtime_error("CUDA backend not enabled at build time.");
        case BackendKind::HIP:
            if (!ham) throw std::invalid_argument("HIPBackend requires a Hamiltonian.");
            return std::make_shared<HIPBackend>(std::move(ham), spec.precision);
    }
    throw std::runtime_error("Unknown backend kind.");
}

}  // namespace qanneal
