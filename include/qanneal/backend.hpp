// Compute-backend selection (mirrors backend.hpp:14-110, plus BackendKind::HIP).
//
// The reference's Backend is a per-spin plugin called from inside the engines' loops. Here the
// per-spin virtuals stay (host FP64, for single-state callers), and a HIP backend additionally
// owns the device copy of the problem: every annealer hands a whole run to the device when its
// backend is_device(). There is no CPU sweep engine in this build, so an annealer on the CPU
// backend refuses to run instead of falling back.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>
#include <stdexcept>
#include <string>
#include <string_view>

#include "qanneal/hamiltonian.hpp"

struct qa_ctx;
struct qa_problem;

namespace qanneal {

enum class BackendKind { CPU, CUDA, HIP };

// Arithmetic of a device run: FP32 production sweep, or the FP64 validation replay of the
// reference trajectory.
enum class DevicePrecision { FP32, FP64Validation };

struct BackendSpec {
    BackendKind kind;
    DevicePrecision precision;
};

// "cpu" | "cuda" | "hip" | "hip_fp64" (case-insensitive first letters as in the reference).
BackendSpec parse_backend(std::string_view name);

inline BackendKind backend_from_string(std::string_view name) { return parse_backend(name).kind; }

inline const char *backend_to_string(BackendKind kind) {
    switch (kind) {
        case BackendKind::CPU: return "cpu";
        case BackendKind::CUDA: return "cuda";
        case BackendKind::HIP: return "hip";
    }
    return "unknown";
}

class Backend {
public:
    virtual ~Backend() = default;
    virtual BackendKind kind() const = 0;
    virtual std::size_t size() const = 0;
    virtual double energy(const int8_t *spins, std::size_t n) const = 0;
    virtual double delta_energy(const int8_t *spins, std::size_t n, std::size_t flip) const = 0;
    virtual void compute_local_fields(const int8_t *spins, std::size_t n, double *fields) const = 0;
    virtual void update_local_fields_after_flip(double *fields, const int8_t *new_spins, std::size_t n,
                                                std::size_t flip, int8_t old_spin) const = 0;
    virtual bool is_device() const { return false; }
    virtual const Hamiltonian &hamiltonian() const = 0;
};

class CPUBackend : public Backend {
public:
    explicit CPUBackend(std::shared_ptr<const Hamiltonian> ham) : ham_(std::move(ham)) {
        if (!ham_) throw std::invalid_argument("CPUBackend requires a Hamiltonian.");
    }
    BackendKind kind() const override { return BackendKind::CPU; }
    std::size_t size() const override { return ham_->size(); }
    double energy(const int8_t *s, std::size_t n) const override { return ham_->energy(s, n); }
    double delta_energy(const int8_t *s, std::size_t n, std::size_t f) const override { return ham_->delta_energy(s, n, f); }
    void compute_local_fields(const int8_t *s, std::size_t n, double *f) const override { ham_->compute_local_fields(s, n, f); }
    void update_local_fields_after_flip(double *f, const int8_t *s, std::size_t n, std::size_t fl, int8_t o) const override {
        ham_->update_local_fields_after_flip(f, s, n, fl, o);
    }
    const Hamiltonian &hamiltonian() const override { return *ham_; }

protected:
    std::shared_ptr<const Hamiltonian> ham_;
};

// HIP backend: the per-spin host methods of CPUBackend plus a device copy of the problem,
// uploaded on first use and reused by every run. Runs take one of a small pool of contexts (one
// HIP stream each) on the selected GPU: each backend prefers its own, assigned round-robin at
// construction, so runs of different backends from different host threads execute concurrently.
class HIPBackend final : public CPUBackend {
public:
    HIPBackend(std::shared_ptr<const Hamiltonian> ham, DevicePrecision precision);
    ~HIPBackend() override;
    BackendKind kind() const override { return BackendKind::HIP; }
    bool is_device() const override { return true; }
    DevicePrecision precision() const { return precision_; }
    void set_precision(DevicePrecision p) { precision_ = p; }

    // A context to run on and the device problem in it, held until the lease ends: the backend's
    // own context when it is free, else another free one (the problem is uploaded there on first
    // use, when it is small), else the backend's own context once it frees up.
    struct Lease {
        std::unique_lock<std::mutex> lock;
        qa_ctx *ctx = nullptr;
        qa_problem *problem = nullptr;
    };
    Lease lease() const;

private:
    qa_problem *problem_on(int slot) const;  // uploads on first use; caller holds the slot
    bool has_problem(int slot) const;
    bool reserve_extra_copy() const;

    DevicePrecision precision_;
    int slot_;                               // own context pool slot
    int64_t bytes_;                          // approximate device bytes of one problem copy
    mutable int64_t extra_bytes_ = 0;        // bytes of this backend's copies in other contexts
    mutable std::vector<qa_problem *> problems_;  // per pool slot
    mutable std::mutex upload_mu_;
};

std::shared_ptr<Backend> make_backend(BackendSpec spec, std::shared_ptr<const Hamiltonian> ham);

// Number of HIP contexts (streams) HIP backends are spread over: QANNEAL_HIP_STREAMS (1..16),
// default 4. No GPU call is made.
int hip_context_count();

inline std::shared_ptr<Backend> make_backend(BackendKind kind, std::shared_ptr<const Hamiltonian> ham) {
    return make_backend(BackendSpec{kind, DevicePrecision::FP32}, std::move(ham));
}

// Non-owning alias, as backend.hpp:107-110: the caller keeps the Hamiltonian alive.
inline std::shared_ptr<Backend> make_backend(BackendKind kind, const Hamiltonian &ham) {
    return make_backend(kind, std::shared_ptr<const Hamiltonian>(&ham, [](const Hamiltonian *) {}));
}

// Raises the reference's exception type for a C-ABI status (EINVAL -> invalid_argument,
// anything else -> runtime_error).
void throw_on_status(int status);

}  // namespace qanneal
