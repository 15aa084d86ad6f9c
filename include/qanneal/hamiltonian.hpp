// Problem-model interface (mirrors hamiltonian.hpp:10-50). The host methods are FP64 and serve
// callers that evaluate single states (energy checks, probe-based schedule tuning); the annealing
// sweeps themselves run on the device.
#pragma once

#include <cstddef>
#include <cstdint>

#include "qanneal/state.hpp"

namespace qanneal {

class Hamiltonian {
public:
    virtual ~Hamiltonian() = default;

    virtual double energy(const int8_t *spins, std::size_t n) const = 0;
    virtual double delta_energy(const int8_t *spins, std::size_t n, std::size_t flip) const = 0;
    virtual std::size_t size() const = 0;

    double energy(const State &s) const { return energy(s.spins.data(), s.size()); }
    double delta_energy(const State &s, std::size_t flip) const { return delta_energy(s.spins.data(), s.size(), flip); }

    // fields[i] = h_i + sum_{j != i} J_ij s_j, so that dE(flip i) = -2 s_i fields[i].
    virtual void compute_local_fields(const int8_t *spins, std::size_t n, double *fields) const {
        for (std::size_t i = 0; i < n; ++i) {
            const double s = static_cast<double>(spins[i]);
            fields[i] = s != 0.0 ? -delta_energy(spins, n, i) / (2.0 * s) : 0.0;
        }
    }
    virtual void update_local_fields_after_flip(double *fields, const int8_t *new_spins, std::size_t n,
                                                std::size_t flip, int8_t old_spin) const {
        (void)flip;
        (void)old_spin;
        compute_local_fields(new_spins, n, fields);
    }
};

}  // namespace qanneal
