// Dense Ising model E = c + sum_i h_i s_i + sum_{i<j} J_ij s_i s_j (mirrors dense_ising.hpp).
// energy() reads the upper triangle; fields and dE read full rows and skip J_ii.
#pragma once

#include <cstddef>
#include <vector>

#include "qanneal/hamiltonian.hpp"

namespace qanneal {

class DenseIsing final : public Hamiltonian {
public:
    DenseIsing() = default;
    DenseIsing(std::vector<double> h, std::vector<double> J, std::size_t n, double c = 0.0);

    using Hamiltonian::delta_energy;
    using Hamiltonian::energy;

    std::size_t size() const override { return n_; }
    double energy(const int8_t *spins, std::size_t n) const override;
    double delta_energy(const int8_t *spins, std::size_t n, std::size_t flip) const override;
    void compute_local_fields(const int8_t *spins, std::size_t n, double *fields) const override;
    void update_local_fields_after_flip(double *fields, const int8_t *new_spins, std::size_t n,
                                        std::size_t flip, int8_t old_spin) const override;

    const std::vector<double> &h() const { return h_; }
    const std::vector<double> &J() const { return J_; }
    double constant() const { return c_; }

private:
    std::vector<double> h_, J_;
    std::size_t n_ = 0;
    double c_ = 0.0;
};

}  // namespace qanneal
