// Sparse Ising model over an undirected edge list (mirrors sparse_ising.hpp).
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

#include "qanneal/hamiltonian.hpp"

namespace qanneal {

struct SparseEdge {
    std::size_t i;
    std::size_t j;
    double value;
};

class SparseIsing final : public Hamiltonian {
public:
    SparseIsing() = default;
    SparseIsing(std::vector<double> h, std::vector<SparseEdge> edges, std::size_t n, double c = 0.0);

    using Hamiltonian::delta_energy;
    using Hamiltonian::energy;

    std::size_t size() const override { return n_; }
    double energy(const int8_t *spins, std::size_t n) const override;
    double delta_energy(const int8_t *spins, std::size_t n, std::size_t flip) const override;
    void compute_local_fields(const int8_t *spins, std::size_t n, double *fields) const override;
    void update_local_fields_after_flip(double *fields, const int8_t *new_spins, std::size_t n,
                                        std::size_t flip, int8_t old_spin) const override;

    const std::vector<double> &h() const { return h_; }
    const std::vector<SparseEdge> &edges() const { return edges_; }
    double constant() const { return c_; }

private:
    std::vector<double> h_;
    std::vector<SparseEdge> edges_;
    std::vector<std::vector<std::pair<std::size_t, double>>> adj_;  // neighbours in edge order
    std::size_t n_ = 0;
    double c_ = 0.0;
};

}  // namespace qanneal
