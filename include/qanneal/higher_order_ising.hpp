// Higher-order Ising (HUBO) model E = c + sum_i h_i s_i + sum_t J_t prod_{i in t} s_i
// (mirrors higher_order_ising.hpp). The constructor reduces repeated variables by parity
// (s^2 = 1), folds degree-0/1 terms into c/h, and merges equal supports.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "qanneal/hamiltonian.hpp"

namespace qanneal {

class HigherOrderIsing final : public Hamiltonian {
public:
    using RawTerm = std::pair<std::vector<std::size_t>, double>;

    HigherOrderIsing() = default;
    HigherOrderIsing(std::size_t n, const std::vector<RawTerm> &terms, double c = 0.0);

    using Hamiltonian::delta_energy;
    using Hamiltonian::energy;

    std::size_t size() const override { return n_; }
    double energy(const int8_t *spins, std::size_t n) const override;
    double delta_energy(const int8_t *spins, std::size_t n, std::size_t flip) const override;
    void compute_local_fields(const int8_t *spins, std::size_t n, double *fields) const override;
    void update_local_fields_after_flip(double *fields, const int8_t *new_spins, std::size_t n,
                                        std::size_t flip, int8_t old_spin) const override;

    std::vector<double> linear_terms() const { return h_; }
    double coupling_energy(const int8_t *spins, std::size_t n) const;
    std::size_t max_degree() const { return max_degree_; }
    std::size_t num_terms() const { return coeff_.size(); }
    const std::vector<double> &h() const { return h_; }
    double constant() const { return c_; }

    // Consolidated terms as CSR (what the device upload takes).
    const std::vector<double> &term_coefficients() const { return coeff_; }
    const std::vector<std::size_t> &term_offsets() const { return offset_; }
    const std::vector<std::uint32_t> &term_variables() const { return vars_; }

private:
    std::vector<double> coeff_;
    std::vector<std::size_t> offset_;
    std::vector<std::uint32_t> vars_;
    std::vector<std::vector<std::uint32_t>> incident_;
    std::vector<double> h_;
    std::size_t n_ = 0;
    double c_ = 0.0;
    std::size_t max_degree_ = 0;
};

}  // namespace qanneal
