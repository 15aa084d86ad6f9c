// Inverse-temperature schedule (mirrors schedule.hpp:9-30).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <vector>

namespace qanneal {

struct AnnealSchedule {
    std::vector<double> betas;

    std::size_t size() const { return betas.size(); }

    // steps points from beta_start to beta_end with a constant increment; one point -> beta_end.
    static AnnealSchedule linear(double beta_start, double beta_end, std::size_t steps) {
        if (steps == 0) throw std::invalid_argument("Schedule steps must be > 0.");
        AnnealSchedule s;
        if (steps == 1) {
            s.betas.push_back(beta_end);
            return s;
        }
        const double inc = (beta_end - beta_start) / static_cast<double>(steps - 1);
        s.betas.reserve(steps);
        for (std::size_t k = 0; k < steps; ++k) s.betas.push_back(beta_start + inc * static_cast<double>(k));
        return s;
    }
};

}  // namespace qanneal
