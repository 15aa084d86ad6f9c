// Umbrella header for the annealing-sweep API.
#pragma once

#include "qanneal/annealer.hpp"
#include "qanneal/backend.hpp"
#include "qanneal/dense_ising.hpp"
#include "qanneal/hamiltonian.hpp"
#include "qanneal/higher_order_ising.hpp"
#include "qanneal/metrics.hpp"
#include "qanneal/qubo.hpp"
#include "qanneal/schedule.hpp"
#include "qanneal/sparse_ising.hpp"
#include "qanneal/sqa_annealer.hpp"
#include "qanneal/sqa_chi_annealer.hpp"
#include "qanneal/sqa_parallel_tempering.hpp"
#include "qanneal/sqa_schedule.hpp"
#include "qanneal/sqa_state.hpp"
#include "qanneal/state.hpp"
#include "qanneal/version.hpp"
