#pragma once

#include <cstddef>
#include <memory>

#include "cmi.h"
#include "cmi/aggregate_search.hpp"
#include "cmi/search.hpp"
#include "hip_backend.hpp"

// The k-child search on the device: the host state machine of
// cmi/aggregate_search.hpp driven with the aggregate bounder (one launch of N
// subregions reduced to k child bounds) and a sampler over the node's
// sampling partition, each role rebuilt only when its launch composition
// changes (reference aggregate/driver.hpp:98-462,
// backend/aggregate/cache.cuh:18-120).
namespace cmi::hip
{

struct AggregateRun
{
  search::SolveOutcome outcome;
  bool rule_engaged = true;
  std::uint64_t roles_rebuilt = 0;
};

/// Checks the options, builds the tape and solves.
AggregateRun aggregate_solve(std::shared_ptr<Context> ctx, const model::Problem<double>& problem,
                             const cmi_aggregate_options& o);

/// The reference's device fit of N and the sample fan-out
/// (source/gams/aggregate_solve.cu:298-315).
void aggregate_fit(std::size_t buffer_nodes, std::size_t free_device_bytes, std::size_t branch_fan_out,
                   std::size_t sample_points, std::size_t& partition_budget, std::size_t& sample_fan_out);

}  // namespace cmi::hip
