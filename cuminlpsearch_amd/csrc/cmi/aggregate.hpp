#pragma once

#include <algorithm>
#include <cstddef>
#include <limits>
#include <span>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/region.hpp"

// The k-child aggregate: one launch over N = k * M subregions of a node,
// reduced on the device to k child bounds (reference aggregate/bound.hpp:25-
// 129, aggregate/partition.hpp:17-196). Host-only types; the device side is
// the Aggregate mode of the item kernels and hip/aggregate.cpp.
namespace cmi::aggregate
{

/// What one launch learned about one child (reference bound.hpp:42-64).
/// `hull_ub` is a sound upper bound of the objective over the child's
/// unexcluded part, attained at no known point: never an incumbent.
struct AggregateBound
{
  double lb = std::numeric_limits<double>::max();
  double hull_ub = std::numeric_limits<double>::max();
  bool feasible = true;
};

/// A node's partition: k children, each covered by M subregions. `launch` is
/// `refine`'s slots followed by `branch`'s, so the branch digits are the most
/// significant and child c owns launch regions [c*M, (c+1)*M) (reference
/// partition.hpp:25-100).
struct AggregatePartition
{
  region::SlotAssignment branch;
  region::SlotAssignment refine;
  region::SlotAssignment launch;
  std::size_t branch_fan_out = 1;
  std::size_t refine_fan_out = 1;
  bool nested = false;

  std::size_t total_fan_out() const { return branch_fan_out * refine_fan_out; }
  bool is_leaf() const { return branch.size() == 0; }

  /// The invariants the launch layout relies on (reference
  /// partition.hpp:102-177); throws InvalidConfiguration naming the clause.
  void validate() const
  {
    using region::composition_fan_out;
    if (is_leaf()) {
      return;
    }
    if (composition_fan_out(branch.widths) != branch_fan_out) {
      throw InvalidConfiguration("AggregatePartition: product of branch widths does not equal the branch fan-out k");
    }
    if (composition_fan_out(launch.widths) != total_fan_out()) {
      throw InvalidConfiguration("AggregatePartition: product of launch widths does not equal N = k * M");
    }
    std::size_t const n_branch = branch.size();
    if (launch.size() < n_branch) {
      throw InvalidConfiguration("AggregatePartition: launch has fewer slots than branch");
    }
    if (nested) {
      if (launch.size() != 1 || branch.size() != 1 || refine.size() != 0) {
        throw InvalidConfiguration(
            "AggregatePartition: a nested partition is exactly one slot shared by branch and refine");
      }
      if (launch.var_ids[0] != branch.var_ids[0] || launch.composition[0] != branch.composition[0]) {
        throw InvalidConfiguration(
            "AggregatePartition: the nested slot does not match the branch slot's variable and kind");
      }
      return;
    }
    if (composition_fan_out(refine.widths) != refine_fan_out) {
      throw InvalidConfiguration("AggregatePartition: product of refine widths does not equal the refine fan-out M");
    }
    if (launch.size() != branch.size() + refine.size()) {
      throw InvalidConfiguration("AggregatePartition: launch is not the concatenation of refine and branch");
    }
    std::size_t const offset = launch.size() - n_branch;
    for (std::size_t j = 0; j < n_branch; ++j) {
      if (launch.var_ids[offset + j] != branch.var_ids[j] || launch.widths[offset + j] != branch.widths[j]
          || launch.composition[offset + j] != branch.composition[j]) {
        throw InvalidConfiguration(
            "AggregatePartition: branch slots are not the most significant slots of launch, so child_id = r / M "
            "does not hold");
      }
    }
    for (std::size_t bv : branch.var_ids) {
      if (std::find(refine.var_ids.begin(), refine.var_ids.end(), bv) != refine.var_ids.end()) {
        throw InvalidConfiguration(
            "AggregatePartition: branch and refine act on the same variable without being nested");
      }
    }
  }
};

/// The child a launch-region index belongs to (reference partition.hpp:180-
/// 184).
inline std::size_t child_of_region(std::size_t r, std::size_t refine_fan_out) { return r / refine_fan_out; }

/// `refine` slots first, `branch` slots last (reference partition.hpp:186-
/// 194); domain sizes default to the widths where a half left them empty.
inline region::SlotAssignment concatenate(const region::SlotAssignment& refine, const region::SlotAssignment& branch)
{
  region::SlotAssignment out;
  auto append = [&out](const region::SlotAssignment& a) {
    for (std::size_t j = 0; j < a.size(); ++j) {
      out.composition.kinds.push_back(a.composition[j]);
      out.var_ids.push_back(a.var_ids[j]);
      out.widths.push_back(a.widths[j]);
      out.domain_sizes.push_back(a.domain_sizes.empty() ? a.widths[j] : a.domain_sizes[j]);
    }
  };
  append(refine);
  append(branch);
  return out;
}

/// Fold one child's reduced (min masked lb, max masked ub) into its bound
/// (reference backend/aggregate/replay.cuh:303-322): every subregion excluded
/// exactly when the max is still -inf; such a child is reported as
/// {+inf, +inf, infeasible}.
inline AggregateBound fold_child(double min_lb, double max_ub)
{
  if (max_ub > -std::numeric_limits<double>::infinity()) {
    return {min_lb, max_ub, true};
  }
  double const inf = std::numeric_limits<double>::infinity();
  return {inf, inf, false};
}

}  // namespace cmi::aggregate
