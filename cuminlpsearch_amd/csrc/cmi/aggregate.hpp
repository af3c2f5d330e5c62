#pragma once

#include <limits>

// The k-child aggregate: one launch covers N = k * M subregions of a node and
// the device reduces them to k child bounds. Child c owns launch regions
// [c*M, (c+1)*M), because the branch slots are the most significant digits of
// the launch composition (reference aggregate/bound.hpp:25-129,
// aggregate/partition.hpp:165).
namespace cmi::aggregate
{

/// One child's bound as the reference's AggregateBound carries it
/// (aggregate/bound.hpp:42-64). `hull_ub` encloses the objective over the
/// child's non-excluded subregions; it is attained at no known point and is
/// never an incumbent.
struct AggregateBound
{
  double lb = std::numeric_limits<double>::max();
  double hull_ub = std::numeric_limits<double>::max();
  bool feasible = true;
};

/// Turns a child's device reduction into its bound. The kernel reduces each
/// child's M subregions to (least lb, greatest ub) over the subregions that
/// were not proven infeasible, with excluded ones contributing +inf and -inf.
/// If nothing survived, the greatest ub is still -inf: the child is excluded,
/// and the reference reports such a child as {+inf, +inf, infeasible}
/// (backend/aggregate/replay.cuh:303-322).
inline AggregateBound fold_child(double least_lb, double greatest_ub)
{
  constexpr double inf = std::numeric_limits<double>::infinity();
  AggregateBound b;
  b.feasible = greatest_ub > -inf;
  b.lb = b.feasible ? least_lb : inf;
  b.hull_ub = b.feasible ? greatest_ub : inf;
  return b;
}

}  // namespace cmi::aggregate
