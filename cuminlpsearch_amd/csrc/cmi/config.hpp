#pragma once

// The memory-fit resolver: which search shape (partition_num, enumerate_cap,
// sample_points, max_cycle_size) a problem gets on a device with a given
// amount of free HBM, and how many device bytes a solve can come to hold.
// Host only, a pure function of the problem's variable counts, the roles'
// affine cost coefficients (backend::RegionCostModel, priced from the
// problem's own tape) and a byte budget, so it runs before any device
// allocation and without a GPU.
//
// Same rules as the reference's resolver (config/footprint.hpp:61-149,
// config/resolve.hpp:95-245, config/catalogue.hpp:1-275,
// config/problem_profile.hpp:21-70): the worst composition any box of the
// search can reach at a slot cap, the widest cap whose worst composition
// fits, and a two-phase fit that first buys coverage (slots) at the cheapest
// fan-out, then widens partition_num as far as it can without giving coverage
// back.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <optional>
#include <string>
#include <string_view>

#include "cmi/backend.hpp"
#include "cmi/model.hpp"
#include "cmi/policy.hpp"
#include "cmi/region.hpp"

namespace cmi::config
{

inline constexpr std::size_t kSizeMax = std::numeric_limits<std::size_t>::max();

inline std::size_t sat_mul(std::size_t a, std::size_t b)
{
  if (a != 0 && b > kSizeMax / a) {
    return kSizeMax;
  }
  return a * b;
}

inline std::size_t sat_add(std::size_t a, std::size_t b) { return b > kSizeMax - a ? kSizeMax : a + b; }

using policy::kMaxSlots;
/// Fraction of free device memory the fit spends (reference resolve.hpp:36).
inline constexpr double kBudgetFraction = 0.67;
/// Upper end of the partition_num scan (reference catalogue.hpp:85).
inline constexpr std::size_t kPartitionCeiling = 2000000;

/// What the resolver needs to know about a problem (reference
/// problem_profile.hpp:21-32).
struct ProblemProfile
{
  std::size_t num_binary = 0;
  std::size_t num_integer = 0;
  std::size_t num_continuous = 0;
  std::size_t largest_integer_domain = 0;  // integer points in the widest integer box
  bool objvar_kept = false;                // set by a GAMS caller, not derivable from a Problem
};

inline ProblemProfile profile_of(const model::Problem<double>& p)
{
  ProblemProfile out;
  for (std::size_t i = 0; i < p.var_kinds.size(); ++i) {
    switch (p.var_kinds[i]) {
      case model::VarKind::Binary: ++out.num_binary; break;
      case model::VarKind::Integer: {
        ++out.num_integer;
        // integer points in [ceil(lb), floor(ub)]; an unbounded or
        // astronomically wide box saturates
        double const lo = std::ceil(p.box_bounds[i].lb);
        double const hi = std::floor(p.box_bounds[i].ub);
        double const span = hi - lo;
        std::size_t const d = !(span >= 0.0) ? 0 : span >= 0x1p62 ? kSizeMax : static_cast<std::size_t>(span) + 1;
        out.largest_integer_domain = std::max(out.largest_integer_domain, d);
        break;
      }
      case model::VarKind::Continuous: ++out.num_continuous; break;
    }
  }
  return out;
}

/// Device bytes one composition's roles hold together: the sampler and the
/// bounder always, the enumerator too when the composition is fully
/// enumerable (reference cost_model.hpp:61-75, with per-role fixed parts).
/// A composition wider than the roles can launch costs everything.
inline std::size_t bundle_bytes(const backend::RegionCostModel& c, std::size_t n_regions, bool enumerable)
{
  if (n_regions > c.max_regions) {
    return kSizeMax;  // wider than one launch indexes: never fits
  }
  std::size_t total = sat_add(c.sampler_fixed, sat_mul(n_regions, c.sampler_per_region));
  total = sat_add(total, sat_add(c.bounder_fixed, sat_mul(n_regions, c.bounder_per_region)));
  if (enumerable) {
    total = sat_add(total, sat_add(c.enumerator_fixed, sat_mul(n_regions, c.enumerator_per_region)));
  }
  return total;
}

/// The largest bundle any box of the search can need with `cap` slots: the
/// policy fills slots integers first, then continuous variables, then
/// binaries, and a box deep in the tree can have handed its discrete slots to
/// wider continuous ones, so the worst composition is the widest prefix of
/// that order; the all-discrete prefix is a separate maximum, since it is
/// narrower but also holds the enumerator (reference footprint.hpp:61-97).
inline std::size_t worst_footprint(std::size_t cap, const ProblemProfile& pp, const backend::RegionCostModel& cost,
                                   std::size_t partition_num, std::size_t enumerate_cap)
{
  std::size_t const int_width = std::max(partition_num, enumerate_cap);
  std::size_t const ni = pp.num_integer, nc = pp.num_continuous, nb = pp.num_binary;
  std::size_t widest = 1, discrete = 1, worst = 0;
  for (std::size_t s = 1; s <= cap; ++s) {
    widest = sat_mul(widest, s <= ni ? int_width : (s <= ni + nc ? partition_num : std::size_t {2}));
    worst = std::max(worst, bundle_bytes(cost, widest, s <= ni || nc == 0));
    if (s <= ni + nb) {
      discrete = sat_mul(discrete, s <= ni ? int_width : std::size_t {2});
      worst = std::max(worst, bundle_bytes(cost, discrete, true));
    }
  }
  return worst;
}

/// The widest slot cap (at most `ceiling` and the number of variables) whose
/// worst composition fits in `budget`; 0 if not even one slot does
/// (reference footprint.hpp:99-149). worst_footprint grows with the cap, so
/// the scan stops at the first cap that does not fit.
inline std::size_t auto_max_cycle_size(const ProblemProfile& pp, const backend::RegionCostModel& cost,
                                       std::size_t partition_num, std::size_t enumerate_cap, std::size_t budget,
                                       std::size_t ceiling)
{
  std::size_t const slots = std::min(ceiling, pp.num_binary + pp.num_integer + pp.num_continuous);
  std::size_t best = 0;
  for (std::size_t cap = 1; cap <= slots; ++cap) {
    if (worst_footprint(cap, pp, cost, partition_num, enumerate_cap) > budget) {
      break;
    }
    best = cap;
  }
  return best;
}

/// How each shape number is chosen (reference catalogue.hpp:12-60).
struct ShapeRules
{
  enum class Partition
  {
    FitToCoverage,
    Pin
  };
  enum class Enumerate
  {
    CoverDomains,     // enumerate_cap = min(largest integer domain, ceiling)
    FollowPartition,  // enumerate_cap = partition_num
    Pin
  };
  enum class Cycle
  {
    FitDevice,  // the widest cap the budget fits (pinned is the fallback without a budget)
    Pin
  };
  std::string_view name;
  Partition partition = Partition::FitToCoverage;
  std::size_t partition_pinned = 0;
  Enumerate enumerate = Enumerate::CoverDomains;
  std::size_t enumerate_ceiling = 16;
  std::size_t enumerate_pinned = 0;
  std::size_t sample_points = 1;
  Cycle cycle = Cycle::FitDevice;
  std::size_t cycle_pinned = 1;
};

/// The rows of the reference's roster (catalogue.hpp:87-149): five
/// GreedyEnumerate rows that select_rules chooses among by position, and
/// width-first, reachable only by name, which runs the WidthFirst policy on
/// one slot per iteration.
inline constexpr ShapeRules kRoster[] = {
    {"all-binary", ShapeRules::Partition::Pin, 2, ShapeRules::Enumerate::Pin, 0, 2, 1,
     ShapeRules::Cycle::FitDevice, 20},
    {"discrete", ShapeRules::Partition::FitToCoverage, 0, ShapeRules::Enumerate::CoverDomains, 16, 0, 5,
     ShapeRules::Cycle::FitDevice, 7},
    {"mixed-binary", ShapeRules::Partition::FitToCoverage, 0, ShapeRules::Enumerate::FollowPartition, 0, 0, 5,
     ShapeRules::Cycle::FitDevice, 4},
    {"mixed-all-small", ShapeRules::Partition::FitToCoverage, 0, ShapeRules::Enumerate::CoverDomains, 16, 0, 10,
     ShapeRules::Cycle::FitDevice, 4},
    {"mixed-all-large", ShapeRules::Partition::FitToCoverage, 0, ShapeRules::Enumerate::CoverDomains, 16, 0, 3,
     ShapeRules::Cycle::FitDevice, 4},
    {"width-first", ShapeRules::Partition::FitToCoverage, 0, ShapeRules::Enumerate::CoverDomains, 16, 0, 10,
     ShapeRules::Cycle::Pin, 1},
};

inline std::optional<ShapeRules> lookup_rules(std::string_view name)
{
  for (const ShapeRules& r : kRoster) {
    if (r.name == name) {
      return r;
    }
  }
  return std::nullopt;
}

/// Continuous variables that count for classification: a kept objective
/// variable is a slot but no degree of freedom (reference catalogue.hpp:176-180).
inline std::size_t classified_continuous(const ProblemProfile& pp)
{
  return pp.num_continuous - (pp.objvar_kept && pp.num_continuous > 0 ? 1 : 0);
}

/// The roster row for a problem's variable kinds (reference catalogue.hpp:182-245).
inline const ShapeRules& select_rules(const ProblemProfile& pp)
{
  std::size_t const cont = classified_continuous(pp);
  std::size_t const live = pp.num_binary + pp.num_integer + pp.num_continuous;
  if (cont == 0 && pp.num_integer == 0) {
    return kRoster[0];
  }
  if (cont == 0) {
    return kRoster[1];
  }
  if (pp.num_binary > 0 && pp.num_integer == 0) {
    return kRoster[2];
  }
  return live <= kMaxSlots ? kRoster[3] : kRoster[4];
}

/// Whether a named row's rules suit the problem's variable kinds (reference
/// catalogue.hpp:247-261).
inline bool is_applicable(const ShapeRules& r, const ProblemProfile& pp)
{
  std::size_t const cont = classified_continuous(pp);
  if (r.name == "all-binary") {
    return cont == 0 && pp.num_integer == 0;
  }
  if (r.name == "discrete") {
    return cont == 0;
  }
  if (r.name == "mixed-binary") {
    return pp.num_integer == 0;
  }
  return true;
}

struct Shape
{
  std::size_t partition_num = 0;
  std::size_t enumerate_cap = 0;
  std::size_t sample_points = 0;
  std::size_t max_cycle_size = 0;
};

/// The shape for `pp` under `rules` with `free_bytes` of free device memory
/// (0 = no budget to fit against: the rules' fallbacks). `cost` prices the
/// roles at rules.sample_points. The budget is kBudgetFraction of the free
/// bytes. Phase 1 fits the widest cap at partition_num 2; phase 2 widens
/// partition_num while the fitted cap stays at phase 1's. A fitted 0 becomes 1
/// at partition_num 2, so an unfittable problem meets the role build's own
/// out-of-memory error for its cheapest composition (reference
/// resolve.hpp:95-245).
inline Shape resolve_shape(const ShapeRules& rules, const ProblemProfile& pp, std::size_t free_bytes,
                           const backend::RegionCostModel& cost)
{
  std::size_t const target = rules.cycle == ShapeRules::Cycle::Pin
      ? rules.cycle_pinned
      : std::min(pp.num_binary + pp.num_integer + pp.num_continuous, kMaxSlots);
  bool const follow = rules.enumerate == ShapeRules::Enumerate::FollowPartition;
  std::size_t enumerate_cap = 0;
  if (rules.enumerate == ShapeRules::Enumerate::CoverDomains) {
    enumerate_cap = std::min(std::max<std::size_t>(pp.largest_integer_domain, 1), rules.enumerate_ceiling);
  } else if (rules.enumerate == ShapeRules::Enumerate::Pin) {
    enumerate_cap = rules.enumerate_pinned;
  }
  std::size_t const budget = free_bytes > 0 ? static_cast<std::size_t>(static_cast<double>(free_bytes) * kBudgetFraction) : 0;
  auto fit = [&](std::size_t q) {
    return auto_max_cycle_size(pp, cost, q, follow ? q : enumerate_cap, budget, target);
  };
  Shape s;
  if (rules.partition == ShapeRules::Partition::Pin) {
    s.partition_num = rules.partition_pinned;
    if (follow) {
      enumerate_cap = s.partition_num;
    }
    s.max_cycle_size = free_bytes > 0 ? fit(s.partition_num) : rules.cycle_pinned;
  } else if (free_bytes == 0) {
    s.partition_num = 2;
    if (follow) {
      enumerate_cap = 2;
    }
    s.max_cycle_size = rules.cycle_pinned;
  } else {
    std::size_t const achieved = fit(2);
    s.partition_num = 2;
    // With nothing fitted at the cheapest fan-out there is no coverage to
    // keep; widening would only make the role build's report name a larger
    // composition, so the fan-out stays at 2 (the reference's loop widens to
    // its ceiling here).
    for (std::size_t q = 3; achieved > 0 && q <= kPartitionCeiling; ++q) {
      if (fit(q) < achieved) {
        break;
      }
      s.partition_num = q;
    }
    if (follow) {
      enumerate_cap = s.partition_num;
    }
    s.max_cycle_size = achieved;
  }
  s.max_cycle_size = std::max<std::size_t>(s.max_cycle_size, 1);
  s.enumerate_cap = enumerate_cap;
  s.sample_points = rules.sample_points;
  region::FanOutSpec {s.partition_num, s.enumerate_cap};  // a pin below the fan-out floors throws here
  return s;
}

/// Where a run's shape came from (reference run_spec.hpp:40-60).
enum class Source
{
  Auto,        // rules selected by the problem's variable kinds
  Named,       // rules named by the caller
  Overridden,  // at least one number given by the caller
};

/// Numbers a caller fixes by hand; 0 leaves a number to the resolver.
struct Overrides
{
  std::size_t partition_num = 0;
  std::size_t enumerate_cap = 0;
  std::size_t sample_points = 0;
  std::size_t max_cycle_size = 0;
  bool any() const { return partition_num || enumerate_cap || sample_points || max_cycle_size; }
};

/// The resolved shape with a caller's overrides laid over it. A given
/// partition_num without an enumerate_cap carries enumerate_cap along with
/// it; the fitted max_cycle_size is not refitted for an overridden fan-out,
/// so an override is the caller's responsibility (reference
/// run_spec.hpp:122-166).
inline Shape apply_overrides(Shape s, const Overrides& o)
{
  if (o.partition_num != 0) {
    s.partition_num = o.partition_num;
    s.enumerate_cap = o.enumerate_cap != 0 ? o.enumerate_cap : o.partition_num;
  } else if (o.enumerate_cap != 0) {
    s.enumerate_cap = o.enumerate_cap;
  }
  if (o.sample_points != 0) {
    s.sample_points = o.sample_points;
  }
  if (o.max_cycle_size != 0) {
    s.max_cycle_size = o.max_cycle_size;
  }
  region::FanOutSpec {s.partition_num, s.enumerate_cap};
  return s;
}

/// A role build that did not fit, explained: the composition's size as a
/// product of widths, and the widest slot cap that would fit this problem
/// with every role a composition holds charged at once (reference
/// resolve.hpp:247-330).
inline std::string explain_over_budget(const backend::OverBudget& f, const ProblemProfile& pp,
                                       const backend::RegionCostModel& cost, std::size_t partition_num,
                                       std::size_t enumerate_cap)
{
  std::string out = "the " + f.role + " for composition " + region::spell(f.composition) + " needs " +
                    std::to_string(f.needed_bytes) + " bytes (" + std::to_string(f.n_regions) + " regions x " +
                    std::to_string(f.bytes_per_element) + " bytes per element) against a budget of " +
                    std::to_string(f.budget_bytes) + " bytes";
  std::size_t const cap = auto_max_cycle_size(pp, cost, partition_num, enumerate_cap, f.budget_bytes, kMaxSlots);
  if (cap == 0) {
    out += "; not even one slot per iteration fits at this fan-out";
  } else {
    out += "; at most " + std::to_string(cap) + " slot(s) per iteration fit (max_cycle_size)";
  }
  return out;
}

}  // namespace cmi::config
