#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <span>
#include <string>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/region.hpp"

// Composition policies: given a box and the variable kinds, which variables
// are split this iteration, how (slot kinds) and how wide (slot widths). A
// policy must be a pure function of its inputs, because a pending node's box
// is rebuilt later by asking the policy again about the parent box.
//
// GreedyEnumerate (reference policy/greedy_enum.hpp:54-186), the default:
// live binaries in variable order, then live integers with the fewest lattice
// points first (enumerated when they have at most enumerate_cap points,
// partitioned otherwise), then live continuous variables widest first, until
// max_cycle_size slots are filled. The order among variables that tie in the
// ranking is whatever std::sort makes of the ascending variable order under
// those two comparisons; it is kept, because it decides which child is which.
namespace cmi::policy
{

/// Run shape parameters a policy reads (reference config/calibration.hpp).
struct SearchCalibration
{
  std::size_t max_cycle_size = 0;  // 0 = kMaxSlots
  std::size_t free_device_bytes = 0;
  std::size_t multiprocessor_count = 0;
  std::size_t enumerate_cap = 0;
};

/// Most slots one composition may have; the device decode holds this many.
inline constexpr std::size_t kMaxSlots = 64;

class CompositionPolicy
{
public:
  explicit CompositionPolicy(SearchCalibration calibration = {})
      : calibration_(calibration)
  {
    std::size_t& cap = calibration_.max_cycle_size;
    if (cap > kMaxSlots) {
      throw InvalidConfiguration("a composition holds at most " + std::to_string(kMaxSlots)
                                 + " slots; max_cycle_size " + std::to_string(cap) + " asks for more");
    }
    if (cap == 0) {
      cap = kMaxSlots;
    }
  }

  virtual ~CompositionPolicy() = default;

  const SearchCalibration& calibration() const { return calibration_; }
  std::size_t max_cycle_size() const { return calibration_.max_cycle_size; }

  virtual region::SlotAssignment choose(std::span<const interval> box,
                                        std::span<const model::VarKind> var_kinds) const = 0;
  virtual std::size_t n_regions(const region::Composition& composition) const = 0;

private:
  SearchCalibration calibration_;
};

/// True when every slot names a different variable and each one is still live
/// (lb < ub) in `box` (reference policy/policy.hpp: the contract every policy
/// keeps).
inline bool assignment_is_distinct_and_live(const region::SlotAssignment& a, std::span<const interval> box)
{
  std::vector<std::size_t> seen(a.var_ids.begin(), a.var_ids.end());
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) {
    return false;
  }
  return std::all_of(a.var_ids.begin(), a.var_ids.end(),
                     [&](std::size_t v) { return v < box.size() && box[v].lb < box[v].ub; });
}

class GreedyEnumCompositionPolicy : public CompositionPolicy
{
public:
  explicit GreedyEnumCompositionPolicy(region::FanOutSpec fan_out, SearchCalibration calibration = {})
      : CompositionPolicy(calibration)
      , fan_out_(fan_out)
  {
  }

  const region::FanOutSpec& fan_out() const { return fan_out_; }

  std::size_t n_regions(const region::Composition& composition) const override
  {
    return region::composition_fan_out(composition, fan_out_);
  }

  /// Lattice points of an integer variable's box, ceil(lb)..floor(ub); 0 for
  /// a box between two integers.
  static std::size_t integer_domain_size(const interval& b)
  {
    double const first = std::ceil(b.lb);
    double const last = std::floor(b.ub);
    return last < first ? 0 : static_cast<std::size_t>(last - first) + 1;
  }

  region::SlotAssignment choose(std::span<const interval> box,
                                std::span<const model::VarKind> var_kinds) const override
  {
    if (var_kinds.empty()) {
      throw ShapeMismatch("a composition cannot be chosen for a problem without variables");
    }
    if (box.size() != var_kinds.size()) {
      throw ShapeMismatch("the box has " + std::to_string(box.size()) + " coordinates for "
                          + std::to_string(var_kinds.size()) + " variables");
    }

    // Live variables of each kind, in variable order.
    std::vector<std::size_t> by_kind[3];
    for (std::size_t v = 0; v < var_kinds.size(); ++v) {
      auto const k = static_cast<std::size_t>(var_kinds[v]);
      if (k < 3 && box[v].ub > box[v].lb) {
        by_kind[k].push_back(v);
      }
    }
    std::vector<std::size_t>& integers = by_kind[static_cast<std::size_t>(model::VarKind::Integer)];
    std::vector<std::size_t>& continuous = by_kind[static_cast<std::size_t>(model::VarKind::Continuous)];
    std::sort(integers.begin(), integers.end(), [&](std::size_t x, std::size_t y) {
      return integer_domain_size(box[x]) < integer_domain_size(box[y]);
    });
    std::sort(continuous.begin(), continuous.end(), [&](std::size_t x, std::size_t y) {
      return (box[x].ub - box[x].lb) > (box[y].ub - box[y].lb);
    });

    region::SlotAssignment out {};
    std::size_t const slots = max_cycle_size();
    auto add = [&](std::size_t v, region::SlotKind kind) {
      out.composition.kinds.push_back(kind);
      out.var_ids.push_back(v);
      out.widths.push_back(region::slot_fan_out(kind, fan_out_));
    };
    for (std::size_t v : by_kind[static_cast<std::size_t>(model::VarKind::Binary)]) {
      if (out.size() == slots) {
        return out;
      }
      add(v, region::SlotKind::BinaryEnumerate);
    }
    for (std::size_t v : integers) {
      if (out.size() == slots) {
        return out;
      }
      bool const small = integer_domain_size(box[v]) <= fan_out_.enumerate_cap();
      add(v, small ? region::SlotKind::IntegerEnumerate : region::SlotKind::IntegerPartition);
    }
    for (std::size_t v : continuous) {
      if (out.size() == slots) {
        return out;
      }
      add(v, region::SlotKind::Continuous);
    }
    return out;
  }

private:
  region::FanOutSpec fan_out_;
};

/// WidthFirst (reference policy/width_first.hpp): every live variable is a
/// candidate, the widest first, until max_cycle_size slots are filled. A
/// continuous variable partitions; an integer enumerates only when it is the
/// unique widest live variable and has at most enumerate_cap lattice points,
/// and partitions otherwise; a binary joins only once every other live
/// variable is at most 1 wide or is an integer that enumerates this
/// iteration. Widths follow the slot kinds (partition_num, enumerate_cap, 2).
/// The order among equal widths is std::sort's over the ascending variable
/// order, kept because it decides which child is which.
class WidthFirstCompositionPolicy : public CompositionPolicy
{
public:
  explicit WidthFirstCompositionPolicy(region::FanOutSpec fan_out, SearchCalibration calibration = {})
      : CompositionPolicy(calibration)
      , fan_out_(fan_out)
  {
  }

  const region::FanOutSpec& fan_out() const { return fan_out_; }

  std::size_t n_regions(const region::Composition& composition) const override
  {
    return region::composition_fan_out(composition, fan_out_);
  }

  region::SlotAssignment choose(std::span<const interval> box,
                                std::span<const model::VarKind> var_kinds) const override
  {
    if (var_kinds.empty()) {
      throw ShapeMismatch("a composition cannot be chosen for a problem without variables");
    }
    if (box.size() != var_kinds.size()) {
      throw ShapeMismatch("the box has " + std::to_string(box.size()) + " coordinates for "
                          + std::to_string(var_kinds.size()) + " variables");
    }
    auto width = [&](std::size_t v) { return box[v].ub - box[v].lb; };
    std::vector<std::size_t> live;
    double widest = 0.0;
    for (std::size_t v = 0; v < box.size(); ++v) {
      if (box[v].ub > box[v].lb) {
        live.push_back(v);
        widest = std::max(widest, width(v));
      }
    }
    std::size_t at_widest = 0;
    for (std::size_t v : live) {
      at_widest += width(v) < widest ? 0 : 1;
    }
    auto enumerates = [&](std::size_t v) {
      return var_kinds[v] == model::VarKind::Integer && at_widest == 1 && !(width(v) < widest)
             && GreedyEnumCompositionPolicy::integer_domain_size(box[v]) <= fan_out_.enumerate_cap();
    };
    auto binary_joins = [&](std::size_t v) {
      return std::all_of(live.begin(), live.end(),
                         [&](std::size_t u) { return u == v || !(width(u) > 1.0) || enumerates(u); });
    };
    struct Candidate
    {
      std::size_t var;
      region::SlotKind kind;
      double width;
    };
    std::vector<Candidate> cand;
    for (std::size_t v : live) {
      switch (var_kinds[v]) {
        case model::VarKind::Continuous: cand.push_back({v, region::SlotKind::Continuous, width(v)}); break;
        case model::VarKind::Integer:
          cand.push_back(
              {v, enumerates(v) ? region::SlotKind::IntegerEnumerate : region::SlotKind::IntegerPartition, width(v)});
          break;
        case model::VarKind::Binary:
          if (binary_joins(v)) {
            cand.push_back({v, region::SlotKind::BinaryEnumerate, width(v)});
          }
          break;
      }
    }
    std::sort(cand.begin(), cand.end(), [](const Candidate& a, const Candidate& b) { return a.width > b.width; });
    region::SlotAssignment out;
    for (const Candidate& c : cand) {
      if (out.size() >= max_cycle_size()) {
        break;
      }
      out.composition.kinds.push_back(c.kind);
      out.var_ids.push_back(c.var);
      out.widths.push_back(region::slot_fan_out(c.kind, fan_out_));
    }
    return out;
  }

private:
  region::FanOutSpec fan_out_;
};

/// BisectionBudget (reference policy/bisection_budget.hpp:40-299): every
/// non-leaf node splits into exactly N = 2^B children, the B bisections shared
/// out per node. Enumerable discrete variables are taken first, cheapest
/// (fewest bisections, ceil(log2(domain)), at least 1) first, while their cost
/// fits the budget; an integer that does not fit is bisected instead. The rest
/// of the budget goes one bisection at a time to the candidate with the widest
/// current width (absolute, or relative to its width at the start of the node
/// when relative_width is set), through std::make_heap / pop_heap / push_heap
/// with a less-than on the key, so ties go where those algorithms put them.
/// Budget left once every candidate is resolved pads the taken enumerate slot
/// with the most bisections (the first such on ties); the padded children are
/// duplicates (domain_sizes < widths). Candidates are gathered in variable
/// order up to max_cycle_size; an integer with more than enumerate_cap points
/// (calibration; 0 = no ceiling) is bisected rather than enumerated.
class BisectionBudgetCompositionPolicy : public CompositionPolicy
{
public:
  explicit BisectionBudgetCompositionPolicy(std::size_t bisection_budget, SearchCalibration calibration = {},
                                            bool relative_width = false)
      : CompositionPolicy(calibration)
      , budget_(bisection_budget)
      , relative_(relative_width)
  {
    if (budget_ < 1 || budget_ > 62) {
      throw InvalidConfiguration("bisection budget " + std::to_string(budget_)
                                 + " is outside [1, 62]: a node needs at least one bisection and 2^B regions must be "
                                   "countable in 64 bits");
    }
  }

  std::size_t bisection_budget() const { return budget_; }
  bool relative_width() const { return relative_; }

  std::size_t n_regions(const region::Composition& composition) const override
  {
    return composition.size() == 0 ? 1 : std::size_t(1) << budget_;
  }

  /// Bisections an enumerate slot over d points costs: ceil(log2(max(d, 1))),
  /// at least 1.
  static std::size_t enumerate_cost(std::size_t d)
  {
    std::size_t bits = 0;
    while ((std::size_t(1) << bits) < std::max<std::size_t>(d, 1)) {
      ++bits;
    }
    return bits == 0 ? 1 : bits;
  }

  region::SlotAssignment choose(std::span<const interval> box,
                                std::span<const model::VarKind> var_kinds) const override
  {
    if (var_kinds.empty()) {
      throw ShapeMismatch("a composition cannot be chosen for a problem without variables");
    }
    if (box.size() != var_kinds.size()) {
      throw ShapeMismatch("the box has " + std::to_string(box.size()) + " coordinates for "
                          + std::to_string(var_kinds.size()) + " variables");
    }
    struct Enumerated
    {
      std::size_t var;
      region::SlotKind kind;
      std::size_t domain;
      std::size_t cost;
    };
    struct Split
    {
      std::size_t var;
      region::SlotKind kind;
      double width;  // halved by each bisection it receives
      double start;  // its width when the node began
      std::size_t bisections;
    };
    std::vector<Enumerated> enumerable;
    std::vector<Split> splittable;
    std::size_t const slots = max_cycle_size();
    std::size_t const ceiling = calibration().enumerate_cap;
    for (std::size_t v = 0; v < var_kinds.size() && enumerable.size() + splittable.size() < slots; ++v) {
      if (!(box[v].ub > box[v].lb)) {
        continue;
      }
      double const w = box[v].ub - box[v].lb;
      if (var_kinds[v] == model::VarKind::Binary) {
        enumerable.push_back({v, region::SlotKind::BinaryEnumerate, 2, 1});
      } else if (var_kinds[v] == model::VarKind::Integer) {
        std::size_t const d = GreedyEnumCompositionPolicy::integer_domain_size(box[v]);
        if (ceiling == 0 || std::max<std::size_t>(d, 1) <= ceiling) {
          enumerable.push_back({v, region::SlotKind::IntegerEnumerate, d, enumerate_cost(d)});
        } else {
          splittable.push_back({v, region::SlotKind::IntegerPartition, w, w, 0});
        }
      } else if (var_kinds[v] == model::VarKind::Continuous) {
        splittable.push_back({v, region::SlotKind::Continuous, w, w, 0});
      }
    }
    region::SlotAssignment out {};
    if (enumerable.empty() && splittable.empty()) {
      return out;
    }

    std::stable_sort(enumerable.begin(), enumerable.end(),
                     [](const Enumerated& x, const Enumerated& y) { return x.cost < y.cost; });
    std::size_t budget = budget_;
    std::size_t taken = 0;
    while (taken < enumerable.size() && enumerable[taken].cost <= budget) {
      budget -= enumerable[taken].cost;
      ++taken;
    }
    for (std::size_t i = taken; i < enumerable.size(); ++i) {
      if (enumerable[i].kind == region::SlotKind::IntegerEnumerate) {
        double const w = box[enumerable[i].var].ub - box[enumerable[i].var].lb;
        splittable.push_back({enumerable[i].var, region::SlotKind::IntegerPartition, w, w, 0});
      }
    }

    auto key = [&](std::size_t i) {
      const Split& c = splittable[i];
      return relative_ ? c.width / c.start : c.width;
    };
    auto narrower = [&](std::size_t x, std::size_t y) { return key(x) < key(y); };
    std::vector<std::size_t> order(splittable.size());
    for (std::size_t i = 0; i < order.size(); ++i) {
      order[i] = i;
    }
    std::make_heap(order.begin(), order.end(), narrower);
    std::size_t given = 0;
    while (given < budget && !order.empty()) {
      std::pop_heap(order.begin(), order.end(), narrower);
      Split& widest = splittable[order.back()];
      widest.width = widest.width / 2.0;
      widest.bisections += 1;
      std::push_heap(order.begin(), order.end(), narrower);
      ++given;
    }
    for (; given < budget && taken > 0; ++given) {
      std::size_t pick = 0;
      for (std::size_t i = 1; i < taken; ++i) {
        if (enumerable[i].cost > enumerable[pick].cost) {
          pick = i;
        }
      }
      enumerable[pick].cost += 1;
    }

    for (std::size_t i = 0; i < taken; ++i) {
      out.composition.kinds.push_back(enumerable[i].kind);
      out.var_ids.push_back(enumerable[i].var);
      out.widths.push_back(std::size_t(1) << enumerable[i].cost);
      out.domain_sizes.push_back(enumerable[i].domain);
    }
    for (const Split& c : splittable) {
      if (c.bisections == 0) {
        continue;
      }
      std::size_t const w = std::size_t(1) << c.bisections;
      out.composition.kinds.push_back(c.kind);
      out.var_ids.push_back(c.var);
      out.widths.push_back(w);
      out.domain_sizes.push_back(w);
    }
    return out;
  }

private:
  std::size_t budget_;
  bool relative_;
};

}  // namespace cmi::policy
