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
// GreedyEnumerate (reference policy/greedy_enum.hpp:54-186) is the one built:
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

}  // namespace cmi::policy
