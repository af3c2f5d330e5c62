#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <memory>
#include <span>
#include <string>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/region.hpp"

// Composition policies: a pure function from a box to a SlotAssignment.
// Only GreedyEnumerate is built; its rules decide node order, so they follow
// the reference (policy/greedy_enum.hpp:54-186) exactly, including the use of
// std::sort (not stable_sort) with the same comparators, which is what fixes
// the order among equal-width candidates.
namespace cmi::policy
{

struct SearchCalibration
{
  std::size_t max_cycle_size = 0;  // 0 = kMaxSlots
  std::size_t free_device_bytes = 0;
  std::size_t multiprocessor_count = 0;
  std::size_t enumerate_cap = 0;
};

inline constexpr std::size_t kMaxSlots = 64;

class CompositionPolicy
{
public:
  explicit CompositionPolicy(SearchCalibration calibration = {})
      : calibration_(calibration)
  {
    if (calibration_.max_cycle_size == 0) {
      calibration_.max_cycle_size = kMaxSlots;
    }
    if (calibration_.max_cycle_size > kMaxSlots) {
      throw InvalidConfiguration("max_cycle_size (" + std::to_string(calibration_.max_cycle_size)
                                 + ") exceeds kMaxSlots (" + std::to_string(kMaxSlots) + ")");
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

/// Binaries first (index order), then integers by smallest remaining domain,
/// then continuous by widest; stops at max_cycle_size slots.
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

  region::SlotAssignment choose(std::span<const interval> box,
                                std::span<const model::VarKind> var_kinds) const override
  {
    if (var_kinds.empty()) {
      throw ShapeMismatch("CompositionPolicy::choose called with an empty var_kinds span");
    }
    if (box.size() != var_kinds.size()) {
      throw ShapeMismatch("CompositionPolicy::choose: box and var_kinds differ in length");
    }
    region::SlotAssignment out {};
    fill_binary(box, var_kinds, out);
    fill_integer(box, var_kinds, out);
    fill_continuous(box, var_kinds, out);
    return out;
  }

  /// Integers in [ceil(lb), floor(ub)]; 0 when there are none.
  static std::size_t integer_domain_size(const interval& b)
  {
    double const lo = std::ceil(b.lb);
    double const hi = std::floor(b.ub);
    if (hi < lo) {
      return 0;
    }
    return static_cast<std::size_t>(hi - lo) + 1;
  }

private:
  static bool unresolved(const interval& b) { return b.ub > b.lb; }

  void append(region::SlotAssignment& out, region::SlotKind kind, std::size_t vid) const
  {
    out.composition.kinds.push_back(kind);
    out.var_ids.push_back(vid);
    out.widths.push_back(region::slot_fan_out(kind, fan_out_));
  }

  bool at_cap(const region::SlotAssignment& out) const { return out.var_ids.size() >= max_cycle_size(); }

  void fill_binary(std::span<const interval> box, std::span<const model::VarKind> kinds,
                   region::SlotAssignment& out) const
  {
    for (std::size_t vid = 0; vid < kinds.size() && !at_cap(out); ++vid) {
      if (kinds[vid] == model::VarKind::Binary && unresolved(box[vid])) {
        append(out, region::SlotKind::BinaryEnumerate, vid);
      }
    }
  }

  void fill_integer(std::span<const interval> box, std::span<const model::VarKind> kinds,
                    region::SlotAssignment& out) const
  {
    std::vector<std::size_t> candidates;
    for (std::size_t vid = 0; vid < kinds.size(); ++vid) {
      if (kinds[vid] == model::VarKind::Integer && unresolved(box[vid])) {
        candidates.push_back(vid);
      }
    }
    std::sort(candidates.begin(), candidates.end(), [&](std::size_t a, std::size_t b) {
      return integer_domain_size(box[a]) < integer_domain_size(box[b]);
    });
    std::size_t const cap = fan_out_.enumerate_cap();
    for (std::size_t vid : candidates) {
      if (at_cap(out)) {
        break;
      }
      region::SlotKind const kind = integer_domain_size(box[vid]) <= cap ? region::SlotKind::IntegerEnumerate
                                                                         : region::SlotKind::IntegerPartition;
      append(out, kind, vid);
    }
  }

  void fill_continuous(std::span<const interval> box, std::span<const model::VarKind> kinds,
                       region::SlotAssignment& out) const
  {
    std::vector<std::size_t> candidates;
    for (std::size_t vid = 0; vid < kinds.size(); ++vid) {
      if (kinds[vid] == model::VarKind::Continuous && unresolved(box[vid])) {
        candidates.push_back(vid);
      }
    }
    std::sort(candidates.begin(), candidates.end(), [&](std::size_t a, std::size_t b) {
      return (box[a].ub - box[a].lb) > (box[b].ub - box[b].lb);
    });
    for (std::size_t vid : candidates) {
      if (at_cap(out)) {
        break;
      }
      append(out, region::SlotKind::Continuous, vid);
    }
  }

  region::FanOutSpec fan_out_;
};

}  // namespace cmi::policy
