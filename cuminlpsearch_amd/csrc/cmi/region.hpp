#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <string>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"

// How a parent box splits into regions. A Composition is the list of slot
// kinds acting this iteration; a SlotAssignment adds which variable fills each
// slot and each slot's width. Region r's digit for slot j is
// (r / prefix[j]) % width[j] with prefix the exclusive product of widths, the
// first slot being the least significant digit (reference
// region/composition.hpp:114-130, region/materialise.hpp:22-40).
namespace cmi::region
{

enum class SlotKind : std::uint8_t
{
  Continuous,
  IntegerPartition,
  IntegerEnumerate,
  BinaryEnumerate,
};

inline char slot_kind_char(SlotKind kind)
{
  switch (kind) {
    case SlotKind::Continuous: return 'C';
    case SlotKind::IntegerPartition: return 'P';
    case SlotKind::IntegerEnumerate: return 'E';
    case SlotKind::BinaryEnumerate: return 'B';
  }
  return '?';
}

struct Composition
{
  std::vector<SlotKind> kinds;

  std::size_t size() const { return kinds.size(); }
  bool operator==(const Composition&) const = default;
  auto begin() const { return kinds.begin(); }
  auto end() const { return kinds.end(); }
  SlotKind operator[](std::size_t i) const { return kinds[i]; }
};

inline std::string spell(const Composition& c)
{
  std::string out(c.size(), '?');
  for (std::size_t i = 0; i < c.size(); ++i) {
    out[i] = slot_kind_char(c[i]);
  }
  return out;
}

struct SlotAssignment
{
  Composition composition;
  std::vector<std::size_t> var_ids;
  std::vector<std::size_t> widths;
  std::vector<std::size_t> domain_sizes;  // empty unless a policy pads enumerate widths

  std::size_t size() const { return composition.size(); }
};

/// The width of a slot as a function of its kind only: every partition slot
/// (continuous or integer) splits into partition_num children, an enumerated
/// integer slot has enumerate_cap children, a binary slot 2 (reference
/// region/fan_out.hpp:22-75).
class FanOutSpec
{
public:
  FanOutSpec(std::size_t partition_num, std::size_t enumerate_cap)
      : partition_num_(partition_num)
      , enumerate_cap_(enumerate_cap)
  {
    if (partition_num_ <= 1) {
      throw InvalidConfiguration("partition_num " + std::to_string(partition_num_)
                                 + " would not split a box; it must be 2 or more");
    }
    if (enumerate_cap_ == 0) {
      throw InvalidConfiguration("enumerate_cap 0 would give an enumerated slot no children; it must be 1 or more");
    }
  }

  /// enumerate_cap defaults to partition_num.
  explicit FanOutSpec(std::size_t partition_num)
      : FanOutSpec(partition_num, partition_num)
  {
  }

  std::size_t partition_num() const { return partition_num_; }
  std::size_t enumerate_cap() const { return enumerate_cap_; }
  bool operator==(const FanOutSpec&) const = default;

private:
  std::size_t partition_num_;
  std::size_t enumerate_cap_;
};

inline std::size_t slot_fan_out(SlotKind kind, const FanOutSpec& f)
{
  if (kind == SlotKind::BinaryEnumerate) {
    return 2;
  }
  if (kind == SlotKind::IntegerEnumerate) {
    return f.enumerate_cap();
  }
  return f.partition_num();
}

/// FNV-1a over (kinds, var_ids, widths): the tripwire that catches a policy
/// that is not a pure function of the box (reference composition.hpp:88-106).
inline std::uint64_t assignment_hash(const SlotAssignment& a)
{
  std::uint64_t h = 1469598103934665603ull;
  auto mix = [&h](std::uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for (SlotKind k : a.composition) {
    mix(static_cast<std::uint64_t>(k));
  }
  for (std::size_t v : a.var_ids) {
    mix(static_cast<std::uint64_t>(v));
  }
  for (std::size_t w : a.widths) {
    mix(static_cast<std::uint64_t>(w));
  }
  return h;
}

/// prefix[j] = product of widths[0..j), saturating at SIZE_MAX.
inline std::vector<std::size_t> slot_prefixes(std::span<const std::size_t> widths)
{
  std::vector<std::size_t> prefix(widths.size());
  std::size_t running = 1;
  for (std::size_t j = 0; j < widths.size(); ++j) {
    prefix[j] = running;
    std::size_t const w = widths[j];
    if (w != 0 && running > std::numeric_limits<std::size_t>::max() / w) {
      running = std::numeric_limits<std::size_t>::max();
    } else {
      running *= w;
    }
  }
  return prefix;
}

inline std::size_t composition_fan_out(std::span<const std::size_t> widths)
{
  std::size_t total = 1;
  for (std::size_t w : widths) {
    if (w != 0 && total > std::numeric_limits<std::size_t>::max() / w) {
      return std::numeric_limits<std::size_t>::max();
    }
    total *= w;
  }
  return total;
}

inline std::size_t composition_fan_out(const Composition& c, const FanOutSpec& f)
{
  std::vector<std::size_t> widths;
  widths.reserve(c.size());
  for (SlotKind k : c) {
    widths.push_back(slot_fan_out(k, f));
  }
  return composition_fan_out(widths);
}

inline bool is_duplicate_child(std::size_t r, const SlotAssignment& a, std::span<const std::size_t> prefix)
{
  if (a.domain_sizes.empty()) {
    return false;
  }
  for (std::size_t j = 0; j < a.composition.size(); ++j) {
    if ((r / prefix[j]) % a.widths[j] >= a.domain_sizes[j]) {
      return true;
    }
  }
  return false;
}

inline bool is_fully_enumerable(const Composition& c)
{
  for (SlotKind k : c) {
    if (k != SlotKind::IntegerEnumerate && k != SlotKind::BinaryEnumerate) {
      return false;
    }
  }
  return true;
}

inline bool can_fathom_without_children(std::size_t live_count, const Composition& c)
{
  return live_count <= c.size() && is_fully_enumerable(c);
}

namespace decode
{

/// One slot's sub-interval of `parent` (reference region/decode.hpp:43-71).
/// Enumerate kinds snap to ceil(parent.lb) + part, clamped to parent.ub;
/// partition kinds split linearly with the top child's ub snapped to the
/// parent's ub. The arithmetic is unfused round-to-nearest; the same
/// expression is evaluated by the device decode (kernels built with
/// -ffp-contract=off), so host and device agree bit for bit.
inline void slot_bounds(SlotKind kind, const interval& parent, std::size_t part, std::size_t fan_out,
                        interval& bound)
{
  switch (kind) {
    case SlotKind::IntegerEnumerate:
    case SlotKind::BinaryEnumerate: {
      double const lo = std::ceil(parent.lb);
      double value = lo + static_cast<double>(part);
      if (value > parent.ub) {
        value = parent.ub;
      }
      bound.lb = value;
      bound.ub = value;
      break;
    }
    case SlotKind::Continuous:
    case SlotKind::IntegerPartition: {
      double const width = (parent.ub - parent.lb) / static_cast<double>(fan_out);
      double const step_lo = width * static_cast<double>(part);
      bound.lb = parent.lb + step_lo;
      if (part + 1 == fan_out) {
        bound.ub = parent.ub;
      } else {
        double const step_hi = width * static_cast<double>(part + 1);
        bound.ub = parent.lb + step_hi;
      }
      break;
    }
  }
}

}  // namespace decode

/// Child `sidx` of `parent` under `assignment`, by repeated division
/// (reference region/materialise.hpp:22-40).
inline void materialise(std::span<const interval> parent, std::size_t sidx, const SlotAssignment& assignment,
                        std::vector<interval>& out)
{
  out.assign(parent.begin(), parent.end());
  std::size_t idx = sidx;
  for (std::size_t j = 0; j < assignment.composition.size(); ++j) {
    std::size_t const width = assignment.widths[j];
    std::size_t const part = idx % width;
    idx /= width;
    std::size_t const dim = assignment.var_ids[j];
    decode::slot_bounds(assignment.composition[j], parent[dim], part, width, out[dim]);
  }
}

}  // namespace cmi::region
