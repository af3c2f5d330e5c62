#pragma once

#include <cstddef>
#include <memory>
#include <span>
#include <vector>

#include "cmi/aggregate.hpp"
#include "cmi/region.hpp"
#include "hip_backend.hpp"

// The aggregate bounder on the device: for each of a batch of regions (a box
// and its launch assignment of N = k * M subregions), the k child bounds,
// reduced in the item kernel's Aggregate mode so only 2k doubles per region
// cross PCIe (reference backend/aggregate/replay.cuh:255-360, whose
// bound_children takes one region per launch; this one takes a batch,
// aggregate/bound.hpp:74-100).
namespace cmi::hip
{

class AggregateBounderImpl
{
public:
  /// launch: the slot kinds of every region's launch assignment; N = product
  /// of its widths (checked per call); k children of M = N / k subregions
  /// each; up to max_batch regions per call. Throws OverBudgetError before
  /// allocating when its device bytes exceed the budget (0 = free memory).
  AggregateBounderImpl(std::shared_ptr<const DeviceTape> tape, const region::Composition& launch,
                       std::size_t n_regions, std::size_t k, std::size_t max_batch, std::size_t budget_bytes);

  struct RegionView
  {
    std::span<const interval> box;
    std::span<const std::size_t> var_ids;
    std::span<const std::size_t> widths;
  };

  /// regions.size() * k bounds, region-major (entry i * k + c is region i's
  /// child c); a view valid until the next call.
  std::span<const aggregate::AggregateBound> bound_children(std::span<const RegionView> regions);

  std::size_t n_subregions() const { return n_; }
  std::size_t branch_fan_out() const { return k_; }
  std::size_t max_batch() const { return max_batch_; }
  std::size_t slot_count() const { return slot_count_; }
  /// Device bytes this bounder holds.
  std::size_t device_bytes() const { return staging_bytes_ + tile_bytes_; }

private:
  std::shared_ptr<const DeviceTape> tape_;
  region::Composition launch_;
  std::size_t n_, k_, m_, max_batch_, slot_count_, n_vars_;
  std::uint32_t tiles_per_region_;
  std::size_t item_stride_ = 0, off_items_ = 0, off_child_ = 0, off_data_ = 0, staging_bytes_ = 0, tile_bytes_ = 0;
  DeviceBuffer staging_d_;
  DeviceBuffer tile_item_d_;
  PinnedBuffer<unsigned char> staging_h_;
  PinnedBuffer<double> child_h_;
  std::vector<aggregate::AggregateBound> bounds_;
};

}  // namespace cmi::hip
