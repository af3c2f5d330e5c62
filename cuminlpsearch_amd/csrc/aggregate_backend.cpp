#include "aggregate_backend.hpp"

#include <cstring>
#include <limits>
#include <string>

namespace cmi::hip
{

namespace
{
std::size_t align_up(std::size_t x, std::size_t a) { return (x + a - 1) / a * a; }
}  // namespace

AggregateBounderImpl::AggregateBounderImpl(std::shared_ptr<const DeviceTape> tape, const region::Composition& launch,
                                           std::size_t n_regions, std::size_t k, std::size_t max_batch,
                                           std::size_t budget_bytes)
    : tape_(std::move(tape))
    , launch_(launch)
    , n_(n_regions)
    , k_(k)
    , m_(1)
    , max_batch_(max_batch)
    , slot_count_(launch.size())
    , n_vars_(tape_->n_vars())
{
  // Same precondition as the reference's build (backend/aggregate/replay.cuh:49-55).
  if (k_ < 1 || n_ == 0 || n_ % k_ != 0) {
    throw InvalidConfiguration(
        "aggregate bounder: the branch fan-out k must divide the total region count N, so that every child owns "
        "exactly M = N / k contiguous subregions");
  }
  if (n_ > (std::size_t(1) << 31)) {
    throw InvalidConfiguration("aggregate bounder: N must be below 2^31");
  }
  if (max_batch_ < 1 || max_batch_ > 65536) {
    throw InvalidConfiguration("aggregate bounder: max_batch must be in [1, 65536]");
  }
  if (slot_count_ > dev::kMaxSlotsPerItem) {
    throw InvalidConfiguration("composition has more slots than the device decode supports");
  }
  m_ = n_ / k_;
  tiles_per_region_ = static_cast<std::uint32_t>((n_ + dev::kTile - 1) / dev::kTile);
  std::uint64_t const total_tiles = std::uint64_t(tiles_per_region_) * max_batch_;
  if (total_tiles > std::numeric_limits<std::uint32_t>::max() / 2) {
    throw InvalidConfiguration("aggregate bounder: max_batch * N is too large for one launch");
  }

  // Staging: items, then 2k child results per region, then each region's
  // parent box and slot arrays. No per-subregion buffer exists: what the
  // bounder holds does not grow with N beyond the tile map.
  off_items_ = 0;
  off_child_ = align_up(max_batch_ * sizeof(dev::Item), 256);
  off_data_ = align_up(off_child_ + max_batch_ * 2 * k_ * sizeof(double), 256);
  item_stride_ = align_up(n_vars_ * 16 + slot_count_ * (4 + 4 + 8 + 1) + 64, 64);
  staging_bytes_ = off_data_ + max_batch_ * item_stride_;
  tile_bytes_ = total_tiles * sizeof(std::uint32_t);

  std::size_t const needed = staging_bytes_ + tile_bytes_;
  Context const& ctx = *tape_->context();
  std::size_t const budget = budget_bytes != 0 ? budget_bytes : ctx.free_bytes();
  if (needed > budget) {
    backend::OverBudget f;
    f.role = "aggregate bounder";
    f.needed_bytes = needed;
    f.budget_bytes = budget;
    f.n_regions = n_;
    f.bytes_per_element = 0;
    f.element_breakdown = "per-batch staging of parent boxes and slot data, plus one ticket per tile";
    f.element_inventory = std::to_string(max_batch_) + " regions per batch";
    f.composition = launch_;
    throw backend::OverBudgetError(f);
  }

  ctx.activate();
  staging_d_.reset(staging_bytes_);
  staging_h_.reset(staging_bytes_);
  std::memset(staging_h_.data(), 0, staging_bytes_);
  child_h_.reset(max_batch_ * 2 * k_);
  bounds_.resize(max_batch_ * k_);

  std::vector<std::uint32_t> tile_item(total_tiles);
  for (std::uint64_t t = 0; t < total_tiles; ++t) {
    tile_item[t] = static_cast<std::uint32_t>(t / tiles_per_region_);
  }
  tile_item_d_.reset(tile_bytes_);
  check(hipMemcpy(tile_item_d_.get(), tile_item.data(), tile_bytes_, hipMemcpyHostToDevice), "hipMemcpy(tile_item)");

  auto* base = staging_d_.as<unsigned char>();
  for (std::size_t i = 0; i < max_batch_; ++i) {
    auto* item = reinterpret_cast<dev::Item*>(staging_h_.data() + off_items_) + i;
    std::size_t const d = off_data_ + i * item_stride_;
    std::size_t const o_parent = d;
    std::size_t const o_var = align_up(o_parent + n_vars_ * 16, 16);
    std::size_t const o_width = align_up(o_var + slot_count_ * 4, 16);
    std::size_t const o_step = align_up(o_width + slot_count_ * 4, 16);
    std::size_t const o_kind = align_up(o_step + slot_count_ * 8, 16);
    tape_->item_code(*item, false);
    item->n_vars = static_cast<std::uint32_t>(n_vars_);
    item->parent = reinterpret_cast<const double*>(base + o_parent);
    item->var_kinds = tape_->var_kinds();
    item->slot_var = reinterpret_cast<const std::uint32_t*>(base + o_var);
    item->slot_width = reinterpret_cast<const std::uint32_t*>(base + o_width);
    item->slot_step = reinterpret_cast<const double*>(base + o_step);
    item->slot_kind = reinterpret_cast<const std::uint8_t*>(base + o_kind);
    item->slot_count = static_cast<std::uint32_t>(slot_count_);
    item->samples = 1;
    item->n_elems = n_;
    item->salt = 0;
    item->tile_begin = static_cast<std::uint32_t>(i * tiles_per_region_);
    item->n_tiles = tiles_per_region_;
    item->child_lb = reinterpret_cast<double*>(base + off_child_) + i * 2 * k_;
    item->child_ub = item->child_lb + k_;
    item->refine = m_;
    auto* kinds = staging_h_.data() + o_kind;
    for (std::size_t j = 0; j < slot_count_; ++j) {
      kinds[j] = static_cast<std::uint8_t>(launch_[j]);
    }
  }
}

std::span<const aggregate::AggregateBound> AggregateBounderImpl::bound_children(std::span<const RegionView> regions)
{
  std::size_t const r_count = regions.size();
  if (r_count == 0) {
    return {};
  }
  if (r_count > max_batch_) {
    throw ShapeMismatch("bound_children: " + std::to_string(r_count) + " regions; this bounder takes at most "
                        + std::to_string(max_batch_) + " per call");
  }
  for (auto const& r : regions) {
    if (r.box.size() != n_vars_) {
      throw ShapeMismatch("bound_children: domain size does not match the problem's variable count");
    }
    if (r.var_ids.size() != slot_count_ || r.widths.size() != slot_count_) {
      throw ShapeMismatch("bound_children: var_ids/widths size does not match this bounder's slot count");
    }
    if (region::composition_fan_out(r.widths) != n_) {
      throw ShapeMismatch("bound_children: product of launch widths does not equal N");
    }
    for (std::size_t v : r.var_ids) {
      if (v >= n_vars_) {
        throw ShapeMismatch("bound_children: slot var_id out of range");
      }
    }
  }
  Context const& ctx = *tape_->context();
  ctx.activate();
  hipStream_t const s = ctx.stream();

  unsigned char* h = staging_h_.data();
  double const inf = std::numeric_limits<double>::infinity();
  // the tape's current function groups (a regrouping replaces their code)
  bool const grouped = tape_->items_grouped_fit(dev::Mode::Aggregate);
  for (std::size_t i = 0; i < r_count; ++i) {
    tape_->item_code(reinterpret_cast<dev::Item*>(h + off_items_)[i], grouped);
  }
  for (std::size_t i = 0; i < r_count; ++i) {
    auto const& r = regions[i];
    auto* child = reinterpret_cast<double*>(h + off_child_) + i * 2 * k_;
    for (std::size_t c = 0; c < k_; ++c) {
      child[c] = inf;        // identity of the min over masked lower bounds
      child[k_ + c] = -inf;  // identity of the max over masked upper bounds
    }
    std::size_t const d = off_data_ + i * item_stride_;
    std::size_t const o_var = align_up(d + n_vars_ * 16, 16);
    std::size_t const o_width = align_up(o_var + slot_count_ * 4, 16);
    std::size_t const o_step = align_up(o_width + slot_count_ * 4, 16);
    std::memcpy(h + d, r.box.data(), n_vars_ * sizeof(interval));
    auto* var = reinterpret_cast<std::uint32_t*>(h + o_var);
    auto* width = reinterpret_cast<std::uint32_t*>(h + o_width);
    auto* step = reinterpret_cast<double*>(h + o_step);
    for (std::size_t j = 0; j < slot_count_; ++j) {
      var[j] = static_cast<std::uint32_t>(r.var_ids[j]);
      width[j] = static_cast<std::uint32_t>(r.widths[j]);
      step[j] = (r.box[r.var_ids[j]].ub - r.box[r.var_ids[j]].lb) / static_cast<double>(r.widths[j]);
    }
  }
  // Only the used regions' items, results and data travel.
  std::size_t const used = off_data_ + r_count * item_stride_;
  check(hipMemcpyAsync(staging_d_.get(), h, used, hipMemcpyHostToDevice, s), "hipMemcpyAsync(aggregate staging)");
  std::uint32_t const n_tiles = static_cast<std::uint32_t>(r_count * tiles_per_region_);
  if (auto const compiled = tape_->items_kernel(true, n_)) {
    compiled->launch(staging_d_.get(), tile_item_d_.as<std::uint32_t>(), n_tiles, s);
  } else {
    dev::launch_items(dev::Mode::Aggregate, tape_->extended(), staging_d_.as<dev::Item>(),
                      tile_item_d_.as<std::uint32_t>(), n_tiles, tape_->item_slots(grouped), tape_->item_waves(grouped),
                      s);
    check(hipGetLastError(), "launch(k_items aggregate)");
  }
  check(hipMemcpyAsync(child_h_.data(), staging_d_.as<unsigned char>() + off_child_, r_count * 2 * k_ * sizeof(double),
                       hipMemcpyDeviceToHost, s),
        "hipMemcpyAsync(child bounds)");
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
  for (std::size_t i = 0; i < r_count; ++i) {
    for (std::size_t c = 0; c < k_; ++c) {
      bounds_[i * k_ + c] = aggregate::fold_child(child_h_[i * 2 * k_ + c], child_h_[i * 2 * k_ + k_ + c]);
    }
  }
  return {bounds_.data(), r_count * k_};
}

}  // namespace cmi::hip
