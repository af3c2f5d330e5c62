#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "cmi/tape.hpp"

// Device-side launch descriptors shared by the kernels (kernels.hip) and the
// host code that fills them (hip_backend.cpp, batch.cpp, capi.cpp).
namespace cmi::dev
{

inline constexpr unsigned kTile = 64;          // elements per workgroup: one wave64
inline constexpr unsigned kMaxSlotsPerItem = 64;
inline constexpr unsigned kMaxGroups = 16;     // waves (function groups) per workgroup

/// Function groups of a grouped tape (tape::compile_groups): wave g of an
/// explicit-box workgroup runs code[begin[g], begin[g+1]).
struct GroupTable
{
  std::uint32_t n_groups;
  std::uint32_t objective_group;
  std::uint32_t begin[kMaxGroups + 1];
};

/// One unit of decode-mode work: a parent box split by a slot assignment,
/// evaluated as a bounder (intervals), sampler (points drawn per region) or
/// enumerator (lattice points). A launch covers any number of items, each a
/// run of consecutive 64-element tiles, possibly from different problems.
struct Item
{
  const tape::Instr* code;  // every group's code, then one Nop word
  GroupTable groups;        // wave g of the item's workgroup runs code[begin[g], begin[g+1])
  std::uint32_t n_instr;
  std::uint32_t n_vars;
  const double* parent;  // n_vars (lb, ub) pairs
  const std::uint8_t* var_kinds;  // n_vars, VarKind
  const std::uint32_t* slot_var;  // slot_count entries each
  const std::uint32_t* slot_width;
  const double* slot_step;  // per slot: (ub - lb) / width of its variable's parent bounds
  const std::uint8_t* slot_kind;
  std::uint32_t slot_count;
  std::uint32_t samples;  // points per region (sampler); 1 otherwise
  std::uint64_t n_elems;  // regions * samples
  std::uint64_t salt;
  std::uint32_t tile_begin;
  std::uint32_t n_tiles;
  // bounder outputs, per region
  std::uint8_t* feasible;
  double* obj_lb;
  // sampler / enumerator outputs
  double* tile_min;  // n_tiles partial minima
  std::int64_t* tile_arg;
  std::uint32_t* done;  // zero between launches; the last tile resets it
  double* cand_value;  // 1
  std::int64_t* cand_index;  // 1
  double* witness;  // n_vars
  // aggregate outputs: child c of this item owns regions [c*refine, (c+1)*refine)
  double* child_lb;  // k entries, +inf before the launch; min of masked obj_lb
  double* child_ub;  // k entries, -inf before the launch; max of masked obj_ub
  std::uint64_t refine;
  // Sliced items: a long tape's function groups spread over n_slices
  // workgroups per tile (1: every group in one workgroup, as `groups` says).
  // Slice s of a tile runs groups [s * W, s * W + W), W = the launch's waves
  // per workgroup; the last slice of a tile to finish combines the verdicts
  // and carries on with the tile's epilogue.
  std::uint32_t n_slices;
  std::uint32_t slice_groups;        // function groups of the sliced code
  const std::uint32_t* slice_begin;  // slice_groups + 1 word offsets into code
  std::uint32_t slice_objective;     // the group that reads out the objective
  std::uint32_t slice_pad;
  std::uint64_t* tile_infeasible;    // kSliceMasks per tile: lanes some slice found infeasible, per point of a
                                     // lane; zero between launches
  std::uint32_t* tile_slices_done;   // per tile: slices finished; zero between launches
  double* tile_obj;                  // kSliceObjDoubles per tile: the objective of each lane and point
                                     // (intervals: (lb, ub) pairs)
};

/// Most points per lane of a sampler or enumerator launch.
inline constexpr unsigned kMaxPointsPerLane = 4;
/// A sliced tile's masks (one per point of a lane) and objective doubles.
inline constexpr std::size_t kSliceMasks = kMaxPointsPerLane;
inline constexpr std::size_t kSliceObjDoubles = std::size_t {kTile} * kMaxPointsPerLane;
/// Device scratch of one sliced tile (tile_infeasible, tile_slices_done, tile_obj).
inline constexpr std::size_t kSliceTileBytes = kSliceMasks * 8 + 4 + kSliceObjDoubles * 8;

enum class Mode : int
{
  Bound = 0,
  Sample = 1,
  Enumerate = 2,
  Aggregate = 3,
};

/// Launch `n_tiles` workgroups of `waves` waves over `items`; `tile_item[t]`
/// names the item of global tile t. `lds_slots` must be >= every item's tape
/// slot count and `waves` >= every item's group count (at most kMaxGroups).
/// `extended`: some item's tape uses the extended op set (then the kernels
/// that interpret it run; results of the other tapes are the same either way).
void launch_items(Mode mode, bool extended, const Item* items, const std::uint32_t* tile_item, std::uint32_t n_tiles,
                  std::uint32_t lds_slots, std::uint32_t waves, hipStream_t stream);

/// What launch_items launches, for a graph's kernel node: the kernel (its
/// arguments are items, tile_item, lds_slots), grid, workgroup size and
/// dynamic LDS bytes. `points_per_lane` (sampler and enumerator: 1, 2 or 4;
/// others 1): each lane evaluates that many points, an item's tiles are then
/// kTile * points_per_lane elements each, and no item may be sliced.
struct ItemsLaunch
{
  const void* func = nullptr;
  unsigned grid = 0;
  unsigned block = 0;
  std::size_t lds_bytes = 0;
};
ItemsLaunch items_launch(Mode mode, bool extended, std::uint32_t n_tiles, std::uint32_t lds_slots,
                         std::uint32_t waves, std::uint32_t points_per_lane = 1);

/// LDS bytes of one decode-mode workgroup of `waves` waves over `lds_slots`
/// slots, with `points_per_lane` points per lane (point modes).
std::size_t items_lds_bytes(Mode mode, std::uint32_t lds_slots, std::uint32_t waves,
                            std::uint32_t points_per_lane = 1);

/// Explicit-box evaluation: box e's variable v is [lb[v*ld+e], ub[v*ld+e]].
/// Writes feasible[e] and obj_lb[e]; obj_ub and roots (n_outputs intervals,
/// output-major: roots[k*n_boxes+e]) when non-null. One workgroup of
/// groups.n_groups waves per 64 boxes; `code` holds every group's code
/// followed by one Nop word.
void launch_boxes(const tape::Instr* code, bool extended, const GroupTable& groups, std::uint32_t n_slots, std::uint32_t n_vars,
                  const double* lb, const double* ub, std::uint64_t ld, std::uint64_t n_boxes,
                  std::uint8_t* feasible, double* obj_lb, double* obj_ub, double* roots, hipStream_t stream);

/// LDS bytes of one explicit-box workgroup over a tape with n_slots slots.
std::size_t boxes_lds_bytes(std::uint32_t n_slots);

/// LDS bytes one workgroup may use on the current device.
std::size_t lds_limit_bytes();

/// Points (round to nearest): x[v*ld+e], one wave per 64 points over an
/// ungrouped tape of n_instr words followed by one Nop word.
void launch_points(const tape::Instr* code, bool extended, std::uint32_t n_instr, std::uint32_t n_slots, std::uint32_t n_vars,
                   const double* x, std::uint64_t ld, std::uint64_t n_points, std::uint8_t* feasible,
                   double* obj, double* roots, hipStream_t stream);

/// Largest per-lane slot count a 64-lane workgroup can hold in LDS.
std::uint32_t max_lds_slots(bool interval);

/// Size of the file-scope device array (kernels.hip, `scratch`) that the
/// no-allocation entry points use instead of allocated scratch. Every device
/// the library runs on holds one copy until the process ends.
inline constexpr std::size_t kStaticScratchBytes = std::size_t(4) << 20;
inline constexpr int kMaxDevices = 64;

/// The current device's copy of that array, resolved with
/// hipGetSymbolAddress the first time a device asks (the context's
/// constructor does, with its device current) and kept per device; host code
/// must use this address, never the array's name. Throws DeviceError if the
/// symbol cannot be resolved.
void* static_scratch();

/// Fill n 32-bit words at p with value, as a kernel (capturable, unlike a
/// memset node).
void launch_fill_u32(std::uint32_t* p, std::size_t n, std::uint32_t value, hipStream_t stream);

}  // namespace cmi::dev
