#pragma once

// Tapes compiled to straight-line gfx950 kernels at run time.
//
// The interpreter (kernels.hip) pays, for every tape word, a scalar fetch, a
// uniform dispatch and an LDS round trip for each operand and result. For a
// problem whose boxes are evaluated many times over, the tape is instead
// turned into HIP source -- one statement per tape word over per-lane
// registers, the same interval_ops.hpp operations in the same order, so the
// results are bit-identical to the interpreter's -- compiled once with hiprtc
// for gfx950 and loaded as a module. The variables are read straight from the
// structure-of-arrays box buffers by the wave of each function group that
// needs them; only the groups' verdicts meet in LDS.

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "cmi/tape.hpp"

namespace cmi::jit
{

/// Boxes per call of a compiled kernel (byte offsets of one variable's
/// column fit in 32 bits).
inline constexpr std::uint64_t kMaxBoxes = std::uint64_t(1) << 29;

/// Name of the kernel in the generated source.
inline constexpr const char* kBoxesKernel = "cmi_boxes_compiled";

/// LDS one workgroup may use on gfx950.
inline constexpr std::uint64_t kLdsBytes = 160 * 1024;

/// Launch shape of a compiled explicit-box kernel.
struct Shape
{
  /// Persistent grid-strided workgroups with double-buffered LDS tiles, or one
  /// workgroup per 64-box tile.
  bool persistent = false;
  /// Waves per workgroup (0 = one per function group). Wave v runs groups
  /// v, v + waves, ...; a wave's groups are interleaved statement by statement.
  std::uint32_t waves = 0;
  /// Compile with the instruction scheduler that maximises instruction-level
  /// parallelism (-amdgpu-sched-strategy=max-ilp) rather than the default,
  /// occupancy-first one. The source, and so every result, is the same.
  bool ilp_schedule = false;
  /// Direct form: 0 issues every variable load at the start of the wave; k > 0
  /// issues them k variables at a time, in the order the statements first
  /// read them, each batch as the statements start reading the one before
  /// it, so a wave's first variables arrive first rather than in whatever
  /// order the memory system serves the whole grid's loads.
  std::uint32_t load_batch = 0;
  /// Write the root intervals when the caller asks for them (the kernel then
  /// computes both bounds of every function). Without them only what the
  /// outputs depend on is computed: the compiler drops the upper bounds of
  /// inequality constraints, which nothing reads.
  bool roots = true;
};

/// Waves per workgroup `shape` gives for `t`.
std::uint32_t shape_waves(const tape::Tape& t, const Shape& shape);

/// Whether the persistent form's two LDS tiles of 64 boxes fit.
bool persistent_fits(const tape::Tape& t, std::uint32_t waves);

/// The full translation unit (interval operations + kernel) for `t`: an
/// extern "C" kernel kBoxesKernel(lb, ub, ld, n, feasible, obj_lb, obj_ub,
/// roots) with the explicit-box contract of dev::launch_boxes; workgroups of
/// shape_waves() waves over 64-box tiles. The direct form has one workgroup
/// per tile, each wave loading the variables it reads; the persistent form has
/// a grid-strided loop over tiles that stages the next tile's boxes into a
/// second LDS buffer while the current one is evaluated.
std::string boxes_source(const tape::Tape& t, const Shape& shape);

/// A compiled, loaded explicit-box kernel for one tape on one device.
class CompiledBoxes
{
public:
  CompiledBoxes(const CompiledBoxes&) = delete;
  CompiledBoxes& operator=(const CompiledBoxes&) = delete;

  void launch(const double* lb, const double* ub, std::uint64_t ld, std::uint64_t n, std::uint8_t* feasible,
              double* obj_lb, double* obj_ub, double* roots, hipStream_t stream) const;

  double compile_seconds() const { return compile_seconds_; }
  std::size_t code_bytes() const { return code_bytes_; }
  bool persistent() const { return persistent_; }
  std::uint32_t groups() const { return groups_; }
  std::uint32_t waves() const { return waves_; }
  bool ilp_schedule() const { return ilp_schedule_; }
  /// Whether this kernel writes root intervals (Shape::roots).
  bool roots() const { return roots_; }
  std::uint32_t load_batch() const { return load_batch_; }
  /// Persistent form: resident workgroups per CU the grid is sized for.
  std::uint32_t wgs_per_cu() const { return wgs_per_cu_; }
  void set_wgs_per_cu(std::uint32_t w) { wgs_per_cu_ = w < 1 ? 1 : w; }

private:
  friend std::shared_ptr<CompiledBoxes> compile_boxes(const tape::Tape& t, int device, const Shape& shape);
  CompiledBoxes() = default;

  struct Module;
  std::shared_ptr<const Module> mod_;
  std::uint32_t groups_ = 1;
  std::uint32_t waves_ = 1;
  bool persistent_ = false;
  bool ilp_schedule_ = false;
  bool roots_ = true;
  std::uint32_t load_batch_ = 0;
  std::uint32_t wgs_per_cu_ = 1;
  std::uint32_t n_cus_ = 1;
  double compile_seconds_ = 0.0;
  std::size_t code_bytes_ = 0;
};

/// Name of the decode-mode kernel in the generated source.
inline constexpr const char* kItemsKernel = "cmi_items_compiled";

/// Whether the decode-mode kernel of `t` fits in LDS (the tile's decoded
/// variables, the groups' verdicts and the objective).
bool items_fit(const tape::Tape& t);

/// The decode-mode counterpart of boxes_source: an extern "C" kernel
/// kItemsKernel(items, tile_item, n_tiles) over dev::Item descriptors that
/// all use this tape; a persistent grid whose workgroups each take a
/// contiguous run of 64-region tiles, one wave per function group. For each
/// tile, wave 0 decodes its sub-boxes into LDS (the
/// interpreter's decode, round to nearest); each wave runs its groups as
/// straight-line code on registers; wave 0 then writes per-region
/// feasibility and objective lower bound (`aggregate` false: the bounder) or
/// folds the tile into its child's bound (`aggregate` true): in registers
/// across the workgroup's consecutive tiles of one child when children cover
/// whole tiles, with one atomic min and max per run, and per tile with a
/// segmented reduction otherwise. Results equal the interpreted k_items
/// kernel's bit for bit.
std::string items_source(const tape::Tape& t, bool aggregate);

/// A compiled, loaded decode-mode kernel for one tape on one device.
class CompiledItems
{
public:
  CompiledItems(const CompiledItems&) = delete;
  CompiledItems& operator=(const CompiledItems&) = delete;

  /// `items` are device dev::Item descriptors (their code and groups fields
  /// are not read); `tile_item[t]` names the item of workgroup t.
  void launch(const void* items, const std::uint32_t* tile_item, std::uint32_t n_tiles, hipStream_t stream) const;

  bool aggregate() const { return aggregate_; }
  std::uint32_t waves() const { return waves_; }
  /// Workgroups per CU the persistent grid is sized for.
  std::uint32_t wgs_per_cu() const { return wgs_per_cu_; }
  double compile_seconds() const { return compile_seconds_; }
  /// The loaded module, for a holder that keeps it alive beyond this object.
  std::shared_ptr<const void> module_handle() const { return mod_; }

private:
  friend std::shared_ptr<CompiledItems> compile_items(const tape::Tape& t, int device, bool aggregate, bool compile);
  CompiledItems() = default;

  struct Module;
  std::shared_ptr<const Module> mod_;
  bool aggregate_ = false;
  std::uint32_t waves_ = 1;
  std::uint32_t n_cus_ = 1;
  std::uint32_t wgs_per_cu_ = 1;
  double compile_seconds_ = 0.0;
};

/// Compile (or fetch from the process-wide cache) the decode-mode kernel of
/// `t` and load it on `device`, which must be current. The cache holds a
/// module while some holder keeps it alive. With compile = false a cache miss
/// returns null instead of compiling. Throws DeviceError with the compiler
/// log on failure, InvalidConfiguration if it does not fit.
std::shared_ptr<CompiledItems> compile_items(const tape::Tape& t, int device, bool aggregate, bool compile = true);

/// Compile HIP source with hiprtc for gfx950 (host only, no device needed),
/// with the max-ILP instruction scheduler when `ilp_schedule`; returns the
/// code object. Throws DeviceError with the compiler log.
std::vector<char> compile_code_object(const std::string& src, bool ilp_schedule = false);

/// Compile (or fetch from the process-wide cache, keyed by device, form and
/// tape contents) the explicit-box kernel of `t` and load it on `device`,
/// which must be current. Throws DeviceError with the compiler log on failure.
/// Each call returns its own handle (launch settings are per handle); the
/// loaded module is shared.
std::shared_ptr<CompiledBoxes> compile_boxes(const tape::Tape& t, int device, const Shape& shape);

}  // namespace cmi::jit
