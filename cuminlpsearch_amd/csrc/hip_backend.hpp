#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <span>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "cmi/backend.hpp"
#include "cmi/model.hpp"
#include "cmi/tape.hpp"
#include "device.hpp"
#include "jit.hpp"

// The HIP backend: device-resident tapes and the three roles implemented over
// the tape kernels. HipBackendFactory is a drop-in RegionBackendFactory
// (reference backend/backend.hpp:193-219, its CUDA-graph implementation
// backend/graph/factory.cuh:58-117).
namespace cmi::hip
{

void check(hipError_t e, const char* what);

/// A pinned host buffer (hipHostMalloc) of trivially copyable T.
template<typename T>
class PinnedBuffer
{
public:
  PinnedBuffer() = default;
  explicit PinnedBuffer(std::size_t n) { reset(n); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  PinnedBuffer(PinnedBuffer&& o) noexcept { swap(o); }
  PinnedBuffer& operator=(PinnedBuffer&& o) noexcept
  {
    swap(o);
    return *this;
  }
  ~PinnedBuffer()
  {
    if (p_ != nullptr) {
      (void)hipHostFree(p_);
    }
  }

  void reset(std::size_t n)
  {
    if (p_ != nullptr) {
      (void)hipHostFree(p_);
      p_ = nullptr;
    }
    n_ = n;
    if (n > 0) {
      check(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    }
  }

  T* data() { return p_; }
  const T* data() const { return p_; }
  std::size_t size() const { return n_; }
  T& operator[](std::size_t i) { return p_[i]; }
  const T& operator[](std::size_t i) const { return p_[i]; }

private:
  void swap(PinnedBuffer& o) noexcept
  {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

/// A device allocation (hipMalloc) owned by value.
class DeviceBuffer
{
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(std::size_t bytes) { reset(bytes); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept
  {
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
  {
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
    return *this;
  }
  ~DeviceBuffer()
  {
    if (p_ != nullptr) {
      (void)hipFree(p_);
    }
  }
  void reset(std::size_t bytes)
  {
    if (p_ != nullptr) {
      (void)hipFree(p_);
      p_ = nullptr;
    }
    bytes_ = bytes;
    if (bytes > 0) {
      check(hipMalloc(&p_, bytes), "hipMalloc");
    }
  }
  template<typename T>
  T* as() const
  {
    return static_cast<T*>(p_);
  }
  void* get() const { return p_; }
  std::size_t bytes() const { return bytes_; }

private:
  void* p_ = nullptr;
  std::size_t bytes_ = 0;
};

/// One device and the stream all of this engine's work for it is ordered on.
class Context
{
public:
  explicit Context(int device);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void activate() const;
  std::uint32_t lds_interval_slots() const { return lds_interval_slots_; }
  std::uint32_t lds_point_slots() const { return lds_point_slots_; }
  /// LDS one workgroup may use.
  std::size_t lds_bytes() const { return lds_bytes_; }
  /// Compute units of the device.
  std::uint32_t n_cus() const { return n_cus_; }
  std::size_t free_bytes() const;
  /// This device's copy of the static scratch array (dev::static_scratch),
  /// resolved when the context was created.
  void* static_scratch() const { return scratch_; }

  /// Keep a loaded run-time-compiled module alive with this context, so that
  /// a later tape of the same contents (a repeated solve of one model) finds
  /// it in the compile cache instead of compiling again. The last
  /// kRetainedModules distinct modules are kept; all are released, with the
  /// device current, when the context is destroyed.
  void retain(std::shared_ptr<const void> module) const;
  static constexpr std::size_t kRetainedModules = 32;

private:
  int device_;
  hipStream_t stream_ = nullptr;
  std::uint32_t lds_interval_slots_ = 0;
  std::uint32_t lds_point_slots_ = 0;
  std::size_t lds_bytes_ = 0;
  std::uint32_t n_cus_ = 1;
  void* scratch_ = nullptr;
  mutable std::mutex retained_mu_;
  mutable std::vector<std::shared_ptr<const void>> retained_;  // oldest first
};

/// A problem compiled to a tape and uploaded once.
class DeviceTape
{
public:
  /// box_groups: waves per explicit-box workgroup (0 = chosen from the tape).
  DeviceTape(std::shared_ptr<Context> ctx, const model::Problem<double>& problem,
             const tape::CompileOptions& opt = {}, std::uint32_t box_groups = 0);

  /// Recompile the explicit-box code of `problem` (the one this tape was
  /// built from) into k function groups (0 = automatic); k is clamped to
  /// [1, min(16, 1 + n_constraints)] and lowered until the workgroup's slots
  /// fit in LDS. Not safe while a launch reads the code.
  void set_box_groups(const model::Problem<double>& problem, std::uint32_t k);
  std::uint32_t box_groups() const { return grouped_.n_groups; }
  const tape::Tape& grouped_tape() const { return grouped_; }

  /// Which kernel evaluates explicit boxes (and decode-mode bounder and
  /// aggregate launches, items_kernel): the tape interpreter, the tape
  /// compiled to a straight-line kernel at run time (jit.hpp), or Auto --
  /// compiled once a single call has at least kAutoCompileBoxes boxes and the
  /// tape has at most kMaxCompiledWords words (the first such call waits for
  /// the compiler), the interpreter otherwise. Results are identical.
  enum class BoxEngine : int
  {
    Auto = 0,
    Interpreted = 1,
    Compiled = 2,
  };
  static constexpr std::size_t kAutoCompileBoxes = 32768;
  static constexpr std::size_t kMaxCompiledWords = 20000;
  /// Largest grouped tape whose decode-mode kernel is compiled: the
  /// persistent tile loop makes its compile grow faster with the tape than
  /// the explicit-box kernel's (about 6 s at 1,400 words).
  static constexpr std::size_t kMaxCompiledItemWords = 2048;
  /// Variables per load batch of a new tape's compiled direct form.
  static constexpr std::uint32_t kDefaultLoadBatch = 8;
  /// Compiled compiles now (and throws if that fails).
  void set_box_engine(BoxEngine e);
  BoxEngine box_engine() const { return engine_; }
  /// The compiled kernel without root-interval stores, if one is loaded.
  std::shared_ptr<const jit::CompiledBoxes> compiled() const;

  /// The compiled decode-mode kernel (jit::compile_items) for a bounder
  /// (aggregate = false) or aggregate launch of n_elems elements, or nullptr
  /// to interpret it. Under Auto a tape is compiled once the decode-mode
  /// launches of that kind on it have reached kCompileAfterWork
  /// element-words (elements times tape words, about a tenth of a second of
  /// interpreted work, which a compile costs about as much as), so short
  /// solves never wait for the compiler and long ones stop interpreting
  /// soon; a kernel the context already holds for the same tape contents (a
  /// repeated solve of one model) is taken at the first launch; a failed
  /// compile keeps the interpreter. Under Compiled always,
  /// under Interpreted never; only for grouped tapes within
  /// kMaxCompiledItemWords words whose tile fits in LDS. Results are identical.
  std::shared_ptr<const jit::CompiledItems> items_kernel(bool aggregate, std::size_t n_elems) const;
  static constexpr double kCompileAfterWork = 17179869184.0;  // 2^34 element-words
  /// The engine of decode-mode launches alone; Auto (the default) follows
  /// the box engine.
  void set_items_engine(BoxEngine e);

  /// Shape of the compiled explicit-box kernel: persistent (double-buffered
  /// tiles, grid sized to wgs_per_cu workgroups per CU; only when its LDS
  /// tiles fit) or direct (one workgroup per tile), and waves per workgroup
  /// (0 = one per function group; fewer waves interleave several groups).
  void set_compiled_shape(bool persistent, std::uint32_t waves, std::uint32_t wgs_per_cu);
  /// Code generation of the compiled kernel: the instruction scheduler
  /// (jit::Shape::ilp_schedule) and the direct form's load batching
  /// (jit::Shape::load_batch, at most 64 variables per batch).
  void set_compiled_codegen(bool ilp_schedule, std::uint32_t load_batch);
  jit::Shape compiled_shape() const;
  std::uint32_t compiled_wgs_per_cu() const { return wgs_per_cu_; }

  /// One configuration of the explicit-box kernel and its measured time.
  struct BoxCandidate
  {
    BoxEngine engine = BoxEngine::Interpreted;
    std::uint32_t groups = 1;
    std::uint32_t waves = 0;
    bool persistent = false;
    std::uint32_t wgs_per_cu = 1;
    bool ilp_schedule = false;
    std::uint32_t load_batch = 0;
    double kernel_ms = 0.0;
  };
  /// Autotuner: time the interpreter and compiled shapes (function groups,
  /// waves per workgroup, persistent or direct, workgroups per CU) on n
  /// synthetic boxes inside the problem's bounds, keep the fastest and
  /// return every candidate with its median kernel time. `problem` must be
  /// the one this tape was built from.
  std::vector<BoxCandidate> autotune(const model::Problem<double>& problem, std::size_t n);

  const tape::Tape& tape() const { return tape_; }
  const tape::Instr* code() const { return code_d_; }
  const std::uint8_t* var_kinds() const { return kinds_d_; }
  std::uint32_t n_vars() const { return tape_.n_vars; }
  std::uint32_t n_instr() const { return static_cast<std::uint32_t>(tape_.code.size()); }
  std::uint32_t n_slots() const { return tape_.n_slots; }
  /// What a decode-mode item runs: the tape in its explicit-box function
  /// groups (one wave per group; grouped = true) or the whole tape on one
  /// wave. The values are the same; the groups shorten each wave's run.
  /// Every group runs in the item's one workgroup (n_slices = 1).
  void item_code(dev::Item& it, bool grouped) const;

  /// How a decode-mode item's function groups spread over workgroups of at
  /// most `waves` waves: `slices` workgroups per 64-element tile, slice s
  /// running groups [s * waves, s * waves + waves) of `groups`. A long tape
  /// (kWideMinWords words or more) also has a wide grouping, one group per
  /// function up to kMaxWideGroups, whose temporaries repeat with period
  /// kWidePeriod so that any workgroup of up to that many of its groups holds
  /// them; it is used when `waves` divides kWidePeriod. The values are those
  /// of the whole tape whatever the plan.
  struct ItemPlan
  {
    const tape::Instr* code = nullptr;
    const std::uint32_t* begin = nullptr;  // device: groups + 1 word offsets
    std::uint32_t n_instr = 0;
    std::uint32_t groups = 1;
    std::uint32_t objective = 0;
    std::uint32_t slots = 0;   // LDS slots per lane
    std::uint32_t waves = 1;   // waves per workgroup
    std::uint32_t slices = 1;  // workgroups per tile
    int code_kind = 0;         // 0 whole tape, 1 grouped, 2 wide
  };
  static constexpr std::size_t kWideMinWords = 4096;
  static constexpr std::uint32_t kMaxWideGroups = 256;
  static constexpr std::uint32_t kWideWordsPerGroup = 128;
  static constexpr std::uint32_t kWidePeriod = 4;
  /// The plan for workgroups of `waves` waves (the launch's); 0: every group
  /// of the grouped code in one workgroup when it fits in LDS, as item_code.
  ItemPlan item_plan(dev::Mode mode, std::uint32_t waves) const;
  /// The whole tape on one wave, one workgroup per tile.
  ItemPlan item_plan_whole() const;
  /// This tape has the wide grouping (a long tape with many functions).
  bool has_wide() const { return wide_.n_groups > 1; }
  /// Most function groups any of this tape's codes has.
  std::uint32_t item_groups_max() const { return std::max(grouped_.n_groups, wide_.n_groups); }
  /// Waves per workgroup for a launch of this tape alone over n_tiles tiles
  /// (slice_waves).
  std::uint32_t item_waves_for(dev::Mode mode, std::size_t n_tiles) const;
  /// Waves per workgroup of a decode-mode launch of `tiles` tiles whose
  /// items have at most `groups` function groups and `words` tape words, on
  /// n_cus compute units: 0 (every item's groups in one workgroup) unless
  /// some tape has kSliceMinWords words; then one wave per SIMD (4, the
  /// width the wide grouping's temporaries repeat with) when there are at
  /// most as many tiles as compute units, so that each tile's groups spread
  /// over several of them, 8 when the tiles cover at most two workgroups per
  /// unit, else 0. Measured per launch kind on a stream of 64 QCQPs of up to
  /// 96 variables at 2,000 req/s (sampler ~256 tiles and bounder ~64 tiles
  /// both at 4: 0.59 ms of kernel time per step; sampler at 8: 1.05 ms; both
  /// at 0: 2.03 ms) and 512 batched ones of up to 48 (bounder, 512 tiles: 8
  /// 0.68 ms, 0 0.71 ms, 4 0.78 ms per step), and on streams of up to 16
  /// variables (tapes of at most 531 words: slicing cost 10 % of the served
  /// rate).
  static std::uint32_t slice_waves(std::uint32_t groups, std::size_t tiles, std::uint32_t n_cus, std::size_t words);
  static constexpr std::size_t kSliceMinWords = 2048;
  /// Fill the Item's code fields from a plan (slice scratch is the caller's).
  void item_code(dev::Item& it, const ItemPlan& plan) const;
  std::uint32_t item_slots(bool grouped) const { return grouped ? grouped_.n_slots : tape_.n_slots; }
  std::uint32_t item_waves(bool grouped) const { return grouped ? groups_.n_groups : 1u; }
  /// Whether a decode-mode workgroup of this tape's groups fits in LDS.
  bool items_grouped_fit(dev::Mode mode) const;
  /// The tape uses the extended op set (Div, PowN, transcendental, ...).
  bool extended() const { return tape_.extended; }
  const std::shared_ptr<Context>& context() const { return ctx_; }

  /// C2-style evaluation of explicit boxes held in device memory.
  void eval_boxes_device(const double* lb, const double* ub, std::size_t ld, std::size_t n, std::uint8_t* feasible,
                         double* obj_lb, double* obj_ub, double* roots, hipStream_t stream) const;
  void eval_points_device(const double* x, std::size_t ld, std::size_t n, std::uint8_t* feasible, double* obj,
                          double* roots, hipStream_t stream) const;

private:
  std::shared_ptr<Context> ctx_;
  tape::CompileOptions opt_;
  tape::Tape tape_;
  tape::Tape grouped_;
  dev::GroupTable groups_ {};
  BoxEngine engine_ = BoxEngine::Auto;
  bool want_persistent_ = true;
  std::uint32_t want_waves_ = 0;
  bool want_ilp_ = false;
  std::uint32_t want_load_batch_ = kDefaultLoadBatch;
  std::uint32_t wgs_per_cu_ = 1;
  mutable std::mutex jit_mu_;
  // compiled kernels without and with root-interval stores (the second is
  // compiled the first time a call asks for root intervals)
  mutable std::shared_ptr<const jit::CompiledBoxes> compiled_;
  mutable std::shared_ptr<const jit::CompiledBoxes> compiled_roots_;
  mutable bool jit_failed_ = false;
  // compiled decode-mode kernels: [0] bounder, [1] aggregate
  mutable std::shared_ptr<const jit::CompiledItems> items_[2];
  mutable bool items_failed_[2] = {false, false};
  mutable double items_work_[2] = {0.0, 0.0};  // element-words launched, interpreted
  mutable bool items_looked_up_[2] = {false, false};  // compile cache searched once
  BoxEngine items_engine_ = BoxEngine::Auto;

  std::shared_ptr<const jit::CompiledBoxes> compile_now(bool roots) const;
  void drop_compiled() const;
  /// The tape in `k` explicit-box function groups (0: tape::auto_groups),
  /// as set_box_groups describes.
  tape::Tape grouping(const model::Problem<double>& problem, std::uint32_t k) const;
  /// Make `g` (its code and group table already on the device) the grouped code.
  void install_groups(tape::Tape g, const tape::Instr* code, const std::uint32_t* begin);
  // The code, the first grouping, the wide grouping and the variable kinds,
  // in one allocation (image_); a later regrouping has buffers of its own.
  DeviceBuffer image_;
  const tape::Instr* code_d_ = nullptr;
  const tape::Instr* grouped_code_d_ = nullptr;
  const std::uint32_t* grouped_begin_d_ = nullptr;  // grouped_.group_begin on the device
  const tape::Instr* wide_code_d_ = nullptr;
  const std::uint32_t* wide_begin_d_ = nullptr;
  const std::uint8_t* kinds_d_ = nullptr;
  DeviceBuffer grouped_code_;   // set_box_groups' code
  DeviceBuffer grouped_begin_;
  tape::Tape wide_;             // n_groups 0: none
};

/// Bytes a role of each kind needs for n_regions (affine in n_regions).
backend::RegionCostModel cost_model_for(const tape::Tape& t, std::size_t samples_per_region);

/// Every check a role build makes before it allocates, in the same order:
/// the shape (InvalidConfiguration), the LDS the point roles need
/// (ResourceExhausted) and the device budget (OverBudgetError with its facts;
/// budget_bytes 0 = the device's free memory now). samples_per_region is the
/// solve-wide sample count; only the sampler allocates per sample. The
/// batched and served paths call it per request, so a solve that would not
/// fit stops exactly as the plain driver stops it.
void check_role_fits(const DeviceTape& tape, dev::Mode mode, const region::Composition& composition,
                     std::size_t n_regions, std::size_t samples_per_region, std::size_t budget_bytes);

/// The sampler, bounder and enumerator over one DeviceTape and a fixed
/// composition shape (slot count, n_regions), reusable across calls; results
/// are views into this object's pinned staging, valid until its next call.
class RoleImpl
{
public:
  RoleImpl(std::shared_ptr<const DeviceTape> tape, dev::Mode mode, const region::Composition& composition,
           std::size_t n_regions, std::size_t samples_per_region, std::size_t budget_bytes);

  dev::Mode mode() const { return mode_; }
  std::size_t n_regions() const { return n_regions_; }
  std::size_t n_elems() const { return n_elems_; }
  std::size_t slot_count() const { return slot_count_; }

  /// Upload the parent box and slot data, launch, and copy results back.
  void run(std::span<const interval> box, std::span<const std::size_t> var_ids, std::span<const std::size_t> widths,
           std::uint64_t salt);

  std::span<const unsigned char> feasible() const { return {feasible_h_.data(), n_regions_}; }
  std::span<const double> obj_lb() const { return {obj_lb_h_.data(), n_regions_}; }
  double cand_value() const { return out_h_[0].value; }
  std::int64_t cand_index() const { return out_h_[0].index; }
  std::span<const double> witness() const { return {witness_h_.data(), n_vars_}; }
  bool has_candidate() const { return out_h_[0].value < 1.7976931348623157e308; }

private:
  struct Candidate
  {
    double value;
    std::int64_t index;
  };

  std::shared_ptr<const DeviceTape> tape_;
  dev::Mode mode_;
  region::Composition composition_;
  std::size_t n_regions_;
  std::size_t samples_;
  std::size_t solve_samples_;
  std::size_t n_elems_;
  std::size_t slot_count_;
  std::size_t n_vars_;
  std::uint32_t n_tiles_;
  std::uint32_t max_slices_ = 1;  // workgroups per tile the tile_item array covers
  DeviceBuffer staging_d_;   // Item + parent box + slot arrays
  DeviceBuffer tile_item_d_;
  DeviceBuffer slice_d_;     // sliced items' per-tile scratch (zero between launches)
  DeviceBuffer feasible_d_, obj_lb_d_;
  DeviceBuffer tile_min_d_, tile_arg_d_, done_d_, out_d_, witness_d_;
  PinnedBuffer<unsigned char> staging_h_;
  PinnedBuffer<unsigned char> feasible_h_;
  PinnedBuffer<double> obj_lb_h_;
  PinnedBuffer<Candidate> out_h_;
  PinnedBuffer<double> witness_h_;
  std::size_t off_parent_ = 0, off_var_ = 0, off_width_ = 0, off_step_ = 0, off_kind_ = 0, staging_bytes_ = 0;
};

/// The sampler's launch without a role and without allocating: staging, tile
/// partials, the done counter and the candidate live in the device's static
/// scratch array (dev::kStaticScratchBytes). Same checks and results as a
/// sampler RoleImpl over the same arguments; a launch whose scratch would
/// exceed the array throws InvalidConfiguration before anything runs. Calls
/// on one device are serialised; value, index and point (n_vars doubles) are
/// host memory the caller owns.
void sample_static(const DeviceTape& tape, const region::Composition& composition, std::span<const interval> box,
                   std::span<const std::size_t> var_ids, std::span<const std::size_t> widths,
                   std::size_t samples_per_region, std::uint64_t salt, double* value, std::int64_t* index,
                   double* point);

/// Static-scratch bytes sample_static needs for these sizes.
std::size_t sample_static_bytes(std::size_t n_vars, std::size_t slot_count, std::size_t n_elems);

class HipBackendFactory : public backend::RegionBackendFactory
{
public:
  explicit HipBackendFactory(std::shared_ptr<Context> ctx, tape::CompileOptions opt = {});

  backend::BackendCapabilities capabilities() const override { return {}; }
  backend::RegionCostModel cost_model(const model::Problem<double>& problem) const override;
  backend::SubdivisionBundle build_subdivision(const model::Problem<double>& problem,
                                               const region::Composition& composition, std::size_t n_regions,
                                               const backend::BuildBudget& budget,
                                               backend::RoleRequest roles) const override;

  /// The tape this factory compiled for `problem` (compiled on first use).
  std::shared_ptr<const DeviceTape> tape_for(const model::Problem<double>& problem) const;

private:
  std::shared_ptr<Context> ctx_;
  tape::CompileOptions opt_;
  mutable const model::Problem<double>* cached_problem_ = nullptr;
  mutable std::shared_ptr<const DeviceTape> cached_tape_;
};

}  // namespace cmi::hip
