#include "hip_backend.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <limits>
#include <stdexcept>

#include "cmi/qcqp.hpp"

namespace cmi::hip
{

void check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    throw DeviceError(static_cast<int>(e), std::string(what) + ": " + hipGetErrorString(e));
  }
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

Context::Context(int device)
    : device_(device)
{
  int count = 0;
  check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (device < 0 || device >= count) {
    throw InvalidConfiguration("device index " + std::to_string(device) + " out of range (" + std::to_string(count)
                               + " devices)");
  }
  check(hipSetDevice(device), "hipSetDevice");
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  lds_interval_slots_ = dev::max_lds_slots(true);
  lds_point_slots_ = dev::max_lds_slots(false);
  lds_bytes_ = dev::lds_limit_bytes();
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) {
    n_cus_ = static_cast<std::uint32_t>(cus);
  }
  // Resolve this device's copy of the static scratch now, with the device
  // current, so the no-allocation entry points never look it up mid-call.
  scratch_ = dev::static_scratch();
}

Context::~Context()
{
  (void)hipSetDevice(device_);
  retained_.clear();
  if (stream_ != nullptr) {
    (void)hipStreamDestroy(stream_);
  }
}

void Context::retain(std::shared_ptr<const void> module) const
{
  if (!module) {
    return;
  }
  std::lock_guard<std::mutex> lock(retained_mu_);
  if (auto it = std::find(retained_.begin(), retained_.end(), module); it != retained_.end()) {
    std::rotate(it, it + 1, retained_.end());  // most recently used last
    return;
  }
  if (retained_.size() >= kRetainedModules) {
    retained_.erase(retained_.begin());
  }
  retained_.push_back(std::move(module));
}

void Context::activate() const { check(hipSetDevice(device_), "hipSetDevice"); }

std::size_t Context::free_bytes() const
{
  activate();
  std::size_t free_b = 0, total_b = 0;
  check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  return free_b;
}

// ---------------------------------------------------------------------------
// DeviceTape
// ---------------------------------------------------------------------------

namespace
{

/// Upload 32-bit words (a group table).
void upload_words(const std::vector<std::uint32_t>& words, DeviceBuffer& buf)
{
  buf.reset(std::max<std::size_t>(words.size(), 1) * sizeof(std::uint32_t));
  if (!words.empty()) {
    check(hipMemcpy(buf.get(), words.data(), words.size() * sizeof(std::uint32_t), hipMemcpyHostToDevice),
          "hipMemcpy(group table)");
  }
}

/// Upload a tape's words followed by one Nop word (the interpreters read one
/// word ahead).
void upload_code(const std::vector<tape::Instr>& code, DeviceBuffer& buf)
{
  std::vector<tape::Instr> padded(code);
  padded.push_back(tape::Instr {});
  buf.reset(padded.size() * sizeof(tape::Instr));
  check(hipMemcpy(buf.get(), padded.data(), padded.size() * sizeof(tape::Instr), hipMemcpyHostToDevice),
        "hipMemcpy(tape)");
}

}  // namespace

DeviceTape::DeviceTape(std::shared_ptr<Context> ctx, const model::Problem<double>& problem,
                       const tape::CompileOptions& opt, std::uint32_t box_groups)
    : ctx_(std::move(ctx))
    , opt_(opt)
{
  problem.validate();
  ctx_->activate();  // the calling thread may be any host thread
  tape_ = tape::compile(problem, opt);
  if (tape_.n_slots > ctx_->lds_interval_slots() || dev::boxes_lds_bytes(tape_.n_slots) > ctx_->lds_bytes()) {
    throw ResourceExhausted("tape needs " + std::to_string(tape_.n_slots) + " slots per lane; LDS holds "
                            + std::to_string(ctx_->lds_interval_slots()) + " interval slots per 64-lane workgroup");
  }
  tape::Tape grouped = grouping(problem, box_groups);
  std::uint32_t const n_roots = 1 + tape_.n_constraints;
  if (tape_.code.size() >= kWideMinWords && n_roots > dev::kMaxGroups) {
    auto const by_size = static_cast<std::uint32_t>((tape_.code.size() + kWideWordsPerGroup - 1) / kWideWordsPerGroup);
    std::uint32_t const k = std::min({n_roots, kMaxWideGroups, std::max(by_size, dev::kMaxGroups + 1)});
    wide_ = tape::compile_groups(problem, k, opt_, kWidePeriod, &tape_);
  }
  // Every array the kernels read, in one allocation filled by one copy (a
  // batch builds hundreds of tapes at once, and each allocation and copy
  // takes the runtime's locks). Codes carry one trailing Nop word.
  std::size_t at = 0;
  auto place = [&at](std::size_t bytes) {
    std::size_t const off = (at + 255) / 256 * 256;
    at = off + bytes;
    return off;
  };
  auto code_bytes = [](const tape::Tape& t) { return (t.code.size() + 1) * sizeof(tape::Instr); };
  bool const wide = wide_.n_groups > 1;
  std::size_t const o_code = place(code_bytes(tape_));
  std::size_t const o_grouped = place(code_bytes(grouped));
  std::size_t const o_grouped_begin = place(grouped.group_begin.size() * sizeof(std::uint32_t));
  std::size_t const o_wide = wide ? place(code_bytes(wide_)) : 0;
  std::size_t const o_wide_begin = wide ? place(wide_.group_begin.size() * sizeof(std::uint32_t)) : 0;
  std::size_t const o_kinds = place(std::max<std::size_t>(problem.var_kinds.size(), 1));
  std::vector<unsigned char> host(at, 0);  // zero bytes are the Nop word
  auto put = [&host](std::size_t off, const void* src, std::size_t bytes) {
    if (bytes > 0) {
      std::memcpy(host.data() + off, src, bytes);
    }
  };
  put(o_code, tape_.code.data(), tape_.code.size() * sizeof(tape::Instr));
  put(o_grouped, grouped.code.data(), grouped.code.size() * sizeof(tape::Instr));
  put(o_grouped_begin, grouped.group_begin.data(), grouped.group_begin.size() * sizeof(std::uint32_t));
  if (wide) {
    put(o_wide, wide_.code.data(), wide_.code.size() * sizeof(tape::Instr));
    put(o_wide_begin, wide_.group_begin.data(), wide_.group_begin.size() * sizeof(std::uint32_t));
  }
  for (std::size_t i = 0; i < problem.var_kinds.size(); ++i) {
    host[o_kinds + i] = static_cast<std::uint8_t>(problem.var_kinds[i]);
  }
  image_.reset(at);
  check(hipMemcpy(image_.get(), host.data(), at, hipMemcpyHostToDevice), "hipMemcpy(tape image)");
  auto* const base = image_.as<unsigned char>();
  code_d_ = reinterpret_cast<const tape::Instr*>(base + o_code);
  kinds_d_ = base + o_kinds;
  if (wide) {
    wide_code_d_ = reinterpret_cast<const tape::Instr*>(base + o_wide);
    wide_begin_d_ = reinterpret_cast<const std::uint32_t*>(base + o_wide_begin);
  }
  install_groups(std::move(grouped), reinterpret_cast<const tape::Instr*>(base + o_grouped),
                 reinterpret_cast<const std::uint32_t*>(base + o_grouped_begin));
}

tape::Tape DeviceTape::grouping(const model::Problem<double>& problem, std::uint32_t k) const
{
  if (k == 0) {
    k = tape::auto_groups(tape_);
  }
  std::uint32_t const n_roots = 1 + tape_.n_constraints;
  k = std::clamp<std::uint32_t>(k, 1, std::min<std::uint32_t>(dev::kMaxGroups, n_roots));
  for (; k > 1; --k) {
    tape::Tape g = tape::compile_groups(problem, k, opt_, 0, &tape_);
    if (dev::boxes_lds_bytes(g.n_slots) <= dev::lds_limit_bytes()) {
      return g;
    }
  }
  return tape_;
}

void DeviceTape::set_box_groups(const model::Problem<double>& problem, std::uint32_t k)
{
  ctx_->activate();
  tape::Tape g = grouping(problem, k);
  upload_code(g.code, grouped_code_);
  upload_words(g.group_begin, grouped_begin_);
  install_groups(std::move(g), grouped_code_.as<tape::Instr>(), grouped_begin_.as<std::uint32_t>());
}

void DeviceTape::install_groups(tape::Tape g, const tape::Instr* code, const std::uint32_t* begin)
{
  grouped_code_d_ = code;
  grouped_begin_d_ = begin;
  grouped_ = std::move(g);
  drop_compiled();  // the compiled kernel follows the grouping; compile again when needed
  if (engine_ == BoxEngine::Compiled) {
    set_box_engine(BoxEngine::Compiled);
  }
  groups_ = {};
  groups_.n_groups = grouped_.n_groups;
  groups_.objective_group = grouped_.objective_group;
  for (std::uint32_t i = 0; i <= grouped_.n_groups; ++i) {
    groups_.begin[i] = grouped_.group_begin[i];
  }
}

void DeviceTape::item_code(dev::Item& it, bool grouped) const
{
  it.n_slices = 1;
  it.slice_groups = 0;
  it.slice_begin = nullptr;
  it.slice_objective = 0;
  if (grouped) {
    it.code = grouped_code_d_;
    it.groups = groups_;
    it.n_instr = static_cast<std::uint32_t>(grouped_.code.size());
  } else {
    it.code = code_d_;
    it.groups = {};
    it.groups.n_groups = 1;
    it.groups.objective_group = 0;
    it.groups.begin[0] = 0;
    it.groups.begin[1] = static_cast<std::uint32_t>(tape_.code.size());
    it.n_instr = static_cast<std::uint32_t>(tape_.code.size());
  }
}

DeviceTape::ItemPlan DeviceTape::item_plan(dev::Mode mode, std::uint32_t waves) const
{
  ItemPlan p;
  std::size_t const lds = ctx_->lds_bytes();
  if (waves == 0) {
    bool const grouped = items_grouped_fit(mode);
    p.code = grouped ? grouped_code_d_ : code_d_;
    p.begin = grouped ? grouped_begin_d_ : nullptr;
    p.n_instr = static_cast<std::uint32_t>(grouped ? grouped_.code.size() : tape_.code.size());
    p.groups = grouped ? grouped_.n_groups : 1;
    p.objective = grouped ? grouped_.objective_group : 0;
    p.slots = item_slots(grouped);
    p.waves = item_waves(grouped);
    p.code_kind = grouped ? 1 : 0;
    return p;
  }
  waves = std::clamp<std::uint32_t>(waves, 1, dev::kMaxGroups);
  if (wide_.n_groups > 1 && kWidePeriod % waves == 0
      && dev::items_lds_bytes(mode, wide_.n_slots, std::max<std::uint32_t>(waves, 2)) <= lds) {
    p.code = wide_code_d_;
    p.begin = wide_begin_d_;
    p.n_instr = static_cast<std::uint32_t>(wide_.code.size());
    p.groups = wide_.n_groups;
    p.objective = wide_.objective_group;
    p.slots = wide_.n_slots;
    p.code_kind = 2;
  } else if (grouped_.n_groups > 1
             && dev::items_lds_bytes(mode, grouped_.n_slots, std::max<std::uint32_t>(std::min(waves, grouped_.n_groups), 2))
                 <= lds) {
    p.code = grouped_code_d_;
    p.begin = grouped_begin_d_;
    p.n_instr = static_cast<std::uint32_t>(grouped_.code.size());
    p.groups = grouped_.n_groups;
    p.objective = grouped_.objective_group;
    p.slots = grouped_.n_slots;
    p.code_kind = 1;
  } else {
    p.code = code_d_;
    p.n_instr = static_cast<std::uint32_t>(tape_.code.size());
    p.slots = tape_.n_slots;
    p.code_kind = 0;
  }
  p.waves = std::min(waves, p.groups);
  p.slices = (p.groups + waves - 1) / waves;
  return p;
}

DeviceTape::ItemPlan DeviceTape::item_plan_whole() const
{
  ItemPlan p;
  p.code = code_d_;
  p.n_instr = static_cast<std::uint32_t>(tape_.code.size());
  p.slots = tape_.n_slots;
  return p;
}

std::uint32_t DeviceTape::slice_waves(std::uint32_t groups, std::size_t tiles, std::uint32_t n_cus, std::size_t words)
{
  if (groups <= kWidePeriod || words < kSliceMinWords) {
    return 0;
  }
  if (tiles <= n_cus) {
    return kWidePeriod;
  }
  if (tiles <= std::size_t {2} * n_cus && groups > 8) {
    return 8;
  }
  return 0;
}

std::uint32_t DeviceTape::item_waves_for(dev::Mode mode, std::size_t n_tiles) const
{
  (void)mode;
  return slice_waves(item_groups_max(), n_tiles, ctx_->n_cus(), tape_.code.size());
}

void DeviceTape::item_code(dev::Item& it, const ItemPlan& plan) const
{
  it.code = plan.code;
  it.n_instr = plan.n_instr;
  it.groups = {};
  if (plan.slices == 1) {
    // every group in one workgroup: the inline table
    if (plan.code_kind == 0) {
      it.groups.n_groups = 1;
      it.groups.objective_group = 0;
      it.groups.begin[0] = 0;
      it.groups.begin[1] = plan.n_instr;
    } else {
      tape::Tape const& t = plan.code_kind == 1 ? grouped_ : wide_;
      it.groups.n_groups = t.n_groups;
      it.groups.objective_group = t.objective_group;
      for (std::uint32_t i = 0; i <= t.n_groups; ++i) {
        it.groups.begin[i] = t.group_begin[i];
      }
    }
  }
  it.n_slices = plan.slices;
  it.slice_groups = plan.groups;
  it.slice_begin = plan.slices > 1 ? plan.begin : nullptr;
  it.slice_objective = plan.objective;
}

bool DeviceTape::items_grouped_fit(dev::Mode mode) const
{
  return dev::items_lds_bytes(mode, item_slots(true), item_waves(true)) <= ctx_->lds_bytes();
}

void DeviceTape::set_box_engine(BoxEngine e)
{
  if (e != BoxEngine::Auto && e != BoxEngine::Interpreted && e != BoxEngine::Compiled) {
    throw InvalidConfiguration("unknown box engine");
  }
  if (e == BoxEngine::Compiled) {
    if (tape_.code.size() > kMaxCompiledWords) {
      throw InvalidConfiguration("tape of " + std::to_string(tape_.code.size()) + " words is above the "
                                 + std::to_string(kMaxCompiledWords) + "-word limit of compiled kernels");
    }
    auto c = compile_now(false);
    std::lock_guard<std::mutex> lock(jit_mu_);
    compiled_ = std::move(c);
    compiled_roots_.reset();
  }
  engine_ = e;
}

void DeviceTape::drop_compiled() const
{
  std::lock_guard<std::mutex> lock(jit_mu_);
  compiled_.reset();
  compiled_roots_.reset();
  jit_failed_ = false;
  for (int i = 0; i < 2; ++i) {
    items_[i].reset();
    items_failed_[i] = false;
    items_looked_up_[i] = false;
    items_work_[i] = 0.0;
  }
}

void DeviceTape::set_items_engine(BoxEngine e)
{
  if (e != BoxEngine::Auto && e != BoxEngine::Interpreted && e != BoxEngine::Compiled) {
    throw InvalidConfiguration("unknown decode-mode engine");
  }
  items_engine_ = e;
}

std::shared_ptr<const jit::CompiledItems> DeviceTape::items_kernel(bool aggregate, std::size_t n_elems) const
{
  BoxEngine const engine = items_engine_ != BoxEngine::Auto ? items_engine_ : engine_;
  if (engine == BoxEngine::Interpreted || grouped_.code.size() > kMaxCompiledItemWords || !jit::items_fit(grouped_)) {
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(jit_mu_);
  int const i = aggregate ? 1 : 0;
  if (engine == BoxEngine::Auto && !items_[i] && !items_failed_[i]) {
    // A kernel already compiled for these tape contents (looked up on the
    // first launch) costs nothing to take; otherwise compile once the
    // interpreted work reaches about what the compile costs.
    if (!items_looked_up_[i]) {
      items_looked_up_[i] = true;
      ctx_->activate();
      items_[i] = jit::compile_items(grouped_, ctx_->device(), aggregate, false);
      if (items_[i]) {
        ctx_->retain(items_[i]->module_handle());
        return items_[i];
      }
    }
    items_work_[i] += static_cast<double>(n_elems) * static_cast<double>(grouped_.code.size());
    if (items_work_[i] < kCompileAfterWork) {
      return nullptr;
    }
  }
  if (!items_[i] && !items_failed_[i]) {
    ctx_->activate();
    try {
      items_[i] = jit::compile_items(grouped_, ctx_->device(), aggregate);
      ctx_->retain(items_[i]->module_handle());
    } catch (const DeviceError&) {
      if (engine == BoxEngine::Compiled) {
        throw;
      }
      items_failed_[i] = true;  // Auto keeps the interpreter, which gives the same results
    }
  }
  return items_[i];
}

jit::Shape DeviceTape::compiled_shape() const
{
  jit::Shape sh;
  sh.waves = jit::shape_waves(grouped_, {false, want_waves_});
  sh.persistent = want_persistent_ && jit::persistent_fits(grouped_, sh.waves);
  sh.ilp_schedule = want_ilp_;
  sh.load_batch = want_load_batch_;
  return sh;
}

std::shared_ptr<const jit::CompiledBoxes> DeviceTape::compile_now(bool roots) const
{
  ctx_->activate();
  jit::Shape sh = compiled_shape();
  sh.roots = roots;
  auto c = jit::compile_boxes(grouped_, ctx_->device(), sh);
  c->set_wgs_per_cu(wgs_per_cu_);
  return c;
}

void DeviceTape::set_compiled_shape(bool persistent, std::uint32_t waves, std::uint32_t wgs_per_cu)
{
  if (wgs_per_cu < 1 || wgs_per_cu > 8) {
    throw InvalidConfiguration("workgroups per CU must be in [1, 8]");
  }
  if (waves > dev::kMaxGroups) {
    throw InvalidConfiguration("at most 16 waves per workgroup");
  }
  want_persistent_ = persistent;
  want_waves_ = waves;
  wgs_per_cu_ = wgs_per_cu;
  drop_compiled();
  if (engine_ == BoxEngine::Compiled) {
    set_box_engine(BoxEngine::Compiled);
  }
}

void DeviceTape::set_compiled_codegen(bool ilp_schedule, std::uint32_t load_batch)
{
  if (load_batch > 64) {
    throw InvalidConfiguration("at most 64 variables per load batch");
  }
  want_ilp_ = ilp_schedule;
  want_load_batch_ = load_batch;
  drop_compiled();
  if (engine_ == BoxEngine::Compiled) {
    set_box_engine(BoxEngine::Compiled);
  }
}

std::shared_ptr<const jit::CompiledBoxes> DeviceTape::compiled() const
{
  std::lock_guard<std::mutex> lock(jit_mu_);
  return compiled_;
}

void DeviceTape::eval_boxes_device(const double* lb, const double* ub, std::size_t ld, std::size_t n,
                                   std::uint8_t* feasible, double* obj_lb, double* obj_ub, double* roots,
                                   hipStream_t stream) const
{
  // Two compiled variants: root intervals cost the kernel the upper bounds
  // of every function; a call without them runs the kernel that skips what
  // no output reads.
  bool const want_roots = roots != nullptr;
  std::shared_ptr<const jit::CompiledBoxes> c;
  if (engine_ == BoxEngine::Compiled) {
    std::lock_guard<std::mutex> lock(jit_mu_);
    auto& slot = want_roots ? compiled_roots_ : compiled_;
    if (!slot) {
      slot = compile_now(want_roots);
    }
    c = slot;
  } else if (engine_ == BoxEngine::Auto && n >= kAutoCompileBoxes && n <= jit::kMaxBoxes
             && tape_.code.size() <= kMaxCompiledWords) {
    std::lock_guard<std::mutex> lock(jit_mu_);
    auto& slot = want_roots ? compiled_roots_ : compiled_;
    if (!slot && !jit_failed_) {
      try {
        slot = compile_now(want_roots);
      } catch (const DeviceError&) {
        jit_failed_ = true;  // Auto keeps the interpreter, which gives the same results
      }
    }
    c = slot;
  }
  if (c) {
    c->launch(lb, ub, ld, n, feasible, obj_lb, obj_ub, roots, stream);
    return;
  }
  dev::launch_boxes(grouped_code_d_, grouped_.extended, groups_, grouped_.n_slots, n_vars(), lb, ub, ld, n, feasible,
                    obj_lb, obj_ub, roots, stream);
  check(hipGetLastError(), "launch(k_boxes)");
}

std::vector<DeviceTape::BoxCandidate> DeviceTape::autotune(const model::Problem<double>& problem, std::size_t n)
{
  if (n == 0 || n > jit::kMaxBoxes) {
    throw InvalidConfiguration("autotune needs between 1 and 2^29 boxes");
  }
  std::uint32_t const nv = n_vars();
  ctx_->activate();
  // Synthetic boxes inside the variable bounds: log-uniform widths, clipped.
  std::vector<double> lb(static_cast<std::size_t>(nv) * n), ub(lb.size());
  qcqp::Rng rng(0x5eedULL);
  for (std::uint32_t v = 0; v < nv; ++v) {
    double const lo = problem.box_bounds[v].lb, hi = problem.box_bounds[v].ub;
    double const span = std::isfinite(hi - lo) && hi > lo ? hi - lo : 1.0;
    for (std::size_t e = 0; e < n; ++e) {
      double const w = span * std::exp(rng.uniform(std::log(1e-5), 0.0));
      double const a = (std::isfinite(lo) ? lo : -1.0) + rng.uniform() * (span - w);
      lb[v * n + e] = a;
      ub[v * n + e] = a + w;
    }
  }
  // The boxes are timed as callers stream them: from HBM. Enough copies of
  // the set are rotated that together they are 2.5 times the 256 MB Infinity
  // Cache, so hardly any launch finds its boxes there (at most 32 copies; one
  // when a set alone is that large).
  std::size_t const set_bytes = lb.size() * sizeof(double) * 2;
  std::size_t const n_sets = std::clamp<std::size_t>((640u << 20) / std::max<std::size_t>(set_bytes, 1) + 1, 1, 32);
  std::vector<DeviceBuffer> d_lb, d_ub;
  for (std::size_t k = 0; k < n_sets; ++k) {
    d_lb.emplace_back(lb.size() * sizeof(double));
    d_ub.emplace_back(ub.size() * sizeof(double));
    check(hipMemcpy(d_lb.back().get(), lb.data(), lb.size() * sizeof(double), hipMemcpyHostToDevice),
          "hipMemcpy(autotune)");
    check(hipMemcpy(d_ub.back().get(), ub.data(), ub.size() * sizeof(double), hipMemcpyHostToDevice),
          "hipMemcpy(autotune)");
  }
  DeviceBuffer d_f(n), d_lo(n * sizeof(double)), d_hi(n * sizeof(double));
  hipStream_t const s = ctx_->stream();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  std::size_t turn = 0;
  auto launch = [&]() {
    std::size_t const k = turn++ % n_sets;
    eval_boxes_device(d_lb[k].as<double>(), d_ub[k].as<double>(), n, n, d_f.as<std::uint8_t>(), d_lo.as<double>(),
                      d_hi.as<double>(), nullptr, s);
  };
  // Per-launch time of runs of 64 back-to-back launches, the median over
  // `runs` runs. Callers streaming boxes keep the device busy for long
  // stretches, and a kernel's sustained rate can differ from its rate in a
  // short burst after an idle gap (measured: up to 8 % on the 600-node DAG),
  // so short runs would rank the shapes for the wrong case.
  constexpr int kRun = 64;
  auto time_it = [&](int runs) {
    for (int k = 0; k < 8; ++k) {
      launch();
    }
    std::vector<float> ms;
    for (int k = 0; k < runs; ++k) {
      check(hipEventRecord(e0, s), "hipEventRecord");
      for (int q = 0; q < kRun; ++q) {
        launch();
      }
      check(hipEventRecord(e1, s), "hipEventRecord");
      check(hipEventSynchronize(e1), "hipEventSynchronize");
      float t = 0.0f;
      check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
      ms.push_back(t / static_cast<float>(kRun));
    }
    std::nth_element(ms.begin(), ms.begin() + ms.size() / 2, ms.end());
    return static_cast<double>(ms[ms.size() / 2]);
  };

  std::uint32_t const n_roots = 1 + tape_.n_constraints;
  std::vector<BoxCandidate> cands;
  for (std::uint32_t k : {4u, 8u}) {
    if (k > n_roots && k != 4u) {
      continue;
    }
    BoxCandidate c;
    c.engine = BoxEngine::Interpreted;
    c.groups = k;
    cands.push_back(c);
  }
  bool const compilable = tape_.code.size() <= kMaxCompiledWords;
  for (std::uint32_t k : {1u, 4u, 8u}) {
    if (!compilable || (k > n_roots && k != 1u)) {
      continue;
    }
    for (std::uint32_t w : {1u, 2u, 4u, 8u}) {
      if (w > k) {
        continue;
      }
      auto add = [&](bool persistent, std::uint32_t wgs, bool ilp, std::uint32_t batch) {
        BoxCandidate c;
        c.engine = BoxEngine::Compiled;
        c.groups = k;
        c.waves = w;
        c.persistent = persistent;
        c.wgs_per_cu = wgs;
        c.ilp_schedule = ilp;
        c.load_batch = batch;
        cands.push_back(c);
      };
      if (w <= 2) {
        // few waves per tile: code generation decides how well one wave
        // overlaps its loads and its FP64 chains
        for (bool ilp : {false, true}) {
          for (std::uint32_t batch : {0u, kDefaultLoadBatch}) {
            add(false, 1, ilp, batch);
          }
        }
      } else {
        add(false, 1, false, kDefaultLoadBatch);
        add(true, 1, false, 0);
        add(true, 2, false, 0);
      }
    }
  }
  auto apply = [&](const BoxCandidate& c) {
    engine_ = BoxEngine::Interpreted;  // no compile while regrouping
    set_box_groups(problem, c.groups);
    if (c.engine == BoxEngine::Compiled) {
      want_persistent_ = c.persistent;
      want_waves_ = c.waves;
      want_ilp_ = c.ilp_schedule;
      want_load_batch_ = c.load_batch;
      wgs_per_cu_ = c.wgs_per_cu;
      set_box_engine(BoxEngine::Compiled);
    }
  };
  std::vector<BoxCandidate> done;
  for (auto c : cands) {
    try {
      apply(c);
      if (c.engine == BoxEngine::Compiled) {
        auto const sh = compiled_shape();
        if (c.persistent && !sh.persistent) {
          continue;  // tiles do not fit: same kernel as the direct form
        }
        c.waves = sh.waves;
      }
      c.groups = grouped_.n_groups;
      c.kernel_ms = time_it(3);
      done.push_back(c);
    } catch (const DeviceError&) {
      // a shape the run-time compiler rejects is left out
    }
  }
  if (done.empty()) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw DeviceError(0, "autotune: no configuration ran");
  }
  // The leaders are often within a few per cent of each other: time the four
  // fastest again, three times as long, and keep the fastest of those.
  std::vector<std::size_t> rank(done.size());
  for (std::size_t i = 0; i < rank.size(); ++i) {
    rank[i] = i;
  }
  std::sort(rank.begin(), rank.end(), [&](std::size_t a, std::size_t b) { return done[a].kernel_ms < done[b].kernel_ms; });
  rank.resize(std::min<std::size_t>(rank.size(), 4));
  std::size_t best = rank.front();
  double best_ms = std::numeric_limits<double>::infinity();
  for (std::size_t i : rank) {
    try {
      apply(done[i]);
      double const ms = time_it(9);
      done[i].kernel_ms = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = i;
      }
    } catch (const DeviceError&) {
      // compiled once above; a failure now leaves the first timing in place
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  apply(done[best]);
  return done;
}

void DeviceTape::eval_points_device(const double* x, std::size_t ld, std::size_t n, std::uint8_t* feasible,
                                    double* obj, double* roots, hipStream_t stream) const
{
  if (tape_.n_slots > ctx_->lds_point_slots()) {
    throw ResourceExhausted("tape needs more point slots than LDS holds");
  }
  dev::launch_points(code(), extended(), n_instr(), n_slots(), n_vars(), x, ld, n, feasible, obj, roots, stream);
  check(hipGetLastError(), "launch(k_points)");
}

// ---------------------------------------------------------------------------
// Roles
// ---------------------------------------------------------------------------

backend::RegionCostModel cost_model_for(const tape::Tape& t, std::size_t samples)
{
  backend::RegionCostModel m;
  std::size_t const fixed = 4096 + t.n_vars * 24 + 64 * 32;
  m.bounder_fixed = fixed;
  m.bounder_per_region = 1 + sizeof(double) + 4 / 64 + 1;
  m.sampler_fixed = fixed + t.n_vars * 8;
  m.sampler_per_region = (samples * 16 + 63) / 64 + 1;
  m.enumerator_fixed = fixed + t.n_vars * 8;
  m.enumerator_per_region = 1;
  m.max_regions = ((std::size_t(1) << 31) - 1) / std::max<std::size_t>(samples, 1);
  return m;
}

namespace
{
std::size_t align_up(std::size_t x, std::size_t a) { return (x + a - 1) / a * a; }
}  // namespace

void check_role_fits(const DeviceTape& tape, dev::Mode mode, const region::Composition& composition,
                     std::size_t n_regions, std::size_t samples_per_region, std::size_t budget_bytes)
{
  std::size_t const samples = mode == dev::Mode::Sample ? samples_per_region : 1;
  if (samples < 1) {
    throw InvalidConfiguration("sample_points must be at least 1");
  }
  if (n_regions == 0 || n_regions > (std::size_t(1) << 31) / samples) {
    throw InvalidConfiguration("n_regions * samples must be in [1, 2^31)");
  }
  if (composition.size() > dev::kMaxSlotsPerItem) {
    throw InvalidConfiguration("composition has more slots than the device decode supports");
  }
  Context const& ctx = *tape.context();
  if (mode != dev::Mode::Bound && tape.n_slots() > ctx.lds_point_slots()) {
    throw ResourceExhausted("tape needs more point slots than LDS holds");
  }

  // Priced before anything is allocated (reference replay.cuh:285-304).
  backend::RegionCostModel const cm = cost_model_for(tape.tape(), samples);
  std::size_t needed = 0;
  std::size_t per_region = 0;
  std::string role;
  switch (mode) {
    case dev::Mode::Bound:
      per_region = cm.bounder_per_region;
      needed = cm.bounder_fixed + per_region * n_regions;
      role = "interval bounder";
      break;
    case dev::Mode::Sample:
      per_region = cm.sampler_per_region;
      needed = cm.sampler_fixed + per_region * n_regions;
      role = "point sampler";
      break;
    case dev::Mode::Enumerate:
      per_region = cm.enumerator_per_region;
      needed = cm.enumerator_fixed + per_region * n_regions;
      role = "exact enumerator";
      break;
  }
  std::size_t const budget = budget_bytes != 0 ? budget_bytes : ctx.free_bytes();
  if (needed <= budget) {
    return;
  }
  backend::OverBudget f;
  f.role = role;
  f.needed_bytes = needed;
  f.budget_bytes = budget;
  f.n_regions = n_regions;
  f.elements_per_region = samples;
  f.bytes_per_element = (per_region + samples - 1) / samples;
  f.element_breakdown = mode == dev::Mode::Bound
      ? "one lower bound (8 B) and one feasibility byte read back per region, plus tile minima"
      : "one 16 B (value, index) partial per 64-element tile, plus one byte per region";
  f.element_inventory = std::to_string(tape.tape().code.size()) + " tape instructions over "
      + std::to_string(tape.n_vars()) + " variables";
  f.cost = cm;
  f.composition = composition;
  f.solve_samples_per_region = samples_per_region;
  f.sampler_bytes = 0;
  throw backend::OverBudgetError(f);
}

RoleImpl::RoleImpl(std::shared_ptr<const DeviceTape> tape, dev::Mode mode, const region::Composition& composition,
                   std::size_t n_regions, std::size_t samples_per_region, std::size_t budget_bytes)
    : tape_(std::move(tape))
    , mode_(mode)
    , composition_(composition)
    , n_regions_(n_regions)
    , samples_(mode == dev::Mode::Sample ? samples_per_region : 1)
    , solve_samples_(samples_per_region)
    , slot_count_(composition.size())
    , n_vars_(tape_->n_vars())
{
  check_role_fits(*tape_, mode_, composition_, n_regions_, solve_samples_, budget_bytes);
  Context const& ctx = *tape_->context();
  n_elems_ = n_regions_ * samples_;
  n_tiles_ = static_cast<std::uint32_t>((n_elems_ + dev::kTile - 1) / dev::kTile);

  ctx.activate();
  off_parent_ = align_up(sizeof(dev::Item), 256);
  off_var_ = align_up(off_parent_ + n_vars_ * 16, 16);
  off_width_ = align_up(off_var_ + slot_count_ * 4, 16);
  off_step_ = align_up(off_width_ + slot_count_ * 4, 16);
  off_kind_ = align_up(off_step_ + slot_count_ * 8, 16);
  staging_bytes_ = align_up(off_kind_ + slot_count_, 16);
  staging_d_.reset(staging_bytes_);
  staging_h_.reset(staging_bytes_);
  std::memset(staging_h_.data(), 0, staging_bytes_);

  // A launch over few tiles may spread each tile's function groups over
  // several workgroups (item_waves_for); every workgroup names item 0, and
  // the slices of a tile meet in a per-tile scratch that each launch leaves
  // zeroed.
  if (tape_->item_waves_for(mode_, n_tiles_) != 0) {
    max_slices_ = (DeviceTape::kMaxWideGroups + DeviceTape::kWidePeriod - 1) / DeviceTape::kWidePeriod;
  }
  std::size_t const blocks = std::size_t {n_tiles_} * max_slices_;
  std::vector<std::uint32_t> tile_item(blocks, 0);
  tile_item_d_.reset(blocks * sizeof(std::uint32_t));
  check(hipMemcpy(tile_item_d_.get(), tile_item.data(), blocks * sizeof(std::uint32_t), hipMemcpyHostToDevice),
        "hipMemcpy(tile_item)");
  if (max_slices_ > 1) {
    slice_d_.reset(std::size_t {n_tiles_} * dev::kSliceTileBytes + 256);
    check(hipMemset(slice_d_.get(), 0, slice_d_.bytes()), "hipMemset(slice scratch)");
  }

  if (mode_ == dev::Mode::Bound) {
    feasible_d_.reset(n_regions_);
    obj_lb_d_.reset(n_regions_ * sizeof(double));
    feasible_h_.reset(n_regions_);
    obj_lb_h_.reset(n_regions_);
  } else {
    tile_min_d_.reset(n_tiles_ * sizeof(double));
    tile_arg_d_.reset(n_tiles_ * sizeof(std::int64_t));
    done_d_.reset(sizeof(std::uint32_t));
    check(hipMemset(done_d_.get(), 0, sizeof(std::uint32_t)), "hipMemset(done)");
    out_d_.reset(sizeof(Candidate));
    witness_d_.reset(std::max<std::size_t>(n_vars_, 1) * sizeof(double));
    out_h_.reset(1);
    out_h_[0] = {std::numeric_limits<double>::max(), -1};
    witness_h_.reset(std::max<std::size_t>(n_vars_, 1));
  }

  // The Item's pointers never change; only the parent box, slot data and salt
  // are rewritten per call.
  auto* item = reinterpret_cast<dev::Item*>(staging_h_.data());
  auto* base = staging_d_.as<unsigned char>();
  tape_->item_code(*item, false);
  item->n_vars = static_cast<std::uint32_t>(n_vars_);
  item->parent = reinterpret_cast<const double*>(base + off_parent_);
  item->var_kinds = tape_->var_kinds();
  item->slot_var = reinterpret_cast<const std::uint32_t*>(base + off_var_);
  item->slot_width = reinterpret_cast<const std::uint32_t*>(base + off_width_);
  item->slot_step = reinterpret_cast<const double*>(base + off_step_);
  item->slot_kind = reinterpret_cast<const std::uint8_t*>(base + off_kind_);
  item->slot_count = static_cast<std::uint32_t>(slot_count_);
  item->samples = static_cast<std::uint32_t>(samples_);
  item->n_elems = n_elems_;
  item->salt = 0;
  item->tile_begin = 0;
  item->n_tiles = n_tiles_;
  item->feasible = feasible_d_.as<std::uint8_t>();
  item->obj_lb = obj_lb_d_.as<double>();
  item->tile_min = tile_min_d_.as<double>();
  item->tile_arg = tile_arg_d_.as<std::int64_t>();
  item->done = done_d_.as<std::uint32_t>();
  item->cand_value = out_d_.as<double>();
  item->cand_index = out_d_.get() != nullptr ? reinterpret_cast<std::int64_t*>(out_d_.as<unsigned char>() + 8) : nullptr;
  item->witness = witness_d_.as<double>();
  auto* kinds = reinterpret_cast<std::uint8_t*>(staging_h_.data() + off_kind_);
  for (std::size_t j = 0; j < slot_count_; ++j) {
    kinds[j] = static_cast<std::uint8_t>(composition_[j]);
  }
}

void RoleImpl::run(std::span<const interval> box, std::span<const std::size_t> var_ids,
                   std::span<const std::size_t> widths, std::uint64_t salt)
{
  if (box.size() != n_vars_) {
    throw ShapeMismatch("set_domain: domain size does not match the problem's variable count");
  }
  if (var_ids.size() != slot_count_ || widths.size() != slot_count_) {
    throw ShapeMismatch("set_domain: var_ids/widths size does not match this role's slot count");
  }
  if (region::composition_fan_out(widths) != n_regions_) {
    throw ShapeMismatch("set_domain: product of widths does not match this role's n_regions");
  }
  for (std::size_t j = 0; j < slot_count_; ++j) {
    if (var_ids[j] >= n_vars_) {
      throw ShapeMismatch("set_domain: slot var_id out of range");
    }
  }
  Context const& ctx = *tape_->context();
  ctx.activate();
  hipStream_t const s = ctx.stream();

  unsigned char* h = staging_h_.data();
  auto* item = reinterpret_cast<dev::Item*>(h);
  item->salt = salt;
  std::memcpy(h + off_parent_, box.data(), n_vars_ * sizeof(interval));
  auto* var = reinterpret_cast<std::uint32_t*>(h + off_var_);
  auto* width = reinterpret_cast<std::uint32_t*>(h + off_width_);
  auto* step = reinterpret_cast<double*>(h + off_step_);
  for (std::size_t j = 0; j < slot_count_; ++j) {
    var[j] = static_cast<std::uint32_t>(var_ids[j]);
    width[j] = static_cast<std::uint32_t>(widths[j]);
    step[j] = (box[var_ids[j]].ub - box[var_ids[j]].lb) / static_cast<double>(widths[j]);
  }
  // the tape's current function groups (a regrouping replaces their code)
  auto const compiled = mode_ == dev::Mode::Bound ? tape_->items_kernel(false, n_elems_) : nullptr;
  std::uint32_t const waves = compiled || max_slices_ == 1 ? 0u : tape_->item_waves_for(mode_, n_tiles_);
  DeviceTape::ItemPlan const plan = tape_->item_plan(mode_, waves);
  if (plan.slices > max_slices_) {
    throw InvalidConfiguration("decode-mode item: more slices per tile than this role's launch table holds");
  }
  tape_->item_code(*item, plan);
  if (plan.slices > 1) {
    auto* const sc = slice_d_.as<unsigned char>();
    std::size_t const masks_bytes = std::size_t {n_tiles_} * dev::kSliceMasks * 8;
    item->tile_infeasible = reinterpret_cast<std::uint64_t*>(sc);
    item->tile_slices_done = reinterpret_cast<std::uint32_t*>(sc + masks_bytes);
    item->tile_obj = reinterpret_cast<double*>(sc + align_up(masks_bytes + std::size_t {n_tiles_} * 4, 16));
  }
  check(hipMemcpyAsync(staging_d_.get(), h, staging_bytes_, hipMemcpyHostToDevice, s), "hipMemcpyAsync(staging)");
  if (compiled) {
    compiled->launch(staging_d_.get(), tile_item_d_.as<std::uint32_t>(), n_tiles_, s);
  } else {
    dev::launch_items(mode_, tape_->extended(), staging_d_.as<dev::Item>(), tile_item_d_.as<std::uint32_t>(),
                      n_tiles_ * plan.slices, plan.slots, plan.waves, s);
  }
  check(hipGetLastError(), "launch(k_items)");
  if (mode_ == dev::Mode::Bound) {
    check(hipMemcpyAsync(feasible_h_.data(), feasible_d_.get(), n_regions_, hipMemcpyDeviceToHost, s),
          "hipMemcpyAsync(feasible)");
    check(hipMemcpyAsync(obj_lb_h_.data(), obj_lb_d_.get(), n_regions_ * sizeof(double), hipMemcpyDeviceToHost, s),
          "hipMemcpyAsync(obj_lb)");
  } else {
    check(hipMemcpyAsync(out_h_.data(), out_d_.get(), sizeof(Candidate), hipMemcpyDeviceToHost, s),
          "hipMemcpyAsync(candidate)");
    check(hipMemcpyAsync(witness_h_.data(), witness_d_.get(), n_vars_ * sizeof(double), hipMemcpyDeviceToHost, s),
          "hipMemcpyAsync(witness)");
  }
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

namespace
{

/// Where each piece of a no-allocation sampler launch sits in the static
/// scratch array.
struct StaticLayout
{
  std::size_t parent, var, width, step, kind, tile_item, tile_min, tile_arg, done, out, witness, total;

  StaticLayout(std::size_t n_vars, std::size_t slots, std::size_t n_tiles)
  {
    parent = align_up(sizeof(dev::Item), 256);
    var = align_up(parent + n_vars * 16, 16);
    width = align_up(var + slots * 4, 16);
    step = align_up(width + slots * 4, 16);
    kind = align_up(step + slots * 8, 16);
    tile_item = align_up(kind + slots, 256);
    tile_min = align_up(tile_item + n_tiles * 4, 256);
    tile_arg = align_up(tile_min + n_tiles * 8, 256);
    done = align_up(tile_arg + n_tiles * 8, 256);
    out = align_up(done + 4, 16);
    witness = align_up(out + 16, 16);
    total = align_up(witness + std::max<std::size_t>(n_vars, 1) * 8, 256);
  }
};

std::mutex g_static_mu[dev::kMaxDevices];

}  // namespace

std::size_t sample_static_bytes(std::size_t n_vars, std::size_t slot_count, std::size_t n_elems)
{
  return StaticLayout(n_vars, slot_count, (n_elems + dev::kTile - 1) / dev::kTile).total;
}

void sample_static(const DeviceTape& tape, const region::Composition& composition, std::span<const interval> box,
                   std::span<const std::size_t> var_ids, std::span<const std::size_t> widths,
                   std::size_t samples_per_region, std::uint64_t salt, double* value, std::int64_t* index,
                   double* point)
{
  std::size_t const n_vars = tape.n_vars();
  std::size_t const slots = composition.size();
  if (box.size() != n_vars) {
    throw ShapeMismatch("set_domain: domain size does not match the problem's variable count");
  }
  if (var_ids.size() != slots || widths.size() != slots) {
    throw ShapeMismatch("set_domain: var_ids/widths size does not match the composition's slot count");
  }
  for (std::size_t j = 0; j < slots; ++j) {
    if (var_ids[j] >= n_vars) {
      throw ShapeMismatch("set_domain: slot var_id out of range");
    }
  }
  std::size_t const n_regions = region::composition_fan_out(widths);
  // Shape and LDS checks of a sampler build; no budget, since nothing is allocated.
  check_role_fits(tape, dev::Mode::Sample, composition, n_regions, samples_per_region,
                  std::numeric_limits<std::size_t>::max());
  std::size_t const n_elems = n_regions * samples_per_region;
  std::size_t const n_tiles = (n_elems + dev::kTile - 1) / dev::kTile;
  StaticLayout const L(n_vars, slots, n_tiles);
  if (L.total > dev::kStaticScratchBytes) {
    throw InvalidConfiguration("this sample needs " + std::to_string(L.total) + " bytes of static scratch ("
                               + std::to_string(n_tiles) + " tiles of 64 points, " + std::to_string(n_vars)
                               + " variables); the array holds " + std::to_string(dev::kStaticScratchBytes));
  }

  Context const& ctx = *tape.context();
  ctx.activate();
  std::lock_guard<std::mutex> lock(g_static_mu[ctx.device()]);
  auto* const base = static_cast<unsigned char*>(ctx.static_scratch());
  hipStream_t const s = ctx.stream();

  // The Item and the slot arrays are assembled on the stack (at most
  // kMaxSlotsPerItem slots), the parent box is copied from the caller's array.
  dev::Item item {};
  bool const grouped = tape.items_grouped_fit(dev::Mode::Sample);
  tape.item_code(item, grouped);
  item.n_vars = static_cast<std::uint32_t>(n_vars);
  item.parent = reinterpret_cast<const double*>(base + L.parent);
  item.var_kinds = tape.var_kinds();
  item.slot_var = reinterpret_cast<const std::uint32_t*>(base + L.var);
  item.slot_width = reinterpret_cast<const std::uint32_t*>(base + L.width);
  item.slot_step = reinterpret_cast<const double*>(base + L.step);
  item.slot_kind = reinterpret_cast<const std::uint8_t*>(base + L.kind);
  item.slot_count = static_cast<std::uint32_t>(slots);
  item.samples = static_cast<std::uint32_t>(samples_per_region);
  item.n_elems = n_elems;
  item.salt = salt;
  item.tile_begin = 0;
  item.n_tiles = static_cast<std::uint32_t>(n_tiles);
  item.tile_min = reinterpret_cast<double*>(base + L.tile_min);
  item.tile_arg = reinterpret_cast<std::int64_t*>(base + L.tile_arg);
  item.done = reinterpret_cast<std::uint32_t*>(base + L.done);
  item.cand_value = reinterpret_cast<double*>(base + L.out);
  item.cand_index = reinterpret_cast<std::int64_t*>(base + L.out + 8);
  item.witness = reinterpret_cast<double*>(base + L.witness);
  std::uint32_t var[dev::kMaxSlotsPerItem] = {};
  std::uint32_t width[dev::kMaxSlotsPerItem] = {};
  double step[dev::kMaxSlotsPerItem] = {};
  std::uint8_t kind[dev::kMaxSlotsPerItem] = {};
  for (std::size_t j = 0; j < slots; ++j) {
    var[j] = static_cast<std::uint32_t>(var_ids[j]);
    width[j] = static_cast<std::uint32_t>(widths[j]);
    step[j] = (box[var_ids[j]].ub - box[var_ids[j]].lb) / static_cast<double>(widths[j]);
    kind[j] = static_cast<std::uint8_t>(composition[j]);
  }
  check(hipMemcpyAsync(base, &item, sizeof(item), hipMemcpyHostToDevice, s), "hipMemcpyAsync(static item)");
  check(hipMemcpyAsync(base + L.parent, box.data(), n_vars * sizeof(interval), hipMemcpyHostToDevice, s),
        "hipMemcpyAsync(static parent)");
  if (slots > 0) {
    check(hipMemcpyAsync(base + L.var, var, slots * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync(static var)");
    check(hipMemcpyAsync(base + L.width, width, slots * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync(static width)");
    check(hipMemcpyAsync(base + L.step, step, slots * 8, hipMemcpyHostToDevice, s),
          "hipMemcpyAsync(static step)");
    check(hipMemcpyAsync(base + L.kind, kind, slots, hipMemcpyHostToDevice, s), "hipMemcpyAsync(static kind)");
  }
  // Every tile belongs to item 0, and the done counter starts at zero: the
  // array keeps whatever the previous call left, so both are cleared here.
  dev::launch_fill_u32(reinterpret_cast<std::uint32_t*>(base + L.tile_item), n_tiles, 0u, s);
  dev::launch_fill_u32(reinterpret_cast<std::uint32_t*>(base + L.done), 1, 0u, s);
  check(hipGetLastError(), "launch(fill static scratch)");
  dev::launch_items(dev::Mode::Sample, tape.extended(), reinterpret_cast<const dev::Item*>(base),
                    reinterpret_cast<const std::uint32_t*>(base + L.tile_item), static_cast<std::uint32_t>(n_tiles),
                    tape.item_slots(grouped), tape.item_waves(grouped), s);
  check(hipGetLastError(), "launch(k_items static)");
  check(hipMemcpyAsync(value, base + L.out, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(static value)");
  check(hipMemcpyAsync(index, base + L.out + 8, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(static index)");
  check(hipMemcpyAsync(point, base + L.witness, n_vars * sizeof(double), hipMemcpyDeviceToHost, s),
        "hipMemcpyAsync(static witness)");
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

namespace
{

class HipBounder final : public backend::RegionBounder
{
public:
  explicit HipBounder(std::unique_ptr<RoleImpl> impl)
      : impl_(std::move(impl))
  {
  }
  backend::BoundResult bound(const backend::Region& r) override
  {
    impl_->run(r.box, r.assignment.var_ids, r.assignment.widths, 0);
    return {impl_->feasible(), impl_->obj_lb(), impl_->n_regions()};
  }
  std::size_t n_regions() const override { return impl_->n_regions(); }

private:
  std::unique_ptr<RoleImpl> impl_;
};

backend::CandidateResult candidate_of(const RoleImpl& impl)
{
  return {impl.has_candidate(), impl.cand_value(), impl.witness(), impl.cand_index()};
}

class HipSampler final : public backend::RegionSampler
{
public:
  explicit HipSampler(std::unique_ptr<RoleImpl> impl)
      : impl_(std::move(impl))
  {
  }
  backend::CandidateResult sample(const backend::Region& r, std::uint64_t salt) override
  {
    impl_->run(r.box, r.assignment.var_ids, r.assignment.widths, salt);
    return candidate_of(*impl_);
  }
  std::size_t n_samples() const override { return impl_->n_elems(); }

private:
  std::unique_ptr<RoleImpl> impl_;
};

class HipEnumerator final : public backend::RegionEnumerator
{
public:
  explicit HipEnumerator(std::unique_ptr<RoleImpl> impl)
      : impl_(std::move(impl))
  {
  }
  backend::CandidateResult enumerate(const backend::Region& r) override
  {
    impl_->run(r.box, r.assignment.var_ids, r.assignment.widths, 0);
    return candidate_of(*impl_);
  }
  std::size_t n_points() const override { return impl_->n_elems(); }

private:
  std::unique_ptr<RoleImpl> impl_;
};

}  // namespace

HipBackendFactory::HipBackendFactory(std::shared_ptr<Context> ctx, tape::CompileOptions opt)
    : ctx_(std::move(ctx))
    , opt_(opt)
{
  if (ctx_ == nullptr) {
    throw InvalidConfiguration("HipBackendFactory requires a device context");
  }
}

std::shared_ptr<const DeviceTape> HipBackendFactory::tape_for(const model::Problem<double>& problem) const
{
  if (cached_problem_ != &problem || cached_tape_ == nullptr) {
    cached_tape_ = std::make_shared<DeviceTape>(ctx_, problem, opt_);
    cached_problem_ = &problem;
  }
  return cached_tape_;
}

backend::RegionCostModel HipBackendFactory::cost_model(const model::Problem<double>& problem) const
{
  return cost_model_for(tape::compile(problem, opt_), 1);
}

backend::SubdivisionBundle HipBackendFactory::build_subdivision(const model::Problem<double>& problem,
                                                                const region::Composition& composition,
                                                                std::size_t n_regions,
                                                                const backend::BuildBudget& budget,
                                                                backend::RoleRequest roles) const
{
  auto t = tape_for(problem);
  backend::SubdivisionBundle out;
  if (roles.sampler) {
    out.sampler = std::make_unique<HipSampler>(std::make_unique<RoleImpl>(
        t, dev::Mode::Sample, composition, n_regions, budget.samples_per_region, budget.bytes));
  }
  if (roles.bounder) {
    out.bounder = std::make_unique<HipBounder>(
        std::make_unique<RoleImpl>(t, dev::Mode::Bound, composition, n_regions, budget.samples_per_region, budget.bytes));
  }
  if (roles.enumerator && region::is_fully_enumerable(composition)) {
    out.enumerator = std::make_unique<HipEnumerator>(
        std::make_unique<RoleImpl>(t, dev::Mode::Enumerate, composition, n_regions, budget.samples_per_region,
                                   budget.bytes));
  }
  return out;
}

}  // namespace cmi::hip
