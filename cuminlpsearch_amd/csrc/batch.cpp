#include "batch.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>

#include "cmi/policy.hpp"

namespace cmi::capi
{
std::shared_ptr<const cmi::policy::CompositionPolicy> make_policy(const cmi_solve_options& o);
}

namespace cmi::batch
{

namespace
{
std::size_t align_up(std::size_t x, std::size_t a) { return (x + a - 1) / a * a; }

/// fn(i) for i in [0, n) on up to min(16, hardware threads) host threads;
/// rethrows the first exception after every thread has stopped.
void parallel_for(std::size_t n, const std::function<void(std::size_t)>& fn)
{
  std::size_t const threads = std::min<std::size_t>(n, std::clamp<std::size_t>(std::thread::hardware_concurrency(), 1, 16));
  if (threads <= 1) {
    for (std::size_t i = 0; i < n; ++i) {
      fn(i);
    }
    return;
  }
  std::atomic<std::size_t> next {0};
  std::mutex mu;
  std::exception_ptr error;
  auto work = [&] {
    for (std::size_t i; (i = next.fetch_add(1)) < n;) {
      try {
        fn(i);
      } catch (...) {
        std::lock_guard<std::mutex> lock(mu);
        if (!error) {
          error = std::current_exception();
        }
        next = n;
      }
    }
  };
  std::vector<std::thread> pool;
  for (std::size_t t = 1; t < threads; ++t) {
    pool.emplace_back(work);
  }
  work();
  for (std::thread& t : pool) {
    t.join();
  }
  if (error) {
    std::rethrow_exception(error);
  }
}

struct Candidate
{
  double value;
  std::int64_t index;
};

/// Steps with fewer live solves than this run their per-solve host work on
/// the calling thread: waking the workers costs more than it saves.
constexpr std::size_t kParallelMembers = 256;

/// Host threads of a runner's workers: CMI_BATCH_HOST_THREADS when it holds
/// a count in [1, 64], else min(16, hardware threads).
/// CMI_ITEM_WAVES: waves per workgroup of the batched launches (0: each
/// item's groups in one workgroup; 1, 2, 4, 8 or 16), one value for every
/// launch or three separated by commas (sampler, enumerator, bounder); -1
/// where unset.
void item_waves_override(int (&out)[3])
{
  out[0] = out[1] = out[2] = -1;
  const char* e = std::getenv("CMI_ITEM_WAVES");
  if (e == nullptr) {
    return;
  }
  int vals[3];
  int n = 0;
  for (const char* c = e; n < 3;) {
    char* end = nullptr;
    long const w = std::strtol(c, &end, 10);
    if (end == c || !(w == 0 || w == 1 || w == 2 || w == 4 || w == 8 || w == 16)) {
      return;
    }
    vals[n++] = static_cast<int>(w);
    if (*end == '\0') {
      break;
    }
    if (*end != ',') {
      return;
    }
    c = end + 1;
  }
  if (n == 1) {
    out[0] = out[1] = out[2] = vals[0];
  } else if (n == 3) {
    out[0] = vals[0];
    out[1] = vals[1];
    out[2] = vals[2];
  }
}

/// CMI_POINTS_PER_LANE: points each lane of a sampler or enumerator launch
/// evaluates (1, 2 or 4; else the default of points_per_lane), or 0 when
/// unset.
std::uint32_t points_per_lane_override()
{
  if (const char* e = std::getenv("CMI_POINTS_PER_LANE")) {
    if (std::strcmp(e, "1") == 0 || std::strcmp(e, "2") == 0 || std::strcmp(e, "4") == 0) {
      return static_cast<std::uint32_t>(e[0] - '0');
    }
  }
  return 0;
}

/// Points per lane of a sampler or enumerator launch of `workgroups`
/// workgroups (at one point per lane) on n_cus compute units: 2 when they
/// number at least 4 per unit, so that 2 per unit remain, else 1. Measured
/// on 512 batched QCQPs of up to 48 variables (sampler ~2,048 workgroups:
/// 2 points 14 % less kernel time per step than 1, 4 points 6 %) and on
/// streams whose sampler launches have at most one workgroup per unit (2
/// points: up to 8 % more kernel time).
std::uint32_t points_per_lane(std::size_t workgroups, std::uint32_t n_cus)
{
  return workgroups >= std::size_t {4} * n_cus ? 2u : 1u;
}

/// CMI_STEP_GRAPH=1 runs each step's copies and launches as one HIP graph
/// (built per shape of step, its nodes' parameters set every step) instead
/// of one by one. Measured on the bench's streams and batch, the graph was
/// no faster: the 2,000 req/s stream of small QCQPs 395-415 against 423-441
/// served solves/s, the 50 req/s stream's p50 94-99 against 60 ms (a graph
/// launch with parameter updates costs more than a 4-operation step), the
/// batch equal, the stream of up to 96 variables 3 % faster.
bool step_graph()
{
  const char* e = std::getenv("CMI_STEP_GRAPH");
  return e != nullptr && std::strcmp(e, "1") == 0;
}

std::size_t host_threads()
{
  if (const char* v = std::getenv("CMI_BATCH_HOST_THREADS")) {
    char* end = nullptr;
    long const t = std::strtol(v, &end, 10);
    if (end != v && *end == '\0' && t >= 1 && t <= 64) {
      return static_cast<std::size_t>(t);
    }
  }
  return std::clamp<std::size_t>(std::thread::hardware_concurrency(), 1, 16);
}
}  // namespace

Workers::Workers(std::size_t threads)
{
  for (std::size_t t = 1; t < threads; ++t) {
    threads_.emplace_back([this] { loop(); });
  }
}

Workers::~Workers()
{
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  wake_.notify_all();
  for (std::thread& t : threads_) {
    t.join();
  }
}

void Workers::work()
{
  for (std::size_t i; (i = next_.fetch_add(1)) < n_;) {
    try {
      (*fn_)(i);
    } catch (...) {
      std::lock_guard<std::mutex> lock(mu_);
      if (!error_) {
        error_ = std::current_exception();
      }
      next_ = n_;
    }
  }
}

void Workers::loop()
{
  std::uint64_t seen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lock(mu_);
      wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
      if (stop_) {
        return;
      }
      seen = generation_;
    }
    work();
    std::lock_guard<std::mutex> lock(mu_);
    if (--busy_ == 0) {
      idle_.notify_one();
    }
  }
}

void Workers::run(std::size_t n, const std::function<void(std::size_t)>& fn)
{
  if (threads_.empty() || n <= 1) {
    for (std::size_t i = 0; i < n; ++i) {
      fn(i);
    }
    return;
  }
  {
    std::lock_guard<std::mutex> lock(mu_);
    fn_ = &fn;
    n_ = n;
    next_ = 0;
    error_ = nullptr;
    busy_ = threads_.size();
    ++generation_;
  }
  wake_.notify_all();
  work();
  std::exception_ptr error;
  {
    std::unique_lock<std::mutex> lock(mu_);
    idle_.wait(lock, [&] { return busy_ == 0; });
    fn_ = nullptr;
    error = error_;
  }
  if (error) {
    std::rethrow_exception(error);
  }
}

Workers& Runner::workers()
{
  if (!workers_) {
    workers_ = std::make_unique<Workers>(host_threads());
  }
  return *workers_;
}

Runner::Runner(std::shared_ptr<hip::Context> ctx)
    : ctx_(std::move(ctx))
{
  ctx_->activate();
  for (hipEvent_t& e : ev_) {
    hip::check(hipEventCreate(&e), "hipEventCreate");
  }
}

void Runner::StepGraph::reset()
{
  if (exec != nullptr) {
    (void)hipGraphExecDestroy(exec);
  }
  if (graph != nullptr) {
    (void)hipGraphDestroy(graph);
  }
  *this = StepGraph {};
}

Runner::~Runner()
{
  ctx_->activate();
  graph_.reset();
  for (hipEvent_t e : ev_) {
    (void)hipEventDestroy(e);
  }
}

void Runner::ensure(hip::DeviceBuffer& d, std::size_t bytes)
{
  if (d.bytes() < bytes) {
    d.reset(std::max(bytes, d.bytes() * 2));
  }
}

template<typename T>
void Runner::ensure_host(hip::PinnedBuffer<T>& h, std::size_t n)
{
  if (h.size() < n) {
    h.reset(std::max(n, h.size() * 2));
  }
}

void Runner::run(std::span<const Request> reqs)
{
  using hip::check;
  ctx_->activate();
  hipStream_t const s = ctx_->stream();
  std::size_t const n = reqs.size();
  slots_.assign(n, Placement {});
  if (n == 0) {
    return;
  }

  // Order items by launch: samplers, enumerators, bounders.
  std::vector<std::size_t> order;
  order.reserve(n);
  std::size_t counts[3] = {0, 0, 0};
  for (int pass = 0; pass < 3; ++pass) {
    dev::Mode const m = pass == 0 ? dev::Mode::Sample : pass == 1 ? dev::Mode::Enumerate : dev::Mode::Bound;
    for (std::size_t i = 0; i < n; ++i) {
      if (reqs[i].mode == m) {
        order.push_back(i);
        ++counts[pass];
      }
    }
  }

  // Waves per workgroup of each launch (DeviceTape::slice_waves): when a
  // pass's tiles leave compute units idle, every item's function groups
  // spread over several workgroups per tile (sliced items), so a long tape's
  // tile no longer runs all its groups on one compute unit; otherwise each
  // item's groups share one workgroup. CMI_ITEM_WAVES overrides (0: one
  // workgroup per tile).
  std::uint32_t pass_waves[3] = {0, 0, 0};
  {
    std::size_t pass_tiles[3] = {0, 0, 0};
    std::size_t pass_words[3] = {0, 0, 0};
    std::uint32_t pass_groups[3] = {1, 1, 1};
    for (std::size_t k = 0; k < n; ++k) {
      int const pass = k < counts[0] ? 0 : k < counts[0] + counts[1] ? 1 : 2;
      Request const& r = reqs[order[k]];
      std::size_t const elems = region::composition_fan_out(r.assignment->widths)
          * (r.mode == dev::Mode::Sample ? r.samples : 1);
      pass_tiles[pass] += (elems + dev::kTile - 1) / dev::kTile;
      pass_groups[pass] = std::max(pass_groups[pass], r.tape->item_groups_max());
      pass_words[pass] = std::max<std::size_t>(pass_words[pass], r.tape->n_instr());
    }
    for (int pass = 0; pass < 3; ++pass) {
      pass_waves[pass] =
          hip::DeviceTape::slice_waves(pass_groups[pass], pass_tiles[pass], ctx_->n_cus(), pass_words[pass]);
    }
  }
  {
    int forced[3];
    item_waves_override(forced);
    for (int pass = 0; pass < 3; ++pass) {
      if (forced[pass] >= 0) {
        pass_waves[pass] = static_cast<std::uint32_t>(forced[pass]);
      }
    }
  }

  // Sizes.
  std::size_t tiles[3] = {0, 0, 0};  // workgroups: tiles times slices
  std::size_t total_regions = 0, n_cand = 0, witness_doubles = 0, data_bytes = 0, sliced_tiles = 0;
  std::uint32_t lds[3] = {1, 1, 1};    // most slots of any item's code
  std::uint32_t waves[3] = {1, 1, 1};  // most waves of any item's workgroup
  bool ext[3] = {false, false, false};
  plans_.resize(n);
  double step_ops = 0.0;
  for (std::size_t k = 0; k < n; ++k) {
    std::size_t const i = order[k];
    Request const& r = reqs[i];
    int const pass = k < counts[0] ? 0 : k < counts[0] + counts[1] ? 1 : 2;
    std::size_t const slots = r.assignment->size();
    std::size_t const n_regions = region::composition_fan_out(r.assignment->widths);
    std::size_t const samples = r.mode == dev::Mode::Sample ? r.samples : 1;
    if (n_regions == 0 || samples == 0 || n_regions > (std::size_t(1) << 31) / samples) {
      throw InvalidConfiguration("batched request: n_regions * samples must be in [1, 2^31)");
    }
    if (slots > dev::kMaxSlotsPerItem) {
      throw InvalidConfiguration("batched request: too many slots");
    }
    if (r.box.size() != r.tape->n_vars()) {
      throw ShapeMismatch("batched request: box size does not match the problem's variable count");
    }
    std::size_t const n_elems = n_regions * samples;
    Placement& p = slots_[i];
    p.item = k;
    p.n_regions = n_regions;
    p.n_elems = n_elems;
    p.n_vars = r.tape->n_vars();
    step_ops += static_cast<double>(n_elems)
        * static_cast<double>(r.mode == dev::Mode::Bound ? r.tape->tape().flops : r.tape->tape().reachable_nodes);
    hip::DeviceTape::ItemPlan& plan = plans_[k];
    plan = r.tape->item_plan(r.mode, pass_waves[pass]);
    std::size_t const n_tiles = (n_elems + dev::kTile - 1) / dev::kTile;
    tiles[pass] += n_tiles * plan.slices;
    lds[pass] = std::max(lds[pass], plan.slots);
    waves[pass] = std::max(waves[pass], plan.waves);
    ext[pass] = ext[pass] || r.tape->extended();
    if (r.mode == dev::Mode::Bound) {
      p.region_off = total_regions;
      total_regions += n_regions;
    } else {
      p.cand_off = n_cand++;
      p.witness_off = witness_doubles;
      witness_doubles += p.n_vars;
    }
    data_bytes += align_up(p.n_vars * 16, 16) + align_up(slots * 4, 16) * 2 + align_up(slots * 8, 16)
        + align_up(slots, 16);
  }
  // A launch whose largest code and widest workgroup together do not fit in
  // LDS runs every item's whole tape on one wave instead.
  for (int pass = 0; pass < 3; ++pass) {
    dev::Mode const m = pass == 0 ? dev::Mode::Sample : pass == 1 ? dev::Mode::Enumerate : dev::Mode::Bound;
    if (waves[pass] > 1 && dev::items_lds_bytes(m, lds[pass], waves[pass]) > ctx_->lds_bytes()) {
      tiles[pass] = 0;
      lds[pass] = 1;
      waves[pass] = 1;
      for (std::size_t k = 0; k < n; ++k) {
        int const kp = k < counts[0] ? 0 : k < counts[0] + counts[1] ? 1 : 2;
        if (kp == pass) {
          plans_[k] = reqs[order[k]].tape->item_plan_whole();
          tiles[pass] += (slots_[order[k]].n_elems + dev::kTile - 1) / dev::kTile;
          lds[pass] = std::max(lds[pass], plans_[k].slots);
        }
      }
    }
  }
  for (std::size_t k = 0; k < n; ++k) {
    if (plans_[k].slices > 1) {
      sliced_tiles += (slots_[order[k]].n_elems + dev::kTile - 1) / dev::kTile;
    }
  }
  // Points per lane of the sampler and enumerator launches (points_per_lane):
  // a launch whose workgroup still fits in LDS evaluates that many points per
  // lane, its items' tiles then covering kTile * ppl elements each.
  std::uint32_t ppl[3] = {1, 1, 1};
  {
    std::uint32_t const forced = points_per_lane_override();
    for (int pass = 0; pass < 2; ++pass) {
      dev::Mode const m = pass == 0 ? dev::Mode::Sample : dev::Mode::Enumerate;
      // four points of the extended op set would spill registers
      std::uint32_t const want
          = std::min<std::uint32_t>(forced != 0 ? forced : points_per_lane(tiles[pass], ctx_->n_cus()), ext[pass] ? 2 : 4);
      if (want <= 1 || counts[pass] == 0 || dev::items_lds_bytes(m, lds[pass], waves[pass], want) > ctx_->lds_bytes()) {
        continue;
      }
      ppl[pass] = want;
      tiles[pass] = 0;
      for (std::size_t k = 0; k < n; ++k) {
        int const kp = k < counts[0] ? 0 : k < counts[0] + counts[1] ? 1 : 2;
        if (kp == pass) {
          tiles[pass] += (slots_[order[k]].n_elems + dev::kTile * want - 1) / (dev::kTile * want) * plans_[k].slices;
        }
      }
    }
  }
  if (lds[2] > ctx_->lds_interval_slots() || std::max(lds[0], lds[1]) > ctx_->lds_point_slots()) {
    throw ResourceExhausted("batched request needs more LDS slots than a workgroup holds");
  }
  std::size_t const total_tiles = tiles[0] + tiles[1] + tiles[2];
  std::size_t const cand_tiles = tiles[0] + tiles[1];

  std::size_t const off_items = 0;
  std::size_t const off_tile_item = align_up(n * sizeof(dev::Item), 256);
  std::size_t const off_data = align_up(off_tile_item + total_tiles * 4, 256);
  std::size_t const staging_bytes = off_data + data_bytes;

  off_obj_lb_ = align_up(total_regions, 16);
  off_cand_ = off_obj_lb_ + total_regions * 8;
  off_witness_ = off_cand_ + n_cand * sizeof(Candidate);
  results_bytes_ = std::max<std::size_t>(off_witness_ + witness_doubles * 8, 16);
  std::size_t const scratch_bytes = std::max<std::size_t>(cand_tiles * 16, 16);

  ensure(staging_d_, staging_bytes);
  ensure(results_d_, results_bytes_);
  ensure(scratch_d_, scratch_bytes);
  if (done_capacity_ < n_cand) {
    std::size_t const cap = std::max<std::size_t>(n_cand, done_capacity_ * 2);
    done_d_.reset(cap * 4);
    check(hipMemsetAsync(done_d_.get(), 0, cap * 4, s), "hipMemsetAsync(done)");
    done_capacity_ = cap;
  }
  ensure_host(staging_h_, staging_bytes);
  ensure_host(results_h_, results_bytes_);
  if (slice_capacity_ < sliced_tiles) {
    std::size_t const cap = std::max(sliced_tiles, slice_capacity_ * 2);
    slice_d_.reset(cap * dev::kSliceTileBytes + 256);
    check(hipMemsetAsync(slice_d_.get(), 0, slice_d_.bytes(), s), "hipMemsetAsync(slice scratch)");
    slice_capacity_ = cap;
  }
  // slice scratch: every tile's masks, then counts, then objectives
  std::size_t const slice_masks_bytes = slice_capacity_ * dev::kSliceMasks * 8;
  auto* const slice_inf = slice_d_.as<std::uint64_t>();
  auto* const slice_done = reinterpret_cast<std::uint32_t*>(slice_d_.as<unsigned char>() + slice_masks_bytes);
  auto* const slice_obj = reinterpret_cast<double*>(slice_d_.as<unsigned char>()
                                                    + align_up(slice_masks_bytes + slice_capacity_ * 4, 16));
  std::size_t slice_cursor = 0;

  unsigned char* const h = staging_h_.data();
  auto* const d = staging_d_.as<unsigned char>();
  auto* const items = reinterpret_cast<dev::Item*>(h + off_items);
  auto* const tile_item = reinterpret_cast<std::uint32_t*>(h + off_tile_item);
  auto* const res = results_d_.as<unsigned char>();
  auto* const scratch = scratch_d_.as<unsigned char>();

  std::size_t data_off = off_data;
  std::size_t tile_cursor = 0;       // global tile index
  std::size_t cand_tile_cursor = 0;  // tile index inside the candidate scratch
  std::size_t launch_first_tile = 0;
  std::size_t launch_first_item = 0;
  int current_pass = -1;
  for (std::size_t k = 0; k < n; ++k) {
    std::size_t const i = order[k];
    Request const& r = reqs[i];
    Placement const& p = slots_[i];
    int const pass = k < counts[0] ? 0 : k < counts[0] + counts[1] ? 1 : 2;
    if (pass != current_pass) {
      current_pass = pass;
      launch_first_tile = tile_cursor;
      launch_first_item = k;
    }
    std::size_t const slots = r.assignment->size();
    std::size_t const off_parent = data_off;
    std::size_t const off_var = off_parent + align_up(p.n_vars * 16, 16);
    std::size_t const off_width = off_var + align_up(slots * 4, 16);
    std::size_t const off_step = off_width + align_up(slots * 4, 16);
    std::size_t const off_kind = off_step + align_up(slots * 8, 16);
    data_off = off_kind + align_up(slots, 16);

    std::memcpy(h + off_parent, r.box.data(), p.n_vars * 16);
    auto* var = reinterpret_cast<std::uint32_t*>(h + off_var);
    auto* width = reinterpret_cast<std::uint32_t*>(h + off_width);
    auto* step = reinterpret_cast<double*>(h + off_step);
    auto* kind = reinterpret_cast<std::uint8_t*>(h + off_kind);
    for (std::size_t j = 0; j < slots; ++j) {
      if (r.assignment->var_ids[j] >= p.n_vars) {
        throw ShapeMismatch("batched request: slot var_id out of range");
      }
      var[j] = static_cast<std::uint32_t>(r.assignment->var_ids[j]);
      width[j] = static_cast<std::uint32_t>(r.assignment->widths[j]);
      step[j] = (r.box[var[j]].ub - r.box[var[j]].lb) / static_cast<double>(width[j]);
      kind[j] = static_cast<std::uint8_t>(r.assignment->composition[j]);
    }

    std::size_t const tile_elems = std::size_t {dev::kTile} * ppl[pass];
    std::uint32_t const n_tiles = static_cast<std::uint32_t>((p.n_elems + tile_elems - 1) / tile_elems);
    hip::DeviceTape::ItemPlan const& plan = plans_[k];
    dev::Item& it = items[k];
    std::memset(&it, 0, sizeof(it));
    r.tape->item_code(it, plan);
    if (plan.slices > 1) {
      it.tile_infeasible = slice_inf + slice_cursor * dev::kSliceMasks;
      it.tile_slices_done = slice_done + slice_cursor;
      it.tile_obj = slice_obj + slice_cursor * dev::kSliceObjDoubles;
      slice_cursor += n_tiles;
    }
    it.n_vars = static_cast<std::uint32_t>(p.n_vars);
    it.parent = reinterpret_cast<const double*>(d + off_parent);
    it.var_kinds = r.tape->var_kinds();
    it.slot_var = reinterpret_cast<const std::uint32_t*>(d + off_var);
    it.slot_width = reinterpret_cast<const std::uint32_t*>(d + off_width);
    it.slot_step = reinterpret_cast<const double*>(d + off_step);
    it.slot_kind = reinterpret_cast<const std::uint8_t*>(d + off_kind);
    it.slot_count = static_cast<std::uint32_t>(slots);
    it.samples = static_cast<std::uint32_t>(r.mode == dev::Mode::Sample ? r.samples : 1);
    it.n_elems = p.n_elems;
    it.salt = r.salt;
    it.tile_begin = static_cast<std::uint32_t>(tile_cursor - launch_first_tile);
    it.n_tiles = n_tiles;
    if (r.mode == dev::Mode::Bound) {
      it.feasible = res + p.region_off;
      it.obj_lb = reinterpret_cast<double*>(res + off_obj_lb_) + p.region_off;
    } else {
      it.tile_min = reinterpret_cast<double*>(scratch) + cand_tile_cursor;
      it.tile_arg = reinterpret_cast<std::int64_t*>(scratch + cand_tiles * 8) + cand_tile_cursor;
      it.done = done_d_.as<std::uint32_t>() + p.cand_off;
      it.cand_value = reinterpret_cast<double*>(res + off_cand_ + p.cand_off * sizeof(Candidate));
      it.cand_index = reinterpret_cast<std::int64_t*>(res + off_cand_ + p.cand_off * sizeof(Candidate) + 8);
      it.witness = reinterpret_cast<double*>(res + off_witness_) + p.witness_off;
      cand_tile_cursor += n_tiles;
    }
    std::size_t const blocks = std::size_t {n_tiles} * plan.slices;
    for (std::size_t t = 0; t < blocks; ++t) {
      tile_item[tile_cursor + t] = static_cast<std::uint32_t>(k - launch_first_item);
    }
    tile_cursor += blocks;
  }

  // the launches of every kTimedEvery-th run are bracketed by events
  bool const timed = runs_++ % kTimedEvery == 0;
  dev::ItemsLaunch launch[3];
  void* args[3][3];
  const dev::Item* items_d[3];
  const std::uint32_t* tile_item_d[3];
  std::uint32_t slots_arg[3];
  {
    std::size_t first_item = 0, first_tile = 0;
    for (int pass = 0; pass < 3; ++pass) {
      dev::Mode const m = pass == 0 ? dev::Mode::Sample : pass == 1 ? dev::Mode::Enumerate : dev::Mode::Bound;
      items_d[pass] = reinterpret_cast<const dev::Item*>(d + off_items) + first_item;
      tile_item_d[pass] = reinterpret_cast<const std::uint32_t*>(d + off_tile_item) + first_tile;
      slots_arg[pass] = lds[pass];
      args[pass][0] = &items_d[pass];
      args[pass][1] = &tile_item_d[pass];
      args[pass][2] = &slots_arg[pass];
      if (counts[pass] > 0) {
        launch[pass] = dev::items_launch(m, ext[pass], static_cast<std::uint32_t>(tiles[pass]), lds[pass],
                                         waves[pass], ppl[pass]);
        ++launches_;
      }
      first_item += counts[pass];
      first_tile += tiles[pass];
    }
  }
  if (step_graph()) {
    // one graph per shape of step (which kinds launch, which of them run the
    // extended op set, and their points per lane): copy in, the launches between two event records,
    // copy out; every step sets its nodes' parameters and launches it
    std::uint32_t key = 0;
    for (int pass = 0; pass < 3; ++pass) {
      key |= (counts[pass] > 0 ? 1u : 0u) << pass;
      key |= (counts[pass] > 0 && ext[pass] ? 1u : 0u) << (pass + 3);
      key |= (counts[pass] > 0 ? ppl[pass] : 0u) << (8 + 4 * pass);
    }
    if (!graph_.exec || graph_.key != key) {
      graph_.reset();
      check(hipGraphCreate(&graph_.graph, 0), "hipGraphCreate");
      check(hipGraphAddMemcpyNode1D(&graph_.h2d, graph_.graph, nullptr, 0, d, h, staging_bytes,
                                    hipMemcpyHostToDevice),
            "hipGraphAddMemcpyNode1D");
      check(hipGraphAddEventRecordNode(&graph_.ev0, graph_.graph, &graph_.h2d, 1, ev_[0]),
            "hipGraphAddEventRecordNode");
      hipGraphNode_t prev = graph_.ev0;
      for (int pass = 0; pass < 3; ++pass) {
        if (counts[pass] == 0) {
          continue;
        }
        hipKernelNodeParams kp {};
        kp.func = const_cast<void*>(launch[pass].func);
        kp.gridDim = dim3(launch[pass].grid);
        kp.blockDim = dim3(launch[pass].block);
        kp.sharedMemBytes = static_cast<unsigned>(launch[pass].lds_bytes);
        kp.kernelParams = args[pass];
        check(hipGraphAddKernelNode(&graph_.kernel[pass], graph_.graph, &prev, 1, &kp), "hipGraphAddKernelNode");
        prev = graph_.kernel[pass];
      }
      check(hipGraphAddEventRecordNode(&graph_.ev1, graph_.graph, &prev, 1, ev_[1]),
            "hipGraphAddEventRecordNode");
      check(hipGraphAddMemcpyNode1D(&graph_.d2h, graph_.graph, &graph_.ev1, 1, results_h_.data(), results_d_.get(),
                                    results_bytes_, hipMemcpyDeviceToHost),
            "hipGraphAddMemcpyNode1D");
      check(hipGraphInstantiate(&graph_.exec, graph_.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      graph_.key = key;
    } else {
      check(hipGraphExecMemcpyNodeSetParams1D(graph_.exec, graph_.h2d, d, h, staging_bytes, hipMemcpyHostToDevice),
            "hipGraphExecMemcpyNodeSetParams1D");
      for (int pass = 0; pass < 3; ++pass) {
        if (counts[pass] == 0) {
          continue;
        }
        hipKernelNodeParams kp {};
        kp.func = const_cast<void*>(launch[pass].func);
        kp.gridDim = dim3(launch[pass].grid);
        kp.blockDim = dim3(launch[pass].block);
        kp.sharedMemBytes = static_cast<unsigned>(launch[pass].lds_bytes);
        kp.kernelParams = args[pass];
        check(hipGraphExecKernelNodeSetParams(graph_.exec, graph_.kernel[pass], &kp),
              "hipGraphExecKernelNodeSetParams");
      }
      check(hipGraphExecMemcpyNodeSetParams1D(graph_.exec, graph_.d2h, results_h_.data(), results_d_.get(),
                                              results_bytes_, hipMemcpyDeviceToHost),
            "hipGraphExecMemcpyNodeSetParams1D");
    }
    check(hipGraphLaunch(graph_.exec, s), "hipGraphLaunch");
  } else {
    check(hipMemcpyAsync(d, h, staging_bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync(batch staging)");
    if (timed) {
      check(hipEventRecord(ev_[0], s), "hipEventRecord");
    }
    for (int pass = 0; pass < 3; ++pass) {
      if (counts[pass] > 0) {
        check(hipLaunchKernel(launch[pass].func, dim3(launch[pass].grid), dim3(launch[pass].block), args[pass],
                              launch[pass].lds_bytes, s),
              "launch(k_items batch)");
      }
    }
    if (timed) {
      check(hipEventRecord(ev_[1], s), "hipEventRecord");
    }
    check(hipMemcpyAsync(results_h_.data(), results_d_.get(), results_bytes_, hipMemcpyDeviceToHost, s),
          "hipMemcpyAsync(batch results)");
  }
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (timed) {
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
    timed_seconds_ += static_cast<double>(ms) * 1e-3;
    timed_ops_ += step_ops;
  }
  fp64_ops_ += step_ops;
}

std::size_t Runner::free_bytes(hip::Context& ctx)
{
  if (free_steps_ == 0) {
    free_bytes_ = ctx.free_bytes();
  }
  free_steps_ = (free_steps_ + 1) % kFreeBytesEvery;
  return free_bytes_;
}

double Runner::kernel_seconds() const
{
  // the timed runs' device time per FP64 operation, over all runs' operations
  return timed_ops_ > 0.0 ? timed_seconds_ * (fp64_ops_ / timed_ops_) : timed_seconds_;
}

std::span<const unsigned char> Runner::feasible(std::size_t i) const
{
  Placement const& p = slots_[i];
  return {results_h_.data() + p.region_off, p.n_regions};
}

std::span<const double> Runner::obj_lb(std::size_t i) const
{
  Placement const& p = slots_[i];
  return {reinterpret_cast<const double*>(results_h_.data() + off_obj_lb_) + p.region_off, p.n_regions};
}

backend::CandidateResult Runner::candidate(std::size_t i) const
{
  Placement const& p = slots_[i];
  Candidate c;
  std::memcpy(&c, results_h_.data() + off_cand_ + p.cand_off * sizeof(Candidate), sizeof(c));
  backend::CandidateResult out;
  out.found = c.value < std::numeric_limits<double>::max();
  out.value = c.value;
  out.index = c.index;
  out.point = {reinterpret_cast<const double*>(results_h_.data() + off_witness_) + p.witness_off, p.n_vars};
  return out;
}

std::size_t shared_host_budget(const cmi_solve_options& o, std::size_t residents)
{
  if (o.host_budget_bytes != 0) {
    return o.host_budget_bytes;
  }
  std::size_t const total = search::resolve_host_budget(0);
  if (total == 0) {
    return 0;  // meminfo unreadable: unbudgeted, as a plain solve would be
  }
  return std::max<std::size_t>(1, total / std::max<std::size_t>(residents, 1));
}

search::SolveOptions search_options(const cmi_solve_options& o, std::size_t host_budget)
{
  search::SolveOptions so;
  so.iter_limit = o.iter_limit;
  so.tolerance = o.tolerance;
  so.sample_points = o.sample_points;
  so.host_budget_bytes = host_budget;
  so.frontier_policy = o.compact_frontier ? search::FrontierPolicy::Compact : search::FrontierPolicy::StopAtBudget;
  if (o.has_known_primal) {
    so.known_primal_bound = o.known_primal;
  }
  return so;
}

std::size_t step(Runner& runner, std::span<Member> members, const cmi_solve_options& o, hip::Context& ctx)
{
  // One device budget for the whole step: with no budget given, the free
  // device memory, read every kFreeBytesEvery steps (a read costs 13-19 us,
  // 5 % of a small served step); a plain solve reads it once per role build.
  std::size_t const budget = o.budget_bytes != 0 ? o.budget_bytes : runner.free_bytes(ctx);
  struct Owner
  {
    Member* member;
    std::size_t first_req;
    bool enumerate;
  };
  // What each live solve asks for this step. next(), the budget check and,
  // after the launch, on_subdivide/on_enumerate touch only that solve's own
  // state, so for a large step they run on the host workers; the requests
  // are packed in member order either way.
  struct Pending
  {
    const search::SolveState::Request* q = nullptr;
    bool enumerate = false;
  };
  std::vector<Pending> pending(members.size());
  auto ask = [&](std::size_t i) {
    Member& m = members[i];
    search::SolveState& st = *m.state;
    if (st.finished()) {
      return;
    }
    const search::SolveState::Request* q = nullptr;
    try {
      q = st.next();
    } catch (const std::bad_alloc&) {
      st.stop(search::StopReason::AllocationFailure);
      return;
    }
    if (q == nullptr) {
      return;
    }
    bool const enumerate = q->kind == search::SolveState::Kind::Enumerate;
    std::size_t const n_regions = region::composition_fan_out(q->assignment->widths);
    // The plain driver builds (and prices) the sampler, then the bounder, or
    // the enumerator on a fathom; a build that does not fit stops the solve.
    try {
      if (m.tape == nullptr) {
        throw ResourceExhausted("the problem's device tape could not be built");
      }
      region::Composition const& c = q->assignment->composition;
      if (enumerate) {
        hip::check_role_fits(*m.tape, dev::Mode::Enumerate, c, n_regions, o.sample_points, budget);
      } else {
        hip::check_role_fits(*m.tape, dev::Mode::Sample, c, n_regions, o.sample_points, budget);
        hip::check_role_fits(*m.tape, dev::Mode::Bound, c, n_regions, o.sample_points, budget);
      }
    } catch (const backend::OverBudgetError& e) {
      st.stop_over_budget(e.facts());
      return;
    } catch (const ResourceExhausted&) {
      st.stop(search::StopReason::DeviceMemory);
      return;
    }
    pending[i] = {q, enumerate};
  };
  bool const parallel = members.size() >= kParallelMembers;
  if (parallel) {
    runner.workers().run(members.size(), ask);
  } else {
    for (std::size_t i = 0; i < members.size(); ++i) {
      ask(i);
    }
  }

  std::vector<Request> reqs;
  std::vector<Owner> owners;
  reqs.reserve(2 * members.size());
  owners.reserve(members.size());
  for (std::size_t i = 0; i < members.size(); ++i) {
    const search::SolveState::Request* q = pending[i].q;
    if (q == nullptr) {
      continue;
    }
    Member& m = members[i];
    owners.push_back({&m, reqs.size(), pending[i].enumerate});
    if (pending[i].enumerate) {
      reqs.push_back({m.tape, dev::Mode::Enumerate, q->box, q->assignment, q->salt, 1});
    } else {
      reqs.push_back({m.tape, dev::Mode::Sample, q->box, q->assignment, q->salt,
                      static_cast<std::uint32_t>(o.sample_points)});
      reqs.push_back({m.tape, dev::Mode::Bound, q->box, q->assignment, q->salt, 1});
    }
  }
  if (reqs.empty()) {
    return 0;
  }
  runner.run(reqs);
  auto answer = [&](std::size_t k) {
    Owner const& w = owners[k];
    search::SolveState& st = *w.member->state;
    try {
      if (w.enumerate) {
        st.on_enumerate(runner.candidate(w.first_req), runner.n_elems(w.first_req));
      } else {
        backend::BoundResult const b {runner.feasible(w.first_req + 1), runner.obj_lb(w.first_req + 1),
                                      runner.n_elems(w.first_req + 1)};
        st.on_subdivide(runner.candidate(w.first_req), runner.n_elems(w.first_req), b);
      }
    } catch (const std::bad_alloc&) {
      st.stop(search::StopReason::AllocationFailure);
    }
  };
  if (parallel) {
    runner.workers().run(owners.size(), answer);
  } else {
    for (std::size_t k = 0; k < owners.size(); ++k) {
      answer(k);
    }
  }
  return reqs.size();
}

std::vector<Result> solve_batch(std::shared_ptr<hip::Context> ctx,
                                std::span<const model::Problem<double>* const> problems,
                                const cmi_solve_options& o, BatchStats* stats)
{
  auto const t0 = std::chrono::steady_clock::now();
  std::size_t const n = problems.size();
  auto pol = capi::make_policy(o);
  search::SolveOptions const so = search_options(o, shared_host_budget(o, n));
  // Each solve's tape (compiled on the host, uploaded once) and search state
  // are independent of the others': build them on several host threads. A
  // tape that cannot be built stops only its own solve, at its first
  // iteration, with DeviceMemory (where a plain solve's first role build
  // would have failed).
  std::vector<std::unique_ptr<hip::DeviceTape>> tapes(n);
  std::vector<std::unique_ptr<search::SolveState>> states(n);
  parallel_for(n, [&](std::size_t i) {
    states[i] = std::make_unique<search::SolveState>(*problems[i], pol, so);
    try {
      tapes[i] = std::make_unique<hip::DeviceTape>(ctx, *problems[i]);
    } catch (const ResourceExhausted&) {
      tapes[i].reset();
    } catch (const std::bad_alloc&) {
      tapes[i].reset();
    }
  });
  std::vector<Member> members(n);
  for (std::size_t i = 0; i < n; ++i) {
    members[i] = Member {states[i].get(), tapes[i].get()};
  }
  std::vector<Result> results(n);
  std::vector<bool> done(n, false);
  Runner runner(ctx);
  std::size_t remaining = n;
  std::uint64_t steps = 0;
  while (remaining > 0) {
    step(runner, members, o, *ctx);
    ++steps;
    double const now = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (std::size_t i = 0; i < n; ++i) {
      if (!done[i] && states[i]->finished()) {
        results[i].outcome = states[i]->finish();
        results[i].seconds = now;
        done[i] = true;
        --remaining;
      }
    }
  }
  if (stats != nullptr) {
    stats->steps = steps;
    stats->launches = runner.launches();
    stats->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats->kernel_seconds = runner.kernel_seconds();
    stats->fp64_ops = runner.fp64_ops();
  }
  return results;
}

}  // namespace cmi::batch
