#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <limits>
#include <memory>
#include <new>
#include <optional>
#include <span>
#include <string>
#include <utility>
#include <vector>

#include "cmi/backend.hpp"
#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/policy.hpp"
#include "cmi/region.hpp"

// Host best-first branch and bound: the frontier heap, the parent-box
// history, the host budget, the epilogue and the driver loop.
//
// Node order is part of the result (iteration counts, the incumbent found
// under an iteration cap), so everything that decides it is restated exactly
// from the reference: the comparator (search/frontier.hpp:66-89), the binary
// heap's sift rules including its tie behaviour (frontier.hpp:152-338), the
// history's LIFO slot reuse that produces pidx values (search/history.hpp:
// 41-104) and the driver's per-iteration protocol (search/driver.hpp:221-407).
namespace cmi::search
{

enum class StopReason
{
  Converged,
  Exhausted,
  IterationLimit,
  HostMemory,
  AllocationFailure,
  DeviceMemory,
};

inline const char* to_string(StopReason r)
{
  switch (r) {
    case StopReason::Converged: return "converged";
    case StopReason::Exhausted: return "exhausted";
    case StopReason::IterationLimit: return "iteration-limit";
    case StopReason::HostMemory: return "host-memory";
    case StopReason::AllocationFailure: return "allocation-failure";
    case StopReason::DeviceMemory: return "device-memory";
  }
  return "unknown";
}

enum class FrontierPolicy
{
  StopAtBudget,
  Compact,
};

/// gub - bound <= tol * max(1, |gub|) (reference search/budget.hpp:215-218).
inline bool gap_closed(double gub, double bound, double tolerance)
{
  return gub - bound <= tolerance * std::max(1.0, std::fabs(gub));
}

inline std::size_t available_host_bytes()
{
  std::ifstream meminfo("/proc/meminfo");
  std::string key;
  while (meminfo >> key) {
    if (key != "MemAvailable:") {
      meminfo.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
      continue;
    }
    unsigned long long value = 0;
    std::string unit;
    if (!(meminfo >> value)) {
      return 0;
    }
    meminfo >> unit;
    return static_cast<std::size_t>(value) * (unit == "kB" ? 1024u : 1u);
  }
  return 0;
}

inline std::size_t resolve_host_budget(std::size_t requested)
{
  return requested != 0 ? requested : available_host_bytes() / 2;
}

inline bool over_host_budget(std::size_t live_bytes, std::size_t budget_bytes)
{
  return budget_bytes != 0 && live_bytes >= budget_bytes;
}

inline std::size_t compact_keep_count(std::size_t budget_bytes, std::size_t history_bytes, std::size_t node_bytes)
{
  std::size_t const room = budget_bytes - std::min(budget_bytes, history_bytes);
  return std::max<std::size_t>(1, room / (4 * node_bytes));
}

struct DropAccounting
{
  double lb_min = std::numeric_limits<double>::infinity();
  std::size_t viable = 0;
  std::size_t dominated = 0;

  void fold(double lb, double gub)
  {
    if (lb <= gub) {
      lb_min = std::min(lb_min, lb);
      ++viable;
    } else {
      ++dominated;
    }
  }

  bool lost_nothing() const { return viable == 0; }
};

/// Box containment plus a 1e-6 integrality check (reference epilogue.hpp:35-59).
inline bool witness_is_admissible(std::span<const double> point, std::span<const model::VarKind> kinds,
                                  std::span<const interval> root_box)
{
  constexpr double kIntegralityTolerance = 1e-6;
  if (point.size() != kinds.size() || point.size() != root_box.size()) {
    return false;
  }
  for (std::size_t i = 0; i < point.size(); ++i) {
    double const v = point[i];
    if (!(v >= root_box[i].lb) || !(v <= root_box[i].ub)) {
      return false;
    }
    if (kinds[i] == model::VarKind::Integer || kinds[i] == model::VarKind::Binary) {
      if (std::fabs(v - std::round(v)) > kIntegralityTolerance) {
        return false;
      }
    }
  }
  return true;
}

struct FinalBounds
{
  double glb;
  bool proven_optimal;
  bool infeasible;
};

/// Final bracket and claims (reference epilogue.hpp:89-127).
inline FinalBounds finalise_bounds(double glb, double gub, double tolerance, bool found_incumbent, bool converged,
                                   bool pending_empty, double frontier_min, std::size_t viable,
                                   const DropAccounting& dropped)
{
  FinalBounds out {glb, false, false};
  out.proven_optimal = found_incumbent && (converged || pending_empty || viable == 0)
      && gap_closed(gub, dropped.lb_min, tolerance);
  if (out.proven_optimal) {
    out.glb = gub;
  } else if (pending_empty && dropped.lost_nothing()) {
    out.infeasible = true;
  } else {
    double const frontier = pending_empty ? std::numeric_limits<double>::max() : frontier_min;
    out.glb = std::min(frontier, dropped.lb_min);
    if (found_incumbent) {
      out.glb = std::min(out.glb, gub);
    }
  }
  return out;
}

/// Parent boxes, refcounted, slots reused LIFO; slot 0 is the root sentinel.
struct IntervalHistory
{
  std::vector<std::vector<interval>> intervals;

  std::size_t enqueue(std::vector<interval> box)
  {
    std::size_t const bytes = entry_bytes(box.size());
    std::size_t idx = 0;
    if (!free_slots_.empty()) {
      idx = free_slots_.back();
      free_slots_.pop_back();
      intervals[idx] = std::move(box);
      refs_[idx] = 1;
    } else {
      idx = intervals.size();
      intervals.push_back(std::move(box));
      refs_.push_back(1);
    }
    live_bytes_ += bytes;
    peak_live_ = std::max(peak_live_, live_count());
    return idx;
  }

  void add_ref(std::size_t idx)
  {
    if (idx != 0) {
      ++refs_[idx];
    }
  }

  void release(std::size_t idx)
  {
    if (idx == 0) {
      return;
    }
    if (--refs_[idx] > 0) {
      return;
    }
    live_bytes_ -= entry_bytes(intervals[idx].size());
    std::vector<interval>().swap(intervals[idx]);
    free_slots_.push_back(idx);
    ++freed_;
  }

  std::size_t live_bytes() const { return live_bytes_; }
  std::size_t live_count() const { return intervals.size() - free_slots_.size(); }
  std::size_t ref_count(std::size_t idx) const { return refs_[idx]; }
  std::uint64_t freed_count() const { return freed_; }
  std::size_t peak_live() const { return peak_live_; }

private:
  static constexpr std::size_t entry_bytes(std::size_t n_vars)
  {
    return n_vars * sizeof(interval) + sizeof(std::vector<interval>);
  }

  std::vector<std::size_t> refs_;
  std::vector<std::size_t> free_slots_;
  std::size_t live_bytes_ = 0;
  std::uint64_t freed_ = 0;
  std::size_t peak_live_ = 0;
};

struct Node
{
  std::size_t sidx;
  std::size_t pidx;
  std::size_t depth;
  double lb;
  std::uint64_t assignment_hash = 0;

  /// Least lb first, then deeper first, then larger pidx first.
  bool operator<(const Node& o) const
  {
    if (lb < o.lb) {
      return true;
    }
    if (o.lb < lb) {
      return false;
    }
    if (depth != o.depth) {
      return depth > o.depth;
    }
    return pidx > o.pidx;
  }

  /// Rebuild this node's box from its parent in `history` by re-running the
  /// policy on the parent and decoding sidx; the hash check rejects a policy
  /// that is not pure (reference frontier.hpp:113-144).
  void materialise(const IntervalHistory& history, std::vector<interval>& out, const policy::CompositionPolicy& pol,
                   std::span<const model::VarKind> kinds, std::span<const interval> root_box) const
  {
    if (pidx == 0) {
      out.assign(root_box.begin(), root_box.end());
      return;
    }
    const std::vector<interval>& parent = history.intervals[pidx];
    region::SlotAssignment const a = pol.choose(parent, kinds);
    if (region::assignment_hash(a) != assignment_hash) {
      throw InvalidConfiguration(
          "CompositionPolicy::choose is not a pure function of (box, var_kinds): re-invoking it on this "
          "node's parent returned an assignment whose hash does not match the one used to encode the "
          "node's sidx");
    }
    region::materialise(parent, sidx, a, out);
  }
};

/// Binary min-heap of pending nodes with the reference's exact sift rules.
class IntervalPQueue
{
public:
  explicit IntervalPQueue(std::size_t capacity)
      : elems_(capacity)
  {
  }

  void enqueue(const Node& e)
  {
    if (n_ == elems_.size()) {
      elems_.resize(elems_.empty() ? 1 : elems_.size() * 2);
    }
    elems_[n_] = e;
    sift_up(n_);
    ++n_;
  }

  Node dequeue()
  {
    Node res = elems_[0];
    --n_;
    if (n_ > 0) {
      elems_[0] = elems_[n_];
      sift_down(0);
    }
    return res;
  }

  bool empty() const { return n_ == 0; }
  std::size_t size() const { return n_; }
  const Node& peek() const { return elems_[0]; }
  std::size_t capacity_bytes() const { return elems_.capacity() * sizeof(Node); }

  std::size_t count_viable(double gub) const
  {
    // iterative form of the reference's downward-closed recursive count
    std::size_t count = 0;
    std::vector<std::size_t> stack;
    if (n_ > 0) {
      stack.push_back(0);
    }
    while (!stack.empty()) {
      std::size_t const i = stack.back();
      stack.pop_back();
      if (i >= n_ || elems_[i].lb > gub) {
        continue;
      }
      ++count;
      stack.push_back(2 * i + 2);
      stack.push_back(2 * i + 1);
    }
    return count;
  }

  template<typename OnEvict>
  void compact(std::size_t n_keep, OnEvict&& on_evict)
  {
    if (n_ <= n_keep) {
      return;
    }
    std::nth_element(elems_.begin(), elems_.begin() + static_cast<std::ptrdiff_t>(n_keep),
                     elems_.begin() + static_cast<std::ptrdiff_t>(n_));
    for (std::size_t i = n_keep; i < n_; ++i) {
      on_evict(static_cast<const Node&>(elems_[i]));
    }
    n_ = n_keep;
    for (std::size_t i = n_ / 2; i-- > 0;) {
      sift_down(i);
    }
    if (elems_.size() > 2 * n_keep) {
      std::vector<Node> shrunk(2 * n_keep);
      std::copy(elems_.begin(), elems_.begin() + static_cast<std::ptrdiff_t>(n_), shrunk.begin());
      elems_.swap(shrunk);
    }
  }

private:
  void sift_up(std::size_t idx)
  {
    Node const e = elems_[idx];
    while (idx != 0 && e < elems_[(idx - 1) / 2]) {
      elems_[idx] = elems_[(idx - 1) / 2];
      idx = (idx - 1) / 2;
    }
    elems_[idx] = e;
  }

  void sift_down(std::size_t idx)
  {
    Node const e = elems_[idx];
    while (2 * idx + 1 < n_) {
      std::size_t const l = 2 * idx + 1;
      std::size_t const r = 2 * idx + 2;
      std::size_t const child = (r < n_ && elems_[r] < elems_[l]) ? r : l;
      if (!(elems_[child] < e)) {
        break;
      }
      elems_[idx] = elems_[child];
      idx = child;
    }
    elems_[idx] = e;
  }

  std::vector<Node> elems_;
  std::size_t n_ = 0;
};

struct RunTelemetry
{
  std::uint64_t iterations = 0;
  std::uint64_t fathomed = 0;
  std::uint64_t fathomed_points = 0;
  std::uint64_t bounded = 0;
  std::uint64_t sampled = 0;
  std::uint64_t enqueued = 0;
  std::uint64_t witness_rejected = 0;
  std::uint64_t pruned_duplicate = 0;
  std::uint64_t pruned_infeasible = 0;
  std::uint64_t pruned_dominated = 0;
  DropAccounting dropped;
  std::size_t pending = 0;
  std::size_t viable = 0;
  std::uint64_t history_freed = 0;
  std::size_t history_peak_live = 0;
  std::uint64_t graphs_built = 0;
};

struct SolveOutcome
{
  double lower_bound = std::numeric_limits<double>::lowest();
  double upper_bound = std::numeric_limits<double>::max();
  std::vector<double> best_point;
  StopReason stop_reason = StopReason::IterationLimit;
  bool proven_optimal = false;
  bool infeasible = false;
  RunTelemetry counters;
  /// Facts of the role build that stopped the search with DeviceMemory, if a
  /// build over its budget did (reference driver.hpp:408-416 reports them).
  std::optional<backend::OverBudget> over_budget;

  bool found_incumbent() const { return upper_bound < std::numeric_limits<double>::max(); }
};

/// Roles per Composition, built on first use (reference search/cache.hpp:40-159).
class BackendCache
{
public:
  BackendCache(std::shared_ptr<const backend::RegionBackendFactory> factory, const model::Problem<double>& problem,
               const policy::CompositionPolicy& pol, std::size_t budget_bytes, std::size_t samples_per_region)
      : factory_(std::move(factory))
      , problem_(problem)
      , policy_(pol)
      , budget_ {budget_bytes, samples_per_region, false}
  {
  }

  backend::SubdivisionBundle& subdivision(const region::Composition& c)
  {
    return find_or_build(subdivisions_, c, backend::RoleRequest::subdividing());
  }

  backend::RegionEnumerator* enumerator(const region::Composition& c)
  {
    if (!region::is_fully_enumerable(c) || !factory_->capabilities().exact_enumeration) {
      return nullptr;
    }
    return find_or_build(enumerators_, c, backend::RoleRequest::enumerating()).enumerator.get();
  }

  std::uint64_t graphs_built() const { return built_; }

private:
  struct Entry
  {
    region::Composition composition;
    backend::SubdivisionBundle bundle;
  };

  backend::SubdivisionBundle& find_or_build(std::vector<std::unique_ptr<Entry>>& entries, const region::Composition& c,
                                            backend::RoleRequest roles)
  {
    for (auto& e : entries) {
      if (e->composition == c) {
        return e->bundle;
      }
    }
    std::size_t const n_regions = policy_.n_regions(c);
    auto built = factory_->build_subdivision(problem_, c, n_regions, budget_, roles);
    built_ += (built.sampler ? 1 : 0) + (built.bounder ? 1 : 0) + (built.enumerator ? 1 : 0);
    entries.push_back(std::make_unique<Entry>(Entry {c, std::move(built)}));
    return entries.back()->bundle;
  }

  std::shared_ptr<const backend::RegionBackendFactory> factory_;
  const model::Problem<double>& problem_;
  const policy::CompositionPolicy& policy_;
  backend::BuildBudget budget_;
  std::vector<std::unique_ptr<Entry>> subdivisions_;
  std::vector<std::unique_ptr<Entry>> enumerators_;
  std::uint64_t built_ = 0;
};

struct SolveOptions
{
  std::uint32_t iter_limit = 1000000;
  double tolerance = 1e-6;
  std::size_t sample_points = 1;
  std::size_t host_budget_bytes = 0;
  FrontierPolicy frontier_policy = FrontierPolicy::StopAtBudget;
  std::optional<double> known_primal_bound;
};

/// One solve's best-first search as a resumable state machine. `next()` runs
/// the host part of an iteration up to the point where device work is needed
/// and returns that request; the caller evaluates it (one role call each, or
/// batched with other solves' requests) and hands the results to
/// `on_enumerate` / `on_subdivide`, which finish the iteration exactly as the
/// reference's loop body does (search/driver.hpp:221-407). Because every
/// decision depends only on this solve's own state and results, a solve
/// produces the same iterates alone and interleaved with others.
class SolveState
{
public:
  enum class Kind
  {
    Enumerate,
    Subdivide
  };

  struct Request
  {
    Kind kind;
    std::span<const interval> box;
    const region::SlotAssignment* assignment;
    std::uint64_t salt;
  };

  SolveState(const model::Problem<double>& problem, std::shared_ptr<const policy::CompositionPolicy> pol,
             SolveOptions opt)
      : problem_(problem)
      , policy_(std::move(pol))
      , opt_(opt)
      , pending_(10000)
  {
    if (policy_ == nullptr) {
      throw InvalidConfiguration("SearchDriver requires a non-null CompositionPolicy");
    }
    if (opt_.sample_points < 1) {
      throw InvalidConfiguration("sample_points must be at least 1");
    }
    gub_ = opt_.known_primal_bound.value_or(std::numeric_limits<double>::max());
    history_.enqueue({});  // slot 0: the root sentinel
    host_budget_ = resolve_host_budget(opt_.host_budget_bytes);
    pending_.enqueue(Node {0, 0, 0, glb_, 0});
  }

  const model::Problem<double>& problem() const { return problem_; }
  bool finished() const { return done_; }
  double gub() const { return gub_; }
  std::uint32_t iteration() const { return iter_; }

  /// Host part of the next iteration; nullptr once the search has stopped.
  const Request* next()
  {
    if (done_) {
      return nullptr;
    }
    if (!(iter_ < opt_.iter_limit && !pending_.empty() && !closed(glb_))) {
      done_ = true;
      return nullptr;
    }
    ++iter_;
    Node const cur = pending_.dequeue();
    if (closed(cur.lb)) {
      converged_ = true;
      stop_ = StopReason::Converged;
      stopped_early_ = true;
      done_ = true;
      return nullptr;
    }
    glb_ = cur.lb;
    std::span<const model::VarKind> const kinds = problem_.var_kinds;
    std::span<const interval> const root_box = problem_.box_bounds;
    cur.materialise(history_, box_, *policy_, kinds, root_box);
    history_.release(cur.pidx);
    assignment_ = policy_->choose(box_, kinds);
    depth_ = cur.depth;

    std::size_t live = 0;
    for (const interval& b : box_) {
      if (b.ub > b.lb) {
        ++live;
      }
    }
    bool const fathom = region::can_fathom_without_children(live, assignment_.composition) && enumerate_capable_;
    request_.kind = fathom ? Kind::Enumerate : Kind::Subdivide;
    request_.box = box_;
    request_.assignment = &assignment_;
    request_.salt = iter_;
    if (!fathom) {
      box_idx_ = history_.enqueue(box_);
    }
    return &request_;
  }

  /// The enumerator's answer for the current (fathom) request.
  void on_enumerate(const backend::CandidateResult& exact, std::size_t n_points)
  {
    if (exact.found && exact.value < gub_) {
      if (witness_is_admissible(exact.point, problem_.var_kinds, problem_.box_bounds)) {
        gub_ = exact.value;
        best_point_.assign(exact.point.begin(), exact.point.end());
      } else {
        ++t_.witness_rejected;
      }
    }
    ++t_.fathomed;
    t_.fathomed_points += n_points;
  }

  /// The sampler's and bounder's answers for the current (subdivide) request.
  void on_subdivide(const backend::CandidateResult& drawn, std::size_t n_samples, const backend::BoundResult& bounds)
  {
    t_.sampled += n_samples;
    double const cand = drawn.value;
    if (cand < gub_) {
      if (witness_is_admissible(drawn.point, problem_.var_kinds, problem_.box_bounds)) {
        gub_ = cand;
        best_point_.assign(drawn.point.begin(), drawn.point.end());
      } else {
        ++t_.witness_rejected;
      }
    }
    t_.bounded += bounds.n_regions;
    bool const may_duplicate = !assignment_.domain_sizes.empty();
    std::vector<std::size_t> const prefix =
        may_duplicate ? region::slot_prefixes(assignment_.widths) : std::vector<std::size_t> {};
    std::uint64_t const hash = region::assignment_hash(assignment_);
    for (std::size_t r = 0; r < bounds.n_regions; ++r) {
      if (may_duplicate && region::is_duplicate_child(r, assignment_, prefix)) {
        ++t_.pruned_duplicate;
        continue;
      }
      if (!bounds.feasible[r]) {
        ++t_.pruned_infeasible;
        continue;
      }
      if (bounds.obj_lb[r] > gub_) {
        ++t_.pruned_dominated;
        continue;
      }
      pending_.enqueue(Node {r, box_idx_, depth_ + 1, bounds.obj_lb[r], hash});
      history_.add_ref(box_idx_);
      ++t_.enqueued;
    }
    history_.release(box_idx_);

    if (over_host_budget(pending_.capacity_bytes() + history_.live_bytes(), host_budget_)) {
      if (opt_.frontier_policy != FrontierPolicy::Compact) {
        stop(StopReason::HostMemory);
        return;
      }
      std::size_t const n_keep = compact_keep_count(host_budget_, history_.live_bytes(), sizeof(Node));
      pending_.compact(n_keep, [&](const Node& e) {
        t_.dropped.fold(e.lb, gub_);
        history_.release(e.pidx);
      });
      if (over_host_budget(pending_.capacity_bytes() + history_.live_bytes(), host_budget_)) {
        stop(StopReason::HostMemory);
      }
    }
  }

  /// End the search early (a device build failed, host allocation failed).
  void stop(StopReason reason)
  {
    stop_ = reason;
    stopped_early_ = true;
    done_ = true;
  }

  /// Disable the fathom path (a backend without exact enumeration).
  void set_enumerate_capable(bool v) { enumerate_capable_ = v; }

  void set_graphs_built(std::uint64_t n) { t_.graphs_built = n; }

  /// Stop with DeviceMemory, keeping the facts of the build that did not fit.
  void stop_over_budget(backend::OverBudget facts)
  {
    over_budget_ = std::move(facts);
    stop(StopReason::DeviceMemory);
  }

  SolveOutcome finish()
  {
    if (!stopped_early_) {
      stop_ = pending_.empty() ? StopReason::Exhausted : closed(glb_) ? StopReason::Converged
                                                                      : StopReason::IterationLimit;
    }
    bool const found = gub_ < std::numeric_limits<double>::max();
    std::size_t const viable = pending_.count_viable(gub_);
    FinalBounds const fin = finalise_bounds(glb_, gub_, opt_.tolerance, found, converged_, pending_.empty(),
                                            pending_.empty() ? std::numeric_limits<double>::max()
                                                             : pending_.peek().lb,
                                            viable, t_.dropped);
    t_.iterations = iter_;
    t_.pending = pending_.size();
    t_.viable = viable;
    t_.history_freed = history_.freed_count();
    t_.history_peak_live = history_.peak_live();
    SolveOutcome out;
    out.lower_bound = fin.glb;
    out.upper_bound = gub_;
    out.best_point = best_point_;
    out.stop_reason = stop_;
    out.proven_optimal = fin.proven_optimal;
    out.infeasible = fin.infeasible;
    out.counters = t_;
    out.over_budget = over_budget_;
    return out;
  }

private:
  bool closed(double bound) const { return gap_closed(gub_, bound, opt_.tolerance); }

  const model::Problem<double>& problem_;
  std::shared_ptr<const policy::CompositionPolicy> policy_;
  SolveOptions opt_;
  IntervalPQueue pending_;
  IntervalHistory history_;
  std::size_t host_budget_ = 0;
  double gub_ = std::numeric_limits<double>::max();
  double glb_ = std::numeric_limits<double>::lowest();
  std::uint32_t iter_ = 0;
  std::vector<double> best_point_;
  RunTelemetry t_;
  bool converged_ = false;
  bool stopped_early_ = false;
  bool done_ = false;
  bool enumerate_capable_ = true;
  StopReason stop_ = StopReason::IterationLimit;
  std::optional<backend::OverBudget> over_budget_;
  std::vector<interval> box_;
  region::SlotAssignment assignment_;
  std::size_t depth_ = 0;
  std::size_t box_idx_ = 0;
  Request request_ {};
};

/// Best-first search over a Problem with one device round trip per role call
/// (the reference's SearchDriver, search/driver.hpp:68-559). Each iteration:
/// dequeue the least-lb node, rebuild its box, pick a SlotAssignment, fathom
/// exactly if every live variable is enumerated, else sample for an incumbent
/// and bound every child region, enqueueing those neither infeasible nor
/// dominated.
class SearchDriver
{
public:
  SearchDriver(std::shared_ptr<const policy::CompositionPolicy> pol,
               std::shared_ptr<const backend::RegionBackendFactory> backend, std::uint32_t iter_limit = 1000000,
               double tolerance = 1e-6, std::size_t sample_points = 1, std::size_t budget_bytes = 0,
               std::size_t host_budget_bytes = 0, FrontierPolicy frontier_policy = FrontierPolicy::StopAtBudget,
               std::optional<double> known_primal_bound = std::nullopt)
      : policy_(std::move(pol))
      , backend_(std::move(backend))
      , budget_bytes_(budget_bytes)
  {
    if (policy_ == nullptr) {
      throw InvalidConfiguration("SearchDriver requires a non-null CompositionPolicy");
    }
    if (backend_ == nullptr) {
      throw InvalidConfiguration("SearchDriver requires a non-null RegionBackendFactory");
    }
    if (sample_points < 1) {
      throw InvalidConfiguration("sample_points must be at least 1");
    }
    opt_.iter_limit = iter_limit;
    opt_.tolerance = tolerance;
    opt_.sample_points = sample_points;
    opt_.host_budget_bytes = host_budget_bytes;
    opt_.frontier_policy = frontier_policy;
    opt_.known_primal_bound = known_primal_bound;
  }

  SolveOutcome solve(const model::Problem<double>& problem)
  {
    SolveState st(problem, policy_, opt_);
    st.set_enumerate_capable(backend_->capabilities().exact_enumeration);
    BackendCache cache(backend_, problem, *policy_, budget_bytes_, opt_.sample_points);
    try {
      while (const SolveState::Request* req = st.next()) {
        backend::Region const reg {req->box, *req->assignment};
        if (req->kind == SolveState::Kind::Enumerate) {
          backend::RegionEnumerator* const e = cache.enumerator(req->assignment->composition);
          backend::CandidateResult const exact = e->enumerate(reg);
          st.on_enumerate(exact, e->n_points());
          continue;
        }
        backend::SubdivisionBundle& roles = cache.subdivision(req->assignment->composition);
        backend::CandidateResult const drawn = roles.sampler->sample(reg, req->salt);
        backend::BoundResult const bounds = roles.bounder->bound(reg);
        st.on_subdivide(drawn, roles.sampler->n_samples(), bounds);
      }
    } catch (const backend::OverBudgetError& e) {
      st.stop_over_budget(e.facts());
    } catch (const ResourceExhausted&) {
      st.stop(StopReason::DeviceMemory);
    } catch (const std::bad_alloc&) {
      st.stop(StopReason::AllocationFailure);
    }
    st.set_graphs_built(cache.graphs_built());
    return st.finish();
  }

private:
  std::shared_ptr<const policy::CompositionPolicy> policy_;
  std::shared_ptr<const backend::RegionBackendFactory> backend_;
  std::size_t budget_bytes_;
  SolveOptions opt_;
};

}  // namespace cmi::search
