#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
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
#include "cmi/frontier.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/policy.hpp"
#include "cmi/region.hpp"

// Host best-first branch and bound: one solve's resumable state machine, the
// per-composition role cache and the plain driver that runs one solve with one
// device round trip per role call. The heap, the parent-box store, the host
// budget and the end-of-run rules are in cmi/frontier.hpp. The per-iteration
// protocol (which node is taken, when the run stops, which children are
// pruned and in which order survivors are enqueued) follows the reference's
// loop (search/driver.hpp:221-407), because it decides the node order.
namespace cmi::search
{

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
    // Only an enumerate slot padded past its domain (domain size < width) has
    // duplicate children; without one the per-region check is skipped.
    bool may_duplicate = false;
    for (std::size_t j = 0; j < assignment_.domain_sizes.size(); ++j) {
      may_duplicate = may_duplicate || assignment_.domain_sizes[j] < assignment_.widths[j];
    }
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
