#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <optional>
#include <span>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "cmi/aggregate.hpp"
#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/policy.hpp"
#include "cmi/region.hpp"
#include "cmi/search.hpp"

// The k-child search: each iteration takes one pending node, splits it into k
// children on its widest live variable, and has the device bound every child
// over M = N / k subregions of it in one launch, so a child's bound is the
// least lower bound over a finer cover than the child itself (reference
// aggregate/driver.hpp:98-462, aggregate/policy.hpp:106-330,
// aggregate/frontier.hpp:106-394, aggregate/selection.hpp:15-125).
//
// Only the k-way branch is ever re-derived (a pending child stores its index
// among its parent's k children); the refine that produced its bound is
// recomputed on the device and discarded. The frontier keeps two orders over
// the same pending children: the one the selection rule explores in and the
// least-lower-bound order the bracket is read from, so a rule that explores
// out of bound order cannot move the reported dual bound.
namespace cmi::aggregate
{

// ---- selection rules ---------------------------------------------------------

/// Which pending child is explored next. LeastLowerBound is the plain
/// best-first order; RejectIndex explores the child whose objective range
/// lies furthest below the incumbent relative to its width,
/// (f_best - lb) / (hull_ub - lb), re-keying the frontier whenever the
/// incumbent improves; RejectIndexStale keys each child once, against the
/// incumbent of its enqueue.
enum class SelectionRule
{
  LeastLowerBound,
  RejectIndex,
  RejectIndexStale,
};

inline const char* to_string(SelectionRule r)
{
  switch (r) {
    case SelectionRule::LeastLowerBound: return "least-lower-bound";
    case SelectionRule::RejectIndex: return "reject-index";
    case SelectionRule::RejectIndexStale: return "reject-index-stale";
  }
  return "unknown";
}

inline std::optional<SelectionRule> parse_selection_rule(std::string_view s)
{
  for (SelectionRule r : {SelectionRule::LeastLowerBound, SelectionRule::RejectIndex, SelectionRule::RejectIndexStale}) {
    if (s == to_string(r)) {
      return r;
    }
  }
  return std::nullopt;
}

/// Whether an improved incumbent reorders the frontier under this rule.
inline bool rekeys_on_incumbent(SelectionRule r) { return r == SelectionRule::RejectIndex; }

/// A child's exploration key, smaller first under every rule. Without an
/// incumbent (f_best not below DBL_MAX) every rule orders by lb. A child
/// whose objective range has collapsed (hull_ub <= lb) sorts first if it can
/// still beat the incumbent and last otherwise.
inline double selection_key(SelectionRule rule, double lb, double hull_ub, double f_best)
{
  if (rule == SelectionRule::LeastLowerBound || !(f_best < std::numeric_limits<double>::max())) {
    return lb;
  }
  double const spread = hull_ub - lb;
  if (!(spread > 0.0)) {
    constexpr double inf = std::numeric_limits<double>::infinity();
    return lb < f_best ? -inf : inf;
  }
  return -((f_best - lb) / spread);
}

// ---- the partition of one node ---------------------------------------------------

/// A node split k ways (`branch`, one slot on one variable) with each child
/// covered by M subregions (`refine`, slots on the other variables). `launch`
/// is what the device decodes: the refine slots followed by the branch slot,
/// which makes the branch digit the most significant one, so child c owns
/// launch regions [c*M, (c+1)*M). When the node has a single live variable
/// the two halves share it (`nested`): `launch` is one slot of width N on it,
/// and because a uniform k-way split of a power-of-two width nests exactly in
/// the N-way split, r / M still names the child.
struct AggregatePartition
{
  region::SlotAssignment branch;
  region::SlotAssignment refine;
  region::SlotAssignment launch;
  std::size_t k = 1;
  std::size_t m = 1;
  bool nested = false;

  bool is_leaf() const { return branch.size() == 0; }
  std::size_t total_fan_out() const { return k * m; }

  /// Throws InvalidConfiguration unless the three assignments agree with the
  /// layout the device reduction assumes.
  void check() const
  {
    if (is_leaf()) {
      return;
    }
    auto bad = [](const char* what) { throw InvalidConfiguration(std::string("aggregate partition: ") + what); };
    if (region::composition_fan_out(branch.widths) != k) {
      bad("the branch does not split the node into k children");
    }
    if (region::composition_fan_out(launch.widths) != total_fan_out()) {
      bad("the launch does not cover k * M subregions");
    }
    if (nested) {
      if (branch.size() != 1 || refine.size() != 0 || launch.size() != 1 || launch.var_ids[0] != branch.var_ids[0]
          || launch.composition[0] != branch.composition[0]) {
        bad("a nested launch must be the branch slot alone, widened to N");
      }
      return;
    }
    if (region::composition_fan_out(refine.widths) != m) {
      bad("the refine does not cover each child with M subregions");
    }
    if (launch.size() != refine.size() + branch.size()) {
      bad("the launch is not the refine slots followed by the branch slots");
    }
    std::size_t const base = refine.size();
    for (std::size_t j = 0; j < branch.size(); ++j) {
      if (launch.var_ids[base + j] != branch.var_ids[j] || launch.widths[base + j] != branch.widths[j]
          || launch.composition[base + j] != branch.composition[j]) {
        bad("the branch slots are not the launch's most significant slots");
      }
      if (std::find(refine.var_ids.begin(), refine.var_ids.end(), branch.var_ids[j]) != refine.var_ids.end()) {
        bad("the refine splits the branch variable as well");
      }
    }
  }
};

/// The child a launch region belongs to.
inline std::size_t child_of(std::size_t r, std::size_t m) { return r / m; }

/// Branch k ways on the widest live variable (lowest index on equal widths,
/// always a uniform split, discrete variables included), refine the rest
/// under a bisection budget of log2(N / k), and give every other live
/// variable a width-1 slot so the launch composition's length follows the
/// live-variable count rather than the bisection allocation (reference
/// aggregate/policy.hpp:106-330).
class WidestBranchPolicy
{
public:
  WidestBranchPolicy(std::size_t k, std::size_t n, policy::SearchCalibration cal = {}, bool relative_width = false)
      : k_(k)
      , n_(n)
      , relative_(relative_width)
      , refine_(bits_of(k, n), cal, relative_width)
  {
    m_ = n_ / k_;
  }

  std::size_t branch_fan_out() const { return k_; }
  std::size_t refine_fan_out() const { return m_; }
  std::size_t total_fan_out() const { return n_; }
  bool relative_width() const { return relative_; }

  /// The k-way split of a box: a pure function of (box, kinds), since a
  /// pending child's box is rebuilt by calling it again on the parent.
  region::SlotAssignment branch(std::span<const interval> box, std::span<const model::VarKind> kinds) const
  {
    if (box.size() != kinds.size()) {
      throw ShapeMismatch("aggregate branch: the box has " + std::to_string(box.size()) + " coordinates for "
                          + std::to_string(kinds.size()) + " variables");
    }
    std::size_t widest = box.size();
    double width = 0.0;
    for (std::size_t v = 0; v < box.size(); ++v) {
      if (box[v].ub > box[v].lb && (widest == box.size() || box[v].ub - box[v].lb > width)) {
        widest = v;
        width = box[v].ub - box[v].lb;
      }
    }
    region::SlotAssignment out;
    if (widest == box.size()) {
      return out;  // nothing live: the box is a point
    }
    add_slot(out, split_kind(kinds[widest]), widest, k_);
    return out;
  }

  AggregatePartition partition(std::span<const interval> box, std::span<const model::VarKind> kinds,
                               std::size_t /*depth*/) const
  {
    AggregatePartition p;
    p.k = k_;
    p.m = m_;
    p.branch = branch(box, kinds);
    if (p.is_leaf()) {
      return p;
    }
    std::size_t const on = p.branch.var_ids[0];
    std::size_t live = 0;
    for (const interval& b : box) {
      live += b.ub > b.lb ? 1 : 0;
    }
    if (live == 1) {
      p.nested = true;
      add_slot(p.launch, p.branch.composition[0], on, n_);
      return p;
    }
    // The refine never sees the branch variable: collapsed to a point it is
    // not live, and the bisection policy skips it.
    std::vector<interval> rest(box.begin(), box.end());
    rest[on].ub = rest[on].lb;
    p.refine = refine_.choose(rest, kinds);
    for (std::size_t v = 0; v < box.size() && p.refine.size() < policy::kMaxSlots - 1; ++v) {
      if (v == on || !(box[v].ub > box[v].lb)
          || std::find(p.refine.var_ids.begin(), p.refine.var_ids.end(), v) != p.refine.var_ids.end()) {
        continue;
      }
      add_slot(p.refine, split_kind(kinds[v]), v, 1);
    }
    for (const region::SlotAssignment* half : {&p.refine, &p.branch}) {
      for (std::size_t j = 0; j < half->size(); ++j) {
        p.launch.composition.kinds.push_back(half->composition[j]);
        p.launch.var_ids.push_back(half->var_ids[j]);
        p.launch.widths.push_back(half->widths[j]);
        p.launch.domain_sizes.push_back(half->domain_sizes.empty() ? half->widths[j] : half->domain_sizes[j]);
      }
    }
    return p;
  }

private:
  static bool power_of_two(std::size_t v) { return v != 0 && (v & (v - 1)) == 0; }

  /// log2(n / k) after checking that k and n can be honoured.
  static std::size_t bits_of(std::size_t k, std::size_t n)
  {
    if (k < 2 || !power_of_two(k)) {
      throw InvalidConfiguration("branch fan-out " + std::to_string(k) + ": it must be a power of two, 2 or more");
    }
    if (n <= k || !power_of_two(n)) {
      throw InvalidConfiguration("total fan-out " + std::to_string(n) + ": it must be a power of two above the branch "
                                 "fan-out " + std::to_string(k) + ", so each child gets at least two subregions");
    }
    std::size_t bits = 0;
    for (std::size_t q = n / k; q > 1; q >>= 1) {
      ++bits;
    }
    return bits;
  }

  /// Uniform splits only: a continuous variable stays continuous, a discrete
  /// one is split as an integer range.
  static region::SlotKind split_kind(model::VarKind kind)
  {
    return kind == model::VarKind::Continuous ? region::SlotKind::Continuous : region::SlotKind::IntegerPartition;
  }

  static void add_slot(region::SlotAssignment& a, region::SlotKind kind, std::size_t var, std::size_t width)
  {
    a.composition.kinds.push_back(kind);
    a.var_ids.push_back(var);
    a.widths.push_back(width);
    a.domain_sizes.push_back(width);
  }

  std::size_t k_;
  std::size_t n_;
  std::size_t m_ = 1;
  bool relative_;
  policy::BisectionBudgetCompositionPolicy refine_;
};

// ---- pending children ------------------------------------------------------------

/// One pending child: its index among its parent's k children, the parent's
/// history slot, its depth, its bound and the upper end of its objective
/// hull (a bound over its unexcluded subregions, never an incumbent).
struct AggregateNode
{
  std::size_t sidx = 0;
  std::size_t pidx = 0;
  std::size_t depth = 0;
  double lb = std::numeric_limits<double>::lowest();
  double hull_ub = std::numeric_limits<double>::max();
  std::uint64_t branch_hash = 0;

  void materialise(const search::IntervalHistory& history, std::vector<interval>& out, const WidestBranchPolicy& pol,
                   std::span<const model::VarKind> kinds, std::span<const interval> root_box) const
  {
    if (pidx == 0) {
      out.assign(root_box.begin(), root_box.end());
      return;
    }
    const std::vector<interval>& parent = history.box(pidx);
    region::SlotAssignment const again = pol.branch(parent, kinds);
    if (region::assignment_hash(again) != branch_hash) {
      throw InvalidConfiguration("the aggregate branch chose a different split for a stored parent box than when its "
                                 "children were enqueued");
    }
    region::materialise(parent, sidx, again, out);
  }
};

/// The pending children, stored once and ordered twice: by the selection
/// key (what to explore) and by lb (the dual bound). Both orders are binary
/// heaps of handles into the store, kept with the standard heap algorithms;
/// a taken child is only marked dead and its handle in the other heap is
/// dropped when it reaches the top or when dead handles outnumber live
/// children by more than two to one (plus 64). Handles carry their own key,
/// depth and parent slot, and a generation that a recycled store slot
/// advances, so a stale handle can neither surface a new child nor disturb
/// the heap order. Ties on the key go to the deeper child, then the larger
/// parent slot, as in the per-region frontier (reference
/// aggregate/frontier.hpp:106-394).
class AggregateFrontier
{
public:
  explicit AggregateFrontier(SelectionRule rule = SelectionRule::LeastLowerBound)
      : rule_(rule)
  {
  }

  SelectionRule rule() const { return rule_; }
  bool empty() const { return live_ == 0; }
  std::size_t size() const { return live_; }

  void enqueue(const AggregateNode& node, double f_best)
  {
    std::size_t slot;
    if (recycled_.empty()) {
      slot = store_.size();
      store_.push_back(Entry {node, 0, true});
    } else {
      slot = recycled_.back();
      recycled_.pop_back();
      Entry& e = store_[slot];
      e = Entry {node, e.generation + 1, true};
    }
    ++live_;
    add(by_key_, handle(slot, selection_key(rule_, node.lb, node.hull_ub, f_best)));
    add(by_lb_, handle(slot, node.lb));
  }

  /// The next child under the selection rule; the frontier must not be empty.
  AggregateNode pop()
  {
    while (!by_key_.empty()) {
      Handle const h = by_key_.front();
      drop_top(by_key_);
      if (!current(h)) {
        continue;
      }
      Entry& e = store_[h.slot];
      e.live = false;
      --live_;
      recycled_.push_back(h.slot);
      AggregateNode const out = e.node;
      sweep_if_stale();
      return out;
    }
    throw InvalidConfiguration("pop from an empty aggregate frontier");
  }

  /// The least pending lower bound; DBL_MAX when nothing is pending.
  double min_lb()
  {
    while (!by_lb_.empty() && !current(by_lb_.front())) {
      drop_top(by_lb_);
    }
    return by_lb_.empty() ? std::numeric_limits<double>::max() : store_[by_lb_.front().slot].node.lb;
  }

  /// A better incumbent: re-key every pending child if the rule depends on it.
  void on_incumbent(double f_best)
  {
    if (!rekeys_on_incumbent(rule_)) {
      return;
    }
    by_key_.clear();
    by_lb_.clear();
    by_key_.reserve(live_);
    by_lb_.reserve(live_);
    for (std::size_t s = 0; s < store_.size(); ++s) {
      if (store_[s].live) {
        const AggregateNode& n = store_[s].node;
        by_key_.push_back(handle(s, selection_key(rule_, n.lb, n.hull_ub, f_best)));
        by_lb_.push_back(handle(s, n.lb));
      }
    }
    std::make_heap(by_key_.begin(), by_key_.end(), later);
    std::make_heap(by_lb_.begin(), by_lb_.end(), later);
  }

  /// Pending children whose bound does not exceed gub.
  std::size_t count_viable(double gub) const
  {
    std::size_t n = 0;
    for (const Entry& e : store_) {
      n += e.live && !(e.node.lb > gub) ? 1 : 0;
    }
    return n;
  }

  /// Host bytes the frontier holds, by its containers' capacities.
  std::size_t capacity_bytes() const
  {
    return store_.capacity() * sizeof(Entry) + (by_key_.capacity() + by_lb_.capacity()) * sizeof(Handle)
           + recycled_.capacity() * sizeof(std::size_t);
  }

  std::vector<AggregateNode> pending() const
  {
    std::vector<AggregateNode> out;
    out.reserve(live_);
    for (const Entry& e : store_) {
      if (e.live) {
        out.push_back(e.node);
      }
    }
    return out;
  }

private:
  struct Entry
  {
    AggregateNode node;
    std::uint64_t generation = 0;
    bool live = false;
  };

  struct Handle
  {
    double key = 0.0;
    std::size_t depth = 0;
    std::size_t pidx = 0;
    std::size_t slot = 0;
    std::uint64_t generation = 0;
  };

  Handle handle(std::size_t slot, double key) const
  {
    const Entry& e = store_[slot];
    return Handle {key, e.node.depth, e.node.pidx, slot, e.generation};
  }

  bool current(const Handle& h) const { return store_[h.slot].live && store_[h.slot].generation == h.generation; }

  /// The heap comparison: a is later than b when b would be explored first.
  static bool later(const Handle& a, const Handle& b)
  {
    if (b.key < a.key || a.key < b.key) {
      return b.key < a.key;
    }
    if (a.depth != b.depth) {
      return b.depth > a.depth;
    }
    return b.pidx > a.pidx;
  }

  static void add(std::vector<Handle>& heap, const Handle& h)
  {
    heap.push_back(h);
    std::push_heap(heap.begin(), heap.end(), later);
  }

  static void drop_top(std::vector<Handle>& heap)
  {
    std::pop_heap(heap.begin(), heap.end(), later);
    heap.pop_back();
  }

  void sweep_if_stale()
  {
    if (by_key_.size() + by_lb_.size() <= 2 * live_ + 64) {
      return;
    }
    for (std::vector<Handle>* heap : {&by_key_, &by_lb_}) {
      std::vector<Handle> kept;
      kept.reserve(live_);
      for (const Handle& h : *heap) {
        if (current(h)) {
          kept.push_back(h);
        }
      }
      std::make_heap(kept.begin(), kept.end(), later);
      heap->swap(kept);
    }
  }

  SelectionRule rule_;
  std::vector<Entry> store_;
  std::vector<std::size_t> recycled_;
  std::vector<Handle> by_key_;
  std::vector<Handle> by_lb_;
  std::size_t live_ = 0;
};

// ---- one solve -------------------------------------------------------------------------

struct AggregateOptions
{
  std::uint32_t iter_limit = 1000000;
  double tolerance = 1e-6;
  std::size_t host_budget_bytes = 0;  // 0: half of MemAvailable
  SelectionRule rule = SelectionRule::LeastLowerBound;
  std::optional<double> known_primal_bound;
};

/// One k-child search as a resumable state machine, like search::SolveState:
/// `next()` runs the host part of iterations until one needs the device and
/// returns that request (the node's box, the bound launch and, unless the
/// sampler's partition is a leaf, the sample launch); `on_results` finishes
/// it. A node whose box is a single point is settled on the host: its
/// objective is evaluated there and it has no children (reference
/// aggregate/driver.hpp:176-330, 412-447).
class AggregateSolveState
{
public:
  struct Request
  {
    std::span<const interval> box;
    const AggregatePartition* bound;
    const AggregatePartition* sample;  // nullptr: no sampling this iteration
    std::uint64_t salt;
  };

  AggregateSolveState(const model::Problem<double>& problem, const WidestBranchPolicy& bound_policy,
                      const WidestBranchPolicy& sample_policy, AggregateOptions opt)
      : problem_(problem)
      , policy_(bound_policy)
      , sample_policy_(sample_policy)
      , opt_(opt)
      , frontier_(opt.rule)
  {
    if (sample_policy_.branch_fan_out() != policy_.branch_fan_out()) {
      throw InvalidConfiguration("the sampler's partition must branch as many ways as the bounder's");
    }
    gub_ = opt_.known_primal_bound.value_or(std::numeric_limits<double>::max());
    history_.enqueue({});  // slot 0: the root sentinel
    host_budget_ = search::resolve_host_budget(opt_.host_budget_bytes);
    frontier_.enqueue(AggregateNode {0, 0, 0, std::numeric_limits<double>::lowest(),
                                     std::numeric_limits<double>::max(), 0},
                      gub_);
  }

  bool finished() const { return done_; }
  double gub() const { return gub_; }

  const Request* next()
  {
    while (!done_) {
      if (!(iter_ < opt_.iter_limit && !frontier_.empty())) {
        done_ = true;
        break;
      }
      double const least = frontier_.min_lb();
      glb_ = least;
      if (closed(least)) {
        converged_ = true;
        stop(search::StopReason::Converged);
        break;
      }
      ++iter_;
      cur_ = frontier_.pop();
      cur_.materialise(history_, box_, policy_, problem_.var_kinds, problem_.box_bounds);
      history_.release(cur_.pidx);
      part_ = policy_.partition(box_, problem_.var_kinds, cur_.depth);
      if (part_.is_leaf()) {
        ++t_.fathomed;
        settle_point();
        continue;
      }
      part_.check();
      sample_part_ = sample_policy_.partition(box_, problem_.var_kinds, cur_.depth);
      box_idx_ = history_.enqueue(box_);
      request_ = Request {box_, &part_, sample_part_.is_leaf() ? nullptr : &sample_part_, iter_};
      return &request_;
    }
    return nullptr;
  }

  /// The sampler's best point (ignored when the request had no sample
  /// launch) and the k child bounds of the current request.
  void on_results(const backend::CandidateResult& drawn, std::size_t n_samples,
                  std::span<const AggregateBound> children)
  {
    if (request_.sample != nullptr) {
      t_.sampled += n_samples;
      if (drawn.found && drawn.value < gub_) {
        if (search::witness_is_admissible(drawn.point, problem_.var_kinds, problem_.box_bounds)) {
          improve(drawn.value, drawn.point);
          engaged_ = true;
        } else {
          ++t_.witness_rejected;
        }
      }
    }
    t_.bounded += children.size();
    std::uint64_t const hash = region::assignment_hash(part_.branch);
    for (std::size_t c = 0; c < children.size(); ++c) {
      if (!children[c].feasible) {
        ++t_.pruned_infeasible;
        continue;
      }
      // A child lies inside its parent, so the parent's bound holds over it
      // too; the child's own cover is a different one and can be weaker.
      double const lb = std::max(children[c].lb, cur_.lb);
      // Above its own hull the child holds no feasible point at all.
      if (lb > children[c].hull_ub) {
        ++t_.pruned_infeasible;
        continue;
      }
      if (lb > gub_) {
        ++t_.pruned_dominated;
        continue;
      }
      frontier_.enqueue(AggregateNode {c, box_idx_, cur_.depth + 1, lb, children[c].hull_ub, hash}, gub_);
      history_.add_ref(box_idx_);
      ++t_.enqueued;
    }
    history_.release(box_idx_);
    if (search::over_host_budget(frontier_.capacity_bytes() + history_.live_bytes(), host_budget_)) {
      stop(search::StopReason::HostMemory);
    }
  }

  void stop(search::StopReason reason)
  {
    stop_ = reason;
    stopped_early_ = true;
    done_ = true;
  }

  void stop_over_budget(backend::OverBudget facts)
  {
    over_budget_ = std::move(facts);
    stop(search::StopReason::DeviceMemory);
  }

  void set_graphs_built(std::uint64_t n) { t_.graphs_built = n; }

  /// False only for a re-keying rule that never had a sampled incumbent, so
  /// explored in lb order throughout.
  bool selection_rule_engaged() const
  {
    return engaged_ || (opt_.rule != SelectionRule::RejectIndex);
  }

  search::SolveOutcome finish()
  {
    if (!stopped_early_) {
      stop_ = frontier_.empty() ? search::StopReason::Exhausted : search::StopReason::IterationLimit;
    }
    bool const found = gub_ < std::numeric_limits<double>::max();
    std::size_t const viable = frontier_.count_viable(gub_);
    double const least = frontier_.empty() ? std::numeric_limits<double>::max() : frontier_.min_lb();
    search::FinalBounds const fin = search::finalise_bounds(glb_, gub_, opt_.tolerance, found, converged_,
                                                            frontier_.empty(), least, viable, t_.dropped);
    t_.iterations = iter_;
    t_.fathomed_points = t_.fathomed;  // a settled leaf is one point
    t_.pending = frontier_.size();
    t_.viable = viable;
    t_.history_freed = history_.freed_count();
    t_.history_peak_live = history_.peak_live();
    search::SolveOutcome out;
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
  bool closed(double bound) const { return search::gap_closed(gub_, bound, opt_.tolerance); }

  void improve(double value, std::span<const double> point)
  {
    gub_ = value;
    best_point_.assign(point.begin(), point.end());
    frontier_.on_incumbent(gub_);
  }

  /// A box with no live variable is a point: evaluate it on the host.
  void settle_point()
  {
    std::vector<double> x(box_.size());
    for (std::size_t i = 0; i < box_.size(); ++i) {
      x[i] = box_[i].lb;
    }
    if (model::first_violated_constraint(problem_, x).has_value()) {
      return;
    }
    double const value = model::evaluate_all(problem_.graph, x)[problem_.objective_root];
    if (!(value < gub_)) {
      return;
    }
    if (!search::witness_is_admissible(x, problem_.var_kinds, problem_.box_bounds)) {
      ++t_.witness_rejected;
      return;
    }
    improve(value, x);
  }

  const model::Problem<double>& problem_;
  const WidestBranchPolicy& policy_;
  const WidestBranchPolicy& sample_policy_;
  AggregateOptions opt_;
  AggregateFrontier frontier_;
  search::IntervalHistory history_;
  std::size_t host_budget_ = 0;
  double gub_ = std::numeric_limits<double>::max();
  double glb_ = std::numeric_limits<double>::lowest();
  std::uint32_t iter_ = 0;
  std::vector<double> best_point_;
  search::RunTelemetry t_;
  bool converged_ = false;
  bool stopped_early_ = false;
  bool done_ = false;
  bool engaged_ = false;
  search::StopReason stop_ = search::StopReason::IterationLimit;
  std::optional<backend::OverBudget> over_budget_;
  AggregateNode cur_ {};
  std::vector<interval> box_;
  AggregatePartition part_;
  AggregatePartition sample_part_;
  std::size_t box_idx_ = 0;
  Request request_ {};
};

}  // namespace cmi::aggregate
