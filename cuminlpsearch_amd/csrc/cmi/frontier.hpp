#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <span>
#include <utility>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/policy.hpp"
#include "cmi/region.hpp"

// The host-side state of one best-first search that the device never sees:
// the pending-node heap, the store of parent boxes the pending nodes point
// into, the host-memory budget over both, and the rules that turn the state at
// the end of a run into a bracket and its claims.
//
// Node order is part of a solve's result (the iteration count, and the
// incumbent found under an iteration cap), so the heap's ordering and tie
// behaviour, the store's slot reuse (which produces the pidx values the
// ordering reads) and the compaction's eviction order follow the reference
// (search/frontier.hpp:66-338, search/history.hpp:20-137,
// search/budget.hpp:60-215, search/epilogue.hpp:35-127).
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

/// The spelling of each stop reason in logs (reference budget.hpp:43-60).
inline const char* to_string(StopReason r)
{
  static constexpr const char* kSpelling[] = {
      "converged", "exhausted", "iteration-limit", "host-memory", "allocation-failure", "device-memory",
  };
  auto const i = static_cast<std::size_t>(r);
  return i < std::size(kSpelling) ? kSpelling[i] : "unknown";
}

/// At the host budget: end the run with nothing discarded, or evict the
/// worst pending nodes and carry on (a different, still sound, search).
enum class FrontierPolicy
{
  StopAtBudget,
  Compact,
};

/// The bracket is closed when gub - bound <= tol * max(1, |gub|).
inline bool gap_closed(double gub, double bound, double tolerance)
{
  double const a = std::fabs(gub);
  double const scale = a > 1.0 ? a : 1.0;
  return gub - bound <= tolerance * scale;
}

/// MemAvailable of /proc/meminfo in bytes (the kernel's estimate of what a new
/// allocation can get), or 0 when it cannot be read.
inline std::size_t host_bytes_available()
{
  std::FILE* f = std::fopen("/proc/meminfo", "r");
  if (f == nullptr) {
    return 0;
  }
  char line[192];
  unsigned long long kib = 0;
  bool seen = false;
  while (!seen && std::fgets(line, sizeof line, f) != nullptr) {
    seen = std::sscanf(line, "MemAvailable: %llu", &kib) == 1;
  }
  std::fclose(f);
  return seen ? static_cast<std::size_t>(kib) * 1024u : 0;
}

/// A requested host budget is used as given; 0 asks for half of what the host
/// has available now. An unreadable meminfo gives 0, i.e. no budget.
inline std::size_t resolve_host_budget(std::size_t requested)
{
  if (requested > 0) {
    return requested;
  }
  return host_bytes_available() / 2;
}

/// Reached at exactly the budget (>=): a compaction sizes the heap so that its
/// next doubling lands on the budget, and that doubling must trip the check.
/// A budget of 0 is no budget.
inline bool over_host_budget(std::size_t live_bytes, std::size_t budget_bytes)
{
  if (budget_bytes == 0) {
    return false;
  }
  return live_bytes >= budget_bytes;
}

/// Nodes a compaction keeps: a quarter of what the budget left after the
/// history holds. The compacted heap keeps a capacity of 2 * keep, so it
/// reaches the whole room only at its next doubling, keep insertions later,
/// rather than straight away. At least 1; a history that fills the budget on
/// its own leaves the caller over budget after compacting, and it stops.
inline std::size_t compact_keep_count(std::size_t budget_bytes, std::size_t history_bytes, std::size_t node_bytes)
{
  std::size_t const room = history_bytes < budget_bytes ? budget_bytes - history_bytes : 0;
  std::size_t const keep = room / node_bytes / 4;
  return keep > 0 ? keep : 1;
}

/// What compactions threw away. Only evictions the incumbent did not already
/// dominate can hide the optimum; the least lower bound among them is a floor
/// below which nothing was lost, and every claim at the end is gated on it.
struct DropAccounting
{
  double lb_min = std::numeric_limits<double>::infinity();
  std::size_t viable = 0;
  std::size_t dominated = 0;

  void fold(double lb, double gub)
  {
    if (!(lb <= gub)) {
      ++dominated;
      return;
    }
    ++viable;
    if (lb < lb_min) {
      lb_min = lb;
    }
  }

  bool lost_nothing() const { return viable == 0; }
};

/// Whether a candidate may become the incumbent: inside the root box, and
/// within 1e-6 of an integer in every discrete coordinate. Constraints are not
/// checked again here; the device predicate already did.
inline bool witness_is_admissible(std::span<const double> point, std::span<const model::VarKind> kinds,
                                  std::span<const interval> root_box)
{
  constexpr double kLatticeTolerance = 1e-6;
  if (point.size() != kinds.size() || kinds.size() != root_box.size()) {
    return false;
  }
  for (std::size_t i = 0; i < point.size(); ++i) {
    double const v = point[i];
    bool const inside = v >= root_box[i].lb && v <= root_box[i].ub;
    if (!inside) {
      return false;
    }
    bool const discrete = kinds[i] != model::VarKind::Continuous;
    if (discrete && std::fabs(v - std::round(v)) > kLatticeTolerance) {
      return false;
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

/// The bracket a run reports and what it may claim.
///  * Proven optimal: an incumbent exists, the search settled (closed, emptied,
///    or nothing left under the incumbent), and nothing evicted lies further
///    than the tolerance below it. The lower bound is then the incumbent.
///  * Infeasible: otherwise, when the frontier emptied and no viable node was
///    ever evicted; glb is left as the loop had it.
///  * Otherwise the least of the pending minimum, the eviction floor and the
///    incumbent.
inline FinalBounds finalise_bounds(double glb, double gub, double tolerance, bool found_incumbent, bool converged,
                                   bool pending_empty, double frontier_min, std::size_t viable,
                                   const DropAccounting& dropped)
{
  bool const settled = converged || pending_empty || viable == 0;
  if (found_incumbent && settled && gap_closed(gub, dropped.lb_min, tolerance)) {
    return FinalBounds {gub, true, false};
  }
  if (pending_empty && dropped.lost_nothing()) {
    return FinalBounds {glb, false, true};
  }
  double bound = pending_empty ? std::numeric_limits<double>::max() : frontier_min;
  bound = std::min(bound, dropped.lb_min);
  if (found_incumbent) {
    bound = std::min(bound, gub);
  }
  return FinalBounds {bound, false, false};
}

/// Parent boxes of pending nodes, reference counted. A node names its parent
/// by slot index (pidx); a freed slot is handed out again last-in first-out.
/// Slot 0 is the root sentinel: it stands for the problem's own box and is
/// never counted, freed or reused.
class IntervalHistory
{
public:
  /// Store a box with one reference (its creator's); returns its slot.
  std::size_t enqueue(std::vector<interval> box)
  {
    live_bytes_ += charge(box.size());
    std::size_t slot;
    if (recycled_.empty()) {
      slot = slots_.size();
      slots_.push_back(Slot {std::move(box), 1});
    } else {
      slot = recycled_.back();
      recycled_.pop_back();
      slots_[slot] = Slot {std::move(box), 1};
    }
    peak_live_ = std::max(peak_live_, live_count());
    return slot;
  }

  void add_ref(std::size_t slot)
  {
    if (slot != 0) {
      slots_[slot].refs += 1;
    }
  }

  void release(std::size_t slot)
  {
    if (slot == 0) {
      return;
    }
    Slot& s = slots_[slot];
    s.refs -= 1;
    if (s.refs != 0) {
      return;
    }
    live_bytes_ -= charge(s.box.size());
    std::vector<interval>().swap(s.box);
    recycled_.push_back(slot);
    freed_ += 1;
  }

  const std::vector<interval>& box(std::size_t slot) const { return slots_[slot].box; }
  std::size_t slot_count() const { return slots_.size(); }
  std::size_t ref_count(std::size_t slot) const { return slots_[slot].refs; }
  std::size_t live_bytes() const { return live_bytes_; }
  std::size_t live_count() const { return slots_.size() - recycled_.size(); }
  std::uint64_t freed_count() const { return freed_; }
  std::size_t peak_live() const { return peak_live_; }

private:
  struct Slot
  {
    std::vector<interval> box;
    std::size_t refs = 0;
  };

  /// A stored box costs its intervals plus the vector that holds them; the
  /// host budget counts this figure.
  static constexpr std::size_t charge(std::size_t n_vars)
  {
    return n_vars * sizeof(interval) + sizeof(std::vector<interval>);
  }

  std::vector<Slot> slots_;
  std::vector<std::size_t> recycled_;
  std::size_t live_bytes_ = 0;
  std::uint64_t freed_ = 0;
  std::size_t peak_live_ = 0;
};

/// A pending region: child `sidx` of the box in history slot `pidx`.
struct Node
{
  std::size_t sidx;
  std::size_t pidx;
  std::size_t depth;
  double lb;
  std::uint64_t assignment_hash = 0;

  /// Priority: the least lower bound; on equal bounds the deeper node, then the
  /// larger parent slot (reference frontier.hpp:66-89).
  bool operator<(const Node& o) const
  {
    if (lb < o.lb || o.lb < lb) {
      return lb < o.lb;
    }
    if (depth != o.depth) {
      return depth > o.depth;
    }
    return pidx > o.pidx;
  }

  /// This node's box: the parent's assignment is recomputed by the policy and
  /// checked against the hash recorded at enqueue time (a policy that is not a
  /// pure function of the box would decode a different child), then sidx is
  /// decoded from it.
  void materialise(const IntervalHistory& history, std::vector<interval>& out, const policy::CompositionPolicy& pol,
                   std::span<const model::VarKind> kinds, std::span<const interval> root_box) const
  {
    if (pidx == 0) {
      out.assign(root_box.begin(), root_box.end());
      return;
    }
    const std::vector<interval>& parent = history.box(pidx);
    region::SlotAssignment const again = pol.choose(parent, kinds);
    if (region::assignment_hash(again) != assignment_hash) {
      throw InvalidConfiguration(
          "the composition policy chose a different assignment for a stored parent box than when its children "
          "were enqueued; a policy must depend on nothing but the box and the variable kinds");
    }
    region::materialise(parent, sidx, again, out);
  }
};

static_assert(sizeof(Node) == 40, "a pending node is five 8-byte fields");

/// Pending nodes in an array-backed binary min-heap under Node::operator<.
/// Ties are broken by the shape of the heap, so the exact sift rules are part
/// of the node order: a new node rises while it strictly outranks its parent;
/// a sinking node follows the right child only when it strictly outranks the
/// left, and stops at the first child that does not strictly outrank it.
class IntervalPQueue
{
public:
  explicit IntervalPQueue(std::size_t initial_capacity)
      : heap_(initial_capacity)
  {
  }

  void enqueue(const Node& n)
  {
    if (count_ == heap_.size()) {
      heap_.resize(heap_.empty() ? 1 : 2 * heap_.size());
    }
    std::size_t hole = count_;
    count_ += 1;
    while (hole > 0) {
      std::size_t const up = (hole - 1) / 2;
      if (!(n < heap_[up])) {
        break;
      }
      heap_[hole] = heap_[up];
      hole = up;
    }
    heap_[hole] = n;
  }

  Node dequeue()
  {
    Node const best = heap_[0];
    count_ -= 1;
    if (count_ > 0) {
      sink(0, heap_[count_]);
    }
    return best;
  }

  bool empty() const { return count_ == 0; }
  std::size_t size() const { return count_; }
  const Node& peek() const { return heap_[0]; }
  std::size_t capacity() const { return heap_.capacity(); }
  std::size_t capacity_bytes() const { return heap_.capacity() * sizeof(Node); }

  /// Pending nodes with lb <= gub. The set is closed under heap ancestry (a
  /// parent's lb is no larger), so the walk stops descending at the first node
  /// above gub.
  std::size_t count_viable(double gub) const
  {
    std::size_t n = 0;
    std::vector<std::size_t> todo;
    todo.push_back(0);
    while (!todo.empty()) {
      std::size_t const i = todo.back();
      todo.pop_back();
      if (i < count_ && !(heap_[i].lb > gub)) {
        n += 1;
        todo.push_back(2 * i + 1);
        todo.push_back(2 * i + 2);
      }
    }
    return n;
  }

  /// Keep the n_keep highest-priority nodes. The rest are handed to on_evict
  /// in the array order std::nth_element leaves them in (the order in which
  /// their history slots are released, which decides later slot reuse). The
  /// survivors are re-heapified bottom-up and the array is cut to 2 * n_keep
  /// entries when it is larger.
  template<typename OnEvict>
  void compact(std::size_t n_keep, OnEvict&& on_evict)
  {
    if (count_ <= n_keep) {
      return;
    }
    auto const first = heap_.begin();
    std::nth_element(first, first + static_cast<std::ptrdiff_t>(n_keep), first + static_cast<std::ptrdiff_t>(count_));
    for (std::size_t i = n_keep; i < count_; ++i) {
      on_evict(std::as_const(heap_[i]));
    }
    count_ = n_keep;
    for (std::size_t i = count_ / 2; i > 0; --i) {
      sink(i - 1, heap_[i - 1]);
    }
    if (heap_.size() > 2 * n_keep) {
      std::vector<Node> smaller(2 * n_keep);
      std::copy_n(heap_.begin(), count_, smaller.begin());
      heap_.swap(smaller);
    }
  }

private:
  /// Place `n` at `hole` and move it down to where no child outranks it.
  void sink(std::size_t hole, Node n)
  {
    for (;;) {
      std::size_t child = 2 * hole + 1;
      if (child >= count_) {
        break;
      }
      if (child + 1 < count_ && heap_[child + 1] < heap_[child]) {
        child += 1;
      }
      if (!(heap_[child] < n)) {
        break;
      }
      heap_[hole] = heap_[child];
      hole = child;
    }
    heap_[hole] = n;
  }

  std::vector<Node> heap_;
  std::size_t count_ = 0;
};

}  // namespace cmi::search
