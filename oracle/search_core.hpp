// CPU oracle -- TEST INFRASTRUCTURE ONLY. The search pieces of the oracle's
// best-first restatement, in a header so that the restated host tests
// (tests/search_pins.cpp) can check them case by case; oracle.cpp's orc_solve
// is built on them. Nothing in the product includes this file.
//
// Restated from the reference: policy/greedy_enum.hpp:54-186 (composition),
// search/frontier.hpp:66-338 (node order, heap, compaction),
// search/history.hpp:20-137 (parent boxes), search/budget.hpp:60-215 (host
// budget), search/epilogue.hpp:35-127 (incumbent admission, final claims),
// and for the k-child search aggregate/policy.hpp:106-330 (partition),
// aggregate/selection.hpp:15-125 (keys) and aggregate/frontier.hpp:106-394
// (the twice-ordered frontier).
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "cmi.h"

namespace orc
{

struct Iv
{
  double lb, ub;
};

// ---- composition (GreedyEnumerate) ----------------------------------------

struct Assign
{
  std::vector<std::uint8_t> kinds;
  std::vector<std::size_t> vars, widths;
  std::vector<std::size_t> domains;  // empty for GreedyEnumerate (no duplicate children)
};

inline std::size_t int_domain(Iv b)
{
  double const lo = std::ceil(b.lb), hi = std::floor(b.ub);
  return hi < lo ? 0 : static_cast<std::size_t>(hi - lo) + 1;
}

inline Assign choose(const std::vector<Iv>& box, const std::vector<std::uint8_t>& vk, std::size_t part,
                     std::size_t cap, std::size_t max_slots)
{
  Assign a;
  auto full = [&] { return a.vars.size() >= max_slots; };
  auto push = [&](std::uint8_t k, std::size_t v, std::size_t w) {
    a.kinds.push_back(k);
    a.vars.push_back(v);
    a.widths.push_back(w);
  };
  for (std::size_t v = 0; v < vk.size() && !full(); ++v) {
    if (vk[v] == CMI_BINARY && box[v].ub > box[v].lb) {
      push(CMI_SLOT_BINARY_ENUMERATE, v, 2);
    }
  }
  std::vector<std::size_t> ints;
  for (std::size_t v = 0; v < vk.size(); ++v) {
    if (vk[v] == CMI_INTEGER && box[v].ub > box[v].lb) {
      ints.push_back(v);
    }
  }
  std::sort(ints.begin(), ints.end(),
            [&](std::size_t x, std::size_t y) { return int_domain(box[x]) < int_domain(box[y]); });
  for (std::size_t v : ints) {
    if (full()) {
      break;
    }
    if (int_domain(box[v]) <= cap) {
      push(CMI_SLOT_INTEGER_ENUMERATE, v, cap);
    } else {
      push(CMI_SLOT_INTEGER_PARTITION, v, part);
    }
  }
  std::vector<std::size_t> cont;
  for (std::size_t v = 0; v < vk.size(); ++v) {
    if (vk[v] == CMI_CONTINUOUS && box[v].ub > box[v].lb) {
      cont.push_back(v);
    }
  }
  std::sort(cont.begin(), cont.end(),
            [&](std::size_t x, std::size_t y) { return (box[x].ub - box[x].lb) > (box[y].ub - box[y].lb); });
  for (std::size_t v : cont) {
    if (full()) {
      break;
    }
    push(CMI_SLOT_CONTINUOUS, v, part);
  }
  return a;
}

// ---- composition (WidthFirst, policy/width_first.hpp) -----------------------

inline Assign choose_width_first(const std::vector<Iv>& box, const std::vector<std::uint8_t>& vk, std::size_t part,
                                 std::size_t cap, std::size_t max_slots)
{
  std::vector<std::size_t> live;
  double top = 0.0;
  for (std::size_t v = 0; v < box.size(); ++v) {
    if (box[v].ub > box[v].lb) {
      live.push_back(v);
      if (box[v].ub - box[v].lb > top) {
        top = box[v].ub - box[v].lb;
      }
    }
  }
  std::size_t n_top = 0;
  for (std::size_t v : live) {
    n_top += !(box[v].ub - box[v].lb < top);
  }
  auto collapses = [&](std::size_t v) {
    return vk[v] == CMI_INTEGER && n_top == 1 && !(box[v].ub - box[v].lb < top) && int_domain(box[v]) <= cap;
  };
  struct C
  {
    std::size_t v;
    std::uint8_t k;
    double w;
  };
  std::vector<C> c;
  for (std::size_t v : live) {
    double const w = box[v].ub - box[v].lb;
    if (vk[v] == CMI_CONTINUOUS) {
      c.push_back({v, CMI_SLOT_CONTINUOUS, w});
    } else if (vk[v] == CMI_INTEGER) {
      c.push_back({v, static_cast<std::uint8_t>(collapses(v) ? CMI_SLOT_INTEGER_ENUMERATE : CMI_SLOT_INTEGER_PARTITION), w});
    } else {
      bool ready = true;
      for (std::size_t u : live) {
        if (u != v && box[u].ub - box[u].lb > 1.0 && !collapses(u)) {
          ready = false;
        }
      }
      if (ready) {
        c.push_back({v, CMI_SLOT_BINARY_ENUMERATE, w});
      }
    }
  }
  std::sort(c.begin(), c.end(), [](const C& a, const C& b) { return a.w > b.w; });
  Assign a;
  for (std::size_t i = 0; i < c.size() && a.vars.size() < max_slots; ++i) {
    a.kinds.push_back(c[i].k);
    a.vars.push_back(c[i].v);
    a.widths.push_back(c[i].k == CMI_SLOT_BINARY_ENUMERATE ? 2 : c[i].k == CMI_SLOT_INTEGER_ENUMERATE ? cap : part);
  }
  return a;
}

// ---- composition (BisectionBudget, policy/bisection_budget.hpp:40-299) -----

inline std::size_t bits_for(std::size_t d)
{
  std::size_t const want = d == 0 ? 1 : d;
  std::size_t b = 0;
  for (std::size_t cap = 1; cap < want; cap <<= 1) {
    ++b;
  }
  return b < 1 ? 1 : b;
}

inline Assign choose_bisection(const std::vector<Iv>& box, const std::vector<std::uint8_t>& vk, std::size_t B,
                               std::size_t max_slots, std::size_t enum_ceiling, bool relative)
{
  struct E
  {
    std::size_t v;
    std::uint8_t k;
    std::size_t d, bits;
  };
  struct S
  {
    std::size_t v;
    std::uint8_t k;
    double live, orig;
    std::size_t bits;
  };
  std::vector<E> en;
  std::vector<S> bi;
  for (std::size_t v = 0; v < vk.size() && en.size() + bi.size() < max_slots; ++v) {
    if (!(box[v].ub > box[v].lb)) {
      continue;
    }
    double const w = box[v].ub - box[v].lb;
    if (vk[v] == CMI_BINARY) {
      en.push_back({v, CMI_SLOT_BINARY_ENUMERATE, 2, bits_for(2)});
    } else if (vk[v] == CMI_INTEGER) {
      std::size_t const d = int_domain(box[v]);
      if (enum_ceiling == 0 || std::max<std::size_t>(d, 1) <= enum_ceiling) {
        en.push_back({v, CMI_SLOT_INTEGER_ENUMERATE, d, bits_for(d)});
      } else {
        bi.push_back({v, CMI_SLOT_INTEGER_PARTITION, w, w, 0});
      }
    } else {
      bi.push_back({v, CMI_SLOT_CONTINUOUS, w, w, 0});
    }
  }
  Assign a;
  if (en.empty() && bi.empty()) {
    return a;
  }
  std::stable_sort(en.begin(), en.end(), [](const E& x, const E& y) { return x.bits < y.bits; });
  std::size_t left = B, taken = 0;
  for (; taken < en.size() && en[taken].bits <= left; ++taken) {
    left -= en[taken].bits;
  }
  for (std::size_t i = taken; i < en.size(); ++i) {
    if (en[i].k == CMI_SLOT_INTEGER_ENUMERATE) {
      double const w = box[en[i].v].ub - box[en[i].v].lb;
      bi.push_back({en[i].v, CMI_SLOT_INTEGER_PARTITION, w, w, 0});
    }
  }
  std::vector<std::size_t> h(bi.size());
  for (std::size_t i = 0; i < h.size(); ++i) {
    h[i] = i;
  }
  auto lt = [&](std::size_t x, std::size_t y) {
    return relative ? bi[x].live / bi[x].orig < bi[y].live / bi[y].orig : bi[x].live < bi[y].live;
  };
  std::make_heap(h.begin(), h.end(), lt);
  std::size_t spent = 0;
  for (; spent < left && !h.empty(); ++spent) {
    std::pop_heap(h.begin(), h.end(), lt);
    bi[h.back()].live /= 2.0;
    ++bi[h.back()].bits;
    std::push_heap(h.begin(), h.end(), lt);
  }
  for (; spent < left && taken > 0; ++spent) {
    std::size_t w = 0;
    for (std::size_t i = 1; i < taken; ++i) {
      w = en[i].bits > en[w].bits ? i : w;
    }
    ++en[w].bits;
  }
  for (std::size_t i = 0; i < taken; ++i) {
    a.kinds.push_back(en[i].k);
    a.vars.push_back(en[i].v);
    a.widths.push_back(std::size_t(1) << en[i].bits);
    a.domains.push_back(en[i].d);
  }
  for (const S& c : bi) {
    if (c.bits > 0) {
      a.kinds.push_back(c.k);
      a.vars.push_back(c.v);
      a.widths.push_back(std::size_t(1) << c.bits);
      a.domains.push_back(std::size_t(1) << c.bits);
    }
  }
  return a;
}

/// Whether child r repeats a sibling: some enumerate digit beyond the domain.
inline bool duplicate_child(std::size_t r, const Assign& a)
{
  std::size_t idx = r;
  for (std::size_t j = 0; j < a.domains.size(); ++j) {
    if (idx % a.widths[j] >= a.domains[j]) {
      return true;
    }
    idx /= a.widths[j];
  }
  return false;
}

// ---- pending nodes ----------------------------------------------------------

struct Node
{
  std::size_t sidx, pidx, depth;
  double lb;
  std::uint64_t hash;  // unused by the oracle; keeps the reference's 40-byte node
};

inline bool before(const Node& a, const Node& b)
{
  if (a.lb < b.lb) {
    return true;
  }
  if (b.lb < a.lb) {
    return false;
  }
  if (a.depth != b.depth) {
    return a.depth > b.depth;
  }
  return a.pidx > b.pidx;
}

struct NodeLess
{
  bool operator()(const Node& a, const Node& b) const { return before(a, b); }
};

struct Heap
{
  std::vector<Node> e;
  std::size_t n = 0;

  explicit Heap(std::size_t initial = 0)
      : e(initial)
  {
  }

  void push(Node x)
  {
    if (n == e.size()) {
      e.resize(e.empty() ? 1 : 2 * e.size());
    }
    std::size_t i = n++;
    while (i > 0 && before(x, e[(i - 1) / 2])) {
      e[i] = e[(i - 1) / 2];
      i = (i - 1) / 2;
    }
    e[i] = x;
  }

  void down(std::size_t i)
  {
    Node const x = e[i];
    while (2 * i + 1 < n) {
      std::size_t c = 2 * i + 1;
      if (c + 1 < n && before(e[c + 1], e[c])) {
        ++c;
      }
      if (!before(e[c], x)) {
        break;
      }
      e[i] = e[c];
      i = c;
    }
    e[i] = x;
  }

  Node pop()
  {
    Node top = e[0];
    --n;
    if (n > 0) {
      e[0] = e[n];
      down(0);
    }
    return top;
  }

  std::size_t viable(double gub) const
  {
    std::size_t c = 0;
    std::vector<std::size_t> st {0};
    while (!st.empty()) {
      std::size_t i = st.back();
      st.pop_back();
      if (i >= n || e[i].lb > gub) {
        continue;
      }
      ++c;
      st.push_back(2 * i + 1);
      st.push_back(2 * i + 2);
    }
    return c;
  }

  std::size_t cap_bytes() const { return e.capacity() * sizeof(Node); }

  // keep the `keep` best; `gone` sees the rest in the order nth_element leaves
  // them; heapify bottom-up; cut the array to 2 * keep entries if larger
  template<typename F>
  void compact(std::size_t keep, F&& gone)
  {
    if (n <= keep) {
      return;
    }
    std::nth_element(e.begin(), e.begin() + static_cast<std::ptrdiff_t>(keep),
                     e.begin() + static_cast<std::ptrdiff_t>(n), NodeLess {});
    for (std::size_t i = keep; i < n; ++i) {
      gone(e[i]);
    }
    n = keep;
    for (std::size_t i = n / 2; i-- > 0;) {
      down(i);
    }
    if (e.size() > 2 * keep) {
      std::vector<Node> cut(2 * keep);
      for (std::size_t i = 0; i < n; ++i) {
        cut[i] = e[i];
      }
      e.swap(cut);
    }
  }
};

// ---- parent boxes -------------------------------------------------------------

struct History
{
  std::vector<std::vector<Iv>> box;
  std::vector<std::size_t> refs, free_list;
  std::size_t bytes = 0, freed = 0, peak = 0;

  static std::size_t cost(std::size_t nv) { return nv * sizeof(Iv) + sizeof(std::vector<Iv>); }
  std::size_t live() const { return box.size() - free_list.size(); }

  std::size_t put(std::vector<Iv> b)
  {
    bytes += cost(b.size());
    std::size_t i;
    if (!free_list.empty()) {
      i = free_list.back();
      free_list.pop_back();
      box[i] = std::move(b);
      refs[i] = 1;
    } else {
      box.push_back(std::move(b));
      refs.push_back(1);
      i = box.size() - 1;
    }
    peak = std::max(peak, live());
    return i;
  }
  void ref(std::size_t i)
  {
    if (i != 0) {
      ++refs[i];
    }
  }
  void unref(std::size_t i)
  {
    if (i != 0 && --refs[i] == 0) {
      bytes -= cost(box[i].size());
      std::vector<Iv>().swap(box[i]);
      free_list.push_back(i);
      ++freed;
    }
  }
};

// ---- host budget and the end-of-run claims ------------------------------------

inline bool closed(double gub, double bound, double tol) { return gub - bound <= tol * std::max(1.0, std::fabs(gub)); }

inline bool over_budget(std::size_t live, std::size_t budget) { return budget != 0 && live >= budget; }

inline std::size_t keep_count(std::size_t budget, std::size_t hist, std::size_t node)
{
  std::size_t const room = budget > hist ? budget - hist : 0;
  return std::max<std::size_t>(1, room / (4 * node));
}

struct Dropped
{
  double floor = INFINITY;
  std::size_t viable = 0, dominated = 0;
  void fold(double lb, double gub)
  {
    if (lb <= gub) {
      floor = std::min(floor, lb);
      ++viable;
    } else {
      ++dominated;
    }
  }
};

struct Final
{
  double glb;
  bool proven, infeasible;
};

inline Final finalise(double glb, double gub, double tol, bool found, bool converged, bool empty, double front_min,
                      std::size_t viable, const Dropped& d)
{
  if (found && (converged || empty || viable == 0) && closed(gub, d.floor, tol)) {
    return {gub, true, false};
  }
  if (empty && d.viable == 0) {
    return {glb, false, true};
  }
  double g = std::min(empty ? DBL_MAX : front_min, d.floor);
  if (found) {
    g = std::min(g, gub);
  }
  return {g, false, false};
}

inline bool admissible(const std::vector<double>& x, const std::vector<std::uint8_t>& kind, const std::vector<Iv>& root)
{
  if (x.size() != kind.size() || x.size() != root.size()) {
    return false;
  }
  for (std::size_t i = 0; i < x.size(); ++i) {
    if (!(x[i] >= root[i].lb) || !(x[i] <= root[i].ub)) {
      return false;
    }
    if (kind[i] != CMI_CONTINUOUS && std::fabs(x[i] - std::round(x[i])) > 1e-6) {
      return false;
    }
  }
  return true;
}

// ---- the k-child search --------------------------------------------------------

/// One node's k-child split: `branch` one slot (k ways on the widest live
/// variable, uniform), `refine` the other variables under log2(n / k)
/// bisections padded with width-1 slots, `launch` = refine then branch, or
/// the branch slot alone widened to n when only one variable is live.
struct AggPart
{
  Assign branch, refine, launch;
  bool nested = false;
  bool leaf() const { return branch.vars.empty(); }
};

inline std::uint8_t split_kind(std::uint8_t var_kind)
{
  return var_kind == CMI_CONTINUOUS ? CMI_SLOT_CONTINUOUS : CMI_SLOT_INTEGER_PARTITION;
}

inline void put(Assign& a, std::uint8_t kind, std::size_t v, std::size_t w)
{
  a.kinds.push_back(kind);
  a.vars.push_back(v);
  a.widths.push_back(w);
  a.domains.push_back(w);
}

inline Assign widest_split(const std::vector<Iv>& box, const std::vector<std::uint8_t>& vk, std::size_t k)
{
  Assign a;
  std::size_t best = box.size();
  for (std::size_t v = 0; v < box.size(); ++v) {
    if (box[v].ub > box[v].lb && (best == box.size() || box[v].ub - box[v].lb > box[best].ub - box[best].lb)) {
      best = v;
    }
  }
  if (best < box.size()) {
    put(a, split_kind(vk[best]), best, k);
  }
  return a;
}

inline std::size_t log2_exact(std::size_t v)
{
  std::size_t b = 0;
  while ((std::size_t(1) << b) < v) {
    ++b;
  }
  return b;
}

inline AggPart agg_partition(const std::vector<Iv>& box, const std::vector<std::uint8_t>& vk, std::size_t k,
                             std::size_t n, bool relative)
{
  AggPart p;
  p.branch = widest_split(box, vk, k);
  if (p.leaf()) {
    return p;
  }
  std::size_t const on = p.branch.vars[0];
  std::size_t live = 0;
  for (const Iv& b : box) {
    live += b.ub > b.lb;
  }
  if (live == 1) {
    p.nested = true;
    put(p.launch, p.branch.kinds[0], on, n);
    return p;
  }
  std::vector<Iv> rest = box;
  rest[on].ub = rest[on].lb;
  p.refine = choose_bisection(rest, vk, log2_exact(n / k), 64, 0, relative);
  for (std::size_t v = 0; v < box.size() && p.refine.vars.size() < 63; ++v) {
    bool const taken = std::find(p.refine.vars.begin(), p.refine.vars.end(), v) != p.refine.vars.end();
    if (v != on && box[v].ub > box[v].lb && !taken) {
      put(p.refine, split_kind(vk[v]), v, 1);
    }
  }
  p.launch = p.refine;
  put(p.launch, p.branch.kinds[0], on, k);
  return p;
}

/// Exploration key, smaller first: 0 least lower bound, 1 and 2 the reject
/// index -(f - lb) / (hub - lb) (lb until an incumbent exists).
inline double agg_key(int rule, double lb, double hub, double f)
{
  if (rule == CMI_SELECT_LEAST_LOWER_BOUND || !(f < DBL_MAX)) {
    return lb;
  }
  if (!(hub - lb > 0.0)) {
    return lb < f ? -INFINITY : INFINITY;
  }
  return -((f - lb) / (hub - lb));
}

struct AggNode
{
  std::size_t sidx, pidx, depth;
  double lb, hub;
  std::uint64_t hash;
};

/// Pending children in a slot store, with a key heap and an lb heap of
/// references (std heap algorithms, lazy deletion, generation-checked slot
/// reuse, a sweep when stale references exceed 2 * live + 64).
struct AggFront
{
  struct Slot
  {
    AggNode n;
    std::uint64_t gen;
    bool alive;
  };
  struct Ref
  {
    double key;
    std::size_t depth, pidx, slot;
    std::uint64_t gen;
  };
  static_assert(sizeof(Slot) == 64 && sizeof(Ref) == 40, "the reference's layout, which the host budget counts");

  int rule;
  std::vector<Slot> slots;
  std::vector<std::size_t> spare;
  std::vector<Ref> keyh, lbh;
  std::size_t live = 0;

  static bool cmp(const Ref& a, const Ref& b)  // "a below b" for std heaps
  {
    if (a.key < b.key || b.key < a.key) {
      return a.key > b.key;
    }
    if (a.depth != b.depth) {
      return a.depth < b.depth;
    }
    return a.pidx < b.pidx;
  }
  bool fresh(const Ref& r) const { return slots[r.slot].alive && slots[r.slot].gen == r.gen; }
  Ref ref(std::size_t s, double key) const { return {key, slots[s].n.depth, slots[s].n.pidx, s, slots[s].gen}; }

  void push(const AggNode& x, double f)
  {
    std::size_t s;
    if (!spare.empty()) {
      s = spare.back();
      spare.pop_back();
      slots[s] = {x, slots[s].gen + 1, true};
    } else {
      s = slots.size();
      slots.push_back({x, 0, true});
    }
    ++live;
    keyh.push_back(ref(s, agg_key(rule, x.lb, x.hub, f)));
    std::push_heap(keyh.begin(), keyh.end(), cmp);
    lbh.push_back(ref(s, x.lb));
    std::push_heap(lbh.begin(), lbh.end(), cmp);
  }

  AggNode pop()
  {
    for (;;) {
      Ref const r = keyh.front();
      std::pop_heap(keyh.begin(), keyh.end(), cmp);
      keyh.pop_back();
      if (!fresh(r)) {
        continue;
      }
      slots[r.slot].alive = false;
      --live;
      spare.push_back(r.slot);
      if (keyh.size() + lbh.size() > 2 * live + 64) {
        for (std::vector<Ref>* h : {&keyh, &lbh}) {
          std::vector<Ref> k2;
          k2.reserve(live);
          for (const Ref& q : *h) {
            if (fresh(q)) {
              k2.push_back(q);
            }
          }
          std::make_heap(k2.begin(), k2.end(), cmp);
          h->swap(k2);
        }
      }
      return slots[r.slot].n;
    }
  }

  double min_lb()
  {
    while (!lbh.empty() && !fresh(lbh.front())) {
      std::pop_heap(lbh.begin(), lbh.end(), cmp);
      lbh.pop_back();
    }
    return lbh.empty() ? DBL_MAX : slots[lbh.front().slot].n.lb;
  }

  void incumbent(double f)
  {
    if (rule != CMI_SELECT_REJECT_INDEX) {
      return;
    }
    keyh.clear();
    lbh.clear();
    keyh.reserve(live);
    lbh.reserve(live);
    for (std::size_t s = 0; s < slots.size(); ++s) {
      if (slots[s].alive) {
        keyh.push_back(ref(s, agg_key(rule, slots[s].n.lb, slots[s].n.hub, f)));
        lbh.push_back(ref(s, slots[s].n.lb));
      }
    }
    std::make_heap(keyh.begin(), keyh.end(), cmp);
    std::make_heap(lbh.begin(), lbh.end(), cmp);
  }

  std::size_t viable(double gub) const
  {
    std::size_t c = 0;
    for (const Slot& s : slots) {
      c += s.alive && !(s.n.lb > gub);
    }
    return c;
  }

  std::size_t cap_bytes() const
  {
    return slots.capacity() * sizeof(Slot) + (keyh.capacity() + lbh.capacity()) * sizeof(Ref)
           + spare.capacity() * sizeof(std::size_t);
  }
};

}  // namespace orc
