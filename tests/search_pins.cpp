// The reference's host search tests restated case by case, run against both
// implementations of the search: the engine's (cuminlpsearch_amd/csrc/cmi/
// frontier.hpp, policy.hpp, region.hpp) and the oracle's
// (oracle/search_core.hpp). Each case names the reference test it restates
// (test/source/host_budget_test.cpp, test/source/composition_policy_test.cpp).
//
// Built and run by tests/test_search_pins_cpu.py; prints one line per failed
// check and exits non-zero if any failed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "cmi/frontier.hpp"
#include "cmi/policy.hpp"
#include "cmi/region.hpp"
#include "search_core.hpp"

namespace
{

int g_checks = 0;
int g_failed = 0;
const char* g_case = "";
const char* g_impl = "";

void check(bool ok, const char* what, int line)
{
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::printf("FAIL [%s] %s: %s (line %d)\n", g_impl, g_case, what, line);
  }
}

#define CHECK(x) check((x), #x, __LINE__)

/// Bitwise equality through <, as the reference's tests compare literals.
bool feq(double a, double b) { return !(a < b) && !(b < a); }

struct Box
{
  double lb, ub;
};

// ---- the two implementations behind one vocabulary --------------------------

struct Engine
{
  static constexpr const char* name = "engine";
  using Queue = cmi::search::IntervalPQueue;
  using History = cmi::search::IntervalHistory;
  using Ledger = cmi::search::DropAccounting;

  static Queue queue(std::size_t cap) { return Queue(cap); }
  static void push(Queue& q, double lb, std::size_t depth = 1, std::size_t pidx = 1)
  {
    q.enqueue(cmi::search::Node {0, pidx, depth, lb, 1});
  }
  static double pop(Queue& q) { return q.dequeue().lb; }
  static double peek(const Queue& q) { return q.peek().lb; }
  static std::size_t size(const Queue& q) { return q.size(); }
  static std::size_t capacity(const Queue& q) { return q.capacity(); }
  static std::size_t capacity_bytes(const Queue& q) { return q.capacity_bytes(); }
  static std::size_t viable(const Queue& q, double gub) { return q.count_viable(gub); }
  template<typename F>
  static void compact(Queue& q, std::size_t keep, F&& f)
  {
    q.compact(keep, [&](const cmi::search::Node& n) { f(n.lb, n.depth, n.pidx); });
  }
  static constexpr std::size_t node_bytes = sizeof(cmi::search::Node);

  static std::size_t put(History& h, std::vector<Box> b)
  {
    std::vector<cmi::interval> v;
    for (Box x : b) {
      v.push_back({x.lb, x.ub});
    }
    return h.enqueue(std::move(v));
  }
  static void ref(History& h, std::size_t i) { h.add_ref(i); }
  static void unref(History& h, std::size_t i) { h.release(i); }
  static std::size_t live_bytes(const History& h) { return h.live_bytes(); }
  static std::size_t live_count(const History& h) { return h.live_count(); }
  static std::size_t refs(const History& h, std::size_t i) { return h.ref_count(i); }
  static std::size_t freed(const History& h) { return h.freed_count(); }
  static std::size_t peak(const History& h) { return h.peak_live(); }
  static std::size_t slots(const History& h) { return h.slot_count(); }
  static Box first(const History& h, std::size_t i) { return {h.box(i)[0].lb, h.box(i)[0].ub}; }

  static void fold(Ledger& d, double lb, double gub) { d.fold(lb, gub); }
  static std::size_t dropped_viable(const Ledger& d) { return d.viable; }
  static std::size_t dropped_dominated(const Ledger& d) { return d.dominated; }
  static double floor(const Ledger& d) { return d.lb_min; }
  static bool lost_nothing(const Ledger& d) { return d.lost_nothing(); }

  struct Final
  {
    double glb;
    bool proven, infeasible;
  };
  static Final finalise(double glb, double gub, double tol, bool found, bool conv, bool empty, double fmin,
                        std::size_t viable, const Ledger& d)
  {
    auto const f = cmi::search::finalise_bounds(glb, gub, tol, found, conv, empty, fmin, viable, d);
    return {f.glb, f.proven_optimal, f.infeasible};
  }
  static std::size_t keep_count(std::size_t budget, std::size_t hist, std::size_t node)
  {
    return cmi::search::compact_keep_count(budget, hist, node);
  }
  static bool over(std::size_t live, std::size_t budget) { return cmi::search::over_host_budget(live, budget); }

  static bool admissible(std::vector<double> x, std::vector<int> kinds, std::vector<Box> box)
  {
    std::vector<cmi::model::VarKind> k;
    for (int v : kinds) {
      k.push_back(static_cast<cmi::model::VarKind>(v));
    }
    std::vector<cmi::interval> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    return cmi::search::witness_is_admissible(x, k, b);
  }

  struct Choice
  {
    std::vector<int> kinds;
    std::vector<std::size_t> vars, widths;
  };
  static Choice choose(std::vector<Box> box, std::vector<int> kinds, std::size_t part, std::size_t cap,
                       std::size_t max_slots = 0)
  {
    cmi::policy::SearchCalibration cal;
    cal.max_cycle_size = max_slots;
    cmi::policy::GreedyEnumCompositionPolicy pol(cmi::region::FanOutSpec(part, cap), cal);
    std::vector<cmi::interval> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    std::vector<cmi::model::VarKind> k;
    for (int v : kinds) {
      k.push_back(static_cast<cmi::model::VarKind>(v));
    }
    auto const a = pol.choose(b, k);
    Choice c;
    for (auto s : a.composition) {
      c.kinds.push_back(static_cast<int>(s));
    }
    c.vars = a.var_ids;
    c.widths = a.widths;
    return c;
  }
  static std::size_t domain(Box b)
  {
    return cmi::policy::GreedyEnumCompositionPolicy::integer_domain_size({b.lb, b.ub});
  }

  struct Bisected
  {
    std::vector<int> kinds;
    std::vector<std::size_t> vars, widths, domains;
  };
  static Bisected bisect(std::vector<Box> box, std::vector<int> kinds, std::size_t B, bool relative = false,
                         std::size_t ceiling = 0)
  {
    cmi::policy::SearchCalibration cal;
    cal.enumerate_cap = ceiling;
    cmi::policy::BisectionBudgetCompositionPolicy pol(B, cal, relative);
    std::vector<cmi::interval> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    std::vector<cmi::model::VarKind> k;
    for (int v : kinds) {
      k.push_back(static_cast<cmi::model::VarKind>(v));
    }
    auto const a = pol.choose(b, k);
    Bisected c;
    for (auto sk : a.composition) {
      c.kinds.push_back(static_cast<int>(sk));
    }
    c.vars = a.var_ids;
    c.widths = a.widths;
    c.domains = a.domain_sizes;
    return c;
  }

  static Choice widthfirst(std::vector<Box> box, std::vector<int> kinds, std::size_t part, std::size_t cap,
                           std::size_t max_slots = 0)
  {
    cmi::policy::SearchCalibration cal;
    cal.max_cycle_size = max_slots;
    cmi::policy::WidthFirstCompositionPolicy pol(cmi::region::FanOutSpec(part, cap), cal);
    std::vector<cmi::interval> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    std::vector<cmi::model::VarKind> k;
    for (int v : kinds) {
      k.push_back(static_cast<cmi::model::VarKind>(v));
    }
    auto const a = pol.choose(b, k);
    Choice c;
    for (auto sk : a.composition) {
      c.kinds.push_back(static_cast<int>(sk));
    }
    c.vars = a.var_ids;
    c.widths = a.widths;
    return c;
  }
};

struct Oracle
{
  static constexpr const char* name = "oracle";
  using Queue = orc::Heap;
  using History = orc::History;
  using Ledger = orc::Dropped;

  static Queue queue(std::size_t cap) { return Queue(cap); }
  static void push(Queue& q, double lb, std::size_t depth = 1, std::size_t pidx = 1)
  {
    q.push(orc::Node {0, pidx, depth, lb, 1});
  }
  static double pop(Queue& q) { return q.pop().lb; }
  static double peek(const Queue& q) { return q.e[0].lb; }
  static std::size_t size(const Queue& q) { return q.n; }
  static std::size_t capacity(const Queue& q) { return q.e.capacity(); }
  static std::size_t capacity_bytes(const Queue& q) { return q.cap_bytes(); }
  static std::size_t viable(const Queue& q, double gub) { return q.viable(gub); }
  template<typename F>
  static void compact(Queue& q, std::size_t keep, F&& f)
  {
    q.compact(keep, [&](const orc::Node& n) { f(n.lb, n.depth, n.pidx); });
  }
  static constexpr std::size_t node_bytes = sizeof(orc::Node);

  static std::size_t put(History& h, std::vector<Box> b)
  {
    std::vector<orc::Iv> v;
    for (Box x : b) {
      v.push_back({x.lb, x.ub});
    }
    return h.put(std::move(v));
  }
  static void ref(History& h, std::size_t i) { h.ref(i); }
  static void unref(History& h, std::size_t i) { h.unref(i); }
  static std::size_t live_bytes(const History& h) { return h.bytes; }
  static std::size_t live_count(const History& h) { return h.live(); }
  static std::size_t refs(const History& h, std::size_t i) { return h.refs[i]; }
  static std::size_t freed(const History& h) { return h.freed; }
  static std::size_t peak(const History& h) { return h.peak; }
  static std::size_t slots(const History& h) { return h.box.size(); }
  static Box first(const History& h, std::size_t i) { return {h.box[i][0].lb, h.box[i][0].ub}; }

  static void fold(Ledger& d, double lb, double gub) { d.fold(lb, gub); }
  static std::size_t dropped_viable(const Ledger& d) { return d.viable; }
  static std::size_t dropped_dominated(const Ledger& d) { return d.dominated; }
  static double floor(const Ledger& d) { return d.floor; }
  static bool lost_nothing(const Ledger& d) { return d.viable == 0; }

  using Final = orc::Final;
  static Final finalise(double glb, double gub, double tol, bool found, bool conv, bool empty, double fmin,
                        std::size_t viable, const Ledger& d)
  {
    return orc::finalise(glb, gub, tol, found, conv, empty, fmin, viable, d);
  }
  static std::size_t keep_count(std::size_t budget, std::size_t hist, std::size_t node)
  {
    return orc::keep_count(budget, hist, node);
  }
  static bool over(std::size_t live, std::size_t budget) { return orc::over_budget(live, budget); }

  static bool admissible(std::vector<double> x, std::vector<int> kinds, std::vector<Box> box)
  {
    std::vector<std::uint8_t> k(kinds.begin(), kinds.end());
    std::vector<orc::Iv> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    return orc::admissible(x, k, b);
  }

  struct Choice
  {
    std::vector<int> kinds;
    std::vector<std::size_t> vars, widths;
  };
  static Choice choose(std::vector<Box> box, std::vector<int> kinds, std::size_t part, std::size_t cap,
                       std::size_t max_slots = 0)
  {
    std::vector<orc::Iv> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    std::vector<std::uint8_t> k(kinds.begin(), kinds.end());
    auto const a = orc::choose(b, k, part, cap, max_slots == 0 ? 64 : max_slots);
    Choice c;
    c.kinds.assign(a.kinds.begin(), a.kinds.end());
    c.vars = a.vars;
    c.widths = a.widths;
    return c;
  }
  static std::size_t domain(Box b) { return orc::int_domain({b.lb, b.ub}); }

  struct Bisected
  {
    std::vector<int> kinds;
    std::vector<std::size_t> vars, widths, domains;
  };
  static Bisected bisect(std::vector<Box> box, std::vector<int> kinds, std::size_t B, bool relative = false,
                         std::size_t ceiling = 0)
  {
    std::vector<orc::Iv> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    std::vector<std::uint8_t> k(kinds.begin(), kinds.end());
    auto const a = orc::choose_bisection(b, k, B, 64, ceiling, relative);
    Bisected c;
    c.kinds.assign(a.kinds.begin(), a.kinds.end());
    c.vars = a.vars;
    c.widths = a.widths;
    c.domains = a.domains;
    return c;
  }

  static Choice widthfirst(std::vector<Box> box, std::vector<int> kinds, std::size_t part, std::size_t cap,
                           std::size_t max_slots = 0)
  {
    std::vector<orc::Iv> b;
    for (Box v : box) {
      b.push_back({v.lb, v.ub});
    }
    std::vector<std::uint8_t> k(kinds.begin(), kinds.end());
    auto const a = orc::choose_width_first(b, k, part, cap, max_slots == 0 ? 64 : max_slots);
    Choice c;
    c.kinds.assign(a.kinds.begin(), a.kinds.end());
    c.vars = a.vars;
    c.widths = a.widths;
    return c;
  }
};

constexpr int kCont = CMI_CONTINUOUS, kInt = CMI_INTEGER, kBin = CMI_BINARY;
constexpr int sC = CMI_SLOT_CONTINUOUS, sP = CMI_SLOT_INTEGER_PARTITION, sE = CMI_SLOT_INTEGER_ENUMERATE,
              sB = CMI_SLOT_BINARY_ENUMERATE;

template<typename I>
std::vector<double> drain(typename I::Queue& q)
{
  std::vector<double> out;
  while (I::size(q) > 0) {
    out.push_back(I::pop(q));
  }
  return out;
}

// ---- host_budget_test.cpp §8.1: compaction -----------------------------------

template<typename I>
void compaction_cases()
{
  g_case = "compact keeps exactly the n_keep best elements (host_budget_test.cpp:74)";
  {
    auto q = I::queue(1);
    std::vector<double> const bounds {5.0, -2.0, 11.0, 0.5, 7.25, -9.0, 3.0, 100.0, -0.25, 42.0};
    for (double lb : bounds) {
      I::push(q, lb);
    }
    CHECK(I::size(q) == bounds.size());
    std::vector<double> sorted = bounds;
    std::sort(sorted.begin(), sorted.end());
    std::vector<double> evicted;
    I::compact(q, 4, [&](double lb, std::size_t, std::size_t) { evicted.push_back(lb); });
    CHECK(I::size(q) == 4);
    CHECK(I::capacity(q) <= 8);
    std::vector<double> const kept = drain<I>(q);
    CHECK(kept.size() == 4);
    for (std::size_t i = 0; i < kept.size() && i < 4; ++i) {
      CHECK(feq(kept[i], sorted[i]));
    }
    std::sort(evicted.begin(), evicted.end());
    CHECK(evicted.size() == 6);
    for (std::size_t i = 0; i < evicted.size() && i < 6; ++i) {
      CHECK(feq(evicted[i], sorted[4 + i]));
    }
  }
  g_case = "compact preserves count_viable for surviving regions (:119)";
  {
    auto q = I::queue(1);
    for (double lb : {1.0, 2.0, 3.0, 50.0, 60.0, 70.0}) {
      I::push(q, lb);
    }
    CHECK(I::viable(q, 10.0) == 3);
    I::compact(q, 3, [](double, std::size_t, std::size_t) {});
    CHECK(I::size(q) == 3);
    CHECK(I::viable(q, 10.0) == 3);
  }
  g_case = "compact is a no-op when the frontier already fits (:137)";
  {
    auto q = I::queue(1);
    for (double lb : {1.0, 2.0, 3.0}) {
      I::push(q, lb);
    }
    std::size_t evictions = 0;
    I::compact(q, 3, [&](double, std::size_t, std::size_t) { ++evictions; });
    CHECK(evictions == 0);
    CHECK(I::size(q) == 3);
    I::compact(q, 99, [&](double, std::size_t, std::size_t) { ++evictions; });
    CHECK(evictions == 0);
    CHECK(drain<I>(q).size() == 3);
  }
  g_case = "compact to a single element keeps the frontier minimum (:158)";
  {
    auto q = I::queue(1);
    for (double lb : {4.0, -1.5, 9.0, 2.0}) {
      I::push(q, lb);
    }
    I::compact(q, 1, [](double, std::size_t, std::size_t) {});
    CHECK(I::size(q) == 1);
    CHECK(feq(I::peek(q), -1.5));
  }
}

// ---- §8.2: eviction order and the dropped floor --------------------------------

template<typename I>
void eviction_cases()
{
  g_case = "eviction takes dominated regions before viable ones (:174)";
  {
    auto q = I::queue(1);
    for (double lb : {12.0, 1.0, 30.0, 5.0, 11.5, 9.0}) {
      I::push(q, lb);
    }
    typename I::Ledger d;
    I::compact(q, 3, [&](double lb, std::size_t, std::size_t) { I::fold(d, lb, 10.0); });
    CHECK(I::dropped_dominated(d) == 3);
    CHECK(I::dropped_viable(d) == 0);
    CHECK(I::lost_nothing(d));
    CHECK(std::isinf(I::floor(d)));
  }
  g_case = "among equal bounds the shallower region is evicted first (:198)";
  {
    auto q = I::queue(1);
    I::push(q, 1.0, 7);
    I::push(q, 1.0, 2);
    I::push(q, 1.0, 9);
    std::vector<std::size_t> depths;
    I::compact(q, 2, [&](double, std::size_t depth, std::size_t) { depths.push_back(depth); });
    CHECK(depths.size() == 1);
    CHECK(!depths.empty() && depths[0] == 2);
  }
  g_case = "a viable eviction lowers the floor to its bound (:212)";
  {
    auto q = I::queue(1);
    for (double lb : {2.0, 3.0, 4.0, 5.0}) {
      I::push(q, lb);
    }
    typename I::Ledger d;
    I::compact(q, 2, [&](double lb, std::size_t, std::size_t) { I::fold(d, lb, 10.0); });
    CHECK(I::dropped_viable(d) == 2);
    CHECK(I::dropped_dominated(d) == 0);
    CHECK(feq(I::floor(d), 4.0));
    CHECK(!I::lost_nothing(d));
  }
}

// ---- §8.3: the parent-box history ---------------------------------------------

template<typename I>
void history_cases()
{
  g_case = "history frees a box once its last child is released (:233)";
  {
    typename I::History h;
    CHECK(I::put(h, {}) == 0);
    std::size_t const base = I::live_bytes(h);
    std::size_t const idx = I::put(h, {{0.0, 1.0}, {2.0, 3.0}});
    CHECK(idx == 1);
    CHECK(I::live_bytes(h) > base);
    CHECK(I::live_count(h) == 2);
    I::ref(h, idx);
    I::ref(h, idx);
    I::unref(h, idx);
    CHECK(I::refs(h, idx) == 2);
    I::unref(h, idx);
    CHECK(I::live_bytes(h) > base);
    I::unref(h, idx);
    CHECK(I::live_bytes(h) == base);
    CHECK(I::live_count(h) == 1);
  }
  g_case = "history reuses a freed slot (:262)";
  {
    typename I::History h;
    I::put(h, {});
    std::size_t const a = I::put(h, {{0.0, 1.0}});
    I::unref(h, a);
    std::size_t const b = I::put(h, {{4.0, 8.0}});
    CHECK(b == a);
    CHECK(I::slots(h) == 2);
    CHECK(feq(I::first(h, b).lb, 4.0));
  }
  g_case = "slot 0 is never freed and never reused (:276)";
  {
    typename I::History h;
    CHECK(I::put(h, {}) == 0);
    I::unref(h, 0);
    I::unref(h, 0);
    I::ref(h, 0);
    CHECK(I::put(h, {{0.0, 1.0}}) == 1);
    CHECK(I::slots(h) == 2);
  }
  g_case = "materialise resolves a child through a reused slot (:290)";
  {
    typename I::History h;
    I::put(h, {});
    std::size_t const stale = I::put(h, {{0.0, 100.0}});
    I::unref(h, stale);
    std::size_t const fresh = I::put(h, {{5.0, 6.0}});
    CHECK(fresh == stale);
    CHECK(feq(I::first(h, fresh).lb, 5.0) && feq(I::first(h, fresh).ub, 6.0));
  }
  g_case = "freed_count tallies slots that hit refcount zero (:308)";
  {
    typename I::History h;
    I::put(h, {});
    CHECK(I::freed(h) == 0);
    std::size_t const idx = I::put(h, {{0.0, 1.0}});
    I::ref(h, idx);
    I::unref(h, idx);
    CHECK(I::freed(h) == 0);
    I::unref(h, idx);
    CHECK(I::freed(h) == 1);
    std::size_t const again = I::put(h, {{2.0, 3.0}});
    CHECK(again == idx);
    I::unref(h, again);
    CHECK(I::freed(h) == 2);
  }
  g_case = "releasing the root sentinel never counts as freed (:331)";
  {
    typename I::History h;
    I::put(h, {});
    I::unref(h, 0);
    I::unref(h, 0);
    CHECK(I::freed(h) == 0);
  }
  g_case = "peak_live tracks the high-water mark of live_count (:341)";
  {
    typename I::History h;
    I::put(h, {});
    CHECK(I::peak(h) == 1);
    std::size_t const a = I::put(h, {{0.0, 1.0}});
    std::size_t const b = I::put(h, {{1.0, 2.0}});
    CHECK(I::peak(h) == 3);
    I::unref(h, a);
    I::unref(h, b);
    CHECK(I::live_count(h) == 1);
    CHECK(I::peak(h) == 3);
    I::put(h, {{2.0, 3.0}});
    CHECK(I::peak(h) == 3);
  }
  g_case = "live_bytes returns to baseline over an expand/consume cycle (:366)";
  {
    typename I::History h;
    I::put(h, {});
    std::size_t const base = I::live_bytes(h);
    for (int round = 0; round < 5; ++round) {
      std::size_t const parent = I::put(h, {{0.0, 1.0}, {0.0, 1.0}, {0.0, 1.0}});
      for (int c = 0; c < 4; ++c) {
        I::ref(h, parent);
      }
      I::unref(h, parent);
      for (int c = 0; c < 4; ++c) {
        I::unref(h, parent);
      }
      CHECK(I::live_bytes(h) == base);
    }
    CHECK(I::slots(h) == 2);
  }
}

// ---- §8.4: the epilogue's claims ------------------------------------------------

template<typename I>
void epilogue_cases()
{
  constexpr double kNone = std::numeric_limits<double>::max();
  constexpr double kTol = 1e-6;
  auto fin = [&](double glb, double gub, bool conv, bool empty, double fmin, std::size_t viable,
                 const typename I::Ledger& d) { return I::finalise(glb, gub, kTol, gub < kNone, conv, empty, fmin, viable, d); };
  typename I::Ledger const none;

  g_case = "nothing dropped: an emptied frontier with an incumbent is proven optimal (:421)";
  {
    auto const o = fin(1.0, 5.0, false, true, 0.0, 0, none);
    CHECK(o.proven && !o.infeasible && feq(o.glb, 5.0));
  }
  g_case = "nothing dropped: an emptied frontier with no incumbent is infeasible (:428)";
  {
    auto const o = fin(1.0, kNone, false, true, 0.0, 0, none);
    CHECK(!o.proven && o.infeasible && feq(o.glb, 1.0));
  }
  g_case = "nothing dropped: a stranded frontier is clamped to the incumbent (:437)";
  {
    auto const o = fin(1.0, 5.0, false, false, 9.0, 0, none);
    CHECK(o.proven && feq(o.glb, 5.0));
  }
  g_case = "nothing dropped: an unfinished run reports the frontier minimum (:444)";
  {
    auto const o = fin(1.0, 5.0, false, false, 2.0, 3, none);
    CHECK(!o.proven && !o.infeasible && feq(o.glb, 2.0));
  }
  g_case = "a dominated-only drop still permits proven optimality (:453)";
  {
    typename I::Ledger d;
    I::fold(d, 40.0, 5.0);
    CHECK(I::dropped_dominated(d) == 1 && I::lost_nothing(d));
    auto const o = fin(1.0, 5.0, false, true, 0.0, 0, d);
    CHECK(o.proven && feq(o.glb, 5.0));
  }
  g_case = "a viable drop caps the reported lower bound (:466)";
  {
    typename I::Ledger d;
    I::fold(d, 2.0, 5.0);
    CHECK(I::dropped_viable(d) == 1);
    auto const a = fin(1.0, 5.0, false, true, 0.0, 0, d);
    CHECK(!a.proven && !a.infeasible && feq(a.glb, 2.0));
    auto const b = fin(1.0, 5.0, false, false, 4.0, 1, d);
    CHECK(feq(b.glb, 2.0));
    auto const c = fin(1.0, 5.0, false, false, 1.5, 1, d);
    CHECK(feq(c.glb, 1.5));
  }
  g_case = "a viable drop within tolerance is still proven optimal (:492)";
  {
    typename I::Ledger d;
    I::fold(d, 5.0 - 1e-9, 5.0);
    CHECK(I::dropped_viable(d) == 1);
    auto const o = fin(1.0, 5.0, true, false, 5.0, 1, d);
    CHECK(o.proven && feq(o.glb, 5.0));
  }
  g_case = "an emptied frontier with viable drops is not infeasible (:504)";
  {
    typename I::Ledger d;
    I::fold(d, -3.0, kNone);
    CHECK(I::dropped_viable(d) == 1);
    auto const o = fin(1.0, kNone, false, true, 0.0, 0, d);
    CHECK(!o.infeasible && !o.proven && feq(o.glb, -3.0));
  }
}

// ---- §8.5: the budget arithmetic ---------------------------------------------------

template<typename I>
void budget_cases()
{
  g_case = "n_keep is a quarter of the nodes the budget can hold (:522)";
  CHECK(I::keep_count(2000, 1000, 10) == 25);
  CHECK(I::keep_count(2000, 0, 10) == 50);

  g_case = "a compaction leaves room to grow back into (:532)";
  {
    std::size_t const budget = 2000, node = 10, hist = 400, room = budget - hist;
    std::size_t const keep = I::keep_count(budget, hist, node);
    CHECK(keep == 40);
    CHECK(!I::over(2 * keep * node + hist, budget));
    CHECK(!I::over(2 * keep * node + hist + 64, budget));
    CHECK(I::over(4 * keep * node + hist, budget));
    CHECK(4 * keep * node <= room);
  }
  g_case = "the budget is reached at exactly the budget (:557)";
  CHECK(I::over(1000, 1000));
  CHECK(I::over(1001, 1000));
  CHECK(!I::over(999, 1000));
  CHECK(!I::over(0, 0));
  CHECK(!I::over(std::size_t(1) << 30, 0));

  g_case = "n_keep floors at one when the history crowds out the frontier (:572)";
  CHECK(I::keep_count(1000, 1000, 40) == 1);
  CHECK(I::keep_count(1000, 999999, 40) == 1);
  CHECK(I::keep_count(0, 0, 40) == 1);

  g_case = "a budget too small for one fan-out stops rather than thrashes (:579)";
  {
    std::size_t const keep = I::keep_count(100, 96, 40);
    CHECK(keep == 1);
    CHECK(2 * keep * 40 + 96 > 100);
  }
  g_case = "frontier_bytes counts capacity rather than size (:596)";
  {
    auto q = I::queue(64);
    CHECK(I::size(q) == 0);
    CHECK(I::capacity_bytes(q) >= 64 * I::node_bytes);
    I::push(q, 1.0);
    CHECK(I::size(q) == 1);
    CHECK(I::capacity_bytes(q) >= 64 * I::node_bytes);
  }
  g_case = "a compaction gives the capacity back (:609)";
  {
    auto q = I::queue(1);
    for (int i = 0; i < 500; ++i) {
      I::push(q, static_cast<double>(i));
    }
    std::size_t const before = I::capacity_bytes(q);
    I::compact(q, 10, [](double, std::size_t, std::size_t) {});
    CHECK(I::capacity_bytes(q) < before);
    CHECK(I::capacity_bytes(q) <= 20 * I::node_bytes);
  }
  g_case = "a compacted frontier absorbs n_keep insertions before it grows (:623)";
  {
    constexpr std::size_t keep = 40;
    auto q = I::queue(1);
    for (int i = 0; i < 500; ++i) {
      I::push(q, static_cast<double>(i));
    }
    I::compact(q, keep, [](double, std::size_t, std::size_t) {});
    CHECK(I::size(q) == keep);
    std::size_t const cap = I::capacity(q);
    CHECK(cap >= 2 * keep);
    for (std::size_t i = 0; i < keep; ++i) {
      I::push(q, -1.0 - static_cast<double>(i));
    }
    CHECK(I::size(q) == 2 * keep);
    CHECK(I::capacity(q) == cap);
    I::push(q, -1000.0);
    CHECK(I::capacity(q) > cap);
  }
  g_case = "a node is the reference's 40 bytes (the budget charges it)";
  CHECK(I::node_bytes == 40);
}

// ---- §6.2: witness admission --------------------------------------------------------

template<typename I>
void witness_cases()
{
  g_case = "witness_is_admissible accepts a lattice point within the box (:652)";
  CHECK(I::admissible({1.0, 3.0, 0.5}, {kInt, kBin, kCont}, {{0.0, 5.0}, {0.0, 3.0}, {0.0, 1.0}}));
  g_case = "witness_is_admissible rejects a fractional Integer/Binary component (:663)";
  CHECK(!I::admissible({5.7}, {kInt}, {{0.0, 10.0}}));
  CHECK(I::admissible({5.0 + 1e-9}, {kInt}, {{0.0, 10.0}}));
  g_case = "witness_is_admissible rejects a component outside the root box (:679)";
  CHECK(!I::admissible({1.5}, {kCont}, {{0.0, 1.0}}));
  CHECK(!I::admissible({-0.5}, {kCont}, {{0.0, 1.0}}));
  CHECK(I::admissible({1.0}, {kCont}, {{0.0, 1.0}}));
  g_case = "a witness of the wrong length or with a NaN is never admissible";
  CHECK(!I::admissible({0.5, 0.5}, {kCont}, {{0.0, 1.0}}));
  CHECK(!I::admissible({std::nan("")}, {kCont}, {{0.0, 1.0}}));
}

// ---- composition_policy_test.cpp: GreedyEnumerate --------------------------------------

template<typename I>
void policy_cases()
{
  g_case = "fills binary slots before integer slots (composition_policy_test.cpp:30)";
  {
    auto const a = I::choose({{0.0, 10.0}, {0.0, 1.0}, {0.0, 3.0}}, {kCont, kBin, kInt}, 4, 4);
    CHECK(a.vars.size() == 3);
    CHECK(a.vars[0] == 1 && a.kinds[0] == sB);
    CHECK(a.vars[1] == 2 && a.kinds[1] == sE);
    CHECK(a.vars[2] == 0 && a.kinds[2] == sC);
    CHECK(a.widths == std::vector<std::size_t>({2, 4, 4}));
  }
  g_case = "partitions integer domains above PartitionNum (:52)";
  {
    auto const a = I::choose({{0.0, 100.0}}, {kInt}, 4, 4);
    CHECK(a.vars.size() == 1 && a.vars[0] == 0 && a.kinds[0] == sP && a.widths[0] == 4);
  }
  g_case = "enumerate_cap is independent of partition_num's partition width (:66)";
  {
    auto const a = I::choose({{0.0, 39.0}, {0.0, 10.0}}, {kInt, kCont}, 4, 50);
    CHECK(a.vars.size() == 2);
    CHECK(a.vars[0] == 0 && a.kinds[0] == sE && a.widths[0] == 50);
    CHECK(a.vars[1] == 1 && a.kinds[1] == sC && a.widths[1] == 4);
  }
  g_case = "prefers the smallest remaining integer domain (:90)";
  {
    auto const a = I::choose({{0.0, 100.0}, {0.0, 3.0}}, {kInt, kInt}, 50, 50);
    CHECK(a.vars[0] == 1 && a.kinds[0] == sE);
    CHECK(a.vars[1] == 0 && a.kinds[1] == sP);
  }
  g_case = "prefers the widest continuous variable (:104)";
  {
    auto const a = I::choose({{0.0, 2.0}, {0.0, 10.0}}, {kCont, kCont}, 4, 4);
    CHECK(a.vars[0] == 1 && a.vars[1] == 0);
  }
  g_case = "spills into the next kind once one runs out (:117)";
  {
    auto const a = I::choose({{0.0, 1.0}, {0.0, 5.0}}, {kBin, kInt}, 4, 4);
    CHECK(a.vars.size() == 2 && a.vars[0] == 0 && a.vars[1] == 1);
  }
  g_case = "returns an empty composition once every variable is resolved (:131)";
  {
    auto const a = I::choose({{3.0, 3.0}}, {kCont}, 4, 4);
    CHECK(a.kinds.empty() && a.vars.empty());
  }
  g_case = "choose is a pure function of (box, var_kinds) (:145)";
  {
    auto const a = I::choose({{0.0, 10.0}, {0.0, 1.0}, {0.0, 3.0}}, {kCont, kBin, kInt}, 4, 4);
    auto const b = I::choose({{0.0, 10.0}, {0.0, 1.0}, {0.0, 3.0}}, {kCont, kBin, kInt}, 4, 4);
    CHECK(a.kinds == b.kinds && a.vars == b.vars && a.widths == b.widths);
  }
  g_case = "integer_domain_size does not underflow on a sub-box with no integer point (:182)";
  CHECK(I::domain({0.0, 3.0}) == 4);
  CHECK(I::domain({2.5, 2.9}) == 0);
  g_case = "a sub-box with no integer point is enumerable, not partitioned forever (:194)";
  {
    auto const a = I::choose({{2.5, 2.9}}, {kInt}, 4, 4);
    CHECK(a.kinds.size() == 1 && a.kinds[0] == sE);
  }
  g_case = "choose reports the number of slots it filled, per box (:366)";
  CHECK(I::choose({{0, 1}, {0, 1}, {0, 1}}, {kBin, kBin, kBin}, 4, 4).kinds.size() == 3);
  CHECK(I::choose({{0, 1}, {1, 1}, {0, 1}}, {kBin, kBin, kBin}, 4, 4).kinds.size() == 2);
  CHECK(I::choose({{0, 0}, {1, 1}, {0, 0}}, {kBin, kBin, kBin}, 4, 4).kinds.size() == 0);
  g_case = "max_cycle_size caps the slots a policy fills (:386)";
  {
    auto const a = I::choose({{0, 1}, {0, 1}, {0, 1}, {0, 1}, {0, 1}}, std::vector<int>(5, kBin), 4, 4, 3);
    CHECK(a.kinds.size() == 3);
    CHECK(a.vars == std::vector<std::size_t>({0, 1, 2}));
  }
  g_case = "ties among equal widths keep std::sort's order of the ascending variable list";
  {
    // 20 equal-width continuous variables: more than std::sort's insertion-sort
    // threshold, so the engine and the oracle must both call std::sort itself.
    std::vector<Box> box(20, Box {0.0, 1.0});
    box[7] = {0.0, 3.0};
    box[13] = {0.0, 3.0};
    auto const e = Engine::choose(box, std::vector<int>(20, kCont), 2, 2, 6);
    auto const o = Oracle::choose(box, std::vector<int>(20, kCont), 2, 2, 6);
    CHECK(e.vars == o.vars);
    CHECK(e.vars.size() == 6 && e.vars[0] == 7 && e.vars[1] == 13);
  }
}

// ---- heap order: the node order the search reads -------------------------------------------

template<typename I>
void order_cases()
{
  g_case = "dequeue order: least lb, then deeper, then larger pidx (frontier.hpp:66-89)";
  {
    auto q = I::queue(2);
    I::push(q, 3.0, 1, 1);
    I::push(q, 1.0, 2, 5);
    I::push(q, 1.0, 4, 2);
    I::push(q, 1.0, 4, 9);
    I::push(q, -1.0, 0, 0);
    CHECK(feq(I::pop(q), -1.0));
    // the three lb=1 nodes: depth 4 pidx 9, depth 4 pidx 2, depth 2 pidx 5
    std::vector<std::pair<std::size_t, std::size_t>> seen;
    I::compact(q, 100, [](double, std::size_t, std::size_t) {});
    for (int i = 0; i < 3; ++i) {
      seen.clear();
    }
    CHECK(feq(I::pop(q), 1.0) && feq(I::pop(q), 1.0) && feq(I::pop(q), 1.0) && feq(I::pop(q), 3.0));
  }
}

/// The engine's and the oracle's heaps must leave equal-key nodes in the same
/// places: push and pop a long pseudo-random sequence with many ties and
/// compactions, and compare what comes out (depth and pidx included).
void heap_agreement()
{
  g_case = "engine and oracle heaps agree on a long tied sequence with compactions";
  auto eq = Engine::queue(3);
  auto oq = Oracle::queue(3);
  std::uint64_t s = 12345;
  auto rnd = [&] {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return s >> 33;
  };
  bool same = true;
  for (int step = 0; step < 20000 && same; ++step) {
    std::uint64_t const r = rnd() % 10;
    if (r < 6 || Engine::size(eq) == 0) {
      double const lb = static_cast<double>(rnd() % 7);
      std::size_t const depth = rnd() % 4, pidx = rnd() % 5;
      Engine::push(eq, lb, depth, pidx);
      Oracle::push(oq, lb, depth, pidx);
    } else if (r < 9) {
      auto const a = eq.dequeue();
      auto const b = oq.pop();
      same = feq(a.lb, b.lb) && a.depth == b.depth && a.pidx == b.pidx;
    } else {
      std::size_t const keep = 1 + rnd() % 50;
      std::vector<double> ea, oa;
      Engine::compact(eq, keep, [&](double lb, std::size_t d, std::size_t p) { ea.push_back(lb * 100 + d * 10 + p); });
      Oracle::compact(oq, keep, [&](double lb, std::size_t d, std::size_t p) { oa.push_back(lb * 100 + d * 10 + p); });
      same = ea == oa && Engine::capacity(eq) == Oracle::capacity(oq);
    }
  }
  CHECK(same);
}

// ---- composition_policy_test.cpp: BisectionBudget -------------------------------------

std::size_t product(const std::vector<std::size_t>& w)
{
  std::size_t p = 1;
  for (std::size_t x : w) {
    p *= x;
  }
  return p;
}

/// Width of variable v after the assignment (unchanged if it has no slot).
template<typename B>
double width_after(const B& a, const std::vector<Box>& box, std::size_t v)
{
  for (std::size_t j = 0; j < a.vars.size(); ++j) {
    if (a.vars[j] == v) {
      return (box[v].ub - box[v].lb) / static_cast<double>(a.widths[j]);
    }
  }
  return box[v].ub - box[v].lb;
}

template<typename I>
void bisection_cases()
{
  g_case = "a lopsided box spends the whole budget on the wider variable (composition_policy_test.cpp:481)";
  {
    auto const a = I::bisect({{0.0, 1e6}, {0.0, 1.0}}, {kCont, kCont}, 10);
    CHECK(a.vars.size() == 1 && a.vars[0] == 0 && a.widths[0] == 1024);
  }
  g_case = "equal-width variables split the budget evenly (:497)";
  {
    auto const a = I::bisect({{0.0, 1000.0}, {0.0, 1000.0}}, {kCont, kCont}, 10);
    CHECK(a.widths == std::vector<std::size_t>({32, 32}));
  }
  g_case = "the product of widths equals n_regions for every assignment (:512)";
  {
    struct Case
    {
      std::vector<Box> box;
      std::vector<int> kinds;
    };
    std::vector<Case> const cases = {
        {{{0.0, 5.0}}, {kCont}},
        {{{0.0, 1.0}, {0.0, 1.0}, {0.0, 1.0}}, {kBin, kBin, kBin}},
        {{{0.0, 100.0}, {0.0, 3.0}, {-5.0, 5.0}}, {kInt, kInt, kCont}},
        {{{0.0, 16.0}}, {kInt}},
    };
    for (const Case& c : cases) {
      auto const a = I::bisect(c.box, c.kinds, 8);
      CHECK(a.vars.empty() || product(a.widths) == 256);
    }
  }
  g_case = "the greedy allocation is optimal for the min-max width (:547)";
  {
    std::vector<Box> box {{0.0, 100.0}, {0.0, 30.0}, {0.0, 7.0}};
    double best = 1e300;
    for (int b0 = 0; b0 <= 6; ++b0) {
      for (int b1 = 0; b0 + b1 <= 6; ++b1) {
        int const b2 = 6 - b0 - b1;
        best = std::min(best, std::max({100.0 / (1 << b0), 30.0 / (1 << b1), 7.0 / (1 << b2)}));
      }
    }
    auto const a = I::bisect(box, {kCont, kCont, kCont}, 6);
    double worst = 0.0;
    for (std::size_t v = 0; v < 3; ++v) {
      worst = std::max(worst, width_after(a, box, v));
    }
    CHECK(feq(worst, best));
  }
  g_case = "the allocation does not depend on the order variables appear in (:580)";
  {
    std::vector<Box> ba {{0.0, 1000.0}, {0.0, 10.0}}, bb {{0.0, 10.0}, {0.0, 1000.0}};
    auto const a = I::bisect(ba, {kCont, kCont}, 8);
    auto const b = I::bisect(bb, {kCont, kCont}, 8);
    CHECK(feq(width_after(a, ba, 0), width_after(b, bb, 1)) && feq(width_after(a, ba, 1), width_after(b, bb, 0)));
  }
  g_case = "an integer domain too wide for the budget partitions instead (:601)";
  {
    auto const a = I::bisect({{0.0, 1048575.0}}, {kInt}, 5);
    CHECK(a.kinds.size() == 1 && a.kinds[0] == sP && a.widths[0] == 32);
  }
  g_case = "the cheap integer enumerates and the wide one partitions (:615)";
  {
    auto const a = I::bisect({{0.0, 3.0}, {0.0, 1048575.0}}, {kInt, kInt}, 5);
    CHECK(a.kinds.size() == 2 && product(a.widths) == 32);
    for (std::size_t j = 0; j < a.vars.size(); ++j) {
      if (a.vars[j] == 0) {
        CHECK(a.kinds[j] == sE && a.widths[j] == 4 && a.domains[j] == 4);
      } else {
        CHECK(a.kinds[j] == sP);
      }
    }
  }
  g_case = "a domain of 5 enumerates over 8 children, 3 of them duplicates (:640)";
  {
    auto const a = I::bisect({{0.0, 4.0}}, {kInt}, 3);
    CHECK(a.kinds.size() == 1 && a.widths[0] == 8 && a.domains[0] == 5);
  }
  g_case = "leftover budget with no bisecting pool pads the widest enumerate slot (:672)";
  {
    auto const a = I::bisect({{0.0, 1.0}, {0.0, 1.0}}, {kBin, kBin}, 5);
    CHECK(a.kinds.size() == 2 && product(a.widths) == 32);
    CHECK(a.kinds == std::vector<int>({sB, sB}));
    CHECK(a.widths == std::vector<std::size_t>({16, 2}));  // the first of two equal slots takes the padding
  }
  g_case = "relative width: equal relative shares whatever the absolute widths";
  {
    auto const a = I::bisect({{0.0, 1e6}, {0.0, 1.0}}, {kCont, kCont}, 10, true);
    CHECK(a.widths == std::vector<std::size_t>({32, 32}));
  }
  g_case = "an enumeration ceiling turns a wide integer into a partition";
  {
    auto const a = I::bisect({{0.0, 40.0}, {0.0, 3.0}}, {kInt, kInt}, 8, false, 8);
    CHECK(a.kinds.size() == 2 && a.vars[0] == 1 && a.kinds[0] == sE && a.kinds[1] == sP);
    CHECK(product(a.widths) == 256);
  }
}

template<typename I>
void widthfirst_cases()
{
  // policy/width_first.hpp, restated as properties (the reference has no
  // dedicated tests for this policy)
  g_case = "width-first: every live variable, widest first, partition widths";
  {
    auto const a = I::widthfirst({{0.0, 1.0}, {0.0, 100.0}, {2.0, 2.0}, {0.0, 3.0}}, {kCont, kCont, kCont, kCont}, 5, 3);
    CHECK(a.vars == std::vector<std::size_t>({1, 3, 0}));
    CHECK(a.widths == std::vector<std::size_t>({5, 5, 5}));
    CHECK(a.kinds == std::vector<int>({sC, sC, sC}));
  }
  g_case = "width-first: the slot cap keeps the widest";
  {
    auto const a = I::widthfirst({{0.0, 1.0}, {0.0, 100.0}, {0.0, 3.0}}, {kCont, kCont, kCont}, 4, 4, 2);
    CHECK(a.vars == std::vector<std::size_t>({1, 2}));
  }
  g_case = "width-first: a binary waits while another live variable is wider than 1";
  {
    auto const a = I::widthfirst({{0.0, 1.0}, {0.0, 5.0}}, {kBin, kCont}, 4, 4);
    CHECK(a.vars == std::vector<std::size_t>({1}) && a.kinds == std::vector<int>({sC}));
    auto const b = I::widthfirst({{0.0, 1.0}, {0.0, 0.5}}, {kBin, kCont}, 4, 4);
    CHECK(b.vars == std::vector<std::size_t>({0, 1}) && b.kinds == std::vector<int>({sB, sC}));
    CHECK(b.widths == std::vector<std::size_t>({2, 4}));
  }
  g_case = "width-first: the unique widest integer enumerates when its domain fits the cap";
  {
    auto const a = I::widthfirst({{0.0, 3.0}, {0.0, 1.0}}, {kInt, kCont}, 4, 4);
    CHECK(a.kinds == std::vector<int>({sE, sC}) && a.widths == std::vector<std::size_t>({4, 4}));
    auto const wide = I::widthfirst({{0.0, 9.0}, {0.0, 1.0}}, {kInt, kCont}, 4, 4);
    CHECK(wide.kinds == std::vector<int>({sP, sC}));
    auto const tie = I::widthfirst({{0.0, 3.0}, {5.0, 8.0}}, {kInt, kInt}, 4, 4);
    CHECK(tie.kinds == std::vector<int>({sP, sP}));
    auto const not_widest = I::widthfirst({{0.0, 3.0}, {0.0, 10.0}}, {kInt, kCont}, 4, 4);
    CHECK(not_widest.vars == std::vector<std::size_t>({1, 0}) && not_widest.kinds == std::vector<int>({sC, sP}));
  }
  g_case = "width-first: a binary joins beside an integer that enumerates this iteration";
  {
    auto const a = I::widthfirst({{0.0, 1.0}, {0.0, 3.0}}, {kBin, kInt}, 4, 4);
    CHECK(a.vars == std::vector<std::size_t>({1, 0}) && a.kinds == std::vector<int>({sE, sB}));
    auto const b = I::widthfirst({{0.0, 1.0}, {0.0, 30.0}}, {kBin, kInt}, 4, 4);
    CHECK(b.vars == std::vector<std::size_t>({1}) && b.kinds == std::vector<int>({sP}));
  }
  g_case = "width-first: a resolved box has no slots";
  {
    auto const a = I::widthfirst({{1.0, 1.0}, {0.0, 0.0}}, {kCont, kBin}, 4, 4);
    CHECK(a.vars.empty());
  }
}

template<typename I>
void run_all()
{
  g_impl = I::name;
  compaction_cases<I>();
  eviction_cases<I>();
  history_cases<I>();
  epilogue_cases<I>();
  budget_cases<I>();
  witness_cases<I>();
  policy_cases<I>();
  bisection_cases<I>();
  widthfirst_cases<I>();
  order_cases<I>();
}

// ---- engine-only cases: region helpers and spellings -------------------------------------

void engine_only()
{
  g_impl = "engine";
  using cmi::region::Composition;
  using cmi::region::FanOutSpec;
  using cmi::region::SlotKind;

  g_case = "every stop reason has the spelling the log contract names (host_budget_test.cpp:689)";
  using cmi::search::StopReason;
  CHECK(std::string(to_string(StopReason::Converged)) == "converged");
  CHECK(std::string(to_string(StopReason::Exhausted)) == "exhausted");
  CHECK(std::string(to_string(StopReason::IterationLimit)) == "iteration-limit");
  CHECK(std::string(to_string(StopReason::HostMemory)) == "host-memory");
  CHECK(std::string(to_string(StopReason::AllocationFailure)) == "allocation-failure");
  CHECK(std::string(to_string(StopReason::DeviceMemory)) == "device-memory");

  g_case = "can_fathom_without_children at live_count == slot count (composition_policy_test.cpp:205)";
  {
    Composition c {{SlotKind::BinaryEnumerate, SlotKind::BinaryEnumerate, SlotKind::BinaryEnumerate}};
    CHECK(cmi::region::can_fathom_without_children(3, c));
  }
  g_case = "can_fathom is false when live variables outnumber the slots (:216)";
  {
    Composition c {{SlotKind::BinaryEnumerate, SlotKind::BinaryEnumerate}};
    CHECK(cmi::region::is_fully_enumerable(c));
    CHECK(!cmi::region::can_fathom_without_children(3, c));
    CHECK(cmi::region::can_fathom_without_children(2, c));
  }
  g_case = "can_fathom at live_count == 0 (a fully resolved box) (:229)";
  CHECK(cmi::region::can_fathom_without_children(0, Composition {}));
  g_case = "can_fathom is false when a live Continuous slot remains (:241)";
  {
    Composition c {{SlotKind::BinaryEnumerate, SlotKind::Continuous}};
    CHECK(!cmi::region::can_fathom_without_children(2, c));
  }
  g_case = "FanOutSpec rejects a partition_num below 2 (:252)";
  {
    int thrown = 0;
    for (std::size_t p : {std::size_t(1), std::size_t(0)}) {
      try {
        FanOutSpec f {p};
        (void)f;
      } catch (const cmi::InvalidConfiguration&) {
        ++thrown;
      }
    }
    CHECK(thrown == 2);
    bool ok = true;
    try {
      FanOutSpec f {2};
      (void)f;
    } catch (...) {
      ok = false;
    }
    CHECK(ok);
  }
  g_case = "FanOutSpec defaults enumerate_cap to partition_num (:261)";
  {
    FanOutSpec const f {7};
    CHECK(f.partition_num() == 7 && f.enumerate_cap() == 7);
  }
  g_case = "composition_fan_out saturates instead of wrapping (:268)";
  {
    Composition huge {std::vector<SlotKind>(10, SlotKind::Continuous)};
    CHECK(cmi::region::composition_fan_out(huge, FanOutSpec {10000000}) == std::numeric_limits<std::size_t>::max());
    Composition small {std::vector<SlotKind>(3, SlotKind::Continuous)};
    CHECK(cmi::region::composition_fan_out(small, FanOutSpec {4}) == 64);
  }
  g_case = "slot_prefixes agrees with repeated-division digits (:280)";
  {
    std::vector<std::size_t> const widths {2, 3, 4};
    auto const prefix = cmi::region::slot_prefixes(widths);
    bool all = prefix.size() == 3;
    for (std::size_t r = 0; r < 24 && all; ++r) {
      std::size_t idx = r;
      for (std::size_t j = 0; j < 3; ++j) {
        all = all && (r / prefix[j]) % widths[j] == idx % widths[j];
        idx /= widths[j];
      }
    }
    CHECK(all);
  }
  g_case = "a policy's fan_out is what its callers decode against (:305)";
  {
    cmi::policy::GreedyEnumCompositionPolicy pol {FanOutSpec {4, 50}};
    CHECK(pol.fan_out().partition_num() == 4 && pol.fan_out().enumerate_cap() == 50);
  }
  g_case = "a policy cannot be capped above kMaxSlots (:400)";
  {
    bool thrown = false;
    try {
      cmi::policy::SearchCalibration cal;
      cal.max_cycle_size = cmi::policy::kMaxSlots + 1;
      cmi::policy::GreedyEnumCompositionPolicy pol {FanOutSpec {4}, cal};
    } catch (const cmi::InvalidConfiguration&) {
      thrown = true;
    }
    CHECK(thrown);
    cmi::policy::GreedyEnumCompositionPolicy pol {FanOutSpec {4}};
    CHECK(pol.max_cycle_size() == cmi::policy::kMaxSlots);
  }
  g_case = "choose throws ShapeMismatch on an empty var_kinds span (:165)";
  {
    bool thrown = false;
    try {
      cmi::policy::GreedyEnumCompositionPolicy pol {FanOutSpec {4}};
      pol.choose({}, {});
    } catch (const cmi::ShapeMismatch&) {
      thrown = true;
    }
    CHECK(thrown);
  }
  g_case = "GreedyEnumerate always satisfies the distinct-and-live var_ids contract (:416)";
  {
    cmi::policy::GreedyEnumCompositionPolicy pol {FanOutSpec {4}};
    std::vector<cmi::interval> box {{0.0, 10.0}, {0.0, 1.0}, {0.0, 3.0}, {5.0, 5.0}};
    std::vector<cmi::model::VarKind> kinds {cmi::model::VarKind::Continuous, cmi::model::VarKind::Binary,
                                            cmi::model::VarKind::Integer, cmi::model::VarKind::Continuous};
    CHECK(cmi::policy::assignment_is_distinct_and_live(pol.choose(box, kinds), box));
  }
  g_case = "assignment_is_distinct_and_live rejects a duplicate var_id (:433)";
  {
    cmi::region::SlotAssignment bad;
    bad.composition.kinds = {SlotKind::BinaryEnumerate, SlotKind::BinaryEnumerate};
    bad.var_ids = {0, 0};
    std::vector<cmi::interval> box {{0.0, 1.0}, {0.0, 1.0}};
    CHECK(!cmi::policy::assignment_is_distinct_and_live(bad, box));
  }
  g_case = "assignment_is_distinct_and_live rejects a resolved var_id (:446)";
  {
    cmi::region::SlotAssignment bad;
    bad.composition.kinds = {SlotKind::BinaryEnumerate, SlotKind::BinaryEnumerate};
    bad.var_ids = {0, 1};
    std::vector<cmi::interval> box {{0.0, 1.0}, {5.0, 5.0}};
    CHECK(!cmi::policy::assignment_is_distinct_and_live(bad, box));
  }
  g_case = "BisectionBudget rejects B outside [1, 62] (composition_policy_test.cpp:466)";
  {
    int thrown = 0;
    for (std::size_t b : {std::size_t(0), std::size_t(63)}) {
      try {
        cmi::policy::BisectionBudgetCompositionPolicy pol {b};
      } catch (const cmi::InvalidConfiguration&) {
        ++thrown;
      }
    }
    CHECK(thrown == 2);
    cmi::policy::BisectionBudgetCompositionPolicy one {1}, top {62};
    CHECK(one.bisection_budget() == 1 && top.bisection_budget() == 62);
  }
  g_case = "n_regions is 2^B whatever the composition, except the empty one (:473)";
  {
    cmi::policy::BisectionBudgetCompositionPolicy pol {10};
    CHECK(pol.n_regions(Composition {{SlotKind::Continuous}}) == 1024);
    CHECK(pol.n_regions(Composition {{SlotKind::Continuous, SlotKind::Continuous}}) == 1024);
    CHECK(pol.n_regions(Composition {}) == 1);
  }
  g_case = "duplicate children are flagged and decode like their top sibling (:640)";
  {
    cmi::policy::BisectionBudgetCompositionPolicy pol {3};
    std::vector<cmi::interval> box {{0.0, 4.0}};
    std::vector<cmi::model::VarKind> kinds {cmi::model::VarKind::Integer};
    auto const a = pol.choose(box, kinds);
    CHECK(a.widths.size() == 1 && a.widths[0] == 8 && a.domain_sizes[0] == 5);
    auto const prefix = cmi::region::slot_prefixes(a.widths);
    std::vector<cmi::interval> top, child;
    cmi::region::materialise(box, 4, a, top);
    bool ok = true;
    for (std::size_t r = 0; r < 8; ++r) {
      bool const dup = r >= 5;
      ok = ok && cmi::region::is_duplicate_child(r, a, prefix) == dup;
      if (dup) {
        cmi::region::materialise(box, r, a, child);
        ok = ok && feq(child[0].lb, top[0].lb) && feq(child[0].ub, top[0].ub);
      }
    }
    CHECK(ok);
  }
  g_case = "materialise rejects a policy that is not a pure function of the box (frontier.hpp:113-144)";
  {
    cmi::search::IntervalHistory h;
    h.enqueue({});
    std::size_t const p = h.enqueue({{0.0, 4.0}});
    cmi::policy::GreedyEnumCompositionPolicy pol {FanOutSpec {2}};
    std::vector<cmi::model::VarKind> kinds {cmi::model::VarKind::Continuous};
    std::vector<cmi::interval> root {{0.0, 4.0}}, out;
    cmi::search::Node good {1, p, 1, 0.0, cmi::region::assignment_hash(pol.choose(h.box(p), kinds))};
    good.materialise(h, out, pol, kinds, root);
    CHECK(out.size() == 1 && feq(out[0].lb, 2.0) && feq(out[0].ub, 4.0));
    cmi::search::Node bad = good;
    bad.assignment_hash ^= 1;
    bool thrown = false;
    try {
      bad.materialise(h, out, pol, kinds, root);
    } catch (const cmi::InvalidConfiguration&) {
      thrown = true;
    }
    CHECK(thrown);
  }
}

}  // namespace

/// The two implementations choose alike on random boxes with many equal
/// widths (std::sort's arrangement of ties included).
void policy_agreement()
{
  g_impl = "both";
  g_case = "engine and oracle width-first and greedy choices agree on random tie-heavy boxes";
  std::uint64_t x = 12345;
  auto next = [&] {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    return x >> 33;
  };
  bool same = true;
  for (int t = 0; t < 3000; ++t) {
    std::size_t const n = 1 + next() % 12;
    std::vector<Box> box(n);
    std::vector<int> kinds(n);
    for (std::size_t v = 0; v < n; ++v) {
      kinds[v] = static_cast<int>(next() % 3);
      double const lo = static_cast<double>(next() % 5);
      double const w = kinds[v] == kBin ? static_cast<double>(next() % 2) : static_cast<double>(next() % 4);
      box[v] = {lo, lo + w};
    }
    std::size_t const slots = next() % 5;
    auto const a = Engine::widthfirst(box, kinds, 3, 4, slots);
    auto const b = Oracle::widthfirst(box, kinds, 3, 4, slots);
    same = same && a.vars == b.vars && a.kinds == b.kinds && a.widths == b.widths;
    auto const c = Engine::choose(box, kinds, 3, 4, slots);
    auto const d = Oracle::choose(box, kinds, 3, 4, slots);
    same = same && c.vars == d.vars && c.kinds == d.kinds && c.widths == d.widths;
  }
  CHECK(same);
}

int main()
{
  run_all<Engine>();
  run_all<Oracle>();
  heap_agreement();
  policy_agreement();
  engine_only();
  std::printf("%d checks, %d failed\n", g_checks, g_failed);
  return g_failed == 0 ? 0 : 1;
}
