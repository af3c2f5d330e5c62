// The reference's host tests of the k-child search restated case by case
// (test/source/aggregate_policy_test.cpp, test/source/aggregate_frontier_test.cpp)
// and run against both implementations: the engine's
// (cuminlpsearch_amd/csrc/cmi/aggregate_search.hpp) and the oracle's
// (oracle/search_core.hpp). Boxes are decoded with the engine's region
// decode, which its own known-answer tests pin, for both.
//
// Built and run by tests/test_aggregate_pins_cpu.py; prints one line per
// failed check and exits non-zero if any failed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <set>
#include <tuple>
#include <vector>

#include "cmi/aggregate_search.hpp"
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

bool feq(double a, double b) { return !(a < b) && !(b < a); }

using Box = std::vector<cmi::interval>;
using Kinds = std::vector<cmi::model::VarKind>;
constexpr double kNone = std::numeric_limits<double>::max();

/// A partition in one vocabulary: slot kinds, variables and widths of the
/// three assignments.
struct Part
{
  cmi::region::SlotAssignment branch, refine, launch;
  std::size_t k = 0, m = 0;
  bool nested = false;
  bool leaf() const { return branch.size() == 0; }
};

cmi::region::SlotAssignment from_oracle(const orc::Assign& a)
{
  cmi::region::SlotAssignment s;
  for (std::size_t j = 0; j < a.vars.size(); ++j) {
    s.composition.kinds.push_back(static_cast<cmi::region::SlotKind>(a.kinds[j]));
    s.var_ids.push_back(a.vars[j]);
    s.widths.push_back(a.widths[j]);
    s.domain_sizes.push_back(a.domains.empty() ? a.widths[j] : a.domains[j]);
  }
  return s;
}

struct Engine
{
  static constexpr const char* name = "engine";
  struct Policy
  {
    cmi::aggregate::WidestBranchPolicy p;
  };
  static Policy policy(std::size_t k, std::size_t n) { return {cmi::aggregate::WidestBranchPolicy(k, n)}; }
  static Part partition(const Policy& pol, const Box& box, const Kinds& kinds)
  {
    auto const a = pol.p.partition(box, kinds, 0);
    return {a.branch, a.refine, a.launch, a.k, a.m, a.nested};
  }
  static cmi::region::SlotAssignment branch(const Policy& pol, const Box& box, const Kinds& kinds)
  {
    return pol.p.branch(box, kinds);
  }
  static bool rejects(std::size_t k, std::size_t n)
  {
    try {
      cmi::aggregate::WidestBranchPolicy(k, n);
    } catch (const cmi::InvalidConfiguration&) {
      return true;
    }
    return false;
  }

  struct Front
  {
    cmi::aggregate::AggregateFrontier f;
  };
  static Front front(int rule) { return {cmi::aggregate::AggregateFrontier(static_cast<cmi::aggregate::SelectionRule>(rule))}; }
  static void push(Front& q, std::size_t sidx, std::size_t pidx, std::size_t depth, double lb, double hub, double f)
  {
    q.f.enqueue(cmi::aggregate::AggregateNode {sidx, pidx, depth, lb, hub, 0}, f);
  }
  struct Popped
  {
    double lb;
    std::size_t depth, pidx, sidx;
  };
  static Popped pop(Front& q)
  {
    auto const n = q.f.pop();
    return {n.lb, n.depth, n.pidx, n.sidx};
  }
  static double min_lb(Front& q) { return q.f.min_lb(); }
  static void incumbent(Front& q, double f) { q.f.on_incumbent(f); }
  static std::size_t size(const Front& q) { return q.f.size(); }
  static bool empty(const Front& q) { return q.f.empty(); }
  static std::size_t viable(const Front& q, double gub) { return q.f.count_viable(gub); }
  static std::size_t bytes(const Front& q) { return q.f.capacity_bytes(); }
  static double key(int rule, double lb, double hub, double f)
  {
    return cmi::aggregate::selection_key(static_cast<cmi::aggregate::SelectionRule>(rule), lb, hub, f);
  }

  using Queue = cmi::search::IntervalPQueue;
  static Queue queue() { return Queue(16); }
  static void qpush(Queue& q, std::size_t sidx, std::size_t pidx, std::size_t depth, double lb)
  {
    q.enqueue(cmi::search::Node {sidx, pidx, depth, lb, 0});
  }
  static Popped qpop(Queue& q)
  {
    auto const n = q.dequeue();
    return {n.lb, n.depth, n.pidx, n.sidx};
  }
  static double qpeek(const Queue& q) { return q.peek().lb; }
  static std::size_t qsize(const Queue& q) { return q.size(); }
};

struct Oracle
{
  static constexpr const char* name = "oracle";
  struct Policy
  {
    std::size_t k, n;
  };
  static Policy policy(std::size_t k, std::size_t n) { return {k, n}; }
  static std::vector<orc::Iv> ivs(const Box& box)
  {
    std::vector<orc::Iv> b;
    for (auto const& x : box) {
      b.push_back({x.lb, x.ub});
    }
    return b;
  }
  static std::vector<std::uint8_t> vks(const Kinds& kinds)
  {
    std::vector<std::uint8_t> k;
    for (auto v : kinds) {
      k.push_back(static_cast<std::uint8_t>(v));
    }
    return k;
  }
  static Part partition(const Policy& pol, const Box& box, const Kinds& kinds)
  {
    auto const a = orc::agg_partition(ivs(box), vks(kinds), pol.k, pol.n, false);
    return {from_oracle(a.branch), from_oracle(a.refine), from_oracle(a.launch), pol.k, pol.n / pol.k, a.nested};
  }
  static cmi::region::SlotAssignment branch(const Policy& pol, const Box& box, const Kinds& kinds)
  {
    return from_oracle(orc::widest_split(ivs(box), vks(kinds), pol.k));
  }
  static bool rejects(std::size_t k, std::size_t n)
  {
    auto pow2 = [](std::size_t v) { return v != 0 && (v & (v - 1)) == 0; };
    return k < 2 || !pow2(k) || n <= k || !pow2(n);  // orc_aggregate_solve's option check
  }

  struct Front
  {
    orc::AggFront f;
  };
  static Front front(int rule)
  {
    Front q;
    q.f.rule = rule;
    return q;
  }
  static void push(Front& q, std::size_t sidx, std::size_t pidx, std::size_t depth, double lb, double hub, double f)
  {
    q.f.push(orc::AggNode {sidx, pidx, depth, lb, hub, 0}, f);
  }
  using Popped = Engine::Popped;
  static Popped pop(Front& q)
  {
    auto const n = q.f.pop();
    return {n.lb, n.depth, n.pidx, n.sidx};
  }
  static double min_lb(Front& q) { return q.f.min_lb(); }
  static void incumbent(Front& q, double f) { q.f.incumbent(f); }
  static std::size_t size(const Front& q) { return q.f.live; }
  static bool empty(const Front& q) { return q.f.live == 0; }
  static std::size_t viable(const Front& q, double gub) { return q.f.viable(gub); }
  static std::size_t bytes(const Front& q) { return q.f.cap_bytes(); }
  static double key(int rule, double lb, double hub, double f) { return orc::agg_key(rule, lb, hub, f); }

  using Queue = orc::Heap;
  static Queue queue() { return Queue(16); }
  static void qpush(Queue& q, std::size_t sidx, std::size_t pidx, std::size_t depth, double lb)
  {
    q.push(orc::Node {sidx, pidx, depth, lb, 0});
  }
  static Popped qpop(Queue& q)
  {
    auto const n = q.pop();
    return {n.lb, n.depth, n.pidx, n.sidx};
  }
  static double qpeek(const Queue& q) { return q.e[0].lb; }
  static std::size_t qsize(const Queue& q) { return q.n; }
};

constexpr int kLLB = CMI_SELECT_LEAST_LOWER_BOUND, kRI = CMI_SELECT_REJECT_INDEX,
              kRIS = CMI_SELECT_REJECT_INDEX_STALE;
using VK = cmi::model::VarKind;

bool contains(const Box& outer, const Box& inner)
{
  if (outer.size() != inner.size()) {
    return false;
  }
  for (std::size_t i = 0; i < outer.size(); ++i) {
    if (inner[i].lb < outer[i].lb || inner[i].ub > outer[i].ub) {
      return false;
    }
  }
  return true;
}

std::size_t live_of(const Box& box)
{
  std::size_t n = 0;
  for (auto const& b : box) {
    n += b.ub > b.lb ? 1 : 0;
  }
  return n;
}

/// Random (box, kinds): 1..8 variables, mixed kinds, one in four resolved,
/// widths over seven decades (aggregate_policy_test.cpp:73-112).
struct Boxes
{
  std::mt19937 rng {20260811};
  std::pair<Box, Kinds> next()
  {
    std::uniform_int_distribution<std::size_t> nv(1, 8);
    std::uniform_int_distribution<int> kind(0, 2), dead(0, 3);
    std::uniform_real_distribution<double> mag(-3.0, 4.0), lo(-50.0, 50.0);
    std::size_t const n = nv(rng);
    Box box(n);
    Kinds kinds(n);
    for (std::size_t i = 0; i < n; ++i) {
      int const k = kind(rng);
      kinds[i] = k == 0 ? VK::Continuous : k == 1 ? VK::Integer : VK::Binary;
      if (kinds[i] == VK::Binary) {
        box[i] = {0.0, dead(rng) == 0 ? 0.0 : 1.0};
        continue;
      }
      double l = lo(rng);
      double w = std::pow(10.0, mag(rng));
      if (kinds[i] == VK::Integer) {
        l = std::round(l);
        w = std::max(1.0, std::round(w));
      }
      box[i] = {l, dead(rng) == 0 ? l : l + w};
    }
    return {box, kinds};
  }
};

// ---- aggregate_policy_test.cpp --------------------------------------------------

template<typename I>
void partition_invariants()
{
  g_case = "a partition satisfies every invariant it claims";
  auto const pol = I::policy(4, 64);
  Boxes gen;
  std::size_t leaves = 0, nested = 0, disjoint = 0;
  for (int t = 0; t < 2000; ++t) {
    auto const [box, kinds] = gen.next();
    Part const p = I::partition(pol, box, kinds);
    if (p.leaf()) {
      CHECK(live_of(box) == 0);
      CHECK(p.launch.size() == 0);
      ++leaves;
      continue;
    }
    CHECK(cmi::region::composition_fan_out(p.branch.widths) == 4);
    CHECK(cmi::region::composition_fan_out(p.launch.widths) == 64);
    if (p.nested) {
      ++nested;
      CHECK(live_of(box) == 1);
      CHECK(p.launch.size() == 1);
      CHECK(p.refine.size() == 0);
    } else {
      ++disjoint;
      CHECK(live_of(box) >= 2);
      CHECK(cmi::region::composition_fan_out(p.refine.widths) == 16);
      CHECK(p.launch.size() == p.branch.size() + p.refine.size());
      CHECK(std::find(p.refine.var_ids.begin(), p.refine.var_ids.end(), p.branch.var_ids[0])
            == p.refine.var_ids.end());
      CHECK(p.launch.var_ids.back() == p.branch.var_ids[0]);  // the most significant slot
    }
  }
  CHECK(leaves > 0);
  CHECK(nested > 0);
  CHECK(disjoint > 0);
}

template<typename I>
void subregions_inside_their_child()
{
  g_case = "every launch subregion sits inside the child r / M assigns it to";
  auto const pol = I::policy(4, 64);
  Boxes gen;
  std::size_t checked = 0;
  for (int t = 0; t < 1200; ++t) {
    auto const [box, kinds] = gen.next();
    Part const p = I::partition(pol, box, kinds);
    if (p.leaf()) {
      continue;
    }
    std::vector<Box> child(p.k);
    bool ok = true;
    for (std::size_t c = 0; c < p.k; ++c) {
      cmi::region::materialise(box, c, p.branch, child[c]);
      ok = ok && contains(box, child[c]);
    }
    for (std::size_t r = 0; r < p.k * p.m; ++r) {
      Box sub;
      cmi::region::materialise(box, r, p.launch, sub);
      std::size_t const c = cmi::aggregate::child_of(r, p.m);
      ok = ok && c < p.k && contains(child[c], sub);
      ++checked;
    }
    CHECK(ok);
  }
  CHECK(checked > 50000);
}

template<typename I>
void branch_is_pure()
{
  g_case = "branch is a pure function of the box";
  auto const pol = I::policy(4, 1024);
  Boxes gen;
  for (int t = 0; t < 500; ++t) {
    auto const [box, kinds] = gen.next();
    CHECK(cmi::region::assignment_hash(I::branch(pol, box, kinds))
          == cmi::region::assignment_hash(I::branch(pol, box, kinds)));
  }
}

template<typename I>
void widest_uniform_branch()
{
  g_case = "the branch splits the widest live variable k ways, uniformly";
  auto const pol = I::policy(4, 1024);
  Kinds const kinds {VK::Continuous, VK::Continuous, VK::Integer};
  auto const b = I::branch(pol, Box {{0.0, 1.0}, {0.0, 100.0}, {0.0, 3.0}}, kinds);
  CHECK(b.size() == 1);
  if (b.size() == 1) {
    CHECK(b.var_ids[0] == 1);
    CHECK(b.widths[0] == 4);
    CHECK(b.composition[0] == cmi::region::SlotKind::Continuous);
  }
  auto const d = I::branch(pol, Box {{0.0, 1.0}, {0.0, 2.0}, {0.0, 500.0}}, kinds);
  CHECK(d.size() == 1);
  if (d.size() == 1) {
    CHECK(d.var_ids[0] == 2);
    CHECK(d.composition[0] == cmi::region::SlotKind::IntegerPartition);
  }
  g_case = "equal widths go to the lowest index";
  auto const tie = I::branch(pol, Box {{5.0, 5.0}, {0.0, 2.0}, {1.0, 3.0}}, kinds);
  CHECK(tie.size() == 1 && tie.var_ids[0] == 1);
}

template<typename I>
void resolved_is_leaf()
{
  g_case = "a fully resolved box is a leaf";
  auto const pol = I::policy(4, 1024);
  Part const p = I::partition(pol, Box {{1.0, 1.0}, {2.0, 2.0}}, Kinds {VK::Continuous, VK::Integer});
  CHECK(p.leaf());
  CHECK(p.launch.size() == 0);
}

template<typename I>
void nested_case()
{
  g_case = "the nested case keeps child_id = r / M exact on one variable";
  auto const pol = I::policy(4, 64);
  Box const box {{0.0, 8.0}, {5.0, 5.0}};
  Part const p = I::partition(pol, box, Kinds {VK::Continuous, VK::Continuous});
  CHECK(p.nested);
  CHECK(p.launch.size() == 1 && p.launch.widths[0] == 64);
  for (std::size_t c = 0; c < 4; ++c) {
    Box child, first, last;
    cmi::region::materialise(box, c, p.branch, child);
    cmi::region::materialise(box, c * 16, p.launch, first);
    cmi::region::materialise(box, c * 16 + 15, p.launch, last);
    CHECK(feq(first[0].lb, child[0].lb));
    CHECK(feq(last[0].ub, child[0].ub));
  }
}

template<typename I>
void rejected_fan_outs()
{
  g_case = "fan-outs that cannot be honoured are rejected";
  CHECK(I::rejects(3, 64));
  CHECK(I::rejects(1, 64));
  CHECK(I::rejects(4, 4));
  CHECK(I::rejects(4, 100));
  CHECK(!I::rejects(4, 8));
}

void check_rejects_misordered_launch()
{
  g_impl = "engine";
  g_case = "check() rejects a launch whose branch slot is not the most significant";
  cmi::aggregate::WidestBranchPolicy const pol(4, 64);
  Box const box {{0.0, 8.0}, {0.0, 4.0}};
  auto p = pol.partition(box, Kinds {VK::Continuous, VK::Continuous}, 0);
  CHECK(!p.nested);
  bool fine = true;
  try {
    p.check();
  } catch (const cmi::InvalidConfiguration&) {
    fine = false;
  }
  CHECK(fine);
  std::swap(p.launch.var_ids.front(), p.launch.var_ids.back());
  std::swap(p.launch.widths.front(), p.launch.widths.back());
  std::swap(p.launch.composition.kinds.front(), p.launch.composition.kinds.back());
  CHECK(cmi::region::composition_fan_out(p.launch.widths) == 64);
  bool thrown = false;
  try {
    p.check();
  } catch (const cmi::InvalidConfiguration&) {
    thrown = true;
  }
  CHECK(thrown);
}

// ---- aggregate_frontier_test.cpp -----------------------------------------------

template<typename I>
void lb_rule_matches_plain_queue()
{
  g_case = "least-lower-bound pops in the per-region queue's order";
  std::mt19937 rng {20260811};
  std::uniform_int_distribution<int> act(0, 2);
  std::uniform_int_distribution<std::size_t> depth_d(0, 12), sidx_d(0, 3);
  std::normal_distribution<double> lb_d(0.0, 50.0);
  std::size_t compared = 0, next_pidx = 1;
  bool ok = true;
  for (int seq = 0; seq < 60; ++seq) {
    auto ref = I::queue();
    auto sub = I::front(kLLB);
    std::size_t live = 0;
    std::set<std::tuple<double, std::size_t, std::size_t>> used;
    for (int step = 0; step < 3200; ++step) {
      if (act(rng) != 0 || live == 0) {
        double lb;
        std::size_t depth, pidx;
        do {
          lb = std::round(lb_d(rng) * 100.0) / 100.0;
          depth = depth_d(rng);
          pidx = next_pidx++;
        } while (!used.insert({lb, depth, pidx}).second);
        std::size_t const sidx = sidx_d(rng);
        I::qpush(ref, sidx, pidx, depth, lb);
        I::push(sub, sidx, pidx, depth, lb, lb + 1.0, kNone);
        ++live;
        continue;
      }
      auto const r = I::qpop(ref);
      auto const s = I::pop(sub);
      ok = ok && feq(r.lb, s.lb) && r.depth == s.depth && r.pidx == s.pidx && r.sidx == s.sidx;
      --live;
      ++compared;
      ok = ok && I::size(sub) == I::qsize(ref);
      if (I::qsize(ref) > 0) {
        ok = ok && feq(I::min_lb(sub), I::qpeek(ref));
      }
    }
    while (live > 0) {
      auto const r = I::qpop(ref);
      auto const s = I::pop(sub);
      ok = ok && feq(r.lb, s.lb) && r.depth == s.depth && r.pidx == s.pidx && r.sidx == s.sidx;
      --live;
      ++compared;
    }
    ok = ok && I::empty(sub);
  }
  CHECK(ok);
  CHECK(compared >= 100000);
}

template<typename I>
void min_lb_monotone()
{
  g_case = "min_lb never decreases under any rule";
  for (int rule : {kLLB, kRI, kRIS}) {
    std::mt19937 rng {99};
    auto q = I::front(rule);
    double f = kNone;
    I::push(q, 0, 0, 0, -1000.0, 1000.0, f);
    double prev = I::min_lb(q);
    std::size_t pidx = 1;
    bool ok = true;
    for (int step = 0; step < 4000 && !I::empty(q); ++step) {
      auto const parent = I::pop(q);
      for (std::size_t c = 0; c < 4; ++c) {
        double const lb = parent.lb + std::uniform_real_distribution<double>(0.0, 5.0)(rng);
        double const ub = lb + std::uniform_real_distribution<double>(0.1, 50.0)(rng);
        if (lb > f) {
          continue;
        }
        I::push(q, c, pidx, parent.depth + 1, lb, ub, f);
      }
      ++pidx;
      if (step % 50 == 0 && parent.lb + 20.0 < f) {
        f = parent.lb + 20.0;
        I::incumbent(q, f);
      }
      if (I::empty(q)) {
        break;
      }
      double const now = I::min_lb(q);
      ok = ok && now >= prev;
      prev = now;
    }
    CHECK(ok);
  }
}

template<typename I>
void lazy_deletion()
{
  g_case = "lazy deletion never returns a child twice or skips a live one";
  std::mt19937 rng {4242};
  auto q = I::front(kRI);
  std::multiset<double> in;
  double f = 100.0;
  std::size_t pidx = 1;
  bool ok = true;
  for (int step = 0; step < 20000; ++step) {
    if (std::uniform_int_distribution<int>(0, 2)(rng) != 0 || I::empty(q)) {
      double const lb = std::uniform_real_distribution<double>(-50.0, 50.0)(rng);
      double const ub = lb + std::uniform_real_distribution<double>(0.0, 20.0)(rng);
      I::push(q, 0, pidx++, 1, lb, ub, f);
      in.insert(lb);
    } else {
      auto const n = I::pop(q);
      auto const it = in.find(n.lb);
      ok = ok && it != in.end();
      if (it != in.end()) {
        in.erase(it);
      }
    }
    if (step % 1000 == 0) {
      f -= 1.0;
      I::incumbent(q, f);
    }
    ok = ok && I::size(q) == in.size();
  }
  while (!I::empty(q)) {
    auto const n = I::pop(q);
    auto const it = in.find(n.lb);
    ok = ok && it != in.end();
    if (it != in.end()) {
      in.erase(it);
    }
  }
  CHECK(ok);
  CHECK(in.empty());
  CHECK(I::bytes(q) < 20000 * 48);  // 48: the reference node's size
}

template<typename I>
void reject_index_order()
{
  g_case = "reject-index pops the most promising child, not the loosest bound";
  auto by_lb = I::front(kLLB);
  I::push(by_lb, 0, 1, 1, -1000.0, 1000.0, 0.0);
  I::push(by_lb, 1, 2, 1, -10.0, -9.0, 0.0);
  CHECK(feq(I::pop(by_lb).lb, -1000.0));
  auto by_ri = I::front(kRI);
  I::push(by_ri, 0, 1, 1, -1000.0, 1000.0, 0.0);
  I::push(by_ri, 1, 2, 1, -10.0, -9.0, 0.0);
  CHECK(feq(I::pop(by_ri).lb, -10.0));
  CHECK(feq(I::min_lb(by_ri), -1000.0));
}

template<typename I>
void affine_invariance()
{
  g_case = "reject-index is invariant to a positive affine rescaling";
  std::mt19937 rng {7};
  std::uniform_real_distribution<double> v(-100.0, 100.0), pos(0.01, 100.0);
  bool ok = true;
  for (int i = 0; i < 5000; ++i) {
    double const lb = v(rng), ub = lb + pos(rng), f = lb + pos(rng), a = pos(rng), b = v(rng);
    double const plain = I::key(kRI, lb, ub, f);
    double const scaled = I::key(kRI, a * lb + b, a * ub + b, a * f + b);
    ok = ok && std::abs(plain - scaled) < 1e-9 * (1.0 + std::abs(plain));
  }
  CHECK(ok);
}

template<typename I>
void degenerate_keys()
{
  g_case = "the selection keys' degenerate cases";
  constexpr double inf = std::numeric_limits<double>::infinity();
  for (double lb : {-5.0, 0.0, 12.5}) {
    CHECK(feq(I::key(kRI, lb, lb + 3.0, kNone), I::key(kLLB, lb, lb + 3.0, kNone)));
  }
  CHECK(feq(I::key(kRI, 1.0, 1.0, 5.0), -inf));
  CHECK(feq(I::key(kRI, 9.0, 9.0, 5.0), inf));
  CHECK(I::key(kRI, 10.0, 20.0, 5.0) > I::key(kRI, 1.0, 20.0, 5.0));
  CHECK(feq(I::key(kRIS, 1.0, 3.0, 5.0), I::key(kRI, 1.0, 3.0, 5.0)));
}

template<typename I>
void rekeying()
{
  g_case = "re-keying on a better incumbent flips the order when it should";
  auto q = I::front(kRI);
  I::push(q, 0, 1, 1, 0.0, 100.0, 1000.0);
  I::push(q, 1, 2, 1, 50.0, 51.0, 1000.0);
  I::incumbent(q, 50.2);
  CHECK(feq(I::pop(q).lb, 0.0));
  auto stale = I::front(kRIS);
  I::push(stale, 0, 1, 1, 0.0, 100.0, 1000.0);
  I::push(stale, 1, 2, 1, 50.0, 51.0, 1000.0);
  I::incumbent(stale, 50.2);
  CHECK(feq(I::pop(stale).lb, 50.0));
  auto unimproved = I::front(kRI);
  I::push(unimproved, 0, 1, 1, 0.0, 100.0, 1000.0);
  I::push(unimproved, 1, 2, 1, 50.0, 51.0, 1000.0);
  CHECK(feq(I::pop(unimproved).lb, 50.0));
}

template<typename I>
void live_set_only()
{
  g_case = "count_viable and min_lb answer over the live set only";
  auto q = I::front(kRI);
  for (int i = 0; i < 10; ++i) {
    I::push(q, 0, static_cast<std::size_t>(i + 1), 1, i, i + 1.0, 100.0);
  }
  CHECK(I::viable(q, 4.0) == 5);
  CHECK(feq(I::min_lb(q), 0.0));
  std::vector<double> gone;
  while (I::size(q) > 5) {
    gone.push_back(I::pop(q).lb);
  }
  CHECK(I::viable(q, -1.0) == 0);
  double expected = kNone;
  for (int i = 0; i < 10; ++i) {
    if (std::find(gone.begin(), gone.end(), double(i)) == gone.end()) {
      expected = std::min(expected, double(i));
    }
  }
  CHECK(feq(I::min_lb(q), expected));
}

template<typename I>
void empty_reports_none()
{
  g_case = "an empty frontier reports no lower bound";
  auto q = I::front(kLLB);
  CHECK(I::empty(q));
  CHECK(feq(I::min_lb(q), kNone));
  I::push(q, 0, 1, 1, -3.0, 2.0, kNone);
  CHECK(feq(I::min_lb(q), -3.0));
  I::pop(q);
  CHECK(I::empty(q));
  CHECK(feq(I::min_lb(q), kNone));
}

template<typename I>
void run_all()
{
  g_impl = I::name;
  partition_invariants<I>();
  subregions_inside_their_child<I>();
  branch_is_pure<I>();
  widest_uniform_branch<I>();
  resolved_is_leaf<I>();
  nested_case<I>();
  rejected_fan_outs<I>();
  lb_rule_matches_plain_queue<I>();
  min_lb_monotone<I>();
  lazy_deletion<I>();
  reject_index_order<I>();
  affine_invariance<I>();
  degenerate_keys<I>();
  rekeying<I>();
  live_set_only<I>();
  empty_reports_none<I>();
}

// ---- the two implementations against each other ----------------------------------

void agreement()
{
  g_impl = "both";
  g_case = "engine and oracle partition every box alike";
  Boxes gen;
  bool ok = true;
  for (std::size_t n : {8u, 64u, 1024u}) {
    auto const e = Engine::policy(4, n);
    auto const o = Oracle::policy(4, n);
    for (int t = 0; t < 1000; ++t) {
      auto const [box, kinds] = gen.next();
      Part const a = Engine::partition(e, box, kinds), b = Oracle::partition(o, box, kinds);
      ok = ok && a.nested == b.nested && cmi::region::assignment_hash(a.branch) == cmi::region::assignment_hash(b.branch)
           && cmi::region::assignment_hash(a.launch) == cmi::region::assignment_hash(b.launch)
           && a.launch.domain_sizes == b.launch.domain_sizes;
    }
  }
  CHECK(ok);

  g_case = "engine and oracle frontiers pop alike under every rule, ties included";
  for (int rule : {kLLB, kRI, kRIS}) {
    std::mt19937 rng {31};
    auto e = Engine::front(rule);
    auto o = Oracle::front(rule);
    double f = kNone;
    bool same = true;
    std::size_t pidx = 1;
    for (int step = 0; step < 30000; ++step) {
      if (std::uniform_int_distribution<int>(0, 2)(rng) != 0 || Engine::empty(e)) {
        // coarse bounds and few depths: many exact ties on every field
        double const lb = std::floor(std::uniform_real_distribution<double>(-5.0, 5.0)(rng));
        double const ub = lb + std::floor(std::uniform_real_distribution<double>(0.0, 3.0)(rng));
        std::size_t const depth = std::uniform_int_distribution<std::size_t>(0, 2)(rng);
        std::size_t const p = pidx / 4;
        Engine::push(e, pidx % 4, p, depth, lb, ub, f);
        Oracle::push(o, pidx % 4, p, depth, lb, ub, f);
        ++pidx;
      } else {
        auto const a = Engine::pop(e);
        auto const b = Oracle::pop(o);
        same = same && feq(a.lb, b.lb) && a.depth == b.depth && a.pidx == b.pidx && a.sidx == b.sidx;
      }
      if (step % 700 == 0 && f > -4.0) {
        f = f == kNone ? 4.0 : f - 0.5;
        Engine::incumbent(e, f);
        Oracle::incumbent(o, f);
      }
      same = same && Engine::size(e) == Oracle::size(o) && feq(Engine::min_lb(e), Oracle::min_lb(o))
             && Engine::bytes(e) == Oracle::bytes(o);
    }
    CHECK(same);
  }
}

}  // namespace

int main()
{
  run_all<Engine>();
  run_all<Oracle>();
  check_rejects_misordered_launch();
  agreement();
  std::printf("%d checks, %d failed\n", g_checks, g_failed);
  return g_failed == 0 ? 0 : 1;
}
