// The reference-side adapter (include/cmi_reference_adapter.hpp) compiled
// against reference-shaped types (reference_shapes.hpp) and run on the device:
// it builds, bounds, samples and enumerates through the reference's
// RegionBackendFactory interface, and maps CMI_OVER_BUDGET to OverBudgetError
// with every fact intact. Bounds and candidates are checked bit for bit
// against the oracle (liboracle.so); the facts against a direct cmi_role_build.
//
// Usage: adapter_test [device]; exit 0 when every check holds. Built by
// cuminlpsearch_amd/build.py, run by tests/test_gpu_adapter.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "reference_shapes.hpp"

#include "cmi_reference_adapter.hpp"

extern "C" {
int orc_bound(const cmi_problem_desc* d, const cmi_ivl* box, const size_t* var_ids, const size_t* widths,
              const uint8_t* kinds, size_t slots, uint8_t* feasible, double* obj_lb);
int orc_candidate(const cmi_problem_desc* d, const cmi_ivl* box, const size_t* var_ids, const size_t* widths,
                  const uint8_t* kinds, size_t slots, size_t samples, uint64_t salt, int* found, double* value,
                  int64_t* index, double* point);
}

namespace
{

using namespace cuminlp;
using model::Op;
using model::VarKind;

int g_failed = 0;
int g_checks = 0;

void check(bool ok, const char* what, int line)
{
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::printf("FAIL line %d: %s\n", line, what);
  }
}
#define CHECK(x) check((x), #x, __LINE__)

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

/// min x0*x1 + sqr(x2) - 2*x0 + 0.5*x1   s.t.  x0 + x1 <= 1.5,  x2*x0 = ... LE 3
model::Problem<double> qcqp()
{
  model::Problem<double> p;
  std::size_t const x0 = p.var(-2.0, 3.0, VarKind::Continuous);
  std::size_t const x1 = p.var(-1.0, 2.5, VarKind::Continuous);
  std::size_t const x2 = p.var(-4.0, 1.0, VarKind::Continuous);
  auto& g = p.graph;
  std::size_t const m01 = g.emit(Op::Mul, {x0, x1});
  std::size_t const s2 = g.emit(Op::Sqr, {x2});
  std::size_t const two = p.constant(2.0);
  std::size_t const half = p.constant(0.5);
  std::size_t const t0 = g.emit(Op::Mul, {two, x0});
  std::size_t const t1 = g.emit(Op::Mul, {half, x1});
  std::size_t const a = g.emit(Op::Add, {m01, s2});
  std::size_t const b = g.emit(Op::Sub, {a, t0});
  p.objective_root = g.emit(Op::Add, {b, t1});
  std::size_t const c0 = g.emit(Op::Add, {x0, x1});
  p.constraints.push_back({c0, model::Cmp::LE, 1.5});
  std::size_t const c1 = g.emit(Op::Mul, {x2, x0});
  p.constraints.push_back({c1, model::Cmp::LE, 3.0});
  return p;
}

/// min (y0 - 2)^2 + y0*y1 - y1  with y0, y1 integer in [0, 4]: all-integer, enumerable.
model::Problem<double> integers()
{
  model::Problem<double> p;
  std::size_t const y0 = p.var(0.0, 4.0, VarKind::Integer);
  std::size_t const y1 = p.var(0.0, 4.0, VarKind::Integer);
  auto& g = p.graph;
  std::size_t const two = p.constant(2.0);
  std::size_t const d = g.emit(Op::Sub, {y0, two});
  std::size_t const d2 = g.emit(Op::Sqr, {d});
  std::size_t const m = g.emit(Op::Mul, {y0, y1});
  std::size_t const s = g.emit(Op::Add, {d2, m});
  p.objective_root = g.emit(Op::Sub, {s, y1});
  return p;
}

/// The engine's flat arrays of `p` (as the adapter hands them over), for the oracle.
struct Flat
{
  std::vector<std::uint8_t> op, kind, cmp;
  std::vector<std::uint64_t> in0, in1, vidx, croot;
  std::vector<double> cst, rhs;
  std::vector<std::int32_t> iexp;
  std::vector<cmi_ivl> box;
  std::uint64_t root = 0;
  cmi_problem_desc d {};

  explicit Flat(const model::Problem<double>& p)
  {
    cmi_problem* e = backend::mi355x::to_engine(p);
    cmi_problem_info info {};
    backend::mi355x::throw_on(cmi_problem_info_get(e, &info));
    op.resize(info.n_nodes);
    in0.resize(info.n_nodes);
    in1.resize(info.n_nodes);
    cst.resize(info.n_nodes);
    vidx.resize(info.n_nodes);
    iexp.resize(info.n_nodes);
    box.resize(info.n_vars);
    kind.resize(info.n_vars);
    croot.resize(info.n_constraints);
    cmp.resize(info.n_constraints);
    rhs.resize(info.n_constraints);
    backend::mi355x::throw_on(cmi_problem_export(e, op.data(), in0.data(), in1.data(), cst.data(), vidx.data(),
                                                 iexp.data(), box.data(), kind.data(), &root, croot.data(),
                                                 cmp.data(), rhs.data()));
    cmi_problem_destroy(e);
    d = cmi_problem_desc {info.n_nodes, op.data(),  in0.data(), in1.data(), cst.data(),
                          vidx.data(),  iexp.data(), info.n_vars, box.data(), kind.data(),
                          root,         info.n_constraints, croot.data(), cmp.data(), rhs.data()};
  }
};

void roles_equal_the_oracle(backend::mi355x::Factory& f)
{
  model::Problem<double> const p = qcqp();
  Flat const flat(p);
  region::SlotAssignment a;
  a.composition.kinds = {region::SlotKind::Continuous, region::SlotKind::Continuous};
  a.var_ids = {0, 2};
  a.widths = {4, 8};
  std::size_t const n_regions = 32, samples = 8;
  auto bundle = f.build_subdivision(p, a.composition, n_regions, backend::BuildBudget {0, samples, false},
                                    backend::RoleRequest {true, true, true});
  CHECK(bundle.sampler != nullptr && bundle.bounder != nullptr);
  CHECK(bundle.enumerator == nullptr);  // not an all-enumerate composition
  CHECK(bundle.bounder->n_regions() == n_regions);
  CHECK(bundle.sampler->n_samples() == n_regions * samples);

  std::vector<cu::interval<double>> box {{-1.0, 2.0}, {-1.0, 2.5}, {-3.0, 0.5}};
  backend::Region<double> const region {box, a};
  auto const b = bundle.bounder->bound(region);
  std::vector<std::uint8_t> feas(n_regions);
  std::vector<double> lb(n_regions);
  std::uint8_t const kinds[] = {CMI_SLOT_CONTINUOUS, CMI_SLOT_CONTINUOUS};
  CHECK(orc_bound(&flat.d, reinterpret_cast<const cmi_ivl*>(box.data()), a.var_ids.data(), a.widths.data(), kinds,
                  2, feas.data(), lb.data()) == 0);
  CHECK(b.n_regions == n_regions);
  bool bound_same = b.feasible.size() == n_regions && b.obj_lb.size() == n_regions;
  for (std::size_t r = 0; bound_same && r < n_regions; ++r) {
    bound_same = b.feasible[r] == feas[r] && same_bits(b.obj_lb[r], lb[r]);
  }
  CHECK(bound_same);

  for (std::uint64_t salt : {1ull, 7ull, 123456789ull}) {
    auto const c = bundle.sampler->sample(region, salt);
    int found = 0;
    double value = 0.0;
    std::int64_t index = 0;
    std::vector<double> point(3);
    CHECK(orc_candidate(&flat.d, reinterpret_cast<const cmi_ivl*>(box.data()), a.var_ids.data(), a.widths.data(),
                        kinds, 2, samples, salt, &found, &value, &index, point.data()) == 0);
    CHECK(c.found == (found != 0));
    CHECK(same_bits(c.value, value) && c.index == index);
    bool pt = c.point.size() == 3;
    for (std::size_t v = 0; pt && v < 3; ++v) {
      pt = same_bits(c.point[v], point[v]);
    }
    CHECK(pt);
  }

  // ShapeMismatch for a box of the wrong length, through the adapter.
  std::vector<cu::interval<double>> short_box {{-1.0, 2.0}, {-1.0, 2.5}};
  bool shape = false;
  try {
    bundle.bounder->bound(backend::Region<double> {short_box, a});
  } catch (const ShapeMismatch&) {
    shape = true;
  }
  CHECK(shape);

  // The cost model in the reference's shape, from the engine's per-role costs.
  auto const cm = f.cost_model(p);
  CHECK(cm.bound_bytes_per_region > 0 && cm.sampler_bytes_per_sample > 0 && cm.fixed_bytes > 0);
}

void enumerator_equals_the_oracle(backend::mi355x::Factory& f)
{
  model::Problem<double> const p = integers();
  Flat const flat(p);
  region::SlotAssignment a;
  a.composition.kinds = {region::SlotKind::IntegerEnumerate, region::SlotKind::IntegerEnumerate};
  a.var_ids = {0, 1};
  a.widths = {5, 5};
  auto bundle = f.build_subdivision(p, a.composition, 25, backend::BuildBudget {0, 4, false},
                                    backend::RoleRequest {false, false, true});
  CHECK(bundle.enumerator != nullptr && bundle.sampler == nullptr && bundle.bounder == nullptr);
  CHECK(bundle.enumerator->n_points() == 25);
  std::vector<cu::interval<double>> box = p.box_bounds;
  auto const c = bundle.enumerator->enumerate(backend::Region<double> {box, a});
  int found = 0;
  double value = 0.0;
  std::int64_t index = 0;
  std::vector<double> point(2);
  std::uint8_t const kinds[] = {CMI_SLOT_INTEGER_ENUMERATE, CMI_SLOT_INTEGER_ENUMERATE};
  CHECK(orc_candidate(&flat.d, reinterpret_cast<const cmi_ivl*>(box.data()), a.var_ids.data(), a.widths.data(),
                      kinds, 2, 0, 0, &found, &value, &index, point.data()) == 0);
  CHECK(c.found && found != 0 && same_bits(c.value, value) && c.index == index);
  // brute force over the 25 lattice points: the minimum and its lowest index
  double best = INFINITY;
  std::int64_t best_i = -1;
  for (int y1 = 0; y1 < 5; ++y1) {
    for (int y0 = 0; y0 < 5; ++y0) {
      double const v = (y0 - 2.0) * (y0 - 2.0) + double(y0) * y1 - y1;
      if (v < best) {
        best = v;
        best_i = y1 * 5 + y0;
      }
    }
  }
  CHECK(c.value == best && c.index == best_i);
  CHECK(c.point.size() == 2 && c.point[0] == double(best_i % 5) && c.point[1] == double(best_i / 5));
}

void over_budget_carries_every_fact(backend::mi355x::Factory& f, int device)
{
  model::Problem<double> const p = qcqp();
  region::Composition comp;
  comp.kinds = {region::SlotKind::Continuous, region::SlotKind::Continuous, region::SlotKind::Continuous};
  std::size_t const n_regions = 64, samples = 16, budget = 100;

  // What the C-ABI reports for the same build.
  cmi_ctx* ctx = nullptr;
  backend::mi355x::throw_on(cmi_ctx_create(device, &ctx));
  cmi_problem* e = backend::mi355x::to_engine(p);
  cmi_tape* t = nullptr;
  backend::mi355x::throw_on(cmi_tape_create(ctx, e, &t));
  std::uint8_t const kinds[] = {CMI_SLOT_CONTINUOUS, CMI_SLOT_CONTINUOUS, CMI_SLOT_CONTINUOUS};
  cmi_role* r = nullptr;
  cmi_over_budget raw {};
  CHECK(cmi_role_build(t, CMI_ROLE_SAMPLER, kinds, 3, n_regions, samples, budget, &r, &raw) == CMI_OVER_BUDGET);
  CHECK(r == nullptr);
  cmi_cost_model cm {};
  backend::mi355x::throw_on(cmi_cost_model_get(t, samples, &cm));

  bool thrown = false;
  try {
    f.build_subdivision(p, comp, n_regions, backend::BuildBudget {budget, samples, false},
                        backend::RoleRequest {true, true, false});
  } catch (const backend::OverBudgetError& err) {
    thrown = true;
    backend::OverBudget const& o = err.facts();
    CHECK(o.role == "point sampler" && o.role == raw.role);
    CHECK(o.needed_bytes == raw.needed_bytes && o.needed_bytes == cm.sampler_fixed + cm.sampler_per_region * n_regions);
    CHECK(o.budget_bytes == budget);
    CHECK(o.n_regions == n_regions);
    CHECK(o.elements_per_region == samples);
    CHECK(o.bytes_per_element == raw.bytes_per_element && o.bytes_per_element > 0);
    CHECK(!o.element_breakdown.empty() && o.element_breakdown == raw.element_breakdown);
    CHECK(!o.element_inventory.empty() && o.element_inventory == raw.element_inventory);
    CHECK(o.cost.bound_bytes_per_region == cm.bounder_per_region);
    CHECK(o.cost.exact_bytes_per_region == cm.enumerator_per_region);
    CHECK(o.cost.sampler_bytes_per_sample == (cm.sampler_per_region + samples - 1) / samples);
    CHECK(o.cost.fixed_bytes == cm.sampler_fixed + cm.bounder_fixed);
    CHECK(o.composition == comp);
    CHECK(!o.fan_out.has_value());
    CHECK(o.solve_samples_per_region == samples);
    CHECK(o.sampler_bytes == 0);
  }
  CHECK(thrown);

  // A bounder-only build over budget names the bounder.
  bool bounder = false;
  try {
    f.build_subdivision(p, comp, n_regions, backend::BuildBudget {budget, samples, false},
                        backend::RoleRequest {false, true, false});
  } catch (const backend::OverBudgetError& err) {
    bounder = err.facts().role == "interval bounder" && err.facts().elements_per_region == 1;
  }
  CHECK(bounder);
  cmi_tape_destroy(t);
  cmi_problem_destroy(e);
  cmi_ctx_destroy(ctx);
}

}  // namespace

int main(int argc, char** argv)
{
  int const device = argc > 1 ? std::atoi(argv[1]) : 0;
  try {
    backend::mi355x::Factory f(device);
    std::printf("adapter: roles\n");
    std::fflush(stdout);
    roles_equal_the_oracle(f);
    std::printf("adapter: enumerator\n");
    std::fflush(stdout);
    enumerator_equals_the_oracle(f);
    std::printf("adapter: over budget\n");
    std::fflush(stdout);
    over_budget_carries_every_fact(f, device);
  } catch (const std::exception& ex) {
    std::printf("FAIL: uncaught %s\n", ex.what());
    return 2;
  }
  std::printf("%d checks, %d failed\n", g_checks, g_failed);
  return g_failed == 0 ? 0 : 1;
}
