// The reference-side binding of the MI355X engine: a
// cuminlp::backend::RegionBackendFactory<double> whose roles call the C-ABI of
// include/cmi.h. A maintainer adds this header to the reference tree and
// constructs mi355x::Factory where the reference constructs
// GraphBackendFactory (source/gams/solve.cu:92-93, and the fixed examples such
// as source/fixed_examples/qps.cu:109-110); SearchDriver and BackendCache stay
// as they are.
//
// Include it after the reference's own headers that declare
// cuminlp::model::Problem (model/problem.hpp), cuminlp::region::Composition
// and SlotAssignment (region/composition.hpp, region/slot.hpp), the backend
// interface (backend/backend.hpp, backend/cost_model.hpp) and the error
// classes (errors.hpp). It names nothing else of the reference, and link it
// against libcmi.so.
//
// Mapping (reference interface -> C-ABI):
//   RegionBackendFactory::cost_model      -> cmi_cost_model_get
//   RegionBackendFactory::build_subdivision -> cmi_role_build per requested role
//   RegionSampler::sample / n_samples     -> cmi_sample / cmi_role_size
//   RegionBounder::bound / n_regions      -> cmi_bound / cmi_role_size
//   RegionEnumerator::enumerate / n_points -> cmi_enumerate / cmi_role_size
//   OverBudgetError{OverBudget}           <- CMI_OVER_BUDGET + cmi_over_budget
// Result spans point into role-owned pinned memory and stay valid until the
// next call on the same role (backend.hpp:49-50, 66-67).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <string>
#include <utility>
#include <vector>

#include "cmi.h"

namespace cuminlp::backend::mi355x
{

/// A C-ABI status as the reference's exception (errors.hpp:11-57). A device
/// failure has no CUDA error code to carry, so it is the base cuminlp::error.
inline void throw_on(cmi_status s)
{
  if (s == CMI_OK) {
    return;
  }
  std::string const why = cmi_last_error();
  switch (s) {
    case CMI_SHAPE_MISMATCH: throw ShapeMismatch(why);
    case CMI_INVALID_CONFIG: throw InvalidConfiguration(why);
    case CMI_OVER_BUDGET: throw ResourceExhausted(why);
    case CMI_INVALID_DAG: throw InvalidDAG(why);
    case CMI_INVALID_PROBLEM: throw InvalidProblem(why);
    default: throw error(why);
  }
}

/// The facts of a build that did not fit, member for member
/// (backend.hpp:236-262). The engine's role cost is affine per role; it is
/// reported in the reference's RegionCostModel shape (cost_model.hpp:19-56).
inline RegionCostModel cost_of(const cmi_cost_model& c, std::size_t samples_per_region)
{
  std::size_t const s = samples_per_region == 0 ? 1 : samples_per_region;
  return RegionCostModel {
      .sampler_bytes_per_sample = (c.sampler_per_region + s - 1) / s,
      .bound_bytes_per_region = c.bounder_per_region,
      .exact_bytes_per_region = c.enumerator_per_region,
      .fixed_bytes = c.sampler_fixed + c.bounder_fixed,
  };
}

inline OverBudget facts_of(const cmi_over_budget& f)
{
  region::Composition composition;
  for (std::size_t j = 0; j < f.composition_size && j < sizeof(f.composition_kinds); ++j) {
    composition.kinds.push_back(static_cast<region::SlotKind>(f.composition_kinds[j]));
  }
  std::optional<region::FanOutSpec> fan_out;
  if (f.has_fan_out != 0) {
    fan_out.emplace(f.fan_out_partition_num, f.fan_out_enumerate_cap);
  }
  return OverBudget {
      .role = f.role,
      .needed_bytes = f.needed_bytes,
      .budget_bytes = f.budget_bytes,
      .n_regions = f.n_regions,
      .elements_per_region = f.elements_per_region,
      .bytes_per_element = f.bytes_per_element,
      .element_breakdown = f.element_breakdown,
      .element_inventory = f.element_inventory,
      .cost = cost_of(f.cost, f.solve_samples_per_region),
      .composition = std::move(composition),
      .fan_out = fan_out,
      .solve_samples_per_region = f.solve_samples_per_region,
      .sampler_bytes = f.sampler_bytes,
  };
}

/// The problem's DAG as the engine's flat arrays (same Op, VarKind and Cmp
/// numbering as model/dag.hpp and model/problem.hpp), validated by the engine.
inline cmi_problem* to_engine(const model::Problem<double>& p)
{
  std::size_t const n = p.graph.nodes.size();
  std::vector<std::uint8_t> op(n);
  std::vector<std::uint64_t> in0(n, 0), in1(n, 0), var_index(n, 0);
  std::vector<double> constant(n, 0.0);
  std::vector<std::int32_t> int_exp(n, 0);
  for (std::size_t i = 0; i < n; ++i) {
    const auto& node = p.graph.nodes[i];
    op[i] = static_cast<std::uint8_t>(node.op);
    if (!node.in.empty()) {
      in0[i] = node.in[0];
    }
    if (node.in.size() > 1) {
      in1[i] = node.in[1];
    }
    switch (node.op) {
      case model::Op::Const: constant[i] = node.payload.constant; break;
      case model::Op::Var: var_index[i] = node.payload.var_index; break;
      case model::Op::PowN: int_exp[i] = node.payload.int_exp; break;
      default: break;
    }
  }
  std::vector<cmi_ivl> box;
  std::vector<std::uint8_t> kinds;
  for (std::size_t v = 0; v < p.box_bounds.size(); ++v) {
    box.push_back(cmi_ivl {p.box_bounds[v].lb, p.box_bounds[v].ub});
    kinds.push_back(static_cast<std::uint8_t>(p.var_kinds[v]));
  }
  std::vector<std::uint64_t> con_root;
  std::vector<std::uint8_t> con_cmp;
  std::vector<double> con_rhs;
  for (const auto& c : p.constraints) {
    con_root.push_back(c.root_id);
    con_cmp.push_back(static_cast<std::uint8_t>(c.cmp));
    con_rhs.push_back(c.rhs);
  }
  cmi_problem_desc d {};
  d.n_nodes = n;
  d.op = op.data();
  d.in0 = in0.data();
  d.in1 = in1.data();
  d.constant = constant.data();
  d.var_index = var_index.data();
  d.int_exp = int_exp.data();
  d.n_vars = box.size();
  d.box = box.data();
  d.var_kind = kinds.data();
  d.objective_root = p.objective_root;
  d.n_constraints = con_root.size();
  d.con_root = con_root.data();
  d.con_cmp = con_cmp.data();
  d.con_rhs = con_rhs.data();
  cmi_problem* out = nullptr;
  throw_on(cmi_problem_create(&d, &out));
  return out;
}

/// One built role; destroyed with its owner.
class RoleHandle
{
public:
  explicit RoleHandle(cmi_role* r)
      : r_(r)
  {
  }
  RoleHandle(const RoleHandle&) = delete;
  RoleHandle& operator=(const RoleHandle&) = delete;
  ~RoleHandle() { cmi_role_destroy(r_); }

  cmi_role* get() const { return r_; }
  std::size_t size() const { return cmi_role_size(r_); }

  /// The role's arguments for one region: the box as the engine's interval
  /// layout (the same 16 bytes as cu::interval<double>) and the slot data.
  template<typename Call>
  void run(const Region<double>& g, Call&& call) const
  {
    static_assert(sizeof(g.box[0]) == sizeof(cmi_ivl), "cu::interval<double> and cmi_ivl must both be 16 bytes");
    call(reinterpret_cast<const cmi_ivl*>(g.box.data()), g.box.size(), g.assignment.var_ids.data(),
         g.assignment.widths.data(), g.assignment.var_ids.size());
  }

private:
  cmi_role* r_;
};

inline CandidateResult<double> candidate_of(const cmi_candidate_view& v, std::size_t n_vars)
{
  return CandidateResult<double> {
      .found = v.found != 0,
      .value = v.value,
      .point = std::span<const double>(v.point, n_vars),
      .index = v.index,
  };
}

class Sampler final : public RegionSampler<double>
{
public:
  explicit Sampler(cmi_role* r)
      : role_(r)
  {
  }
  CandidateResult<double> sample(const Region<double>& g, std::uint64_t salt) override
  {
    cmi_candidate_view v {};
    role_.run(g, [&](const cmi_ivl* box, std::size_t nv, const std::size_t* var, const std::size_t* w, std::size_t k) {
      throw_on(cmi_sample(role_.get(), box, nv, var, w, k, salt, &v));
    });
    return candidate_of(v, g.box.size());
  }
  std::size_t n_samples() const override { return role_.size(); }

private:
  RoleHandle role_;
};

class Bounder final : public RegionBounder<double>
{
public:
  explicit Bounder(cmi_role* r)
      : role_(r)
  {
  }
  BoundResult<double> bound(const Region<double>& g) override
  {
    cmi_bound_view v {};
    role_.run(g, [&](const cmi_ivl* box, std::size_t nv, const std::size_t* var, const std::size_t* w, std::size_t k) {
      throw_on(cmi_bound(role_.get(), box, nv, var, w, k, &v));
    });
    return BoundResult<double> {
        .feasible = std::span<const unsigned char>(v.feasible, v.n_regions),
        .obj_lb = std::span<const double>(v.obj_lb, v.n_regions),
        .n_regions = v.n_regions,
    };
  }
  std::size_t n_regions() const override { return role_.size(); }

private:
  RoleHandle role_;
};

class Enumerator final : public RegionEnumerator<double>
{
public:
  explicit Enumerator(cmi_role* r)
      : role_(r)
  {
  }
  CandidateResult<double> enumerate(const Region<double>& g) override
  {
    cmi_candidate_view v {};
    role_.run(g, [&](const cmi_ivl* box, std::size_t nv, const std::size_t* var, const std::size_t* w, std::size_t k) {
      throw_on(cmi_enumerate(role_.get(), box, nv, var, w, k, &v));
    });
    return candidate_of(v, g.box.size());
  }
  std::size_t n_points() const override { return role_.size(); }

private:
  RoleHandle role_;
};

/// The factory. One engine context per factory; one device tape per problem
/// (compiled on first use and kept while the same problem is asked for).
class Factory final : public RegionBackendFactory<double>
{
public:
  explicit Factory(int device = 0)
  {
    throw_on(cmi_ctx_create(device, &ctx_));
  }
  Factory(const Factory&) = delete;
  Factory& operator=(const Factory&) = delete;
  ~Factory() override
  {
    release_tape();
    cmi_ctx_destroy(ctx_);
  }

  BackendCapabilities capabilities() const override { return BackendCapabilities {}; }

  RegionCostModel cost_model(const model::Problem<double>& p) const override
  {
    cmi_cost_model c {};
    throw_on(cmi_cost_model_get(tape_for(p), 1, &c));
    return cost_of(c, 1);
  }

  SubdivisionBundle<double> build_subdivision(const model::Problem<double>& p, const region::Composition& c,
                                              std::size_t n_regions, const BuildBudget& budget,
                                              RoleRequest roles) const override
  {
    cmi_tape* const t = tape_for(p);
    std::vector<std::uint8_t> kinds;
    for (region::SlotKind k : c.kinds) {
      kinds.push_back(static_cast<std::uint8_t>(k));
    }
    // The solve-wide sample count reaches every role, so an over-budget
    // report can price the sampler too; only the sampler allocates per sample.
    auto build = [&](int kind) {
      cmi_role* r = nullptr;
      cmi_over_budget facts {};
      cmi_status const s = cmi_role_build(t, kind, kinds.data(), kinds.size(), n_regions, budget.samples_per_region,
                                          budget.bytes, &r, &facts);
      if (s == CMI_OVER_BUDGET && facts.needed_bytes != 0) {
        throw OverBudgetError(facts_of(facts));
      }
      throw_on(s);
      return r;
    };
    SubdivisionBundle<double> out;
    if (roles.sampler) {
      out.sampler = std::make_unique<Sampler>(build(CMI_ROLE_SAMPLER));
    }
    if (roles.bounder) {
      out.bounder = std::make_unique<Bounder>(build(CMI_ROLE_BOUNDER));
    }
    if (roles.enumerator && region::is_fully_enumerable(c)) {
      out.enumerator = std::make_unique<Enumerator>(build(CMI_ROLE_ENUMERATOR));
    }
    return out;
  }

private:
  cmi_tape* tape_for(const model::Problem<double>& p) const
  {
    if (cached_ != &p || tape_ == nullptr) {
      release_tape();
      problem_ = to_engine(p);
      throw_on(cmi_tape_create(ctx_, problem_, &tape_));
      cached_ = &p;
    }
    return tape_;
  }

  void release_tape() const
  {
    if (tape_ != nullptr) {
      cmi_tape_destroy(tape_);
      tape_ = nullptr;
    }
    if (problem_ != nullptr) {
      cmi_problem_destroy(problem_);
      problem_ = nullptr;
    }
    cached_ = nullptr;
  }

  cmi_ctx* ctx_ = nullptr;
  mutable const model::Problem<double>* cached_ = nullptr;
  mutable cmi_problem* problem_ = nullptr;
  mutable cmi_tape* tape_ = nullptr;
};

}  // namespace cuminlp::backend::mi355x
