// Stand-ins for the reference types that include/cmi_reference_adapter.hpp is
// written against: the names, members and virtual signatures of
// include/cuminlp/{errors.hpp, model/dag.hpp, model/problem.hpp,
// region/slot.hpp, region/composition.hpp, region/fan_out.hpp,
// backend/backend.hpp, backend/cost_model.hpp} (and cu::interval), with no
// behaviour beyond what the adapter test needs. They let the adapter compile
// and run here the way the reference's own cache_test.cpp fakes a factory:
// test infrastructure only.
#pragma once

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <limits>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace cu
{
template<typename T>
struct interval
{
  T lb;
  T ub;
};
}  // namespace cu

namespace cuminlp
{
struct error : std::runtime_error
{
  using std::runtime_error::runtime_error;
};
struct InvalidDAG : error
{
  using error::error;
};
struct InvalidProblem : error
{
  using error::error;
};
struct ShapeMismatch : error
{
  using error::error;
};
struct InvalidConfiguration : error
{
  using error::error;
};
struct ResourceExhausted : error
{
  using error::error;
};

namespace model
{
enum class Op
{
  Var, Const, Add, Sub, Mul, Div, Sqr, Neg, Exp, Log, Sqrt, Sin, Cos, Tanh, PowN, Pow, Abs, Min, Max
};
enum class VarKind
{
  Continuous, Integer, Binary
};
enum class Cmp
{
  LE, EQ
};
template<typename T>
union DAGNodePayload
{
  T constant;
  std::size_t var_index;
  int int_exp;
};
template<typename T>
struct DAGNode
{
  std::size_t id;
  Op op;
  std::vector<std::size_t> in;
  DAGNodePayload<T> payload;
};
template<typename T>
struct ExprDAG
{
  std::vector<DAGNode<T>> nodes;
  std::size_t emit(Op op, std::initializer_list<std::size_t> in, DAGNodePayload<T> payload = {})
  {
    nodes.push_back(DAGNode<T> {nodes.size(), op, std::vector<std::size_t>(in), payload});
    return nodes.size() - 1;
  }
};
template<typename T>
struct ConstraintRef
{
  std::size_t root_id;
  Cmp cmp;
  T rhs;
};
template<typename T>
struct Problem
{
  ExprDAG<T> graph;
  std::vector<cu::interval<T>> box_bounds;
  std::vector<VarKind> var_kinds;
  std::size_t objective_root = std::numeric_limits<std::size_t>::max();
  std::vector<ConstraintRef<T>> constraints;

  std::size_t var(T lb, T ub, VarKind kind)
  {
    DAGNodePayload<T> p;
    p.var_index = box_bounds.size();
    box_bounds.push_back({lb, ub});
    var_kinds.push_back(kind);
    return graph.emit(Op::Var, {}, p);
  }
  std::size_t constant(T c)
  {
    DAGNodePayload<T> p;
    p.constant = c;
    return graph.emit(Op::Const, {}, p);
  }
};
}  // namespace model

namespace region
{
enum class SlotKind : std::uint8_t
{
  Continuous, IntegerPartition, IntegerEnumerate, BinaryEnumerate
};
struct Composition
{
  std::vector<SlotKind> kinds;
  std::size_t size() const { return kinds.size(); }
  bool operator==(const Composition&) const = default;
};
struct SlotAssignment
{
  Composition composition;
  std::vector<std::size_t> var_ids;
  std::vector<std::size_t> widths;
  std::vector<std::size_t> domain_sizes;
};
class FanOutSpec
{
public:
  FanOutSpec(std::size_t partition_num, std::size_t enumerate_cap)
      : p_(partition_num)
      , e_(enumerate_cap)
  {
  }
  std::size_t partition_num() const { return p_; }
  std::size_t enumerate_cap() const { return e_; }

private:
  std::size_t p_, e_;
};
inline bool is_fully_enumerable(const Composition& c)
{
  for (SlotKind k : c.kinds) {
    if (k != SlotKind::IntegerEnumerate && k != SlotKind::BinaryEnumerate) {
      return false;
    }
  }
  return true;
}
}  // namespace region

namespace backend
{
template<typename T>
struct Region
{
  std::span<const cu::interval<T>> box;
  const region::SlotAssignment& assignment;
};
template<typename T>
struct CandidateResult
{
  bool found = false;
  T value {};
  std::span<const T> point;
  std::int64_t index = -1;
};
template<typename T>
struct BoundResult
{
  std::span<const unsigned char> feasible;
  std::span<const T> obj_lb;
  std::size_t n_regions = 0;
};
template<typename T>
struct RegionSampler
{
  virtual ~RegionSampler() = default;
  virtual CandidateResult<T> sample(const Region<T>& region, std::uint64_t salt) = 0;
  virtual std::size_t n_samples() const = 0;
};
template<typename T>
struct RegionBounder
{
  virtual ~RegionBounder() = default;
  virtual BoundResult<T> bound(const Region<T>& region) = 0;
  virtual std::size_t n_regions() const = 0;
};
template<typename T>
struct RegionEnumerator
{
  virtual ~RegionEnumerator() = default;
  virtual CandidateResult<T> enumerate(const Region<T>& region) = 0;
  virtual std::size_t n_points() const = 0;
};
struct BackendCapabilities
{
  bool exact_enumeration = true;
  bool deterministic_sampling = true;
};
template<typename T>
struct SubdivisionBundle
{
  std::unique_ptr<RegionSampler<T>> sampler;
  std::unique_ptr<RegionBounder<T>> bounder;
  std::unique_ptr<RegionEnumerator<T>> enumerator;
};
struct BuildBudget
{
  std::size_t bytes = 0;
  std::size_t samples_per_region = 1;
  bool report_build = false;
};
struct RoleRequest
{
  bool sampler = false;
  bool bounder = false;
  bool enumerator = false;
};
struct RegionCostModel
{
  std::size_t sampler_bytes_per_sample = 0;
  std::size_t bound_bytes_per_region = 0;
  std::size_t exact_bytes_per_region = 0;
  std::size_t fixed_bytes = 0;
};
struct OverBudget
{
  std::string role;
  std::size_t needed_bytes = 0;
  std::size_t budget_bytes = 0;
  std::size_t n_regions = 0;
  std::size_t elements_per_region = 1;
  std::size_t bytes_per_element = 0;
  std::string element_breakdown;
  std::string element_inventory;
  RegionCostModel cost;
  region::Composition composition;
  std::optional<region::FanOutSpec> fan_out;
  std::size_t solve_samples_per_region = 1;
  std::size_t sampler_bytes = 0;
};
class OverBudgetError : public ResourceExhausted
{
public:
  explicit OverBudgetError(OverBudget facts)
      : ResourceExhausted(facts.role + " over budget")
      , facts_(std::make_shared<const OverBudget>(std::move(facts)))
  {
  }
  const OverBudget& facts() const { return *facts_; }

private:
  std::shared_ptr<const OverBudget> facts_;
};
template<typename T>
struct RegionBackendFactory
{
  virtual ~RegionBackendFactory() = default;
  virtual BackendCapabilities capabilities() const = 0;
  virtual RegionCostModel cost_model(const model::Problem<T>& problem) const = 0;
  virtual SubdivisionBundle<T> build_subdivision(const model::Problem<T>& problem,
                                                 const region::Composition& composition, std::size_t n_regions,
                                                 const BuildBudget& budget, RoleRequest roles) const = 0;
};
}  // namespace backend
}  // namespace cuminlp
