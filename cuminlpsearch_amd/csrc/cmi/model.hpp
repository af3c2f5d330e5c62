#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <limits>
#include <optional>
#include <stdexcept>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"

// The problem IR: an SSA-numbered, topologically ordered expression DAG plus
// the box, variable kinds, objective root and constraints.
//
// The public surface (Op, DAGNode, ExprDAG::emit, Expr operators, Problem::var
// / int_var / bin_var / fixed / set_objective / add_constraint / validate)
// keeps the reference's names and meaning (model/dag.hpp:12-93,
// model/problem.hpp:35-358) so a caller of the reference builds problems the
// same way here.  Two construction rules matter for bit parity and are kept:
// `a * a` on the same node emits Sqr (problem.hpp:62-72), and a scalar operand
// becomes a Const node consumed by value at its use (problem.hpp:92-106).
namespace cmi::model
{

enum class Op : std::uint8_t
{
  Var,
  Const,
  Add,
  Sub,
  Mul,
  Div,
  Sqr,
  Neg,
  Exp,
  Log,
  Sqrt,
  Sin,
  Cos,
  Tanh,
  PowN,
  Pow,
  Abs,
  Min,
  Max
};

inline const char* op_name(Op op)
{
  switch (op) {
    case Op::Var: return "Var";
    case Op::Const: return "Const";
    case Op::Add: return "Add";
    case Op::Sub: return "Sub";
    case Op::Mul: return "Mul";
    case Op::Div: return "Div";
    case Op::Sqr: return "Sqr";
    case Op::Neg: return "Neg";
    case Op::Exp: return "Exp";
    case Op::Log: return "Log";
    case Op::Sqrt: return "Sqrt";
    case Op::Sin: return "Sin";
    case Op::Cos: return "Cos";
    case Op::Tanh: return "Tanh";
    case Op::PowN: return "PowN";
    case Op::Pow: return "Pow";
    case Op::Abs: return "Abs";
    case Op::Min: return "Min";
    case Op::Max: return "Max";
  }
  return "?";
}

inline std::size_t op_arity(Op op)
{
  switch (op) {
    case Op::Var:
    case Op::Const:
      return 0;
    case Op::Add:
    case Op::Sub:
    case Op::Mul:
    case Op::Div:
    case Op::Pow:
    case Op::Min:
    case Op::Max:
      return 2;
    default:
      return 1;
  }
}

template<typename T>
union DAGNodePayload
{
  T constant;
  std::size_t var_index;
  int int_exp;
};

/// One DAG node. Invariant: every id in `in` is strictly smaller than `id`.
template<typename T>
struct DAGNode
{
  std::size_t id;
  Op op;
  std::vector<std::size_t> in;
  DAGNodePayload<T> payload;
  std::size_t op_count;  // size of this node's expression tree (with repeats)
  std::size_t live_set_estimate;  // filled in by the tape compiler's liveness pass
};

template<typename T>
class ExprDAG
{
public:
  std::size_t emit(Op op,
                   std::initializer_list<std::size_t> inputs,
                   DAGNodePayload<T> payload = {})
  {
    std::size_t const id = nodes.size();
    std::size_t op_count = 1;
    for (std::size_t in_id : inputs) {
      op_count += nodes[in_id].op_count;
    }
    nodes.push_back(DAGNode<T> {id, op, std::vector<std::size_t>(inputs), payload, op_count, 0});
    return id;
  }

  std::vector<DAGNode<T>> nodes;
};

enum class VarKind : std::uint8_t
{
  Continuous,
  Integer,
  Binary
};

enum class Cmp : std::uint8_t
{
  LE,
  EQ
};

template<typename T>
class Expr
{
public:
  Expr(ExprDAG<T>* g, std::size_t id)
      : graph(g)
      , node_id(id)
  {
  }

  friend Expr operator+(Expr a, Expr b) { return a.binary(Op::Add, b); }

  friend Expr operator-(Expr a, Expr b) { return a.binary(Op::Sub, b); }

  friend Expr operator-(Expr a)
  {
    if (a.is_const()) {
      return a.constant(-a.const_value());
    }
    return {a.graph, a.graph->emit(Op::Neg, {a.node_id})};
  }

  friend Expr operator*(Expr a, Expr b)
  {
    if (a.node_id == b.node_id) {
      if (a.is_const()) {
        return a.constant(a.const_value() * a.const_value());
      }
      return {a.graph, a.graph->emit(Op::Sqr, {a.node_id})};
    }
    return a.binary(Op::Mul, b);
  }

  friend Expr operator/(Expr a, Expr b) { return a.binary(Op::Div, b); }

  friend Expr operator+(Expr a, T b) { return a + a.constant(b); }
  friend Expr operator+(T a, Expr b) { return b.constant(a) + b; }
  friend Expr operator-(Expr a, T b) { return a - a.constant(b); }
  friend Expr operator-(T a, Expr b) { return b.constant(a) - b; }
  friend Expr operator*(Expr a, T b) { return a * a.constant(b); }
  friend Expr operator*(T a, Expr b) { return b.constant(a) * b; }
  friend Expr operator/(Expr a, T b) { return a / a.constant(b); }
  friend Expr operator/(T a, Expr b) { return b.constant(a) / b; }

  friend Expr sqr(Expr a) { return a * a; }

  friend Expr sqrt(Expr a) { return a.unary(Op::Sqrt, [](T v) { return std::sqrt(v); }); }
  friend Expr exp(Expr a) { return a.unary(Op::Exp, [](T v) { return std::exp(v); }); }
  friend Expr log(Expr a) { return a.unary(Op::Log, [](T v) { return std::log(v); }); }
  friend Expr sin(Expr a) { return a.unary(Op::Sin, [](T v) { return std::sin(v); }); }
  friend Expr cos(Expr a) { return a.unary(Op::Cos, [](T v) { return std::cos(v); }); }
  friend Expr tanh(Expr a) { return a.unary(Op::Tanh, [](T v) { return std::tanh(v); }); }
  friend Expr abs(Expr a) { return a.unary(Op::Abs, [](T v) { return std::fabs(v); }); }

  friend Expr pow(Expr a, int n)
  {
    if (a.is_const()) {
      return a.constant(static_cast<T>(std::pow(a.const_value(), n)));
    }
    DAGNodePayload<T> p;
    p.int_exp = n;
    return {a.graph, a.graph->emit(Op::PowN, {a.node_id}, p)};
  }

  friend Expr pow(Expr a, Expr b)
  {
    if (a.is_const() && b.is_const()) {
      return a.constant(static_cast<T>(std::pow(a.const_value(), b.const_value())));
    }
    return a.binary(Op::Pow, b);
  }

  friend Expr min(Expr a, Expr b) { return a.binary(Op::Min, b); }
  friend Expr max(Expr a, Expr b) { return a.binary(Op::Max, b); }

  std::size_t id() const { return node_id; }

private:
  ExprDAG<T>* graph;
  std::size_t node_id;

  bool is_const() const { return graph->nodes[node_id].op == Op::Const; }

  T const_value() const { return graph->nodes[node_id].payload.constant; }

  Expr constant(T c) const
  {
    DAGNodePayload<T> p;
    p.constant = c;
    return {graph, graph->emit(Op::Const, {}, p)};
  }

  Expr binary(Op op, Expr b) const { return {graph, graph->emit(op, {node_id, b.node_id})}; }

  template<class F>
  Expr unary(Op op, F fold) const
  {
    if (is_const()) {
      return constant(fold(const_value()));
    }
    return {graph, graph->emit(op, {node_id})};
  }
};

template<typename T>
struct ConstraintRef
{
  std::size_t root_id;
  Cmp cmp;
  T rhs;
};

/// The full problem IR. Move-only, for the reason the reference gives
/// (problem.hpp:195-207): every Expr points at this object's graph.
template<typename T = double>
struct Problem
{
  Problem() = default;
  Problem(const Problem&) = delete;
  Problem& operator=(const Problem&) = delete;
  Problem(Problem&&) = default;
  Problem& operator=(Problem&&) = default;

  Expr<T> var(T lb, T ub, VarKind kind = VarKind::Continuous)
  {
    std::size_t const idx = box_bounds.size();
    box_bounds.push_back(basic_interval<T> {lb, ub});
    var_kinds.push_back(kind);
    DAGNodePayload<T> p;
    p.var_index = idx;
    return Expr<T>(&graph, graph.emit(Op::Var, {}, p));
  }

  Expr<T> var(VarKind kind = VarKind::Continuous)
  {
    return var(std::numeric_limits<T>::lowest(), std::numeric_limits<T>::max(), kind);
  }

  Expr<T> int_var(T lb, T ub) { return var(lb, ub, VarKind::Integer); }

  Expr<T> bin_var() { return var(T(0), T(1), VarKind::Binary); }

  Expr<T> fixed(T value)
  {
    DAGNodePayload<T> p;
    p.constant = value;
    return Expr<T>(&graph, graph.emit(Op::Const, {}, p));
  }

  Expr<T> expr(std::size_t id) { return Expr<T>(&graph, id); }

  void set_objective(Expr<T> e) { objective_root = e.id(); }

  void add_constraint(Expr<T> lhs, Cmp cmp, T rhs) { constraints.push_back({lhs.id(), cmp, rhs}); }

  std::size_t n_vars() const { return box_bounds.size(); }

  /// Structural checks, host-only, throwing on the first violation. Same
  /// checks and exception types as the reference (problem.hpp:264-351), plus
  /// operand arity and id-range checks that a flat DAG handed over the C-ABI
  /// needs and an Expr-built one satisfies by construction.
  void validate() const
  {
    std::size_t const n = graph.nodes.size();
    for (std::size_t i = 0; i < n; ++i) {
      DAGNode<T> const& node = graph.nodes[i];
      if (node.id != i) {
        throw InvalidDAG("DAG node id does not match its position in the node list");
      }
      if (node.in.size() != op_arity(node.op)) {
        throw InvalidDAG("DAG node has the wrong number of operands for its Op");
      }
      for (std::size_t in_id : node.in) {
        if (in_id >= node.id) {
          throw InvalidDAG(
              "DAG node operand id is not strictly less than the node's own id; "
              "topological-order invariant violated");
        }
      }
      std::size_t expected = 1;
      for (std::size_t in_id : node.in) {
        expected += graph.nodes[in_id].op_count;
      }
      if (node.op_count != expected) {
        throw InvalidDAG("DAG node op_count does not match the size of its subtree");
      }
      if (node.op == Op::Var && node.payload.var_index >= box_bounds.size()) {
        throw InvalidDAG("Op::Var node names a variable index past the box");
      }
    }

    std::size_t var_nodes = 0;
    for (DAGNode<T> const& node : graph.nodes) {
      if (node.op == Op::Var) {
        ++var_nodes;
      }
    }
    if (var_nodes != box_bounds.size() || box_bounds.size() != var_kinds.size()) {
      throw InvalidDAG("box_bounds/var_kinds are out of sync with the number of Op::Var nodes");
    }

    if (objective_root < n && graph.nodes[objective_root].op == Op::Const) {
      throw InvalidDAG("objective root is a bare Const");
    }
    for (ConstraintRef<T> const& c : constraints) {
      if (c.root_id >= n) {
        throw InvalidDAG("constraint root id out of range");
      }
      if (graph.nodes[c.root_id].op == Op::Const) {
        throw InvalidDAG("constraint root is a bare Const");
      }
    }

    if (var_kinds.empty()) {
      throw InvalidProblem("Problem has zero variables");
    }
    if (objective_root == std::numeric_limits<std::size_t>::max()) {
      throw InvalidProblem("Problem's objective was never set (set_objective was never called)");
    }
    if (objective_root >= n) {
      throw InvalidDAG("objective root id out of range");
    }
    for (std::size_t i = 0; i < box_bounds.size(); ++i) {
      basic_interval<T> const& b = box_bounds[i];
      if (!(b.lb <= b.ub)) {
        throw InvalidProblem("variable's lower bound exceeds its upper bound");
      }
      if (var_kinds[i] == VarKind::Integer || var_kinds[i] == VarKind::Binary) {
        if (std::floor(b.lb) < b.lb || std::floor(b.ub) < b.ub) {
          throw InvalidProblem("Integer/Binary variable has non-integer-valued bounds");
        }
      }
      if (var_kinds[i] == VarKind::Binary && (b.lb < T(0) || b.lb > T(0) || b.ub < T(1) || b.ub > T(1))) {
        throw InvalidProblem("Binary variable's bounds are not exactly [0,1]");
      }
    }
    for (ConstraintRef<T> const& c : constraints) {
      if (!std::isfinite(c.rhs)) {
        throw InvalidProblem("constraint rhs must be finite");
      }
    }
  }

  ExprDAG<T> graph;
  std::vector<basic_interval<T>> box_bounds;
  std::vector<VarKind> var_kinds;
  std::size_t objective_root = std::numeric_limits<std::size_t>::max();
  std::vector<ConstraintRef<T>> constraints;
};

/// Round-to-nearest value of every node at `point` (host, not a bound).
/// Same op semantics as the reference's evaluate_prefix (model/eval.hpp:38-116).
template<typename T>
std::vector<T> evaluate_all(ExprDAG<T> const& graph, std::vector<T> const& point)
{
  std::vector<T> value(graph.nodes.size(), T {});
  for (std::size_t i = 0; i < graph.nodes.size(); ++i) {
    DAGNode<T> const& node = graph.nodes[i];
    auto arg = [&](std::size_t k) { return value[node.in[k]]; };
    switch (node.op) {
      case Op::Var:
        if (node.payload.var_index >= point.size()) {
          throw ShapeMismatch("evaluate: point has too few components");
        }
        value[i] = point[node.payload.var_index];
        break;
      case Op::Const: value[i] = node.payload.constant; break;
      case Op::Add: value[i] = arg(0) + arg(1); break;
      case Op::Sub: value[i] = arg(0) - arg(1); break;
      case Op::Mul: value[i] = arg(0) * arg(1); break;
      case Op::Div: value[i] = arg(0) / arg(1); break;
      case Op::Min: value[i] = std::fmin(arg(0), arg(1)); break;
      case Op::Max: value[i] = std::fmax(arg(0), arg(1)); break;
      case Op::Neg: value[i] = -arg(0); break;
      case Op::Sqr: value[i] = arg(0) * arg(0); break;
      case Op::Exp: value[i] = std::exp(arg(0)); break;
      case Op::Log: value[i] = std::log(arg(0)); break;
      case Op::Sqrt: value[i] = std::sqrt(arg(0)); break;
      case Op::Sin: value[i] = std::sin(arg(0)); break;
      case Op::Cos: value[i] = std::cos(arg(0)); break;
      case Op::Tanh: value[i] = std::tanh(arg(0)); break;
      case Op::Abs: value[i] = std::fabs(arg(0)); break;
      case Op::PowN: value[i] = std::pow(arg(0), node.payload.int_exp); break;
      case Op::Pow: value[i] = std::pow(arg(0), arg(1)); break;
    }
  }
  return value;
}

/// `lhs cmp rhs` with the device point predicate's tolerances: LE exact, EQ
/// absolute 1e-8 or relative 1e-6 (reference kernels.cuh:283-320, eval.hpp:186-200).
template<typename T>
bool satisfies(T lhs, Cmp cmp, T rhs, T abs_tol = T(1e-8), T rel_tol = T(1e-6))
{
  if (cmp == Cmp::LE) {
    return lhs <= rhs;
  }
  if (cmp == Cmp::EQ) {
    T const diff = std::fabs(lhs - rhs);
    if (diff <= abs_tol) {
      return true;
    }
    return diff <= rel_tol * std::fmax(std::fabs(lhs), std::fabs(rhs));
  }
  return false;
}

template<typename T>
std::optional<std::size_t> first_violated_constraint(Problem<T> const& problem,
                                                     std::vector<T> const& point)
{
  std::vector<T> const value = evaluate_all(problem.graph, point);
  for (std::size_t i = 0; i < problem.constraints.size(); ++i) {
    ConstraintRef<T> const& c = problem.constraints[i];
    if (!satisfies(value[c.root_id], c.cmp, c.rhs)) {
      return i;
    }
  }
  return std::nullopt;
}

}  // namespace cmi::model
