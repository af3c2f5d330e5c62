#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/model.hpp"

// The DAG tape: a topologically ordered, register-allocated instruction list
// that one device lane interprets per box (or per sample point). It replaces
// the reference's one-kernel-per-DAG-node CUDA graph (backend/graph/
// builder.cuh:261-514) while evaluating exactly the same node operations:
// every reachable non-Const node is one interval operation applied to that
// node's operands, a Const operand is an immediate consumed by value (as
// wire_binary passes it, builder.cuh:412-455), and Const (+) Const is
// evaluated on point intervals (binary_op_scalar_scalar_kernel,
// kernels.cuh:240-253).
//
// Evaluation order among independent nodes does not change any node's value,
// so the compiler schedules for the interpreter:
//
//  * Left-deep sums and differences (the shape every polynomial takes) run in
//    an accumulator register: the chain's running value never touches LDS,
//    and each term -- a variable or shared value, its square, a product of two,
//    optionally times a constant -- is evaluated inside the accumulate
//    instruction that consumes it. Each term instruction performs exactly the
//    node operations of the DAG nodes it covers, in DAG order.
//  * "Leaf" nodes (all operands Var/Const) are re-evaluated at each use
//    instead of being held live, which bounds the live set by the variables
//    plus a few temporaries.
//  * Values are emitted into virtual registers and mapped onto LDS slots by a
//    linear scan afterwards (variables are pinned to slots [0, n_vars)).
//  * Constraint checks and the objective read-out are tape instructions, so a
//    root dies as soon as it has been checked.
namespace cmi::tape
{

enum class TOp : std::uint8_t
{
  Nop = 0,
  LoadConst,  // dst <- point [imm, imm]
  Add,        // dst <- a + b
  AddImm,     // dst <- a + imm            (Const on either side; IEEE + commutes)
  Sub,        // dst <- a - b
  SubImm,     // dst <- a - imm
  ImmSub,     // dst <- imm - a
  Mul,        // dst <- a * b
  MulImm,     // dst <- a * imm            (Const on either side)
  Sqr,        // dst <- a * a
  Neg,        // dst <- -a
  CheckLE,    // feasible &= a <= imm (interval: lb <= imm)
  CheckEQ,    // feasible &= a == imm (interval: lb <= imm <= ub; point: almost_equal)
  Objective,  // objective <- a
  Output,     // roots_out[b] <- a (verification mode only)
  // Accumulator register (ACC) instructions; the term instructions follow.
  AccStore = 16,  // dst <- ACC
  AccAddImm,      // ACC <- ACC + imm
  AccSubImm,      // ACC <- ACC - imm
  AccImmSub,      // ACC <- imm - ACC
  AccScale,       // ACC <- ACC * imm
  AccNeg,         // ACC <- -ACC
  AccCheckLE,     // feasible &= ACC <= imm
  AccCheckEQ,     // feasible &= ACC == imm
  AccObjective,   // objective <- ACC
  AccOutput,      // roots_out[b] <- ACC
  // Extended op set (interpreted by the extended kernels only).
  Div = 32,   // dst <- a / b (IEEE 1788; division by [0, 0] empties the box)
  DivImm,     // dst <- a / imm
  ImmDiv,     // dst <- imm / a
  PowN,       // dst <- a^n, n = imm (an integer)
  Sqrt,       // dst <- sqrt(a) on a's part in [0, inf)
  Exp,        // dst <- exp(a)
  Log,        // dst <- log(a) on a's part in (0, inf)
  Sin,        // dst <- sin(a)
  Cos,        // dst <- cos(a)
  Tanh,       // dst <- tanh(a)
  Abs,        // dst <- |a|
  Min,        // dst <- min(a, b)
  MinImm,     // dst <- min(a, imm)
  Max,        // dst <- max(a, b)
  MaxImm,     // dst <- max(a, imm)
  Count
};

/// First and one-past-last op code of the extended op set.
inline constexpr std::uint8_t kExtBase = 32;
inline constexpr std::uint8_t kExtEnd = 47;
inline bool is_ext_op(std::uint8_t op) { return op >= kExtBase && op < kExtEnd; }

// Term instructions: op = kTermBase + mode * 8 + term.
inline constexpr std::uint8_t kTermBase = 64;
enum class Term : std::uint8_t
{
  S = 0,    // a
  SC = 1,   // a * imm
  SS = 2,   // a * b
  Q = 3,    // a * a (Sqr)
  SSC = 4,  // (a * b) * imm
  QC = 5,   // (a * a) * imm
};
inline constexpr std::uint8_t kTerms = 6;
enum class AccMode : std::uint8_t
{
  Set = 0,  // ACC <- term
  Add = 1,  // ACC <- ACC + term
  Sub = 2,  // ACC <- ACC - term
};

inline std::uint8_t term_op(AccMode m, Term t)
{
  return static_cast<std::uint8_t>(kTermBase + static_cast<std::uint8_t>(m) * 8 + static_cast<std::uint8_t>(t));
}

inline bool is_term_op(std::uint8_t op) { return op >= kTermBase && op < kTermBase + 24 && (op & 7) < kTerms; }

inline std::string op_name(std::uint8_t op)
{
  static const char* base[] = {"nop", "ldc", "add", "addi", "sub", "subi", "isub", "mul", "muli", "sqr",
                               "neg", "chkle", "chkeq", "obj", "out"};
  static const char* acc[] = {"acc.st", "acc.addi", "acc.subi", "acc.isub", "acc.scale",
                              "acc.neg", "acc.chkle", "acc.chkeq", "acc.obj", "acc.out"};
  static const char* modes[] = {"set", "add", "sub"};
  static const char* terms[] = {"s", "sc", "ss", "q", "ssc", "qc"};
  if (op < 15) {
    return base[op];
  }
  if (op >= 16 && op < 26) {
    return acc[op - 16];
  }
  if (is_term_op(op)) {
    unsigned const k = op - kTermBase;
    return std::string("acc.") + modes[k >> 3] + "." + terms[k & 7];
  }
  static const char* ext[] = {"div", "divi", "idiv", "pown", "sqrt", "exp", "log", "sin",
                              "cos", "tanh", "abs", "min", "mini", "max", "maxi"};
  if (is_ext_op(op)) {
    return ext[op - kExtBase];
  }
  return "?";
}

/// One 16-byte tape word: read whole by a scalar load, uniform across a wave.
struct Instr
{
  std::uint8_t op;
  std::uint8_t pad;
  std::uint16_t dst;
  std::uint16_t a;
  std::uint16_t b;
  double imm;
};
static_assert(sizeof(Instr) == 16, "tape words are 16 bytes");

/// Algorithmic FP64 work of one instruction: each directed-rounded product or
/// sum counts once (an interval + is 2, an interval x interval product is 8, a
/// scalar product or a square is 2, a feasibility compare 1); selects, case
/// compares and the rounding-mode handling are not counted.
inline std::uint32_t interval_flops(std::uint8_t op)
{
  if (is_term_op(op)) {
    unsigned const k = op - kTermBase;
    unsigned const mode = k >> 3;
    Term const t = static_cast<Term>(k & 7);
    std::uint32_t f = mode == 0 ? 0 : 2;
    switch (t) {
      case Term::S: break;
      case Term::SC: f += 2; break;
      case Term::SS: f += 8; break;
      case Term::Q: f += 2; break;
      case Term::SSC: f += 10; break;
      case Term::QC: f += 4; break;
    }
    return f;
  }
  switch (static_cast<TOp>(op)) {
    case TOp::Add:
    case TOp::AddImm:
    case TOp::Sub:
    case TOp::SubImm:
    case TOp::ImmSub:
    case TOp::MulImm:
    case TOp::Sqr:
    case TOp::AccAddImm:
    case TOp::AccSubImm:
    case TOp::AccImmSub:
    case TOp::AccScale:
      return 2;
    case TOp::Mul:
      return 8;
    case TOp::CheckLE:
    case TOp::AccCheckLE:
      return 1;
    case TOp::CheckEQ:
    case TOp::AccCheckEQ:
      return 2;
    // extended ops: the directed FP64 operations of their enclosures
    case TOp::Div:
    case TOp::DivImm:
    case TOp::ImmDiv:
      return 24;
    case TOp::Sqrt:
      return 30;
    case TOp::PowN:
      return 12;
    case TOp::Exp:
      return 44;
    case TOp::Log:
      return 100;
    case TOp::Sin:
    case TOp::Cos:
      return 110;
    case TOp::Tanh:
      return 80;
    default:
      return 0;
  }
}

struct Tape
{
  std::vector<Instr> code;
  std::uint32_t n_vars = 0;      // slots [0, n_vars) hold the variables
  std::uint32_t n_slots = 0;     // per-lane slots the tape needs
  std::uint32_t n_outputs = 0;   // Output instructions: 1 + n_constraints
  std::uint32_t n_constraints = 0;
  std::uint32_t reachable_nodes = 0;  // non-Const DAG nodes the roots reach
  std::uint32_t node_ops = 0;         // DAG node operations the tape performs (leaves counted per use)
  std::uint64_t flops = 0;            // interval_flops summed over the tape
  // Function groups: group g is code[group_begin[g], group_begin[g+1]) and
  // uses the temporary slots [n_vars + g*temps_per_group, ...+temps_per_group).
  std::uint32_t n_groups = 1;
  std::vector<std::uint32_t> group_begin;
  std::uint32_t temps_per_group = 0;
  std::uint32_t objective_group = 0;  // the group holding the Objective instruction
  bool extended = false;              // uses the extended op set
};

struct CompileOptions
{
  bool rematerialise_leaves = true;
  bool accumulate = true;    // chains in the accumulator register, fused terms
  bool emit_outputs = true;  // Output instructions for the objective and each constraint root
};

namespace detail
{

/// Instruction with virtual register operands (-1: unused).
struct VInstr
{
  std::uint8_t op;
  std::int64_t dst = -1, a = -1, b = -1;
  double imm = 0.0;
};

class Compiler
{
public:
  Compiler(const model::Problem<double>& p, const CompileOptions& opt)
      : p_(p)
      , nodes_(p.graph.nodes)
      , opt_(opt)
  {
  }

  /// Compile the functions whose root indices (0 = objective, 1 + j =
  /// constraint j) are listed in `selected`, in that order; all of them when
  /// `selected` is empty.
  Tape run(std::vector<std::size_t> selected = {})
  {
    using model::Op;
    std::size_t const n = nodes_.size();
    std::size_t const n_vars = p_.box_bounds.size();
    if (n_vars > 30000) {
      throw InvalidConfiguration("problem too large for the 16-bit tape slot encoding");
    }
    roots_.push_back(p_.objective_root);
    for (auto const& c : p_.constraints) {
      roots_.push_back(c.root_id);
    }
    if (selected.empty()) {
      for (std::size_t k = 0; k < roots_.size(); ++k) {
        selected.push_back(k);
      }
    }
    selected_ = selected;
    analyse();

    next_vreg_ = static_cast<std::int64_t>(n_vars);
    cache_.assign(n, -1);
    for (std::size_t k : selected_) {
      std::size_t const r = roots_[k];
      bool const is_obj = k == 0;
      double const rhs = is_obj ? 0.0 : p_.constraints[k - 1].rhs;
      bool const le = is_obj || p_.constraints[k - 1].cmp == model::Cmp::LE;
      if (opt_.accumulate && acc_capable(r) && parents_[r] == 1 && cache_[r] < 0) {
        emit_acc(r);
        if (opt_.emit_outputs) {
          emit({static_cast<std::uint8_t>(TOp::AccOutput), -1, -1, -1, 0.0}, static_cast<std::int64_t>(k));
        }
        if (is_obj) {
          emit({static_cast<std::uint8_t>(TOp::AccObjective), -1, -1, -1, 0.0});
        } else {
          emit({static_cast<std::uint8_t>(le ? TOp::AccCheckLE : TOp::AccCheckEQ), -1, -1, -1, rhs});
        }
        continue;
      }
      std::int64_t const v = emit_slot(r);
      if (opt_.emit_outputs) {
        emit({static_cast<std::uint8_t>(TOp::Output), -1, v, -1, 0.0}, static_cast<std::int64_t>(k));
      }
      if (is_obj) {
        emit({static_cast<std::uint8_t>(TOp::Objective), -1, v, -1, 0.0});
      } else {
        emit({static_cast<std::uint8_t>(le ? TOp::CheckLE : TOp::CheckEQ), -1, v, -1, rhs});
      }
    }
    return allocate(n_vars);
  }

private:
  using Op = model::Op;

  bool is_var(std::size_t id) const { return nodes_[id].op == Op::Var; }
  /// Ops cheap enough to re-evaluate at every use rather than keep live.
  static bool cheap(Op op)
  {
    return op == Op::Add || op == Op::Sub || op == Op::Mul || op == Op::Sqr || op == Op::Neg;
  }
  bool is_const(std::size_t id) const { return nodes_[id].op == Op::Const; }
  double cval(std::size_t id) const { return nodes_[id].payload.constant; }

  /// Whether node `nd` (reachable, not a Const) is an op the device tape
  /// evaluates; throws the error the compile reports for it otherwise.
  static void check_op(const model::DAGNode<double>& nd, const std::vector<model::DAGNode<double>>& nodes)
  {
    switch (nd.op) {
      case Op::Var:
      case Op::Add:
      case Op::Sub:
      case Op::Mul:
      case Op::Div:
      case Op::Min:
      case Op::Max:
      case Op::Pow:
        return;
      case Op::Sqr:
      case Op::Neg:
      case Op::Sqrt:
      case Op::Exp:
      case Op::Log:
      case Op::Sin:
      case Op::Cos:
      case Op::Tanh:
      case Op::Abs:
      case Op::PowN:
        if (nodes[nd.in[0]].op == Op::Const) {
          throw InvalidDAG("unary op applied directly to an Op::Const");
        }
        return;
      default:
        throw InvalidConfiguration(std::string("Op::") + model::op_name(nd.op)
                                   + " is not implemented by the device tape evaluator");
    }
  }

  /// Reachability and use counts of the selected functions' cone. The cone
  /// is found from the roots down, so a function group's compile costs its
  /// cone, not the whole DAG; the counts do not depend on the visiting
  /// order. A DAG with an op the tape cannot evaluate reports the one with
  /// the highest node index, as a scan from the top would.
  void analyse()
  {
    std::size_t const n = nodes_.size();
    reach_.assign(n, 0);
    parents_.assign(n, 0);
    leaf_.assign(n, 0);
    cone_.clear();
    for (std::size_t k : selected_) {
      std::size_t const r = roots_[k];
      ++parents_[r];
      if (!reach_[r]) {
        reach_[r] = 1;
        cone_.push_back(r);
      }
    }
    for (std::size_t q = 0; q < cone_.size(); ++q) {
      for (std::size_t in : nodes_[cone_[q]].in) {
        ++parents_[in];
        if (!reach_[in]) {
          reach_[in] = 1;
          cone_.push_back(in);
        }
      }
    }
    bool valid = true;
    for (std::size_t i : cone_) {
      auto const& nd = nodes_[i];
      if (nd.op == Op::Const) {
        continue;
      }
      ++reachable_;
      try {
        check_op(nd, nodes_);
      } catch (...) {
        valid = false;
      }
      if (nd.op == Op::Var) {
        continue;
      }
      bool all = true;
      for (std::size_t in : nd.in) {
        all = all && (is_var(in) || is_const(in));
      }
      leaf_[i] = all && opt_.rematerialise_leaves && cheap(nd.op);
    }
    if (!valid) {
      std::sort(cone_.begin(), cone_.end(), std::greater<>());
      for (std::size_t i : cone_) {
        if (nodes_[i].op != Op::Const) {
          check_op(nodes_[i], nodes_);
        }
      }
    }
  }

  void emit(VInstr v, std::int64_t output_index = -1)
  {
    if (output_index >= 0) {
      v.b = -2 - output_index;  // encoded literal, not a register
    }
    code_.push_back(v);
  }

  std::int64_t fresh() { return next_vreg_++; }

  // -- generic slot evaluation -------------------------------------------

  /// Virtual register holding node `id`'s value (emitting it if needed).
  std::int64_t emit_slot(std::size_t id)
  {
    auto const& nd = nodes_[id];
    if (nd.op == Op::Var) {
      return static_cast<std::int64_t>(nd.payload.var_index);
    }
    if (cache_[id] >= 0) {
      return cache_[id];
    }
    std::int64_t v;
    if (opt_.accumulate && !leaf_[id] && acc_capable(id)) {
      emit_acc(id);
      v = fresh();
      emit({static_cast<std::uint8_t>(TOp::AccStore), v, -1, -1, 0.0});
    } else {
      v = emit_generic(id);
    }
    ++node_ops_;
    if (!leaf_[id]) {
      cache_[id] = v;
    }
    return v;
  }

  std::int64_t emit_generic(std::size_t id)
  {
    auto const& nd = nodes_[id];
    std::size_t const arity = nd.in.size();
    bool const c0 = is_const(nd.in[0]);
    bool const c1 = arity > 1 && is_const(nd.in[1]);
    std::int64_t a = -1, b = -1;
    if ((c0 && c1) || (c0 && nd.op == Op::Pow)) {
      a = fresh();
      emit({static_cast<std::uint8_t>(TOp::LoadConst), a, -1, -1, cval(nd.in[0])});
    } else if (!c0) {
      a = emit_slot(nd.in[0]);
    }
    if (arity > 1 && !c1) {
      b = emit_slot(nd.in[1]);
    }
    std::int64_t const d = fresh();
    auto out = [&](TOp op, std::int64_t x, std::int64_t y, double imm) {
      emit({static_cast<std::uint8_t>(op), d, x, y, imm});
    };
    switch (nd.op) {
      case Op::Add:
        if (c1) {
          out(TOp::AddImm, a, -1, cval(nd.in[1]));
        } else if (c0) {
          out(TOp::AddImm, b, -1, cval(nd.in[0]));
        } else {
          out(TOp::Add, a, b, 0.0);
        }
        break;
      case Op::Sub:
        if (c1) {
          out(TOp::SubImm, a, -1, cval(nd.in[1]));
        } else if (c0) {
          out(TOp::ImmSub, b, -1, cval(nd.in[0]));
        } else {
          out(TOp::Sub, a, b, 0.0);
        }
        break;
      case Op::Mul:
        if (c1) {
          out(TOp::MulImm, a, -1, cval(nd.in[1]));
        } else if (c0) {
          out(TOp::MulImm, b, -1, cval(nd.in[0]));
        } else {
          out(TOp::Mul, a, b, 0.0);
        }
        break;
      case Op::Sqr: out(TOp::Sqr, a, -1, 0.0); break;
      case Op::Neg: out(TOp::Neg, a, -1, 0.0); break;
      case Op::Div:
        if (c1) {
          out(TOp::DivImm, a, -1, cval(nd.in[1]));
        } else if (c0) {
          out(TOp::ImmDiv, b, -1, cval(nd.in[0]));
        } else {
          out(TOp::Div, a, b, 0.0);
        }
        break;
      case Op::Min:
      case Op::Max: {
        bool const mn = nd.op == Op::Min;
        if (c1) {
          out(mn ? TOp::MinImm : TOp::MaxImm, a, -1, cval(nd.in[1]));
        } else if (c0) {
          out(mn ? TOp::MinImm : TOp::MaxImm, b, -1, cval(nd.in[0]));
        } else {
          out(mn ? TOp::Min : TOp::Max, a, b, 0.0);
        }
        break;
      }
      case Op::PowN: out(TOp::PowN, a, -1, static_cast<double>(nd.payload.int_exp)); break;
      case Op::Sqrt: out(TOp::Sqrt, a, -1, 0.0); break;
      case Op::Exp: out(TOp::Exp, a, -1, 0.0); break;
      case Op::Log: out(TOp::Log, a, -1, 0.0); break;
      case Op::Sin: out(TOp::Sin, a, -1, 0.0); break;
      case Op::Cos: out(TOp::Cos, a, -1, 0.0); break;
      case Op::Tanh: out(TOp::Tanh, a, -1, 0.0); break;
      case Op::Abs: out(TOp::Abs, a, -1, 0.0); break;
      case Op::Pow: {
        // x**y = exp(y log x), in three words through two temporaries
        std::int64_t const l = fresh();
        emit({static_cast<std::uint8_t>(TOp::Log), l, a, -1, 0.0});
        std::int64_t const m = fresh();
        if (c1) {
          emit({static_cast<std::uint8_t>(TOp::MulImm), m, l, -1, cval(nd.in[1])});
        } else {
          emit({static_cast<std::uint8_t>(TOp::Mul), m, l, b, 0.0});
        }
        out(TOp::Exp, m, -1, 0.0);
        break;
      }
      default: throw InvalidConfiguration("unsupported op reached the tape emitter");
    }
    return d;
  }

  // -- accumulator chains -----------------------------------------------

  /// A node whose value can be produced in ACC by accumulator instructions.
  bool acc_capable(std::size_t id) const
  {
    auto const& nd = nodes_[id];
    switch (nd.op) {
      case Op::Add:
      case Op::Sub:
        return !(is_const(nd.in[0]) && is_const(nd.in[1]));
      case Op::Mul:
        return is_const(nd.in[0]) != is_const(nd.in[1]);
      case Op::Neg:
        return true;
      default:
        return false;
    }
  }

  /// Consumed only here and not yet materialised: may live in ACC or be folded
  /// into a term.
  bool private_value(std::size_t id) const { return parents_[id] == 1 && cache_[id] < 0 && !is_var(id); }

  struct TermRef
  {
    Term kind;
    std::int64_t a = -1, b = -1;
    double imm = 0.0;
    std::uint32_t ops = 0;  // DAG node operations the term performs
  };

  /// Operand of a term: a variable or an already materialised value. Emits
  /// nothing unless `allow_emit` (then the value is materialised first).
  bool term_operand(std::size_t id, bool allow_emit, std::int64_t& v)
  {
    if (is_const(id)) {
      return false;
    }
    if (is_var(id)) {
      v = static_cast<std::int64_t>(nodes_[id].payload.var_index);
      return true;
    }
    if (cache_[id] >= 0) {
      v = cache_[id];
      return true;
    }
    if (allow_emit) {
      v = emit_slot(id);
      return true;
    }
    return false;
  }

  /// Product pattern (a*b or a*a) on term operands; emits operand values first
  /// when they are neither variables nor materialised (ACC must be free).
  bool product_term(std::size_t id, TermRef& t)
  {
    auto const& nd = nodes_[id];
    if (!(leaf_[id] || private_value(id))) {
      return false;
    }
    if (nd.op == Op::Sqr) {
      std::int64_t a;
      if (!term_operand(nd.in[0], true, a)) {
        return false;
      }
      t = {Term::Q, a, -1, 0.0, 1};
      return true;
    }
    if (nd.op == Op::Mul && !is_const(nd.in[0]) && !is_const(nd.in[1])) {
      std::int64_t a, b;
      if (!term_operand(nd.in[0], true, a) || !term_operand(nd.in[1], true, b)) {
        return false;
      }
      t = {Term::SS, a, b, 0.0, 1};
      return true;
    }
    return false;
  }

  /// The value of node `id` as a term of an accumulate instruction. Must be
  /// called while ACC is free: it may emit (materialise) operands.
  TermRef prepare_term(std::size_t id)
  {
    auto const& nd = nodes_[id];
    TermRef t;
    if (is_var(id) || cache_[id] >= 0) {
      std::int64_t v;
      term_operand(id, false, v);
      return {Term::S, v, -1, 0.0, 0};
    }
    if (product_term(id, t)) {
      return t;
    }
    if ((leaf_[id] || private_value(id)) && nd.op == Op::Mul && is_const(nd.in[0]) != is_const(nd.in[1])) {
      std::size_t const x = is_const(nd.in[0]) ? nd.in[1] : nd.in[0];
      double const c = is_const(nd.in[0]) ? cval(nd.in[0]) : cval(nd.in[1]);
      TermRef inner;
      if (is_var(x) || cache_[x] >= 0) {
        std::int64_t v;
        term_operand(x, false, v);
        return {Term::SC, v, -1, c, 1};
      }
      if (product_term(x, inner)) {
        return {inner.kind == Term::Q ? Term::QC : Term::SSC, inner.a, inner.b, c, 2};
      }
    }
    std::int64_t const v = emit_slot(id);
    return {Term::S, v, -1, 0.0, 0};
  }

  void emit_term(AccMode m, const TermRef& t)
  {
    emit({term_op(m, t.kind), -1, t.a, t.b, t.imm});
    node_ops_ += t.ops;
  }

  /// ACC <- value of `x` (ACC is free on entry).
  void acc_from(std::size_t x)
  {
    if (acc_capable(x) && private_value(x)) {
      emit_acc(x);
      return;
    }
    emit_term(AccMode::Set, prepare_term(x));
  }

  /// ACC <- value of acc-capable node `id` (ACC is free on entry).
  void emit_acc(std::size_t id)
  {
    auto const& nd = nodes_[id];
    auto simple = [&](TOp op, double imm) {
      emit({static_cast<std::uint8_t>(op), -1, -1, -1, imm});
      ++node_ops_;
    };
    switch (nd.op) {
      case Op::Add:
      case Op::Sub: {
        std::size_t const l = nd.in[0], r = nd.in[1];
        if (is_const(r)) {
          acc_from(l);
          simple(nd.op == Op::Add ? TOp::AccAddImm : TOp::AccSubImm, cval(r));
          return;
        }
        if (is_const(l)) {
          acc_from(r);
          simple(nd.op == Op::Add ? TOp::AccAddImm : TOp::AccImmSub, cval(l));
          return;
        }
        // The chain runs through the left operand; for a sum whose left
        // operand is a term and right operand a chain, run the right one
        // (IEEE addition commutes bit for bit).
        std::size_t chain = l, other = r;
        if (nd.op == Op::Add && !(acc_capable(l) && private_value(l)) && acc_capable(r) && private_value(r)) {
          chain = r;
          other = l;
        }
        TermRef const t = prepare_term(other);  // may materialise; ACC still free
        acc_from(chain);
        emit_term(nd.op == Op::Add ? AccMode::Add : AccMode::Sub, t);
        ++node_ops_;
        return;
      }
      case Op::Mul: {
        std::size_t const x = is_const(nd.in[0]) ? nd.in[1] : nd.in[0];
        double const c = is_const(nd.in[0]) ? cval(nd.in[0]) : cval(nd.in[1]);
        TermRef inner;
        if (is_var(x) || cache_[x] >= 0) {
          std::int64_t v;
          term_operand(x, false, v);
          emit_term(AccMode::Set, {Term::SC, v, -1, c, 1});
          return;
        }
        if (!(acc_capable(x) && private_value(x)) && product_term(x, inner)) {
          emit_term(AccMode::Set, {inner.kind == Term::Q ? Term::QC : Term::SSC, inner.a, inner.b, c, 2});
          return;
        }
        acc_from(x);
        simple(TOp::AccScale, c);
        return;
      }
      case Op::Neg:
        acc_from(nd.in[0]);
        simple(TOp::AccNeg, 0.0);
        return;
      default:
        throw InvalidConfiguration("emit_acc on a node that cannot run in the accumulator");
    }
  }

  // -- register allocation ------------------------------------------------

  Tape allocate(std::size_t n_vars)
  {
    std::size_t const n_v = static_cast<std::size_t>(next_vreg_);
    std::vector<std::int64_t> last(n_v, -1);
    for (std::size_t i = 0; i < code_.size(); ++i) {
      for (std::int64_t r : {code_[i].a, code_[i].b}) {
        if (r >= 0) {
          last[static_cast<std::size_t>(r)] = static_cast<std::int64_t>(i);
        }
      }
    }
    std::vector<std::int64_t> phys(n_v, -1);
    for (std::size_t v = 0; v < n_vars; ++v) {
      phys[v] = static_cast<std::int64_t>(v);
    }
    std::vector<std::uint32_t> free_slots;  // sorted descending: back() is the lowest
    std::uint32_t next_slot = static_cast<std::uint32_t>(n_vars);
    Tape t;
    t.n_vars = static_cast<std::uint32_t>(n_vars);
    t.n_constraints = static_cast<std::uint32_t>(p_.constraints.size());
    t.reachable_nodes = reachable_;
    t.node_ops = node_ops_;
    for (std::size_t i = 0; i < code_.size(); ++i) {
      VInstr const& v = code_[i];
      Instr ins {};
      ins.op = v.op;
      ins.imm = v.imm;
      auto map = [&](std::int64_t r) -> std::uint16_t {
        if (r < 0) {
          return 0;
        }
        return static_cast<std::uint16_t>(phys[static_cast<std::size_t>(r)]);
      };
      ins.a = map(v.a);
      ins.b = v.b <= -2 ? static_cast<std::uint16_t>(-2 - v.b) : map(v.b);
      // operands whose last use is here are free before the destination is
      // chosen (the interpreter reads every operand before writing)
      for (std::int64_t r : {v.a, v.b}) {
        if (r >= static_cast<std::int64_t>(n_vars) && last[static_cast<std::size_t>(r)] == static_cast<std::int64_t>(i)) {
          std::uint32_t const s = static_cast<std::uint32_t>(phys[static_cast<std::size_t>(r)]);
          if (std::find(free_slots.begin(), free_slots.end(), s) == free_slots.end()) {
            free_slots.insert(std::upper_bound(free_slots.begin(), free_slots.end(), s, std::greater<>()), s);
          }
        }
      }
      if (v.dst >= 0) {
        std::uint32_t s;
        if (!free_slots.empty()) {
          s = free_slots.back();
          free_slots.pop_back();
        } else {
          s = next_slot++;
        }
        phys[static_cast<std::size_t>(v.dst)] = s;
        ins.dst = static_cast<std::uint16_t>(s);
        if (last[static_cast<std::size_t>(v.dst)] < 0) {
          // never read: the slot is free again after this instruction
          free_slots.insert(std::upper_bound(free_slots.begin(), free_slots.end(), s, std::greater<>()), s);
        }
      }
      t.code.push_back(ins);
      t.flops += interval_flops(ins.op);
      t.extended = t.extended || is_ext_op(ins.op);
    }
    if (next_slot > 65535) {
      throw InvalidConfiguration("tape needs more slots than the 16-bit encoding holds");
    }
    t.n_slots = std::max<std::uint32_t>(next_slot, 1);
    t.n_outputs = opt_.emit_outputs ? static_cast<std::uint32_t>(roots_.size()) : 0;
    return t;
  }

  const model::Problem<double>& p_;
  const std::vector<model::DAGNode<double>>& nodes_;
  CompileOptions opt_;
  std::vector<std::size_t> roots_;
  std::vector<std::size_t> selected_;
  std::vector<std::uint8_t> reach_;
  std::vector<std::uint8_t> leaf_;
  std::vector<std::size_t> cone_;  // reachable nodes, highest index first
  std::vector<std::uint32_t> parents_;
  std::vector<std::int64_t> cache_;
  std::vector<VInstr> code_;
  std::int64_t next_vreg_ = 0;
  std::uint32_t reachable_ = 0;
  std::uint32_t node_ops_ = 0;
};

}  // namespace detail

/// Compile a validated Problem into a Tape. Throws InvalidDAG for the shapes the
/// reference's builder rejects (a unary op on a Const) and InvalidConfiguration
/// for ops the device evaluator does not implement.
inline Tape compile(const model::Problem<double>& p, const CompileOptions& opt = {})
{
  Tape t = detail::Compiler(p, opt).run();
  t.group_begin = {0, static_cast<std::uint32_t>(t.code.size())};
  t.temps_per_group = t.n_slots - t.n_vars;
  return t;
}

namespace detail
{

/// Work of each function (0 = objective, 1 + j = constraint j) for balancing
/// function groups: the non-constant DAG nodes it reaches, a node shared by
/// several functions counted once for each (as each group evaluates it). One
/// backward sweep carries a bit per function over the topologically ordered
/// nodes, instead of one compile per function.
inline std::vector<std::size_t> root_work(const model::Problem<double>& p)
{
  auto const& nodes = p.graph.nodes;
  std::size_t const n_roots = 1 + p.constraints.size();
  std::size_t const words = (n_roots + 63) / 64;
  std::vector<std::uint64_t> bits(nodes.size() * words, 0);
  auto mark = [&](std::size_t node, std::size_t k) { bits[node * words + k / 64] |= std::uint64_t(1) << (k % 64); };
  mark(p.objective_root, 0);
  for (std::size_t j = 0; j < p.constraints.size(); ++j) {
    mark(p.constraints[j].root_id, 1 + j);
  }
  std::vector<std::size_t> work(n_roots, 0);
  for (std::size_t i = nodes.size(); i-- > 0;) {
    std::uint64_t const* const b = &bits[i * words];
    bool any = false;
    for (std::size_t w = 0; w < words; ++w) {
      any = any || b[w] != 0;
    }
    if (!any) {
      continue;
    }
    for (std::size_t in : nodes[i].in) {
      for (std::size_t w = 0; w < words; ++w) {
        bits[in * words + w] |= b[w];
      }
    }
    if (nodes[i].op == model::Op::Const) {
      continue;
    }
    for (std::size_t w = 0; w < words; ++w) {
      for (std::uint64_t x = b[w]; x != 0; x &= x - 1) {
        ++work[w * 64 + static_cast<std::size_t>(__builtin_ctzll(x))];
      }
    }
  }
  return work;
}

}  // namespace detail

/// Compile the objective and constraints into `n_groups` independent groups
/// (each one wave's program over the same box), balancing instruction counts
/// by longest-processing-time assignment. A node that several groups need is
/// evaluated by each of them. The groups' temporaries get disjoint slot ranges
/// so the waves of one workgroup share the variable slots and nothing else.
/// slot_period (0 = n_groups): group g's temporaries take range g mod
/// slot_period, for code whose groups run slot_period (or a divisor of it) to
/// a workgroup. `whole`: the problem's tape compiled whole with the same
/// options, when the caller has it (its counts are the grouped tape's).
inline Tape compile_groups(const model::Problem<double>& p, std::uint32_t n_groups, const CompileOptions& opt = {},
                           std::uint32_t slot_period = 0, const Tape* whole_tape = nullptr)
{
  std::size_t const n_roots = 1 + p.constraints.size();
  n_groups = std::max<std::uint32_t>(1, std::min<std::uint32_t>(n_groups, static_cast<std::uint32_t>(n_roots)));
  if (n_groups == 1) {
    return compile(p, opt);
  }
  std::vector<std::vector<std::size_t>> members(n_groups);
  if (n_groups == n_roots) {
    // one function per group
    for (std::size_t k = 0; k < n_roots; ++k) {
      members[k].push_back(k);
    }
  } else {
    std::vector<std::pair<std::size_t, std::size_t>> cost;  // (work, root index)
    std::vector<std::size_t> const work = detail::root_work(p);
    for (std::size_t k = 0; k < n_roots; ++k) {
      cost.push_back({work[k], k});
    }
    std::stable_sort(cost.begin(), cost.end(), [](auto const& a, auto const& b) { return a.first > b.first; });
    std::vector<std::size_t> load(n_groups, 0);
    for (auto const& [c, k] : cost) {
      std::size_t const g = static_cast<std::size_t>(std::min_element(load.begin(), load.end()) - load.begin());
      members[g].push_back(k);
      load[g] += c;
    }
  }
  std::vector<Tape> parts;
  Tape t;
  for (std::uint32_t g = 0; g < n_groups; ++g) {
    auto& m = members[g];
    std::sort(m.begin(), m.end());
    if (!m.empty() && m.front() == 0) {
      t.objective_group = g;
    }
    parts.push_back(detail::Compiler(p, opt).run(m));
  }
  t.n_vars = parts[0].n_vars;
  t.n_constraints = static_cast<std::uint32_t>(p.constraints.size());
  t.n_outputs = parts[0].n_outputs;
  t.n_groups = n_groups;
  for (auto const& q : parts) {
    t.temps_per_group = std::max(t.temps_per_group, q.n_slots - q.n_vars);
  }
  Tape const whole = whole_tape != nullptr ? Tape {} : detail::Compiler(p, opt).run();
  Tape const& w = whole_tape != nullptr ? *whole_tape : whole;
  t.reachable_nodes = w.reachable_nodes;
  std::uint32_t const period = slot_period == 0 ? n_groups : std::min(slot_period, n_groups);
  for (std::uint32_t g = 0; g < n_groups; ++g) {
    t.group_begin.push_back(static_cast<std::uint32_t>(t.code.size()));
    std::uint32_t const shift = (g % period) * t.temps_per_group;
    auto remap = [&](std::uint16_t s) -> std::uint16_t {
      return s >= t.n_vars ? static_cast<std::uint16_t>(s + shift) : s;
    };
    for (Instr ins : parts[g].code) {
      bool const out = ins.op == static_cast<std::uint8_t>(TOp::Output)
          || ins.op == static_cast<std::uint8_t>(TOp::AccOutput);
      bool const writes = (ins.op >= static_cast<std::uint8_t>(TOp::LoadConst) && ins.op <= static_cast<std::uint8_t>(TOp::Neg))
          || is_ext_op(ins.op);
      if (writes || ins.op == static_cast<std::uint8_t>(TOp::AccStore)) {
        ins.dst = remap(ins.dst);
      }
      ins.a = remap(ins.a);
      if (!out) {
        ins.b = remap(ins.b);
      }
      t.code.push_back(ins);
    }
    t.node_ops += parts[g].node_ops;
  }
  t.flops = w.flops;  // algorithmic work: every DAG operation once
  t.extended = w.extended;
  t.group_begin.push_back(static_cast<std::uint32_t>(t.code.size()));
  std::uint64_t const slots = std::uint64_t(t.n_vars) + std::uint64_t(period) * t.temps_per_group;
  if (slots > 65535) {
    throw InvalidConfiguration("grouped tape needs more slots than the 16-bit encoding holds");
  }
  t.n_slots = static_cast<std::uint32_t>(std::max<std::uint64_t>(slots, 1));
  return t;
}

/// Function groups chosen for explicit boxes and decode-mode items: about one
/// per 32 tape words and at most one per function; at most 8, or 16 for a
/// tape of 4,096 words or more, whose per-wave share is otherwise the
/// critical path of a small decode-mode item (one 64-lane tile).
inline std::uint32_t auto_groups(const Tape& t)
{
  std::size_t const by_size = (t.code.size() + 31) / 32;
  std::size_t const roots = 1 + static_cast<std::size_t>(t.n_constraints);
  std::size_t const cap = t.code.size() >= 4096 ? 16 : 8;
  return static_cast<std::uint32_t>(std::clamp<std::size_t>(std::min(by_size, roots), 1, cap));
}

/// Human-readable listing, one instruction per line.
inline std::string disassemble(const Tape& t)
{
  std::string out;
  char buf[192];
  for (std::size_t i = 0; i < t.code.size(); ++i) {
    Instr const& in = t.code[i];
    std::snprintf(buf, sizeof(buf), "%5zu %-12s d=%u a=%u b=%u imm=%.17g\n", i, op_name(in.op).c_str(), in.dst, in.a,
                  in.b, in.imm);
    out += buf;
  }
  return out;
}

}  // namespace cmi::tape
