#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <functional>
#include <cstdint>
#include <string>
#include <vector>

#include "cmi/errors.hpp"
#include "cmi/model.hpp"

// The DAG tape: a topologically ordered, register-allocated instruction list
// that one device lane interprets per box (or per sample point). It replaces
// the reference's one-kernel-per-DAG-node CUDA graph (backend/graph/
// builder.cuh:261-514) while evaluating exactly the same node operations:
// every reachable non-Const node becomes one instruction applying that node's
// Op to its operands, a Const operand is an immediate consumed by value (as
// wire_binary passes it, builder.cuh:412-455), and Const (+) Const is
// evaluated on point intervals (binary_op_scalar_scalar_kernel,
// kernels.cuh:240-253).
//
// Evaluation order among independent nodes does not change any node's value,
// so the compiler is free to schedule: roots are walked depth-first in
// objective-then-constraint order, which turns each left-deep sum into an
// accumulator chain, and "leaf" nodes (all operands Var/Const) are
// re-evaluated at each use instead of being held live, which bounds the live
// set by the variables plus a few temporaries. Slots are recycled on last use.
// Constraint checks and the objective read-out are tape instructions too, so a
// root slot dies as soon as it has been checked.
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
  Count
};

inline const char* top_name(TOp op)
{
  switch (op) {
    case TOp::Nop: return "nop";
    case TOp::LoadConst: return "ldc";
    case TOp::Add: return "add";
    case TOp::AddImm: return "addi";
    case TOp::Sub: return "sub";
    case TOp::SubImm: return "subi";
    case TOp::ImmSub: return "isub";
    case TOp::Mul: return "mul";
    case TOp::MulImm: return "muli";
    case TOp::Sqr: return "sqr";
    case TOp::Neg: return "neg";
    case TOp::CheckLE: return "chkle";
    case TOp::CheckEQ: return "chkeq";
    case TOp::Objective: return "obj";
    case TOp::Output: return "out";
    case TOp::Count: break;
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

/// FP64 operations one instruction costs as algorithmic work: each
/// directed-rounded product or sum counts once; selects, compares used for
/// case analysis and the rounding-mode emulation are not counted.
inline std::uint32_t interval_flops(TOp op)
{
  switch (op) {
    case TOp::Add:
    case TOp::AddImm:
    case TOp::Sub:
    case TOp::SubImm:
    case TOp::ImmSub:
    case TOp::MulImm:
    case TOp::Sqr:
      return 2;
    case TOp::Mul:
      return 8;
    case TOp::CheckLE:
      return 1;
    case TOp::CheckEQ:
      return 2;
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
  std::uint32_t remat_extra = 0;      // instructions added by re-evaluating leaves
  std::uint64_t flops = 0;            // interval_flops summed over the tape
};

struct CompileOptions
{
  bool rematerialise_leaves = true;
  bool emit_outputs = true;  // Output instructions for the objective and each constraint root
};

/// Compile a validated Problem into a Tape. Throws InvalidDAG for the shapes the
/// reference's builder rejects (a unary op on a Const, a Const root) and
/// InvalidConfiguration for ops the device evaluator does not implement.
inline Tape compile(const model::Problem<double>& p, const CompileOptions& opt = {})
{
  using model::Op;
  auto const& nodes = p.graph.nodes;
  std::size_t const n = nodes.size();
  std::size_t const n_vars = p.box_bounds.size();
  if (n_vars > 30000 || n > (std::size_t(1) << 31)) {
    throw InvalidConfiguration("problem too large for the 16-bit tape slot encoding");
  }

  std::vector<std::size_t> roots;
  roots.push_back(p.objective_root);
  for (auto const& c : p.constraints) {
    roots.push_back(c.root_id);
  }

  // Reachability and the op set the device implements.
  std::vector<bool> reach(n, false);
  for (std::size_t r : roots) {
    reach[r] = true;
  }
  std::uint32_t reachable = 0;
  for (std::size_t i = n; i-- > 0;) {
    if (!reach[i]) {
      continue;
    }
    auto const& nd = nodes[i];
    for (std::size_t in : nd.in) {
      reach[in] = true;
    }
    if (nd.op == Op::Const) {
      continue;
    }
    ++reachable;
    switch (nd.op) {
      case Op::Var:
      case Op::Add:
      case Op::Sub:
      case Op::Mul:
        break;
      case Op::Sqr:
      case Op::Neg:
        if (nodes[nd.in[0]].op == Op::Const) {
          throw InvalidDAG("unary op applied directly to an Op::Const");
        }
        break;
      default:
        throw InvalidConfiguration(std::string("Op::") + model::op_name(nd.op)
                                   + " is not implemented by the device tape evaluator");
    }
  }

  auto is_leafish = [&](std::size_t id) {
    Op const op = nodes[id].op;
    return op == Op::Var || op == Op::Const;
  };
  // A leaf: an op node all of whose operands are Var or Const.
  std::vector<bool> leaf(n, false);
  for (std::size_t i = 0; i < n; ++i) {
    if (!reach[i] || is_leafish(i)) {
      continue;
    }
    bool all = true;
    for (std::size_t in : nodes[i].in) {
      all = all && is_leafish(in);
    }
    leaf[i] = all && opt.rematerialise_leaves;
  }

  // Remaining uses of each shared (non-leaf, non-Var) node, counted over the
  // instruction stream the walk below will produce.
  std::vector<std::uint32_t> uses(n, 0);
  {
    std::vector<std::uint8_t> seen(n, 0);
    std::vector<std::size_t> stack;
    auto count_walk = [&](std::size_t root) {
      stack.push_back(root);
      while (!stack.empty()) {
        std::size_t const id = stack.back();
        stack.pop_back();
        auto const& nd = nodes[id];
        if (nd.op == Op::Var || nd.op == Op::Const) {
          continue;
        }
        if (!leaf[id]) {
          if (seen[id]) {
            continue;
          }
          seen[id] = 1;
        }
        for (std::size_t in : nd.in) {
          if (nodes[in].op != Op::Const && nodes[in].op != Op::Var && !leaf[in]) {
            ++uses[in];
          }
          stack.push_back(in);
        }
      }
    };
    for (std::size_t r : roots) {
      count_walk(r);
      if (nodes[r].op != Op::Var && !leaf[r]) {
        ++uses[r];  // the root's own check/objective read
      }
    }
    // A root referenced as the objective and as (several) constraints is read
    // once per reference, and Output reads it once more.
    if (opt.emit_outputs) {
      for (std::size_t r : roots) {
        if (nodes[r].op != Op::Var && !leaf[r]) {
          ++uses[r];
        }
      }
    }
  }

  Tape t;
  t.n_vars = static_cast<std::uint32_t>(n_vars);
  t.n_constraints = static_cast<std::uint32_t>(p.constraints.size());
  t.reachable_nodes = reachable;
  std::vector<std::uint32_t> free_slots;  // kept sorted descending: back() is the lowest free slot
  std::uint32_t next_slot = static_cast<std::uint32_t>(n_vars);
  auto alloc = [&]() -> std::uint32_t {
    if (!free_slots.empty()) {
      std::uint32_t const s = free_slots.back();
      free_slots.pop_back();
      return s;
    }
    return next_slot++;
  };
  auto release = [&](std::uint32_t s) {
    if (s < n_vars) {
      return;
    }
    free_slots.insert(std::upper_bound(free_slots.begin(), free_slots.end(), s, std::greater<>()), s);
  };

  std::vector<std::int64_t> slot_of(n, -1);  // computed shared nodes
  std::vector<std::uint32_t> left(uses);     // remaining reads

  auto emit = [&](TOp op, std::uint32_t dst, std::uint32_t a, std::uint32_t b, double imm) {
    Instr ins {};
    ins.op = static_cast<std::uint8_t>(op);
    ins.dst = static_cast<std::uint16_t>(dst);
    ins.a = static_cast<std::uint16_t>(a);
    ins.b = static_cast<std::uint16_t>(b);
    ins.imm = imm;
    t.code.push_back(ins);
    t.flops += interval_flops(op);
  };

  // Slot holding the value of operand `id` for one read; `temp` reports whether
  // that slot must be released after this read (a re-evaluated leaf copy).
  struct Frame
  {
    std::size_t id;
    bool expanded;
  };
  std::vector<Frame> stack;
  std::vector<std::uint32_t> value_stack;  // slots of evaluated operands, in walk order
  std::vector<bool> value_temp;

  auto read_done = [&](std::size_t id, std::uint32_t slot, bool temp) {
    if (temp) {
      release(slot);
      return;
    }
    if (nodes[id].op == Op::Var) {
      return;
    }
    if (--left[id] == 0) {
      release(slot);
    }
  };

  std::uint32_t leaf_instrs = 0;
  auto evaluate = [&](std::size_t root) {
    stack.push_back({root, false});
    while (!stack.empty()) {
      Frame f = stack.back();
      stack.pop_back();
      std::size_t const id = f.id;
      auto const& nd = nodes[id];
      if (nd.op == Op::Var) {
        value_stack.push_back(static_cast<std::uint32_t>(nd.payload.var_index));
        value_temp.push_back(false);
        continue;
      }
      if (nd.op == Op::Const) {
        value_stack.push_back(UINT32_MAX);  // placeholder, consumed as an immediate
        value_temp.push_back(false);
        continue;
      }
      if (!leaf[id] && slot_of[id] >= 0) {
        value_stack.push_back(static_cast<std::uint32_t>(slot_of[id]));
        value_temp.push_back(false);
        continue;
      }
      if (!f.expanded) {
        stack.push_back({id, true});
        for (std::size_t k = nd.in.size(); k-- > 0;) {
          stack.push_back({nd.in[k], false});
        }
        continue;
      }
      // Operands are on value_stack in order.
      std::size_t const arity = nd.in.size();
      std::uint32_t sl[2] = {0, 0};
      bool tmp[2] = {false, false};
      for (std::size_t k = arity; k-- > 0;) {
        sl[k] = value_stack.back();
        tmp[k] = value_temp.back();
        value_stack.pop_back();
        value_temp.pop_back();
      }
      bool const c0 = nodes[nd.in[0]].op == Op::Const;
      bool const c1 = arity > 1 && nodes[nd.in[1]].op == Op::Const;
      double const k0 = c0 ? nodes[nd.in[0]].payload.constant : 0.0;
      double const k1 = c1 ? nodes[nd.in[1]].payload.constant : 0.0;
      std::uint32_t const before = static_cast<std::uint32_t>(t.code.size());

      // Both operands Const: materialise the left one as a point interval and
      // apply the op with the right one as an immediate.
      if (c0 && c1) {
        std::uint32_t const s = alloc();
        emit(TOp::LoadConst, s, 0, 0, k0);
        sl[0] = s;
        tmp[0] = true;
      }
      bool const a_const = c0 && !c1;
      bool const b_const = c1;
      // Release operand slots before allocating the destination so an
      // operand's slot can be reused in place (operands are read first).
      if (arity > 0 && !(a_const)) {
        read_done(nd.in[0], sl[0], tmp[0]);
      }
      if (arity > 1 && !b_const) {
        read_done(nd.in[1], sl[1], tmp[1]);
      }
      std::uint32_t const dst = alloc();
      switch (nd.op) {
        case Op::Add:
          if (b_const) {
            emit(TOp::AddImm, dst, sl[0], 0, k1);
          } else if (a_const) {
            emit(TOp::AddImm, dst, sl[1], 0, k0);
          } else {
            emit(TOp::Add, dst, sl[0], sl[1], 0.0);
          }
          break;
        case Op::Sub:
          if (b_const) {
            emit(TOp::SubImm, dst, sl[0], 0, k1);
          } else if (a_const) {
            emit(TOp::ImmSub, dst, sl[1], 0, k0);
          } else {
            emit(TOp::Sub, dst, sl[0], sl[1], 0.0);
          }
          break;
        case Op::Mul:
          if (b_const) {
            emit(TOp::MulImm, dst, sl[0], 0, k1);
          } else if (a_const) {
            emit(TOp::MulImm, dst, sl[1], 0, k0);
          } else {
            emit(TOp::Mul, dst, sl[0], sl[1], 0.0);
          }
          break;
        case Op::Sqr:
          emit(TOp::Sqr, dst, sl[0], 0, 0.0);
          break;
        case Op::Neg:
          emit(TOp::Neg, dst, sl[0], 0, 0.0);
          break;
        default:
          throw InvalidConfiguration("unsupported op reached the tape emitter");
      }
      if (leaf[id]) {
        leaf_instrs += static_cast<std::uint32_t>(t.code.size()) - before;
        value_stack.push_back(dst);
        value_temp.push_back(true);
      } else {
        slot_of[id] = dst;
        value_stack.push_back(dst);
        value_temp.push_back(false);
      }
    }
  };

  auto root_value = [&](std::size_t r, std::uint32_t& slot, bool& temp) {
    evaluate(r);
    slot = value_stack.back();
    temp = value_temp.back();
    value_stack.pop_back();
    value_temp.pop_back();
  };

  std::uint32_t output_index = 0;
  for (std::size_t k = 0; k < roots.size(); ++k) {
    std::size_t const r = roots[k];
    std::uint32_t slot = 0;
    bool temp = false;
    root_value(r, slot, temp);
    if (opt.emit_outputs) {
      emit(TOp::Output, 0, slot, output_index, 0.0);
      if (!temp) {
        read_done(r, slot, false);  // a re-evaluated leaf copy stays alive for the check below
      }
    }
    ++output_index;
    if (k == 0) {
      emit(TOp::Objective, 0, slot, 0, 0.0);
    } else {
      auto const& c = p.constraints[k - 1];
      emit(c.cmp == model::Cmp::LE ? TOp::CheckLE : TOp::CheckEQ, 0, slot, 0, c.rhs);
    }
    read_done(r, slot, temp);
  }
  t.n_outputs = opt.emit_outputs ? output_index : 0;
  t.n_slots = std::max<std::uint32_t>(next_slot, 1);
  std::uint32_t distinct_leaves = 0;
  for (std::size_t i = 0; i < n; ++i) {
    if (leaf[i]) {
      distinct_leaves += (nodes[i].in.size() == 2 && nodes[nodes[i].in[0]].op == Op::Const
                          && nodes[nodes[i].in[1]].op == Op::Const)
          ? 2
          : 1;
    }
  }
  t.remat_extra = leaf_instrs > distinct_leaves ? leaf_instrs - distinct_leaves : 0;
  return t;
}

/// Human-readable listing, one instruction per line.
inline std::string disassemble(const Tape& t)
{
  std::string out;
  char buf[160];
  for (std::size_t i = 0; i < t.code.size(); ++i) {
    Instr const& in = t.code[i];
    std::snprintf(buf, sizeof(buf), "%5zu %-6s d=%u a=%u b=%u imm=%.17g\n", i, top_name(static_cast<TOp>(in.op)),
                  in.dst, in.a, in.b, in.imm);
    out += buf;
  }
  return out;
}

}  // namespace cmi::tape
