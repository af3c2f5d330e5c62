// C-ABI over the C++ engine: status codes instead of exceptions, opaque
// handles instead of objects. Declarations and the reference interfaces they
// replace are in include/cmi.h.

#include "cmi.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "aggregate_solve.hpp"
#include "batch.hpp"
#include "jit.hpp"
#include "cmi/config.hpp"
#include "cmi/gams.hpp"
#include "cmi/model.hpp"
#include "cmi/policy.hpp"
#include "cmi/qcqp.hpp"
#include "cmi/search.hpp"
#include "handles.hpp"
#include "hip_backend.hpp"

namespace
{

thread_local std::string g_last_error;

cmi_status fail(cmi_status s, const char* what)
{
  g_last_error = what;
  return s;
}

template<class F>
cmi_status guarded(F&& f)
{
  try {
    f();
    g_last_error.clear();
    return CMI_OK;
  } catch (const cmi::backend::OverBudgetError& e) {
    return fail(CMI_OVER_BUDGET, e.what());
  } catch (const cmi::ResourceExhausted& e) {
    return fail(CMI_OVER_BUDGET, e.what());
  } catch (const cmi::ShapeMismatch& e) {
    return fail(CMI_SHAPE_MISMATCH, e.what());
  } catch (const cmi::InvalidConfiguration& e) {
    return fail(CMI_INVALID_CONFIG, e.what());
  } catch (const cmi::InvalidDAG& e) {
    return fail(CMI_INVALID_DAG, e.what());
  } catch (const cmi::InvalidProblem& e) {
    return fail(CMI_INVALID_PROBLEM, e.what());
  } catch (const cmi::DeviceError& e) {
    return fail(CMI_DEVICE_ERROR, e.what());
  } catch (const cmi::gams::ParseError& e) {
    return fail(CMI_PARSE_ERROR, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CMI_ERROR, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(CMI_ERROR, e.what());
  } catch (...) {
    return fail(CMI_ERROR, "unknown error");
  }
}

cmi::model::Problem<double> problem_from_desc(const cmi_problem_desc& d)
{
  using namespace cmi::model;
  Problem<double> p;
  if (d.n_nodes > 0 && d.op == nullptr) {
    throw cmi::InvalidDAG("problem description has nodes but no op array");
  }
  p.graph.nodes.reserve(d.n_nodes);
  for (std::size_t i = 0; i < d.n_nodes; ++i) {
    if (d.op[i] > static_cast<std::uint8_t>(Op::Max)) {
      throw cmi::InvalidDAG("unknown op code in problem description");
    }
    Op const op = static_cast<Op>(d.op[i]);
    DAGNode<double> node {};
    node.id = i;
    node.op = op;
    std::size_t const arity = op_arity(op);
    if (arity >= 1) {
      if (d.in0 == nullptr) {
        throw cmi::InvalidDAG("problem description lacks in0 for a node with operands");
      }
      node.in.push_back(d.in0[i]);
    }
    if (arity >= 2) {
      if (d.in1 == nullptr) {
        throw cmi::InvalidDAG("problem description lacks in1 for a binary node");
      }
      node.in.push_back(d.in1[i]);
    }
    for (std::size_t in : node.in) {
      if (in >= i) {
        throw cmi::InvalidDAG(
            "DAG node operand id is not strictly less than the node's own id; topological-order invariant "
            "violated");
      }
    }
    node.op_count = 1;
    for (std::size_t in : node.in) {
      node.op_count += p.graph.nodes[in].op_count;
    }
    if (op == Op::Const) {
      if (d.constant == nullptr) {
        throw cmi::InvalidDAG("problem description lacks constant payloads");
      }
      node.payload.constant = d.constant[i];
    } else if (op == Op::Var) {
      if (d.var_index == nullptr) {
        throw cmi::InvalidDAG("problem description lacks var_index payloads");
      }
      node.payload.var_index = d.var_index[i];
    } else if (op == Op::PowN) {
      if (d.int_exp == nullptr) {
        throw cmi::InvalidDAG("problem description lacks int_exp payloads");
      }
      node.payload.int_exp = d.int_exp[i];
    } else {
      node.payload.constant = 0.0;
    }
    p.graph.nodes.push_back(std::move(node));
  }
  if (d.n_vars > 0 && (d.box == nullptr || d.var_kind == nullptr)) {
    throw cmi::InvalidProblem("problem description lacks the box or the variable kinds");
  }
  for (std::size_t v = 0; v < d.n_vars; ++v) {
    p.box_bounds.push_back({d.box[v].lb, d.box[v].ub});
    if (d.var_kind[v] > static_cast<std::uint8_t>(VarKind::Binary)) {
      throw cmi::InvalidProblem("unknown variable kind");
    }
    p.var_kinds.push_back(static_cast<VarKind>(d.var_kind[v]));
  }
  p.objective_root = d.objective_root;
  if (d.n_constraints > 0 && (d.con_root == nullptr || d.con_cmp == nullptr || d.con_rhs == nullptr)) {
    throw cmi::InvalidProblem("problem description lacks constraint arrays");
  }
  for (std::size_t k = 0; k < d.n_constraints; ++k) {
    if (d.con_cmp[k] > static_cast<std::uint8_t>(Cmp::EQ)) {
      throw cmi::InvalidProblem("unknown constraint comparison");
    }
    p.constraints.push_back({d.con_root[k], static_cast<Cmp>(d.con_cmp[k]), d.con_rhs[k]});
  }
  p.validate();
  return p;
}

cmi::region::Composition composition_of(const std::uint8_t* kinds, std::size_t n)
{
  cmi::region::Composition c;
  for (std::size_t j = 0; j < n; ++j) {
    if (kinds[j] > static_cast<std::uint8_t>(cmi::region::SlotKind::BinaryEnumerate)) {
      throw cmi::InvalidConfiguration("unknown slot kind");
    }
    c.kinds.push_back(static_cast<cmi::region::SlotKind>(kinds[j]));
  }
  return c;
}

void copy_text(char* dst, std::size_t cap, const std::string& src)
{
  std::size_t const k = std::min(cap - 1, src.size());
  std::memcpy(dst, src.data(), k);
  dst[k] = '\0';
}

void fill_facts(const cmi::backend::OverBudget& f, cmi_over_budget* out)
{
  std::memset(out, 0, sizeof(*out));
  copy_text(out->role, sizeof(out->role), f.role);
  out->needed_bytes = f.needed_bytes;
  out->budget_bytes = f.budget_bytes;
  out->n_regions = f.n_regions;
  out->elements_per_region = f.elements_per_region;
  out->bytes_per_element = f.bytes_per_element;
  copy_text(out->element_breakdown, sizeof(out->element_breakdown), f.element_breakdown);
  copy_text(out->element_inventory, sizeof(out->element_inventory), f.element_inventory);
  out->cost.sampler_fixed = f.cost.sampler_fixed;
  out->cost.sampler_per_region = f.cost.sampler_per_region;
  out->cost.bounder_fixed = f.cost.bounder_fixed;
  out->cost.bounder_per_region = f.cost.bounder_per_region;
  out->cost.enumerator_fixed = f.cost.enumerator_fixed;
  out->cost.enumerator_per_region = f.cost.enumerator_per_region;
  out->cost.max_regions = f.cost.max_regions;
  std::size_t const n = std::min<std::size_t>(f.composition.size(), sizeof(out->composition_kinds));
  for (std::size_t j = 0; j < n; ++j) {
    out->composition_kinds[j] = static_cast<std::uint8_t>(f.composition[j]);
  }
  out->composition_size = n;
  copy_text(out->composition, sizeof(out->composition), cmi::region::spell(f.composition));
  out->has_fan_out = f.fan_out ? 1 : 0;
  out->fan_out_partition_num = f.fan_out ? f.fan_out->partition_num() : 0;
  out->fan_out_enumerate_cap = f.fan_out ? f.fan_out->enumerate_cap() : 0;
  out->solve_samples_per_region = f.solve_samples_per_region;
  out->sampler_bytes = f.sampler_bytes;
}

/// The inverse of fill_facts, for the explain call.
cmi::backend::OverBudget facts_from(const cmi_over_budget& in)
{
  cmi::backend::OverBudget f;
  f.role.assign(in.role, strnlen(in.role, sizeof(in.role)));
  f.needed_bytes = in.needed_bytes;
  f.budget_bytes = in.budget_bytes;
  f.n_regions = in.n_regions;
  f.elements_per_region = in.elements_per_region;
  f.bytes_per_element = in.bytes_per_element;
  f.element_breakdown.assign(in.element_breakdown, strnlen(in.element_breakdown, sizeof(in.element_breakdown)));
  f.element_inventory.assign(in.element_inventory, strnlen(in.element_inventory, sizeof(in.element_inventory)));
  f.cost.sampler_fixed = in.cost.sampler_fixed;
  f.cost.sampler_per_region = in.cost.sampler_per_region;
  f.cost.bounder_fixed = in.cost.bounder_fixed;
  f.cost.bounder_per_region = in.cost.bounder_per_region;
  f.cost.enumerator_fixed = in.cost.enumerator_fixed;
  f.cost.enumerator_per_region = in.cost.enumerator_per_region;
  f.cost.max_regions = in.cost.max_regions;
  std::size_t const n = std::min<std::size_t>(in.composition_size, sizeof(in.composition_kinds));
  for (std::size_t j = 0; j < n; ++j) {
    if (in.composition_kinds[j] > CMI_SLOT_BINARY_ENUMERATE) {
      throw cmi::InvalidConfiguration("over-budget facts name an unknown slot kind");
    }
    f.composition.kinds.push_back(static_cast<cmi::region::SlotKind>(in.composition_kinds[j]));
  }
  if (in.has_fan_out) {
    f.fan_out = cmi::region::FanOutSpec(in.fan_out_partition_num, in.fan_out_enumerate_cap);
  }
  f.solve_samples_per_region = in.solve_samples_per_region;
  f.sampler_bytes = in.sampler_bytes;
  return f;
}

void fill_outcome(const cmi::search::SolveOutcome& o, double seconds, cmi_solve_outcome* out, double* best,
                  std::size_t n_vars)
{
  out->lower_bound = o.lower_bound;
  out->upper_bound = o.upper_bound;
  out->stop_reason = static_cast<int>(o.stop_reason);
  out->proven_optimal = o.proven_optimal ? 1 : 0;
  out->infeasible = o.infeasible ? 1 : 0;
  out->found_incumbent = o.found_incumbent() ? 1 : 0;
  out->iterations = o.counters.iterations;
  out->fathomed = o.counters.fathomed;
  out->bounded = o.counters.bounded;
  out->sampled = o.counters.sampled;
  out->enqueued = o.counters.enqueued;
  out->pruned_infeasible = o.counters.pruned_infeasible;
  out->pruned_dominated = o.counters.pruned_dominated;
  out->pending = o.counters.pending;
  out->fathomed_points = o.counters.fathomed_points;
  out->witness_rejected = o.counters.witness_rejected;
  out->pruned_duplicate = o.counters.pruned_duplicate;
  out->viable = o.counters.viable;
  out->dropped_viable = o.counters.dropped.viable;
  out->dropped_dominated = o.counters.dropped.dominated;
  out->dropped_floor = o.counters.dropped.lb_min;
  out->history_freed = o.counters.history_freed;
  out->history_peak_live = o.counters.history_peak_live;
  out->wall_seconds = seconds;
  out->has_over_budget = o.over_budget ? 1 : 0;
  if (o.over_budget) {
    fill_facts(*o.over_budget, &out->over_budget);
  } else {
    std::memset(&out->over_budget, 0, sizeof(out->over_budget));
  }
  if (best != nullptr) {
    for (std::size_t v = 0; v < n_vars; ++v) {
      best[v] = v < o.best_point.size() ? o.best_point[v] : __builtin_nan("");
    }
  }
}

}  // namespace

namespace cmi::capi
{

std::shared_ptr<const cmi::policy::CompositionPolicy> make_policy(const cmi_solve_options& o)
{
  cmi::policy::SearchCalibration cal;
  cal.max_cycle_size = o.max_cycle_size;
  switch (o.policy) {
    case CMI_POLICY_GREEDY_ENUMERATE:
      return std::make_shared<cmi::policy::GreedyEnumCompositionPolicy>(
          cmi::region::FanOutSpec(o.partition_num, o.enumerate_cap), cal);
    case CMI_POLICY_BISECTION_BUDGET:
    case CMI_POLICY_RELATIVE_BISECTION_BUDGET:
      cal.enumerate_cap = o.enumerate_cap;
      return std::make_shared<cmi::policy::BisectionBudgetCompositionPolicy>(
          o.bisection_budget, cal, o.policy == CMI_POLICY_RELATIVE_BISECTION_BUDGET);
    case CMI_POLICY_WIDTH_FIRST:
      return std::make_shared<cmi::policy::WidthFirstCompositionPolicy>(
          cmi::region::FanOutSpec(o.partition_num, o.enumerate_cap), cal);
    default: break;
  }
  throw cmi::InvalidConfiguration("unknown composition policy " + std::to_string(o.policy));
}

cmi_status run_guarded(void (*fn)(void*), void* arg)
{
  return guarded([&] { fn(arg); });
}

void fill_outcome_public(const cmi::search::SolveOutcome& o, double seconds, cmi_solve_outcome* out, double* best,
                         std::size_t n_vars)
{
  fill_outcome(o, seconds, out, best, n_vars);
}

}  // namespace cmi::capi

extern "C" {

const char* cmi_last_error(void) { return g_last_error.c_str(); }

const char* cmi_version(void) { return "cuminlpsearch_amd 0.1 (gfx950)"; }

cmi_status cmi_problem_create(const cmi_problem_desc* desc, cmi_problem** out)
{
  if (desc == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] { *out = new cmi_problem {problem_from_desc(*desc)}; });
}

cmi_status cmi_problem_qcqp(const cmi_qcqp_spec* spec, uint64_t seed, cmi_problem** out)
{
  if (spec == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    cmi::qcqp::Spec s;
    s.n = spec->n;
    s.m = spec->m;
    s.n_int = spec->n_int;
    s.cont_lb = spec->cont_lb;
    s.cont_ub = spec->cont_ub;
    s.int_lb = spec->int_lb;
    s.int_ub = spec->int_ub;
    s.pair_density = spec->pair_density;
    s.term_prob = spec->term_prob;
    s.lin_prob = spec->lin_prob;
    s.symmetric_pairs = spec->symmetric_pairs != 0;
    s.rhs_lo = spec->rhs_lo;
    s.rhs_hi = spec->rhs_hi;
    if (s.n == 0 || s.n_int > s.n) {
      throw cmi::InvalidConfiguration("qcqp spec needs n >= 1 and n_int <= n");
    }
    if (spec->target_nodes > 0) {
      s = cmi::qcqp::tune_term_prob(s, seed, spec->target_nodes);
    }
    auto p = cmi::qcqp::generate(s, seed);
    p.validate();
    *out = new cmi_problem {std::move(p)};
  });
}

cmi_status cmi_problem_info_get(const cmi_problem* p, cmi_problem_info* out)
{
  if (p == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    out->n_nodes = p->problem.graph.nodes.size();
    out->n_vars = p->problem.box_bounds.size();
    out->n_constraints = p->problem.constraints.size();
    out->reachable_nodes = cmi::qcqp::reachable_nodes(p->problem);
  });
}

cmi_status cmi_problem_export(const cmi_problem* p, uint8_t* op, uint64_t* in0, uint64_t* in1, double* constant,
                              uint64_t* var_index, int32_t* int_exp, cmi_ivl* box, uint8_t* var_kind,
                              uint64_t* objective_root, uint64_t* con_root, uint8_t* con_cmp, double* con_rhs)
{
  if (p == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null problem");
  }
  return guarded([&] {
    auto const& pr = p->problem;
    for (std::size_t i = 0; i < pr.graph.nodes.size(); ++i) {
      auto const& nd = pr.graph.nodes[i];
      if (op != nullptr) {
        op[i] = static_cast<uint8_t>(nd.op);
      }
      if (in0 != nullptr) {
        in0[i] = nd.in.size() > 0 ? nd.in[0] : 0;
      }
      if (in1 != nullptr) {
        in1[i] = nd.in.size() > 1 ? nd.in[1] : 0;
      }
      if (constant != nullptr) {
        constant[i] = nd.op == cmi::model::Op::Const ? nd.payload.constant : 0.0;
      }
      if (var_index != nullptr) {
        var_index[i] = nd.op == cmi::model::Op::Var ? nd.payload.var_index : 0;
      }
      if (int_exp != nullptr) {
        int_exp[i] = nd.op == cmi::model::Op::PowN ? nd.payload.int_exp : 0;
      }
    }
    for (std::size_t v = 0; v < pr.box_bounds.size(); ++v) {
      if (box != nullptr) {
        box[v] = {pr.box_bounds[v].lb, pr.box_bounds[v].ub};
      }
      if (var_kind != nullptr) {
        var_kind[v] = static_cast<uint8_t>(pr.var_kinds[v]);
      }
    }
    if (objective_root != nullptr) {
      *objective_root = pr.objective_root;
    }
    for (std::size_t k = 0; k < pr.constraints.size(); ++k) {
      if (con_root != nullptr) {
        con_root[k] = pr.constraints[k].root_id;
      }
      if (con_cmp != nullptr) {
        con_cmp[k] = static_cast<uint8_t>(pr.constraints[k].cmp);
      }
      if (con_rhs != nullptr) {
        con_rhs[k] = pr.constraints[k].rhs;
      }
    }
  });
}

void cmi_problem_destroy(cmi_problem* p) { delete p; }

cmi_status cmi_ctx_create(int device, cmi_ctx** out)
{
  if (out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] { *out = new cmi_ctx {std::make_shared<cmi::hip::Context>(device)}; });
}

void cmi_ctx_destroy(cmi_ctx* ctx) { delete ctx; }

cmi_status cmi_tape_create(cmi_ctx* ctx, const cmi_problem* p, cmi_tape** out)
{
  if (ctx == nullptr || p == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] { *out = new cmi_tape {std::make_shared<cmi::hip::DeviceTape>(ctx->ctx, p->problem)}; });
}

cmi_status cmi_tape_info_get(const cmi_tape* t, cmi_tape_info* out)
{
  if (t == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    auto const& tp = t->tape->tape();
    out->n_instr = tp.code.size();
    out->n_slots = tp.n_slots;
    out->n_vars = tp.n_vars;
    out->n_outputs = tp.n_outputs;
    out->reachable_nodes = tp.reachable_nodes;
    out->flops_per_box = tp.flops;
    out->bytes_per_box = 16 * static_cast<std::size_t>(tp.n_vars) + 16 + 1;
    auto const& gt = t->tape->grouped_tape();
    out->box_groups = gt.n_groups;
    out->box_instr = gt.code.size();
    out->box_slots = gt.n_slots;
    out->box_engine = static_cast<int>(t->tape->box_engine());
    auto const c = t->tape->compiled();
    out->box_compiled = c ? 1 : 0;
    out->compile_seconds = c ? c->compile_seconds() : 0.0;
    auto const sh = t->tape->compiled_shape();
    out->compiled_persistent = sh.persistent ? 1 : 0;
    out->compiled_wgs_per_cu = t->tape->compiled_wgs_per_cu();
    out->compiled_waves = sh.waves;
    out->extended = tp.extended ? 1 : 0;
    out->compiled_schedule = sh.ilp_schedule ? CMI_SCHED_MAX_ILP : CMI_SCHED_DEFAULT;
    out->compiled_load_batch = sh.persistent ? 0 : sh.load_batch;
  });
}

cmi_status cmi_tape_autotune(cmi_tape* t, const cmi_problem* p, size_t n_boxes, cmi_box_candidate* results, size_t cap,
                             size_t* n_results)
{
  if (t == nullptr || p == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    if (p->problem.n_vars() != t->tape->n_vars()) {
      throw cmi::ShapeMismatch("problem does not match the tape");
    }
    auto const all = t->tape->autotune(p->problem, n_boxes);
    if (n_results != nullptr) {
      *n_results = all.size();
    }
    for (std::size_t i = 0; i < all.size() && results != nullptr && i < cap; ++i) {
      results[i].engine = static_cast<int>(all[i].engine);
      results[i].groups = all[i].groups;
      results[i].waves = all[i].waves;
      results[i].persistent = all[i].persistent ? 1 : 0;
      results[i].wgs_per_cu = all[i].wgs_per_cu;
      results[i].schedule = all[i].ilp_schedule ? CMI_SCHED_MAX_ILP : CMI_SCHED_DEFAULT;
      results[i].load_batch = all[i].persistent ? 0 : all[i].load_batch;
      results[i].kernel_ms = all[i].kernel_ms;
    }
  });
}

cmi_status cmi_tape_set_compiled_shape(cmi_tape* t, int persistent, size_t waves, size_t wgs_per_cu)
{
  if (t == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    t->tape->set_compiled_shape(persistent != 0, static_cast<std::uint32_t>(std::min<size_t>(waves, 1u << 16)),
                                static_cast<std::uint32_t>(std::min<size_t>(wgs_per_cu, 1u << 16)));
  });
}

cmi_status cmi_tape_set_compiled_codegen(cmi_tape* t, int schedule, size_t load_batch)
{
  if (t == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  if (schedule != CMI_SCHED_DEFAULT && schedule != CMI_SCHED_MAX_ILP) {
    return fail(CMI_INVALID_CONFIG, "unknown compiled-kernel schedule");
  }
  if (load_batch > 64) {
    return fail(CMI_INVALID_CONFIG, "at most 64 variables per load batch");
  }
  return guarded([&] {
    t->tape->set_compiled_codegen(schedule == CMI_SCHED_MAX_ILP, static_cast<std::uint32_t>(load_batch));
  });
}

cmi_status cmi_tape_set_box_engine(cmi_tape* t, int engine)
{
  if (t == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] { t->tape->set_box_engine(static_cast<cmi::hip::DeviceTape::BoxEngine>(engine)); });
}

namespace
{
/// The grouping a new device tape uses for explicit boxes.
/// The function grouping a tape given `groups` (0 = automatic) evaluates
/// explicit boxes with: the same clamping and LDS fit as set_box_groups, with
/// gfx950's 160 KB of LDS.
cmi::tape::Tape grouped_for_boxes(const cmi::model::Problem<double>& p, size_t groups)
{
  cmi::tape::Tape whole = cmi::tape::compile(p);
  std::uint32_t k = groups == 0 ? cmi::tape::auto_groups(whole) : static_cast<std::uint32_t>(std::min<size_t>(groups, 8));
  k = std::clamp<std::uint32_t>(k, 1, std::min<std::uint32_t>(cmi::dev::kMaxGroups, 1 + whole.n_constraints));
  for (; k > 1; --k) {
    cmi::tape::Tape g = cmi::tape::compile_groups(p, k, {}, 0, &whole);
    if (cmi::dev::boxes_lds_bytes(g.n_slots) <= cmi::jit::kLdsBytes) {
      return g;
    }
  }
  return whole;
}

cmi::jit::Shape export_shape(const cmi::tape::Tape& g, const cmi_compiled_shape& in)
{
  if (in.waves > 8) {
    throw cmi::InvalidConfiguration("at most 16 waves per workgroup");
  }
  if (in.schedule != CMI_SCHED_DEFAULT && in.schedule != CMI_SCHED_MAX_ILP) {
    throw cmi::InvalidConfiguration("unknown compiled-kernel schedule");
  }
  if (in.load_batch > 64) {
    throw cmi::InvalidConfiguration("at most 64 variables per load batch");
  }
  cmi::jit::Shape sh;
  sh.waves = cmi::jit::shape_waves(g, {false, static_cast<std::uint32_t>(in.waves)});
  sh.persistent = in.persistent < 0 ? cmi::jit::persistent_fits(g, sh.waves) : in.persistent != 0;
  sh.ilp_schedule = in.schedule == CMI_SCHED_MAX_ILP;
  sh.load_batch = static_cast<std::uint32_t>(in.load_batch);
  sh.roots = in.roots != 0;
  return sh;
}
}  // namespace

cmi_status cmi_tape_compiled_object(const cmi_problem* p, const cmi_compiled_shape* shape, void* buf, size_t cap,
                                    size_t* len)
{
  if (p == nullptr || shape == nullptr || len == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    p->problem.validate();
    auto const g = grouped_for_boxes(p->problem, shape->groups);
    auto const sh = export_shape(g, *shape);
    auto const code = cmi::jit::compile_code_object(cmi::jit::boxes_source(g, sh), sh.ilp_schedule);
    *len = code.size();
    if (buf != nullptr) {
      if (cap < code.size()) {
        throw cmi::ShapeMismatch("buffer of " + std::to_string(cap) + " bytes; the code object has "
                                 + std::to_string(code.size()));
      }
      std::memcpy(buf, code.data(), code.size());
    }
  });
}

cmi_status cmi_tape_compiled_source(const cmi_problem* p, const cmi_compiled_shape* shape, char* buf, size_t cap,
                                    size_t* len)
{
  if (p == nullptr || shape == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    p->problem.validate();
    auto const g = grouped_for_boxes(p->problem, shape->groups);
    std::string const src = cmi::jit::boxes_source(g, export_shape(g, *shape));
    if (len != nullptr) {
      *len = src.size();
    }
    if (buf != nullptr && cap > 0) {
      std::size_t const k = std::min(cap - 1, src.size());
      std::memcpy(buf, src.data(), k);
      buf[k] = '\0';
    }
  });
}

void cmi_tape_destroy(cmi_tape* t) { delete t; }

cmi_status cmi_tape_set_box_groups(cmi_tape* t, const cmi_problem* p, size_t k)
{
  if (t == nullptr || p == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    if (p->problem.n_vars() != t->tape->n_vars()) {
      throw cmi::ShapeMismatch("problem does not match the tape");
    }
    t->tape->set_box_groups(p->problem, static_cast<std::uint32_t>(std::min<size_t>(k, 1u << 16)));
  });
}

cmi_status cmi_tape_compile_words(const cmi_problem* p, size_t k, void* words, size_t cap_words,
                                  uint32_t* group_begin, cmi_tape_info* out)
{
  if (p == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    p->problem.validate();
    auto const whole = cmi::tape::compile(p->problem);
    auto const tp = k <= 1 ? whole
                           : cmi::tape::compile_groups(p->problem,
                                                       static_cast<std::uint32_t>(std::min<size_t>(k, 1u << 16)),
                                                       {}, 0, &whole);
    out->n_instr = whole.code.size();
    out->n_slots = whole.n_slots;
    out->n_vars = whole.n_vars;
    out->n_outputs = whole.n_outputs;
    out->reachable_nodes = whole.reachable_nodes;
    out->flops_per_box = whole.flops;
    out->bytes_per_box = 16 * static_cast<std::size_t>(whole.n_vars) + 16 + 1;
    out->box_groups = tp.n_groups;
    out->box_instr = tp.code.size();
    out->box_slots = tp.n_slots;
    out->box_engine = CMI_BOX_AUTO;
    out->box_compiled = 0;
    out->compile_seconds = 0.0;
    out->compiled_persistent = 0;
    out->compiled_wgs_per_cu = 1;
    out->compiled_waves = 1;
    out->extended = whole.extended ? 1 : 0;
    out->compiled_schedule = CMI_SCHED_DEFAULT;
    out->compiled_load_batch = 0;
    if (words != nullptr) {
      if (cap_words < tp.code.size()) {
        throw cmi::ShapeMismatch("word buffer holds " + std::to_string(cap_words) + " words; the tape has "
                                 + std::to_string(tp.code.size()));
      }
      std::memcpy(words, tp.code.data(), tp.code.size() * sizeof(cmi::tape::Instr));
    }
    if (group_begin != nullptr) {
      for (std::uint32_t g = 0; g <= tp.n_groups; ++g) {
        group_begin[g] = tp.group_begin[g];
      }
    }
  });
}

cmi_status cmi_tape_compile_info(const cmi_problem* p, cmi_tape_info* out, char* listing, size_t cap)
{
  if (p == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    p->problem.validate();
    auto const tp = cmi::tape::compile(p->problem);
    out->n_instr = tp.code.size();
    out->n_slots = tp.n_slots;
    out->n_vars = tp.n_vars;
    out->n_outputs = tp.n_outputs;
    out->reachable_nodes = tp.reachable_nodes;
    out->flops_per_box = tp.flops;
    out->bytes_per_box = 16 * static_cast<std::size_t>(tp.n_vars) + 16 + 1;
    out->box_groups = 1;
    out->box_instr = tp.code.size();
    out->box_slots = tp.n_slots;
    out->box_engine = CMI_BOX_AUTO;
    out->box_compiled = 0;
    out->compile_seconds = 0.0;
    out->compiled_persistent = 0;
    out->compiled_wgs_per_cu = 1;
    out->compiled_waves = 1;
    out->extended = tp.extended ? 1 : 0;
    out->compiled_schedule = CMI_SCHED_DEFAULT;
    out->compiled_load_batch = 0;
    if (listing != nullptr && cap > 0) {
      std::string const text = cmi::tape::disassemble(tp);
      std::size_t const k = std::min(cap - 1, text.size());
      std::memcpy(listing, text.data(), k);
      listing[k] = '\0';
    }
  });
}

cmi_status cmi_eval_boxes_device(cmi_tape* t, const double* lb, const double* ub, size_t ld, size_t n_boxes,
                                 uint8_t* feasible, double* obj_lb, double* obj_ub, cmi_ivl* roots, void* stream)
{
  if (t == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null tape");
  }
  return guarded([&] {
    if (n_boxes == 0) {
      return;
    }
    if (ld < n_boxes) {
      throw cmi::ShapeMismatch("eval_boxes: leading dimension smaller than n_boxes");
    }
    if (feasible == nullptr || obj_lb == nullptr || lb == nullptr || ub == nullptr) {
      throw cmi::ShapeMismatch("eval_boxes: lb, ub, feasible and obj_lb are required");
    }
    if (n_boxes > (std::size_t(1) << 31) * 64) {
      throw cmi::InvalidConfiguration("eval_boxes: too many boxes for one launch");
    }
    auto const& ctx = *t->tape->context();
    ctx.activate();
    hipStream_t const s = stream != nullptr ? static_cast<hipStream_t>(stream) : ctx.stream();
    t->tape->eval_boxes_device(lb, ub, ld, n_boxes, feasible, obj_lb, obj_ub, reinterpret_cast<double*>(roots), s);
  });
}

cmi_status cmi_eval_boxes(cmi_tape* t, const double* lb, const double* ub, size_t ld, size_t n_boxes,
                          uint8_t* feasible, double* obj_lb, double* obj_ub, cmi_ivl* roots)
{
  if (t == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null tape");
  }
  return guarded([&] {
    using cmi::hip::check;
    using cmi::hip::DeviceBuffer;
    if (n_boxes == 0) {
      return;
    }
    if (ld < n_boxes) {
      throw cmi::ShapeMismatch("eval_boxes: leading dimension smaller than n_boxes");
    }
    if (feasible == nullptr || obj_lb == nullptr || lb == nullptr || ub == nullptr) {
      throw cmi::ShapeMismatch("eval_boxes: lb, ub, feasible and obj_lb are required");
    }
    auto const& dt = *t->tape;
    auto const& ctx = *dt.context();
    ctx.activate();
    hipStream_t const s = ctx.stream();
    std::size_t const nv = dt.n_vars();
    std::size_t const n_out = dt.tape().n_outputs;
    DeviceBuffer dlb(std::max<std::size_t>(nv, 1) * n_boxes * 8), dub(std::max<std::size_t>(nv, 1) * n_boxes * 8);
    for (std::size_t v = 0; v < nv; ++v) {
      check(hipMemcpyAsync(dlb.as<double>() + v * n_boxes, lb + v * ld, n_boxes * 8, hipMemcpyHostToDevice, s),
            "hipMemcpyAsync(lb)");
      check(hipMemcpyAsync(dub.as<double>() + v * n_boxes, ub + v * ld, n_boxes * 8, hipMemcpyHostToDevice, s),
            "hipMemcpyAsync(ub)");
    }
    DeviceBuffer dfeas(n_boxes), dlo(n_boxes * 8), dhi(obj_ub != nullptr ? n_boxes * 8 : 0),
        droots(roots != nullptr ? std::max<std::size_t>(n_out, 1) * n_boxes * 16 : 0);
    dt.eval_boxes_device(dlb.as<double>(), dub.as<double>(), n_boxes, n_boxes, dfeas.as<uint8_t>(), dlo.as<double>(),
                         dhi.as<double>(), droots.as<double>(), s);
    check(hipMemcpyAsync(feasible, dfeas.get(), n_boxes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(feasible)");
    check(hipMemcpyAsync(obj_lb, dlo.get(), n_boxes * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(obj_lb)");
    if (obj_ub != nullptr) {
      check(hipMemcpyAsync(obj_ub, dhi.get(), n_boxes * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(obj_ub)");
    }
    if (roots != nullptr && n_out > 0) {
      check(hipMemcpyAsync(roots, droots.get(), n_out * n_boxes * 16, hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync(roots)");
    }
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

cmi_status cmi_eval_points(cmi_tape* t, const double* x, size_t ld, size_t n_points, uint8_t* feasible, double* obj,
                           double* roots)
{
  if (t == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null tape");
  }
  return guarded([&] {
    using cmi::hip::check;
    using cmi::hip::DeviceBuffer;
    if (n_points == 0) {
      return;
    }
    if (ld < n_points || x == nullptr || feasible == nullptr || obj == nullptr) {
      throw cmi::ShapeMismatch("eval_points: bad shape or missing output");
    }
    auto const& dt = *t->tape;
    auto const& ctx = *dt.context();
    ctx.activate();
    hipStream_t const s = ctx.stream();
    std::size_t const nv = dt.n_vars();
    std::size_t const n_out = dt.tape().n_outputs;
    DeviceBuffer dx(std::max<std::size_t>(nv, 1) * n_points * 8);
    for (std::size_t v = 0; v < nv; ++v) {
      check(hipMemcpyAsync(dx.as<double>() + v * n_points, x + v * ld, n_points * 8, hipMemcpyHostToDevice, s),
            "hipMemcpyAsync(x)");
    }
    DeviceBuffer dfeas(n_points), dobj(n_points * 8),
        droots(roots != nullptr ? std::max<std::size_t>(n_out, 1) * n_points * 8 : 0);
    dt.eval_points_device(dx.as<double>(), n_points, n_points, dfeas.as<uint8_t>(), dobj.as<double>(),
                          droots.as<double>(), s);
    check(hipMemcpyAsync(feasible, dfeas.get(), n_points, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(feasible)");
    check(hipMemcpyAsync(obj, dobj.get(), n_points * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(obj)");
    if (roots != nullptr && n_out > 0) {
      check(hipMemcpyAsync(roots, droots.get(), n_out * n_points * 8, hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync(roots)");
    }
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

cmi_status cmi_cost_model_get(const cmi_tape* t, size_t samples_per_region, cmi_cost_model* out)
{
  if (t == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    auto const m = cmi::hip::cost_model_for(t->tape->tape(), samples_per_region);
    out->sampler_fixed = m.sampler_fixed;
    out->sampler_per_region = m.sampler_per_region;
    out->bounder_fixed = m.bounder_fixed;
    out->bounder_per_region = m.bounder_per_region;
    out->enumerator_fixed = m.enumerator_fixed;
    out->enumerator_per_region = m.enumerator_per_region;
    out->max_regions = m.max_regions;
  });
}

cmi_status cmi_role_build(cmi_tape* t, int role_kind, const uint8_t* slot_kinds, size_t slot_count, size_t n_regions,
                          size_t samples_per_region, size_t budget_bytes, cmi_role** out, cmi_over_budget* facts)
{
  if (t == nullptr || out == nullptr || (slot_count > 0 && slot_kinds == nullptr)) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  try {
    auto comp = composition_of(slot_kinds, slot_count);
    cmi::dev::Mode mode;
    switch (role_kind) {
      case CMI_ROLE_SAMPLER: mode = cmi::dev::Mode::Sample; break;
      case CMI_ROLE_BOUNDER: mode = cmi::dev::Mode::Bound; break;
      case CMI_ROLE_ENUMERATOR:
        if (!cmi::region::is_fully_enumerable(comp)) {
          throw cmi::InvalidConfiguration("an enumerator needs a fully enumerable composition");
        }
        mode = cmi::dev::Mode::Enumerate;
        break;
      default: throw cmi::InvalidConfiguration("unknown role kind");
    }
    auto impl = std::make_unique<cmi::hip::RoleImpl>(t->tape, mode, comp, n_regions, samples_per_region,
                                                     budget_bytes);
    *out = new cmi_role {std::move(impl), comp, nullptr};
    g_last_error.clear();
    return CMI_OK;
  } catch (const cmi::backend::OverBudgetError& e) {
    if (facts != nullptr) {
      fill_facts(e.facts(), facts);
    }
    return fail(CMI_OVER_BUDGET, e.what());
  } catch (...) {
    return guarded([] { throw; });
  }
}

static cmi_status run_role(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids,
                           const size_t* widths, size_t slot_count, uint64_t salt)
{
  if (r == nullptr || r->impl == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null or destroyed role");
  }
  if (box == nullptr || (slot_count > 0 && (var_ids == nullptr || widths == nullptr))) {
    return fail(CMI_SHAPE_MISMATCH, "null box or slot arrays");
  }
  return guarded([&] {
    std::span<const cmi::interval> b(reinterpret_cast<const cmi::interval*>(box), n_vars);
    r->impl->run(b, {var_ids, slot_count}, {widths, slot_count}, salt);
  });
}

cmi_status cmi_bound(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids, const size_t* widths,
                     size_t slot_count, cmi_bound_view* out)
{
  if (r != nullptr && r->impl != nullptr && r->impl->mode() != cmi::dev::Mode::Bound) {
    return fail(CMI_INVALID_HANDLE, "role is not a bounder");
  }
  cmi_status const s = run_role(r, box, n_vars, var_ids, widths, slot_count, 0);
  if (s == CMI_OK && out != nullptr) {
    out->feasible = r->impl->feasible().data();
    out->obj_lb = r->impl->obj_lb().data();
    out->n_regions = r->impl->n_regions();
  }
  return s;
}

static void fill_candidate(const cmi::hip::RoleImpl& impl, cmi_candidate_view* out)
{
  if (out == nullptr) {
    return;
  }
  out->found = impl.has_candidate() ? 1 : 0;
  out->value = impl.cand_value();
  out->point = impl.witness().data();
  out->index = impl.cand_index();
}

cmi_status cmi_sample(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids, const size_t* widths,
                      size_t slot_count, uint64_t salt, cmi_candidate_view* out)
{
  if (r != nullptr && r->impl != nullptr && r->impl->mode() != cmi::dev::Mode::Sample) {
    return fail(CMI_INVALID_HANDLE, "role is not a sampler");
  }
  cmi_status const s = run_role(r, box, n_vars, var_ids, widths, slot_count, salt);
  if (s == CMI_OK) {
    fill_candidate(*r->impl, out);
  }
  return s;
}

cmi_status cmi_enumerate(cmi_role* r, const cmi_ivl* box, size_t n_vars, const size_t* var_ids,
                         const size_t* widths, size_t slot_count, cmi_candidate_view* out)
{
  if (r != nullptr && r->impl != nullptr && r->impl->mode() != cmi::dev::Mode::Enumerate) {
    return fail(CMI_INVALID_HANDLE, "role is not an enumerator");
  }
  cmi_status const s = run_role(r, box, n_vars, var_ids, widths, slot_count, 0);
  if (s == CMI_OK) {
    fill_candidate(*r->impl, out);
  }
  return s;
}

cmi_status cmi_sample_noalloc(cmi_tape* t, const uint8_t* slot_kinds, size_t slot_count, const cmi_ivl* box,
                              size_t n_vars, const size_t* var_ids, const size_t* widths, size_t samples_per_region,
                              uint64_t salt, int* found, double* value, int64_t* index, double* point)
{
  if (t == nullptr || box == nullptr || found == nullptr || value == nullptr || index == nullptr || point == nullptr
      || (slot_count > 0 && (slot_kinds == nullptr || var_ids == nullptr || widths == nullptr))) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    auto const comp = composition_of(slot_kinds, slot_count);
    std::span<const cmi::interval> const b(reinterpret_cast<const cmi::interval*>(box), n_vars);
    cmi::hip::sample_static(*t->tape, comp, b, {var_ids, slot_count}, {widths, slot_count}, samples_per_region, salt,
                            value, index, point);
    *found = *value < std::numeric_limits<double>::max() ? 1 : 0;
  });
}

size_t cmi_role_size(const cmi_role* r)
{
  if (r == nullptr) {
    return 0;
  }
  if (r->aggregate != nullptr) {
    return r->aggregate->n_subregions();
  }
  return r->impl != nullptr ? r->impl->n_elems() : 0;
}

cmi_status cmi_aggregate_build(cmi_tape* t, const uint8_t* slot_kinds, size_t slot_count, size_t n_regions, size_t k,
                               size_t max_batch, size_t budget_bytes, cmi_role** out, cmi_over_budget* facts)
{
  if (t == nullptr || out == nullptr || (slot_count > 0 && slot_kinds == nullptr)) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  try {
    auto comp = composition_of(slot_kinds, slot_count);
    auto agg = std::make_unique<cmi::hip::AggregateBounderImpl>(t->tape, comp, n_regions, k, max_batch, budget_bytes);
    *out = new cmi_role {nullptr, comp, std::move(agg)};
    g_last_error.clear();
    return CMI_OK;
  } catch (const cmi::backend::OverBudgetError& e) {
    if (facts != nullptr) {
      fill_facts(e.facts(), facts);
    }
    return fail(CMI_OVER_BUDGET, e.what());
  } catch (...) {
    return guarded([] { throw; });
  }
}

cmi_status cmi_bound_children(cmi_role* r, size_t n_batch, const cmi_ivl* boxes, size_t n_vars, const size_t* var_ids,
                              const size_t* widths, size_t slot_count, cmi_aggregate_bound* out)
{
  if (r == nullptr || r->aggregate == nullptr) {
    return fail(CMI_INVALID_HANDLE, "role is not an aggregate bounder");
  }
  if (n_batch > 0 && (boxes == nullptr || out == nullptr || (slot_count > 0 && (var_ids == nullptr || widths == nullptr)))) {
    return fail(CMI_SHAPE_MISMATCH, "null boxes, slot arrays or output");
  }
  return guarded([&] {
    std::vector<cmi::hip::AggregateBounderImpl::RegionView> views(n_batch);
    for (size_t i = 0; i < n_batch; ++i) {
      views[i].box = {reinterpret_cast<const cmi::interval*>(boxes) + i * n_vars, n_vars};
      views[i].var_ids = {var_ids + i * slot_count, slot_count};
      views[i].widths = {widths + i * slot_count, slot_count};
    }
    auto const res = r->aggregate->bound_children(views);
    for (size_t i = 0; i < res.size(); ++i) {
      out[i].lb = res[i].lb;
      out[i].hull_ub = res[i].hull_ub;
      out[i].feasible = res[i].feasible ? 1 : 0;
    }
  });
}

void cmi_role_destroy(cmi_role* r) { delete r; }

void cmi_solve_options_default(cmi_solve_options* o)
{
  if (o == nullptr) {
    return;
  }
  std::memset(o, 0, sizeof(*o));
  o->partition_num = 2;
  o->enumerate_cap = 2;
  o->max_cycle_size = 0;
  o->sample_points = 1;
  o->iter_limit = 1000000;
  o->tolerance = 1e-6;
}

cmi_status cmi_solve(cmi_ctx* ctx, const cmi_problem* p, const cmi_solve_options* o, cmi_solve_outcome* out,
                     double* best_point)
{
  if (ctx == nullptr || p == nullptr || o == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    auto const t0 = std::chrono::steady_clock::now();
    auto pol = cmi::capi::make_policy(*o);
    auto factory = std::make_shared<cmi::hip::HipBackendFactory>(ctx->ctx);
    std::optional<double> known;
    if (o->has_known_primal) {
      known = o->known_primal;
    }
    cmi::search::SearchDriver drv(pol, factory, o->iter_limit, o->tolerance, o->sample_points, o->budget_bytes,
                                  o->host_budget_bytes,
                                  o->compact_frontier ? cmi::search::FrontierPolicy::Compact
                                                      : cmi::search::FrontierPolicy::StopAtBudget,
                                  known);
    auto const res = drv.solve(p->problem);
    double const secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_outcome(res, secs, out, best_point, p->problem.box_bounds.size());
  });
}

void cmi_aggregate_options_default(cmi_aggregate_options* o)
{
  if (o == nullptr) {
    return;
  }
  std::memset(o, 0, sizeof(*o));
  o->branch_fan_out = 4;
  o->partition_budget = 1024;
  o->sample_fan_out = 8;
  o->sample_points = 1;
  o->iter_limit = 1000000;
  o->tolerance = 1e-6;
  o->selection = CMI_SELECT_LEAST_LOWER_BOUND;
}

cmi_status cmi_aggregate_fit(const cmi_problem* p, size_t free_device_bytes, size_t branch_fan_out,
                             size_t sample_points, size_t* partition_budget, size_t* sample_fan_out)
{
  if (p == nullptr || partition_budget == nullptr || sample_fan_out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    cmi::hip::aggregate_fit(cmi::qcqp::reachable_nodes(p->problem), free_device_bytes, branch_fan_out, sample_points,
                            *partition_budget, *sample_fan_out);
  });
}

cmi_status cmi_aggregate_solve(cmi_ctx* ctx, const cmi_problem* p, const cmi_aggregate_options* o,
                               cmi_solve_outcome* out, double* best_point, int* rule_engaged)
{
  if (ctx == nullptr || p == nullptr || o == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    auto const t0 = std::chrono::steady_clock::now();
    auto const run = cmi::hip::aggregate_solve(ctx->ctx, p->problem, *o);
    double const secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_outcome(run.outcome, secs, out, best_point, p->problem.box_bounds.size());
    if (rule_engaged != nullptr) {
      *rule_engaged = run.rule_engaged ? 1 : 0;
    }
  });
}

cmi_status cmi_solve_batch(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems, const cmi_solve_options* o,
                           cmi_solve_outcome* outs, double* const* best_points)
{
  return cmi_solve_batch_stats(ctx, n, problems, o, outs, best_points, nullptr);
}

cmi_status cmi_solve_batch_stats(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems,
                                 const cmi_solve_options* o, cmi_solve_outcome* outs, double* const* best_points,
                                 cmi_serve_stats* stats)
{
  if (ctx == nullptr || (n > 0 && (problems == nullptr || outs == nullptr)) || o == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    std::vector<const cmi::model::Problem<double>*> ps(n);
    for (std::size_t i = 0; i < n; ++i) {
      if (problems[i] == nullptr) {
        throw cmi::InvalidConfiguration("null problem in batch");
      }
      ps[i] = &problems[i]->problem;
    }
    cmi::batch::BatchStats bs;
    auto results = cmi::batch::solve_batch(ctx->ctx, ps, *o, &bs);
    for (std::size_t i = 0; i < n; ++i) {
      fill_outcome(results[i].outcome, results[i].seconds, &outs[i],
                   best_points != nullptr ? best_points[i] : nullptr, ps[i]->box_bounds.size());
    }
    if (stats != nullptr) {
      *stats = cmi_serve_stats {};
      stats->steps = bs.steps;
      stats->launches = bs.launches;
      stats->peak_active = n;
      stats->wall_seconds = bs.wall_seconds;
      stats->kernel_seconds = bs.kernel_seconds;
      stats->kernel_fp64_ops = bs.fp64_ops;
    }
  });
}

// -- GAMS scalar-format models ---------------------------------------------------

void cmi_gams_options_default(cmi_gams_options* o)
{
  if (o == nullptr) {
    return;
  }
  cmi::gams::ParseOptions const d;
  o->default_bound = d.default_bound;
  o->fold_fixed_to_const = d.fold_fixed_to_const ? 1 : 0;
  o->reject_discrete = d.reject_discrete ? 1 : 0;
}

namespace
{

cmi_status gams_parse_into(const std::function<cmi::gams::ParsedModel(const cmi::gams::ParseOptions&)>& parse,
                           const cmi_gams_options* options, cmi_problem** problem, cmi_gams_model** model,
                           cmi_gams_error* err)
{
  if (problem == nullptr || model == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  cmi::gams::ParseOptions opt;
  if (options != nullptr) {
    opt.default_bound = options->default_bound;
    opt.fold_fixed_to_const = options->fold_fixed_to_const != 0;
    opt.reject_discrete = options->reject_discrete != 0;
  }
  if (err != nullptr) {
    err->line = 0;
    err->kind = CMI_GAMS_SYNTAX;
  }
  try {
    cmi::gams::ParsedModel parsed = parse(opt);
    auto p = std::make_unique<cmi_problem>();
    auto m = std::make_unique<cmi_gams_model>();
    p->problem = std::move(parsed.problem);
    m->model = std::move(parsed);
    *problem = p.release();
    *model = m.release();
    g_last_error.clear();
    return CMI_OK;
  } catch (const cmi::gams::ParseError& e) {
    if (err != nullptr) {
      err->line = e.line;
      err->kind = static_cast<int>(e.kind);
    }
    return fail(CMI_PARSE_ERROR, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CMI_ERROR, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(CMI_ERROR, e.what());
  }
}

}  // namespace

cmi_status cmi_gams_parse(const char* text, size_t len, const cmi_gams_options* options, cmi_problem** problem,
                          cmi_gams_model** model, cmi_gams_error* err)
{
  if (text == nullptr && len > 0) {
    return fail(CMI_INVALID_HANDLE, "null text");
  }
  std::string_view const src(text == nullptr ? "" : text, len);
  return gams_parse_into([src](const cmi::gams::ParseOptions& o) { return cmi::gams::parse(src, o); }, options,
                         problem, model, err);
}

cmi_status cmi_gams_parse_file(const char* path, const cmi_gams_options* options, cmi_problem** problem,
                               cmi_gams_model** model, cmi_gams_error* err)
{
  if (path == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null path");
  }
  std::string const file(path);
  return gams_parse_into([&file](const cmi::gams::ParseOptions& o) { return cmi::gams::parse_file(file, o); },
                         options, problem, model, err);
}

cmi_status cmi_gams_info_get(const cmi_gams_model* m, cmi_gams_info* out)
{
  if (m == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  const cmi::gams::ParsedModel& pm = m->model;
  out->sense = pm.sense == cmi::gams::Sense::Maximise ? CMI_SENSE_MAX : CMI_SENSE_MIN;
  out->n_vars = pm.var_names.size();
  out->n_constraints = pm.constraint_names.size();
  out->n_warnings = pm.warnings.size();
  out->objvar_kept = pm.objvar_kept ? 1 : 0;
  out->has_initial_point = pm.initial_point.empty() ? 0 : 1;
  return CMI_OK;
}

const char* cmi_gams_var_name(const cmi_gams_model* m, size_t i)
{
  return (m != nullptr && i < m->model.var_names.size()) ? m->model.var_names[i].c_str() : nullptr;
}

const char* cmi_gams_constraint_name(const cmi_gams_model* m, size_t i)
{
  return (m != nullptr && i < m->model.constraint_names.size()) ? m->model.constraint_names[i].c_str() : nullptr;
}

const char* cmi_gams_warning(const cmi_gams_model* m, size_t i, int* line)
{
  if (m == nullptr || i >= m->model.warnings.size()) {
    return nullptr;
  }
  if (line != nullptr) {
    *line = m->model.warnings[i].line;
  }
  return m->model.warnings[i].message.c_str();
}

cmi_status cmi_gams_initial_point(const cmi_gams_model* m, double* out, size_t n)
{
  if (m == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  const std::vector<double>& x = m->model.initial_point;
  if (x.empty()) {
    return fail(CMI_INVALID_CONFIG, "the model gave no initial levels");
  }
  if (n != x.size()) {
    return fail(CMI_SHAPE_MISMATCH, "initial point length differs from the number of variables");
  }
  std::copy(x.begin(), x.end(), out);
  return CMI_OK;
}

void cmi_gams_destroy(cmi_gams_model* m) { delete m; }

cmi_status cmi_gams_result_line(int sense, double lower, double upper, char* buf, size_t cap, size_t* len)
{
  if (sense != CMI_SENSE_MIN && sense != CMI_SENSE_MAX) {
    return fail(CMI_INVALID_CONFIG, "sense must be CMI_SENSE_MIN or CMI_SENSE_MAX");
  }
  std::string const line = cmi::gams::result_line(
      sense == CMI_SENSE_MAX ? cmi::gams::Sense::Maximise : cmi::gams::Sense::Minimise, lower, upper);
  if (len != nullptr) {
    *len = line.size();
  }
  if (buf != nullptr && cap > 0) {
    std::size_t const k = std::min(cap - 1, line.size());
    std::memcpy(buf, line.data(), k);
    buf[k] = '\0';
  }
  return CMI_OK;
}

// -- memory-fit resolver --------------------------------------------------------

namespace
{
cmi::config::ProblemProfile profile_arg(const cmi_problem* p, const cmi_problem_profile* pp)
{
  if (pp == nullptr) {
    return cmi::config::profile_of(p->problem);
  }
  cmi::config::ProblemProfile out;
  out.num_binary = pp->num_binary;
  out.num_integer = pp->num_integer;
  out.num_continuous = pp->num_continuous;
  out.largest_integer_domain = pp->largest_integer_domain;
  out.objvar_kept = pp->objvar_kept != 0;
  return out;
}

/// The device roles' cost coefficients, priced from the problem's host tape.
cmi::backend::RegionCostModel host_cost(const cmi_problem* p, std::size_t sample_points)
{
  p->problem.validate();
  return cmi::hip::cost_model_for(cmi::tape::compile(p->problem), std::max<std::size_t>(sample_points, 1));
}
}  // namespace

cmi_status cmi_problem_profile_get(const cmi_problem* p, cmi_problem_profile* out)
{
  if (p == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    auto const pp = cmi::config::profile_of(p->problem);
    out->num_binary = pp.num_binary;
    out->num_integer = pp.num_integer;
    out->num_continuous = pp.num_continuous;
    out->largest_integer_domain = pp.largest_integer_domain;
    out->objvar_kept = 0;
  });
}

cmi_status cmi_ctx_memory(cmi_ctx* ctx, size_t* free_bytes, size_t* total_bytes)
{
  if (ctx == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    ctx->ctx->activate();
    std::size_t f = 0, t = 0;
    cmi::hip::check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes != nullptr) {
      *free_bytes = f;
    }
    if (total_bytes != nullptr) {
      *total_bytes = t;
    }
  });
}

cmi_status cmi_worst_footprint(const cmi_problem* p, const cmi_problem_profile* profile, size_t cap,
                               size_t partition_num, size_t enumerate_cap, size_t sample_points, size_t* bytes)
{
  if (p == nullptr || bytes == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    cmi::region::FanOutSpec {partition_num, enumerate_cap};
    *bytes = cmi::config::worst_footprint(cap, profile_arg(p, profile), host_cost(p, sample_points), partition_num,
                                          enumerate_cap);
  });
}

cmi_status cmi_auto_max_cycle_size(const cmi_problem* p, const cmi_problem_profile* profile, size_t partition_num,
                                   size_t enumerate_cap, size_t sample_points, size_t budget_bytes, size_t ceiling,
                                   size_t* out)
{
  if (p == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    cmi::region::FanOutSpec {partition_num, enumerate_cap};
    *out = cmi::config::auto_max_cycle_size(profile_arg(p, profile), host_cost(p, sample_points), partition_num,
                                            enumerate_cap, budget_bytes, ceiling);
  });
}

cmi_status cmi_explain_over_budget(const cmi_problem* p, int objvar_kept, const cmi_over_budget* facts,
                                   size_t partition_num, size_t enumerate_cap, size_t sample_points, char* buf,
                                   size_t cap, size_t* len)
{
  if (p == nullptr || facts == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    cmi::region::FanOutSpec {partition_num, enumerate_cap};
    cmi::backend::OverBudget const f = facts_from(*facts);
    cmi::config::ProblemProfile pp = cmi::config::profile_of(p->problem);
    pp.objvar_kept = objvar_kept != 0;
    std::string const msg = cmi::config::explain_over_budget(f, pp, host_cost(p, sample_points), partition_num,
                                                             enumerate_cap);
    if (len != nullptr) {
      *len = msg.size();
    }
    if (buf != nullptr && cap > 0) {
      std::size_t const k = std::min(cap - 1, msg.size());
      std::memcpy(buf, msg.data(), k);
      buf[k] = '\0';
    }
  });
}

cmi_status cmi_resolve_shape(const cmi_problem* p, const cmi_resolve_request* req, cmi_resolved_shape* out)
{
  if (p == nullptr || req == nullptr || out == nullptr) {
    return fail(CMI_INVALID_HANDLE, "null argument");
  }
  return guarded([&] {
    cmi::config::ProblemProfile pp = cmi::config::profile_of(p->problem);
    pp.objvar_kept = req->objvar_kept != 0;
    cmi::config::ShapeRules rules;
    cmi::config::Source source = cmi::config::Source::Auto;
    if (req->rules != nullptr) {
      auto const found = cmi::config::lookup_rules(req->rules);
      if (!found) {
        std::string names;
        for (const auto& r : cmi::config::kRoster) {
          names += (names.empty() ? "" : ", ") + std::string(r.name);
        }
        throw cmi::InvalidConfiguration("unknown rules '" + std::string(req->rules) + "'; the roster has " + names);
      }
      if (!cmi::config::is_applicable(*found, pp)) {
        throw cmi::InvalidConfiguration(
            "rules '" + std::string(req->rules) + "' do not apply to a problem with " + std::to_string(pp.num_binary)
            + " binary, " + std::to_string(pp.num_integer) + " integer and " + std::to_string(pp.num_continuous)
            + " continuous variable(s)" + (pp.objvar_kept ? " (one continuous is the kept objective variable)" : ""));
      }
      rules = *found;
      source = cmi::config::Source::Named;
    } else {
      rules = cmi::config::select_rules(pp);
    }
    cmi::config::Overrides ov;
    ov.partition_num = req->partition_num;
    ov.enumerate_cap = req->enumerate_cap;
    ov.sample_points = req->sample_points;
    ov.max_cycle_size = req->max_cycle_size;
    auto const fitted =
        cmi::config::resolve_shape(rules, pp, req->free_device_bytes, host_cost(p, rules.sample_points));
    auto const s = cmi::config::apply_overrides(fitted, ov);
    out->partition_num = s.partition_num;
    out->enumerate_cap = s.enumerate_cap;
    out->sample_points = s.sample_points;
    out->max_cycle_size = s.max_cycle_size;
    out->budget_bytes = req->free_device_bytes > 0
        ? static_cast<std::size_t>(static_cast<double>(req->free_device_bytes) * cmi::config::kBudgetFraction)
        : 0;
    out->footprint_bytes = cmi::config::worst_footprint(std::min(s.max_cycle_size, cmi::config::kMaxSlots), pp,
                                                        host_cost(p, s.sample_points), s.partition_num,
                                                        s.enumerate_cap);
    out->source = ov.any() ? CMI_SOURCE_OVERRIDDEN
                           : (source == cmi::config::Source::Named ? CMI_SOURCE_NAMED : CMI_SOURCE_AUTO);
    std::memset(out->rules, 0, sizeof(out->rules));
    std::memcpy(out->rules, rules.name.data(), std::min(rules.name.size(), sizeof(out->rules) - 1));
  });
}

}  // extern "C"
