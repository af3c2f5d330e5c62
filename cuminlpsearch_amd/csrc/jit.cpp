#include "jit.hpp"

#include <hip/hiprtc.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <functional>
#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <tuple>
#include <vector>

#include "cmi/errors.hpp"
#include "device.hpp"

// Text of interval_ops.hpp, embedded at build time.
extern "C" const char cmi_interval_ops_source[];

namespace cmi::jit
{

namespace
{

using tape::AccMode;
using tape::Instr;
using tape::Term;
using tape::TOp;

/// An exact C++ literal for a double (hexadecimal significand).
std::string lit(double x)
{
  if (std::isnan(x)) {
    return "__builtin_nan(\"\")";
  }
  if (std::isinf(x)) {
    return x > 0 ? "__builtin_huge_val()" : "(-__builtin_huge_val())";
  }
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return std::string("(") + buf + ")";
}

std::string slot(std::uint32_t s) { return "s" + std::to_string(s); }

/// [a] * c with the sign of c known when the source is written.
std::string mul_imm(const std::string& a, double c)
{
  if (c > 0.0) {
    return "iv_mul_pos(" + a + ", " + lit(c) + ", t)";
  }
  if (c < 0.0) {
    return "iv_mul_neg(" + a + ", " + lit(c) + ", t)";
  }
  if (std::isnan(c)) {
    return "iv_mul_imm(" + a + ", " + lit(c) + ", t)";
  }
  return "Iv{0.0, 0.0}";
}

/// The product a term word forms from variable slots before scaling and
/// accumulating it, iv_mul(a, b) or iv_sqr(a), as a key (0: the word forms
/// none, or reads a temporary slot, which the tape may rewrite). Variable
/// slots are never written, so words with equal keys compute bit-identical
/// intervals.
std::uint64_t product_key(const Instr& w, std::uint32_t n_vars)
{
  if (!tape::is_term_op(w.op) || w.a >= n_vars) {
    return 0;
  }
  switch (static_cast<Term>(w.op & 7u)) {
    case Term::SS:
    case Term::SSC: return w.b < n_vars ? (std::uint64_t {2} << 40) | (std::uint64_t {w.a} << 20) | w.b : 0;
    case Term::Q:
    case Term::QC: return (std::uint64_t {1} << 40) | w.a;
    default: return 0;
  }
}

/// The expression of the product a term word forms (product_key).
std::string product_expr(const Instr& w)
{
  auto const kind = static_cast<Term>(w.op & 7u);
  if (kind == Term::SS || kind == Term::SSC) {
    return "iv_mul(" + slot(w.a) + ", " + slot(w.b) + ", t)";
  }
  return "iv_sqr(" + slot(w.a) + ", t)";
}

/// One tape word as one statement; `acc` names the group's accumulator. A
/// term word whose product is already held in a variable takes its name as
/// `prod` instead of forming the product again.
std::string emit_word(const Instr& w, const std::string& acc, bool roots, const std::string& prod = {})
{
  std::string const a = slot(w.a), b = slot(w.b), d = slot(w.dst), c = lit(w.imm);
  auto set = [](const std::string& lhs, const std::string& rhs) { return "  " + lhs + " = " + rhs + ";\n"; };
  auto call = [](const std::string& fn, const std::string& x, const std::string& y) {
    return fn + "(" + x + ", " + y + ", t)";
  };
  auto output = [&](const std::string& x) -> std::string {
    if (!roots) {
      return {};
    }
    return "  if (rt != nullptr) out_store(rt, " + std::to_string(w.b) + ", n, e, to_output(" + x + ", t));\n";
  };
  if (tape::is_term_op(w.op)) {
    std::string x;
    switch (static_cast<Term>(w.op & 7u)) {
      case Term::S: x = a; break;
      case Term::SC: x = mul_imm(a, w.imm); break;
      case Term::SS: x = prod.empty() ? call("iv_mul", a, b) : prod; break;
      case Term::Q: x = prod.empty() ? "iv_sqr(" + a + ", t)" : prod; break;
      case Term::SSC: x = mul_imm(prod.empty() ? call("iv_mul", a, b) : prod, w.imm); break;
      case Term::QC: x = mul_imm(prod.empty() ? "iv_sqr(" + a + ", t)" : prod, w.imm); break;
    }
    switch (static_cast<AccMode>((w.op - tape::kTermBase) >> 3)) {
      case AccMode::Set: return set(acc, x);
      case AccMode::Add: return set(acc, call("iv_add", acc, x));
      case AccMode::Sub: return set(acc, call("iv_sub", acc, x));
    }
    throw InvalidConfiguration("tape word with an unknown accumulate mode");
  }
  switch (static_cast<TOp>(w.op)) {
    case TOp::Nop: return {};
    case TOp::LoadConst: return set(d, "iv_const(" + c + ", t)");
    case TOp::Add: return set(d, call("iv_add", a, b));
    case TOp::AddImm: return set(d, call("iv_add_imm", a, c));
    case TOp::Sub: return set(d, call("iv_sub", a, b));
    case TOp::SubImm: return set(d, call("iv_sub_imm", a, c));
    case TOp::ImmSub: return set(d, call("iv_imm_sub", c, a));
    case TOp::Mul: return set(d, call("iv_mul", a, b));
    case TOp::MulImm: return set(d, mul_imm(a, w.imm));
    case TOp::Sqr: return set(d, "iv_sqr(" + a + ", t)");
    case TOp::Neg: return set(d, "iv_neg(" + a + ")");
    case TOp::CheckLE: return set("f", "f && check_le(" + a + ", " + c + ")");
    case TOp::CheckEQ: return set("f", "f && check_eq(" + a + ", " + c + ")");
    case TOp::Objective: return set("obj", a);
    case TOp::Output: return output(a);
    case TOp::AccStore: return set(d, acc);
    case TOp::AccAddImm: return set(acc, call("iv_add_imm", acc, c));
    case TOp::AccSubImm: return set(acc, call("iv_sub_imm", acc, c));
    case TOp::AccImmSub: return set(acc, call("iv_imm_sub", c, acc));
    case TOp::AccScale: return set(acc, mul_imm(acc, w.imm));
    case TOp::AccNeg: return set(acc, "iv_neg(" + acc + ")");
    case TOp::AccCheckLE: return set("f", "f && check_le(" + acc + ", " + c + ")");
    case TOp::AccCheckEQ: return set("f", "f && check_eq(" + acc + ", " + c + ")");
    case TOp::AccObjective: return set("obj", acc);
    case TOp::AccOutput: return output(acc);
    // extended op set; `f` (the group's feasibility) takes the empty results
    case TOp::Div: return set(d, "iv_div(" + a + ", " + b + ", t, f)");
    case TOp::DivImm: return set(d, "iv_div_imm(" + a + ", " + c + ", t, f)");
    case TOp::ImmDiv: return set(d, "iv_imm_div(" + c + ", " + a + ", t, f)");
    case TOp::PowN: return set(d, "iv_pown(" + a + ", " + std::to_string(static_cast<int>(w.imm)) + ", t, f)");
    case TOp::Sqrt: return set(d, "iv_sqrt(" + a + ", t, f)");
    case TOp::Exp: return set(d, "iv_exp(" + a + ", t)");
    case TOp::Log: return set(d, "iv_log(" + a + ", t, f)");
    case TOp::Sin: return set(d, "iv_sincos(" + a + ", 0, t)");
    case TOp::Cos: return set(d, "iv_sincos(" + a + ", 1, t)");
    case TOp::Tanh: return set(d, "iv_tanh(" + a + ", t)");
    case TOp::Abs: return set(d, "iv_abs(" + a + ")");
    case TOp::Min: return set(d, "iv_min(" + a + ", " + b + ")");
    case TOp::MinImm: return set(d, "iv_min(" + a + ", iv_const(" + c + ", t))");
    case TOp::Max: return set(d, "iv_max(" + a + ", " + b + ")");
    case TOp::MaxImm: return set(d, "iv_max(" + a + ", iv_const(" + c + ", t))");
    default: throw InvalidConfiguration("tape word with an op the compiled kernel does not know");
  }
}

std::uint64_t fnv1a(const void* p, std::size_t n, std::uint64_t h = 0xcbf29ce484222325ULL)
{
  auto const* b = static_cast<const unsigned char*>(p);
  for (std::size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 0x100000001b3ULL;
  }
  return h;
}

void rtc_check(hiprtcResult r, const char* what)
{
  if (r != HIPRTC_SUCCESS) {
    throw DeviceError(static_cast<int>(r), std::string(what) + ": " + hiprtcGetErrorString(r));
  }
}

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    throw DeviceError(static_cast<int>(e), std::string(what) + ": " + hipGetErrorString(e));
  }
}

// Same code-generation flags as the library's own kernels: no contraction,
// packed FP32 off; gfx950 only.
const char* const kOptions[] = {
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
};
// The scheduler that orders each block for instruction-level parallelism
// rather than for occupancy (Shape::ilp_schedule).
const char* const kIlpOptions[] = {"-mllvm", "-amdgpu-sched-strategy=max-ilp"};

}  // namespace

namespace
{

/// Which slot fields of one tape word are read and written.
struct WordUse
{
  bool reads_a, reads_b, writes;
};

WordUse word_use(const Instr& w)
{
  bool const term = tape::is_term_op(w.op);
  auto const op = static_cast<TOp>(w.op);
  unsigned const kind = w.op & 7u;
  bool const ext = tape::is_ext_op(w.op);
  WordUse u;
  u.reads_a =
      term || ext || (w.op >= static_cast<std::uint8_t>(TOp::Add) && w.op <= static_cast<std::uint8_t>(TOp::Output));
  u.reads_b = (term && (kind == static_cast<unsigned>(Term::SS) || kind == static_cast<unsigned>(Term::SSC)))
      || op == TOp::Add || op == TOp::Sub || op == TOp::Mul || op == TOp::Div || op == TOp::Min || op == TOp::Max;
  u.writes = (w.op >= static_cast<std::uint8_t>(TOp::LoadConst) && w.op <= static_cast<std::uint8_t>(TOp::Neg))
      || op == TOp::AccStore || ext;
  return u;
}

/// Slots group code [lo, hi) reads before writing (its variables) and writes.
void slot_use(const tape::Tape& t, std::uint32_t lo, std::uint32_t hi, std::set<std::uint32_t>& vars,
              std::set<std::uint32_t>& temps)
{
  for (std::uint32_t i = lo; i < hi; ++i) {
    Instr const& w = t.code[i];
    auto note = [&](std::uint32_t s) { (s < t.n_vars ? vars : temps).insert(s); };
    WordUse const u = word_use(w);
    if (u.reads_a) {
      note(w.a);
    }
    if (u.reads_b) {
      note(w.b);
    }
    if (u.writes) {
      note(w.dst);
    }
  }
}

/// The words [lo, hi) of one function group in the order a wave whose
/// variables arrive in batches should run them. A group is a run of function
/// blocks, each starting at a term word that sets the accumulator. When no
/// word touches a temporary slot, the blocks share nothing but variables
/// (read-only) and products of variables (recomputed or held by value), so
/// any block order gives the same values; the blocks then go greedily, each
/// next the one that reads the fewest variables no earlier block read (ties:
/// tape order), so the first statements need the fewest loads. Otherwise the
/// tape order is kept.
std::vector<std::uint32_t> arrival_order(const tape::Tape& t, std::uint32_t lo, std::uint32_t hi)
{
  std::vector<std::uint32_t> same;
  for (std::uint32_t i = lo; i < hi; ++i) {
    same.push_back(i);
  }
  std::uint32_t const nv = t.n_vars;
  std::vector<std::pair<std::uint32_t, std::uint32_t>> blocks;  // [begin, end)
  std::vector<std::set<std::uint32_t>> reads;
  for (std::uint32_t i = lo; i < hi; ++i) {
    Instr const& w = t.code[i];
    WordUse const u = word_use(w);
    if (u.writes || (u.reads_a && w.a >= nv) || (u.reads_b && w.b >= nv)) {
      return same;
    }
    bool const starts =
        tape::is_term_op(w.op) && static_cast<AccMode>((w.op - tape::kTermBase) >> 3) == AccMode::Set;
    if (starts) {
      blocks.emplace_back(i, i);
      reads.emplace_back();
    }
    if (blocks.empty()) {
      return same;
    }
    blocks.back().second = i + 1;
    if (u.reads_a) {
      reads.back().insert(w.a);
    }
    if (u.reads_b) {
      reads.back().insert(w.b);
    }
  }
  std::vector<std::uint32_t> out;
  std::vector<bool> used(blocks.size(), false);
  std::set<std::uint32_t> have;
  for (std::size_t step = 0; step < blocks.size(); ++step) {
    std::size_t best = blocks.size();
    std::size_t best_new = 0;
    for (std::size_t b = 0; b < blocks.size(); ++b) {
      if (used[b]) {
        continue;
      }
      std::size_t fresh = 0;
      for (std::uint32_t x : reads[b]) {
        fresh += have.count(x) == 0 ? 1 : 0;
      }
      if (best == blocks.size() || fresh < best_new) {
        best = b;
        best_new = fresh;
      }
    }
    used[best] = true;
    have.insert(reads[best].begin(), reads[best].end());
    for (std::uint32_t i = blocks[best].first; i < blocks[best].second; ++i) {
      out.push_back(i);
    }
  }
  return out;
}

}  // namespace

std::uint32_t shape_waves(const tape::Tape& t, const Shape& shape)
{
  std::uint32_t const k = std::max<std::uint32_t>(t.n_groups, 1);
  return shape.waves == 0 ? k : std::clamp<std::uint32_t>(shape.waves, 1, k);
}

bool persistent_fits(const tape::Tape& t, std::uint32_t waves)
{
  std::uint64_t const w = std::max<std::uint32_t>(waves, 1);
  std::uint64_t const bytes = 2ull * t.n_vars * 64 * 16 + 2 * w * 64 * 4 + 2 * 128 * 8;
  return t.n_vars > 0 && bytes <= kLdsBytes;
}

namespace
{

/// The wave blocks of a tape kernel: "if (wave == 0u) {...} else if ..."; wave
/// v runs groups v, v + w, ... over its variables, each bound by the
/// expression `load(slot, readers)` gives (readers: how many waves read that
/// variable). `batch` > 0 issues those bindings `batch` variables at a time
/// (see Shape::load_batch); `roots` writes root intervals. Each wave block
/// ends with the objective group's output in obj_s[cur * 128 + lane] (lb)
/// and [cur * 128 + 64 + lane] (ub).
void emit_waves(std::ostringstream& o, const tape::Tape& t, std::uint32_t w, const std::string& ind, bool roots,
                std::uint32_t batch, const std::function<std::string(std::uint32_t, std::uint32_t)>& load)
{
  std::uint32_t const k = std::max<std::uint32_t>(t.n_groups, 1);
  std::uint32_t const nv = t.n_vars;
  // Wave v runs groups v, v + w, v + 2w, ...; their statements are
  // interleaved one by one, each group keeping its own order and accumulator.
  // The groups share nothing they write, so the results are those of the
  // groups run one after another, while neighbouring statements are
  // independent (FP64 results take several issue slots to come back).
  // A product that several term words of the wave form (a product node the
  // DAG shares between functions) is formed once, at its first use in that
  // order, and held for the others: the same operation on the same slots, so
  // the same bits.
  // How many waves of the workgroup read each variable: one reader streams
  // it past the caches, several let the later ones find it in L2.
  std::vector<std::uint32_t> readers(nv, 0);
  for (std::uint32_t v = 0; v < w; ++v) {
    std::set<std::uint32_t> vars, temps;
    for (std::uint32_t g = v; g < k; g += w) {
      std::uint32_t const lo = t.group_begin.size() > g ? t.group_begin[g] : 0;
      std::uint32_t const hi =
          t.group_begin.size() > g + 1 ? t.group_begin[g + 1] : static_cast<std::uint32_t>(t.code.size());
      slot_use(t, lo, hi, vars, temps);
    }
    for (std::uint32_t x : vars) {
      ++readers[x];
    }
  }
  for (std::uint32_t v = 0; v < w; ++v) {
    std::set<std::uint32_t> vars, temps;
    std::vector<std::vector<std::uint32_t>> words;
    bool has_obj = false;
    for (std::uint32_t g = v; g < k; g += w) {
      std::uint32_t const lo = t.group_begin.size() > g ? t.group_begin[g] : 0;
      std::uint32_t const hi =
          t.group_begin.size() > g + 1 ? t.group_begin[g + 1] : static_cast<std::uint32_t>(t.code.size());
      slot_use(t, lo, hi, vars, temps);
      if (batch != 0) {
        words.push_back(arrival_order(t, lo, hi));
      } else {
        std::vector<std::uint32_t> st;
        for (std::uint32_t i = lo; i < hi; ++i) {
          st.push_back(i);
        }
        words.push_back(std::move(st));
      }
      has_obj = has_obj || g == t.objective_group;
    }
    std::vector<std::pair<std::size_t, std::uint32_t>> order;  // (stream, word) in emission order
    for (std::size_t i = 0;; ++i) {
      bool any = false;
      for (std::size_t j = 0; j < words.size(); ++j) {
        if (i < words[j].size()) {
          order.emplace_back(j, words[j][i]);
          any = true;
        }
      }
      if (!any) {
        break;
      }
    }
    std::map<std::uint64_t, std::uint32_t> uses;
    for (auto const& [j, i] : order) {
      if (std::uint64_t const key = product_key(t.code[i], nv)) {
        ++uses[key];
      }
    }
    // variables in the order the wave's statements first read them, so the
    // waits for their loads come due in the order the loads were issued
    std::vector<std::uint32_t> first_use;
    std::map<std::uint32_t, std::size_t> rank;  // variable -> position in first_use
    {
      auto see = [&](std::uint32_t x) {
        if (vars.count(x) != 0 && rank.emplace(x, first_use.size()).second) {
          first_use.push_back(x);
        }
      };
      for (auto const& [j, i] : order) {
        WordUse const u = word_use(t.code[i]);
        if (u.reads_a) {
          see(t.code[i].a);
        }
        if (u.reads_b) {
          see(t.code[i].b);
        }
      }
      for (std::uint32_t x : vars) {
        see(x);
      }
    }
    // The statements, each with 1 + the first-use position of the last
    // variable it reads (0: it reads none).
    std::map<std::uint64_t, std::string> held;
    std::vector<std::pair<std::string, std::size_t>> body;
    for (auto const& [j, i] : order) {
      Instr const& word = t.code[i];
      std::string const acc = "acc" + std::to_string(j);
      std::size_t need = 0;
      WordUse const u = word_use(word);
      if (u.reads_a && word.a < nv) {
        need = std::max(need, rank.at(word.a) + 1);
      }
      if (u.reads_b && word.b < nv) {
        need = std::max(need, rank.at(word.b) + 1);
      }
      std::uint64_t const key = product_key(word, nv);
      std::string line;
      if (key != 0 && uses[key] > 1) {
        auto it = held.find(key);
        if (it == held.end()) {
          std::string const name = "p" + std::to_string(held.size());
          body.emplace_back(ind + "  const Iv " + name + " = " + product_expr(word) + ";\n", need);
          it = held.emplace(key, name).first;
        }
        line = emit_word(word, acc, roots, it->second);
      } else {
        line = emit_word(word, acc, roots);
      }
      if (!line.empty()) {
        body.emplace_back(ind + line, need);
      }
    }
    o << ind << (v == 0 ? "if" : "else if") << " (wave == " << v << "u) {\n";
    // Variable loads. Batched (direct form): the first two batches up front,
    // then batch b+1 just before the first statement that reads batch b; a
    // scheduling barrier on each side keeps the compiler from hoisting them.
    std::size_t const n_batches = batch == 0 ? 1 : (first_use.size() + batch - 1) / batch;
    std::size_t issued = 0;
    auto issue_until = [&](std::size_t upto) {
      for (upto = std::min(upto, n_batches); issued < upto; ++issued) {
        std::size_t const lo = batch == 0 ? 0 : issued * batch;
        std::size_t const hi = batch == 0 ? first_use.size() : std::min(first_use.size(), lo + batch);
        if (batch != 0) {
          o << ind << "  __builtin_amdgcn_sched_barrier(0);\n";
        }
        for (std::size_t q = lo; q < hi; ++q) {
          std::uint32_t const x = first_use[q];
          o << ind << "  const Iv " << slot(x) << " = " << load(x, readers[x]) << ";\n";
        }
        if (batch != 0) {
          o << ind << "  __builtin_amdgcn_sched_barrier(0);\n";
        }
      }
    };
    issue_until(batch == 0 ? 1 : 2);
    for (std::uint32_t x : temps) {
      o << ind << "  Iv " << slot(x) << ";\n";
    }
    for (std::size_t j = 0; j < words.size(); ++j) {
      o << ind << "  Iv acc" << j << " = {0.0, 0.0};\n";
    }
    o << ind << "  Iv obj = {0.0, 0.0};\n";
    for (auto const& [line, need] : body) {
      if (batch != 0 && need > 0) {
        issue_until((need - 1) / batch + 2);
      }
      o << line;
    }
    issue_until(n_batches);  // variables no statement reads: dead loads the compiler drops
    if (has_obj) {
      o << ind << "  const Bnd ob = to_output(obj, t);\n";
      o << ind << "  obj_s[cur * 128u + lane] = ob.lb;\n" << ind << "  obj_s[cur * 128u + 64u + lane] = ob.ub;\n";
    }
    o << ind << "}\n";
  }
}

}  // namespace

std::string boxes_source(const tape::Tape& t, const Shape& shape)
{
  std::uint32_t const k = std::max<std::uint32_t>(t.n_groups, 1);
  std::uint32_t const w = shape_waves(t, shape);
  std::uint32_t const nv = t.n_vars;
  bool const persistent = shape.persistent;
  if (persistent && !persistent_fits(t, w)) {
    throw InvalidConfiguration("the double-buffered box tiles of this tape do not fit in LDS");
  }
  std::ostringstream o;
  o << cmi_interval_ops_source << "\n";
  o << "using namespace cmi_iv;\n\n";
  o << "__device__ __forceinline__ void out_store(double* roots, unsigned long k, unsigned long n, unsigned long e,\n"
       "                                          Bnd v)\n{\n"
       "  roots[2 * (k * n + e)] = v.lb;\n  roots[2 * (k * n + e) + 1] = v.ub;\n}\n\n";
  o << "extern \"C\" __global__ void __launch_bounds__(" << 64 * w << ") " << kBoxesKernel
    << "(const double* __restrict__ lb, const double* __restrict__ ub, unsigned long ld, unsigned long n,\n"
       "    unsigned char* __restrict__ feasible, double* __restrict__ obj_lb, double* __restrict__ obj_ub,\n"
       "    double* __restrict__ roots)\n{\n";
  o << "  const unsigned lane = threadIdx.x & 63u;\n";
  o << "  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n";
  if (persistent) {
    // Persistent workgroups walk the 64-box tiles with a stride of the grid.
    // While the waves evaluate tile i from one LDS buffer, each wave's share
    // of tile i+1's variables is already in flight from HBM into registers,
    // and goes to the other buffer before the barrier that ends the step.
    std::uint32_t const share = (nv + w - 1) / w;
    o << "  __shared__ double2 xs[" << 2 * nv * 64 << "];\n";
    o << "  __shared__ unsigned feas_s[" << 2 * 64 * w << "];\n";
    o << "  __shared__ double obj_s[256];\n";
    o << "  const unsigned long n_tiles = (n + 63ul) / 64ul;\n";
    o << "  unsigned long tile = blockIdx.x;\n";
    o << "  if (tile >= n_tiles) {\n    return;\n  }\n";
    o << "  {\n    const unsigned long e0 = tile * 64ul + lane;\n"
         "    const unsigned off = 8u * (unsigned)(e0 < n ? e0 : n - 1);\n";
    o << "    for (unsigned v = wave; v < " << nv << "u; v += " << w << "u) {\n"
         "      const Iv x = load_box(lb + v * ld, ub + v * ld, off);\n"
         "      xs[v * 64u + lane] = make_double2(x.nl, x.ub);\n    }\n  }\n";
    o << "  __syncthreads();\n";
    o << "  for (unsigned step = 0;; ++step) {\n";
    o << "    const unsigned cur = step & 1u;\n";
    o << "    const unsigned long e = tile * 64ul + lane;\n";
    o << "    const bool active = e < n;\n";
    o << "    const unsigned long next = tile + gridDim.x;\n";
    o << "    const bool has_next = next < n_tiles;\n";
    o << "    const unsigned long en = next * 64ul + lane;\n";
    o << "    const unsigned offn = 8u * (unsigned)(en < n ? en : n - 1);\n";
    o << "    Iv pre[" << share << "];\n";
    o << "    if (has_next) {\n";
    for (std::uint32_t j = 0; j < share; ++j) {
      o << "      if (wave + " << j * w << "u < " << nv << "u) {\n"
        << "        pre[" << j << "] = load_box(lb + (wave + " << j * w << "ul) * ld, ub + (wave + " << j * w
        << "ul) * ld, offn);\n      }\n";
    }
    o << "    }\n";
    o << "    const double2* const X = xs + cur * " << nv * 64 << "u;\n";
  } else {
    // One workgroup per 64 boxes; each wave reads the variables its groups
    // need straight from the box arrays.
    o << "  __shared__ unsigned feas_s[" << 64 * w << "];\n";
    o << "  __shared__ double obj_s[128];\n";
    o << "  const unsigned cur = 0u;\n";
    o << "  const unsigned long e = (unsigned long)blockIdx.x * 64ul + lane;\n";
    o << "  const bool active = e < n;\n";
    o << "  const unsigned eo = 8u * (unsigned)(active ? e : n - 1);\n";
  }
  std::string const ind = persistent ? "    " : "  ";
  if (shape.roots) {
    o << ind << "double* const rt = active ? roots : nullptr;\n";
  }
  o << ind << "bool f = true;\n";
  o << ind << "const int t = enter_round_up();\n";
  std::uint32_t const batch = persistent ? 0u : shape.load_batch;
  emit_waves(o, t, w, ind, shape.roots, batch, [&](std::uint32_t x, std::uint32_t readers) -> std::string {
    if (persistent) {
      return "{X[" + std::to_string(x * 64) + "u + lane].x, X[" + std::to_string(x * 64) + "u + lane].y}";
    }
    return std::string(readers > 1 ? "load_box_shared" : "load_box") + "(lb + " + std::to_string(x) + "ul * ld, ub + "
           + std::to_string(x) + "ul * ld, eo)";
  });
  o << ind << "leave_round_up(t, f);\n";
  o << ind << "feas_s[cur * " << 64 * w << "u + wave * 64u + lane] = f ? 1u : 0u;\n";
  if (persistent) {
    std::uint32_t const share = (nv + w - 1) / w;
    o << "    if (has_next) {\n";
    for (std::uint32_t j = 0; j < share; ++j) {
      o << "      if (wave + " << j * w << "u < " << nv << "u) {\n"
        << "        xs[(1u - cur) * " << nv * 64 << "u + (wave + " << j * w << "u) * 64u + lane] = make_double2(pre["
        << j << "].nl, pre[" << j << "].ub);\n      }\n";
    }
    o << "    }\n";
  }
  o << ind << "__syncthreads();\n";
  o << ind << "if (wave == 0u && active) {\n";
  o << ind << "  unsigned all = 1u;\n";
  for (std::uint32_t v = 0; v < w; ++v) {
    o << ind << "  all &= feas_s[cur * " << 64 * w << "u + " << v * 64 << "u + lane];\n";
  }
  o << ind << "  feasible[e] = all ? 1 : 0;\n" << ind << "  obj_lb[e] = obj_s[cur * 128u + lane];\n"
    << ind << "  if (obj_ub != nullptr) {\n" << ind << "    obj_ub[e] = obj_s[cur * 128u + 64u + lane];\n" << ind
    << "  }\n" << ind << "}\n";
  if (persistent) {
    o << "    if (!has_next) {\n      break;\n    }\n    tile = next;\n  }\n";
  }
  o << "}\n";
  return o.str();
}

bool items_fit(const tape::Tape& t)
{
  std::uint64_t const w = std::max<std::uint32_t>(t.n_groups, 1);
  return t.n_vars > 0 && w <= dev::kMaxGroups
         && std::uint64_t {t.n_vars} * 64 * 16 + w * 64 * 4 + 128 * 8 <= kLdsBytes;
}

std::string items_source(const tape::Tape& t, bool aggregate)
{
  if (!items_fit(t)) {
    throw InvalidConfiguration("the decoded variables of this tape's 64-region tile do not fit in LDS");
  }
  std::uint32_t const w = std::max<std::uint32_t>(t.n_groups, 1);
  std::uint32_t const nv = t.n_vars;
  std::ostringstream o;
  o << cmi_interval_ops_source << "\n";
  o << "using namespace cmi_iv;\n\n";
  // dev::Item, field for field; the layout is checked against the host's.
  o << "struct CmiGroups\n{\n  unsigned n_groups, objective_group, begin[" << dev::kMaxGroups + 1 << "];\n};\n"
       "struct CmiItem\n{\n"
       "  const void* code;\n  CmiGroups groups;\n  unsigned n_instr, n_vars;\n  const double* parent;\n"
       "  const unsigned char* var_kinds;\n  const unsigned* slot_var;\n  const unsigned* slot_width;\n"
       "  const double* slot_step;\n  const unsigned char* slot_kind;\n  unsigned slot_count, samples;\n"
       "  unsigned long n_elems, salt;\n  unsigned tile_begin, n_tiles;\n  unsigned char* feasible;\n"
       "  double* obj_lb;\n  double* tile_min;\n  long* tile_arg;\n  unsigned* done;\n  double* cand_value;\n"
       "  long* cand_index;\n  double* witness;\n  double* child_lb;\n  double* child_ub;\n"
       "  unsigned long refine;\n  unsigned n_slices, slice_groups;\n  const unsigned* slice_begin;\n"
       "  unsigned slice_objective, slice_pad;\n  unsigned long* tile_infeasible;\n"
       "  unsigned* tile_slices_done;\n  double* tile_obj;\n};\n";
  o << "static_assert(sizeof(CmiItem) == " << sizeof(dev::Item) << ", \"dev::Item layout\");\n";
  auto field = [&](const char* name, std::size_t off) {
    o << "static_assert(__builtin_offsetof(CmiItem, " << name << ") == " << off << ", \"dev::Item::" << name
      << "\");\n";
  };
  field("parent", offsetof(dev::Item, parent));
  field("slot_var", offsetof(dev::Item, slot_var));
  field("slot_width", offsetof(dev::Item, slot_width));
  field("slot_step", offsetof(dev::Item, slot_step));
  field("slot_kind", offsetof(dev::Item, slot_kind));
  field("slot_count", offsetof(dev::Item, slot_count));
  field("n_elems", offsetof(dev::Item, n_elems));
  field("tile_begin", offsetof(dev::Item, tile_begin));
  field("feasible", offsetof(dev::Item, feasible));
  field("obj_lb", offsetof(dev::Item, obj_lb));
  field("child_lb", offsetof(dev::Item, child_lb));
  field("child_ub", offsetof(dev::Item, child_ub));
  field("refine", offsetof(dev::Item, refine));
  // The interpreter's decode of one slot (kernels.hip slot_bounds), round to
  // nearest: an enumerated slot is the point ceil(lb) + part clamped to ub; a
  // split slot is part `part` of `fan_out` equal parts, the last one ending
  // at the parent's ub. The part width (ub - lb) / fan_out is the item's
  // slot_step, the same correctly rounded quotient formed once on the host.
  o << "__device__ __forceinline__ double lane_value(double x, unsigned j)\n{\n"
       "  const unsigned long u = __builtin_bit_cast(unsigned long, x);\n"
       "  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, j);\n"
       "  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), j);\n"
       "  return __builtin_bit_cast(double, ((unsigned long)hi << 32) | lo);\n}\n\n";
  o << "__device__ __forceinline__ void slot_bounds(unsigned char kind, double plo, double phi, double step,\n"
       "                                            unsigned part, unsigned fan_out, double& lo, double& hi)\n{\n"
       "  if (kind >= 2) {\n    double v = __builtin_ceil(plo) + (double)part;\n    if (v > phi) {\n"
       "      v = phi;\n    }\n    lo = v;\n    hi = v;\n    return;\n  }\n"
       "  lo = plo + step * (double)part;\n"
       "  hi = part + 1u == fan_out ? phi : plo + step * (double)(part + 1u);\n}\n\n";
  // Persistent: workgroup g takes the contiguous tiles [g * chunk, (g + 1) *
  // chunk) one after another. A child of M >= 64 subregions (M a multiple of
  // 64) covers whole tiles, so wave 0 folds each tile into a running bound of
  // its child held in registers and issues the atomics only when the child
  // (or the item) changes; smaller children take the per-tile segmented
  // reduction. Min and max are exact, so neither the fold order nor the
  // atomics' order changes a bit.
  if (aggregate) {
    // one child's running bound, reduced over the wave and folded into the
    // child with one atomic min and max
    o << "__device__ __forceinline__ void flush(const CmiItem* items, unsigned item, unsigned long child, double lo,\n"
         "                                      double hi, unsigned lane)\n{\n"
         "#pragma unroll\n  for (int off = 32; off > 0; off >>= 1) {\n"
         "    lo = fmin(lo, __shfl_xor(lo, off, 64));\n    hi = fmax(hi, __shfl_xor(hi, off, 64));\n  }\n"
         "  if (lane == 0u) {\n    atomicMin(items[item].child_lb + child, lo);\n"
         "    atomicMax(items[item].child_ub + child, hi);\n  }\n}\n\n";
  }
  o << "extern \"C\" __global__ void __launch_bounds__(" << 64 * w << ") " << kItemsKernel
    << "(const CmiItem* __restrict__ items, const unsigned* __restrict__ tile_item, unsigned n_tiles)\n{\n";
  o << "  __shared__ double2 X[" << nv * 64 << "];\n";
  o << "  __shared__ unsigned feas_s[" << 64 * w << "];\n";
  o << "  __shared__ double obj_s[128];\n";
  o << "  const unsigned lane = threadIdx.x & 63u;\n";
  o << "  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n";
  o << "  const unsigned chunk = (n_tiles + gridDim.x - 1u) / gridDim.x;\n";
  o << "  const unsigned first = blockIdx.x * chunk;\n";
  o << "  const unsigned last = first + chunk < n_tiles ? first + chunk : n_tiles;\n";
  o << "  const unsigned cur = 0u;\n";
  o << "  unsigned meta_item = 0xffffffffu, m_count = 0u, m_var = 0u, m_width = 1u, m_kind = 0u;\n"
       "  double m_step = 0.0, m_plo = 0.0, m_phi = 0.0;\n";
  if (aggregate) {
    o << "  unsigned run_item = 0xffffffffu;\n  unsigned long run_child = 0ul;\n"
         "  double run_lo = __builtin_huge_val(), run_hi = -__builtin_huge_val();\n";
  }
  o << "  for (unsigned tile = first; tile < last; ++tile) {\n";
  o << "    const unsigned item = tile_item[tile];\n";
  o << "    const CmiItem& it = items[item];\n";
  o << "    const unsigned long e = (unsigned long)(tile - it.tile_begin) * 64ul + lane;\n";
  o << "    const bool active = e < it.n_elems;\n";
  o << "    const unsigned region = (unsigned)(active ? e : it.n_elems - 1ul);\n";
  // Wave 0 decodes. When the item changes it writes every variable's parent
  // bounds and loads the item's slots, lane j holding slot j; each tile
  // then rewrites only the slot variables, reading the slot data across
  // lanes rather than from memory.
  o << "    if (wave == 0u) {\n";
  o << "      if (item != meta_item) {\n        meta_item = item;\n        m_count = it.slot_count;\n";
  o << "        for (unsigned v = 0; v < " << nv << "u; ++v) {\n"
       "          X[v * 64u + lane] = make_double2(-it.parent[2u * v], it.parent[2u * v + 1u]);\n        }\n";
  o << "        if (lane < m_count) {\n          m_var = it.slot_var[lane];\n          m_width = it.slot_width[lane];\n"
       "          m_kind = it.slot_kind[lane];\n          m_step = it.slot_step[lane];\n"
       "          m_plo = it.parent[2u * m_var];\n          m_phi = it.parent[2u * m_var + 1u];\n        }\n      }\n";
  o << "      unsigned idx = region;\n";
  o << "      for (unsigned j = 0; j < m_count; ++j) {\n"
       "        const unsigned fw = __builtin_amdgcn_readlane(m_width, j);\n        unsigned part;\n"
       "        if ((fw & (fw - 1u)) == 0u) {\n          part = idx & (fw - 1u);\n          idx >>= __builtin_ctz(fw);\n"
       "        } else {\n          part = idx % fw;\n          idx /= fw;\n        }\n"
       "        const unsigned vid = __builtin_amdgcn_readlane(m_var, j);\n        double lo, hi;\n"
       "        slot_bounds((unsigned char)__builtin_amdgcn_readlane(m_kind, j), lane_value(m_plo, j),\n"
       "                    lane_value(m_phi, j), lane_value(m_step, j), part, fw, lo, hi);\n"
       "        X[vid * 64u + lane] = make_double2(-lo, hi);\n      }\n    }\n";
  if (w > 1) {
    o << "    __syncthreads();\n";
  }
  o << "    bool f = true;\n";
  o << "    const int t = enter_round_up();\n";
  emit_waves(o, t, w, "    ", false, 0, [](std::uint32_t x, std::uint32_t) -> std::string {
    return "{X[" + std::to_string(x * 64) + "u + lane].x, X[" + std::to_string(x * 64) + "u + lane].y}";
  });
  o << "    leave_round_up(t, f);\n";
  o << "    feas_s[wave * 64u + lane] = f ? 1u : 0u;\n";
  if (w > 1) {
    o << "    __syncthreads();\n";
  }
  o << "    if (wave == 0u) {\n";
  o << "      unsigned all = 1u;\n";
  for (std::uint32_t v = 0; v < w; ++v) {
    o << "      all &= feas_s[" << v * 64 << "u + lane];\n";
  }
  if (!aggregate) {
    o << "      if (active) {\n        it.feasible[region] = all ? 1 : 0;\n        it.obj_lb[region] = obj_s[lane];\n"
         "      }\n";
  } else {
    o << "      const bool live = active && all != 0u;\n";
    o << "      double lo = live ? obj_s[lane] : __builtin_huge_val();\n";
    o << "      double hi = live ? obj_s[64u + lane] : -__builtin_huge_val();\n";
    // Each lane keeps its own running bound of the current child; the lanes
    // meet only when the child changes.
    o << "      if ((it.refine & 63ul) == 0ul) {\n";
    o << "        const unsigned long child = (unsigned long)(tile - it.tile_begin) * 64ul / it.refine;\n";
    o << "        if (item != run_item || child != run_child) {\n"
         "          if (run_item != 0xffffffffu) {\n            flush(items, run_item, run_child, run_lo, run_hi, lane);\n"
         "          }\n"
         "          run_item = item;\n          run_child = child;\n          run_lo = __builtin_huge_val();\n"
         "          run_hi = -__builtin_huge_val();\n        }\n"
         "        run_lo = fmin(run_lo, lo);\n        run_hi = fmax(run_hi, hi);\n";
    o << "      } else {\n";
    o << "        const unsigned long child = (unsigned long)region / it.refine;\n";
    o << "#pragma unroll\n        for (int off = 1; off < 64; off <<= 1) {\n"
         "          const double olo = __shfl_down(lo, off, 64);\n          const double ohi = __shfl_down(hi, off, 64);\n"
         "          const unsigned long oc = __shfl_down(child, off, 64);\n"
         "          if (lane + off < 64u && oc == child) {\n            lo = fmin(lo, olo);\n"
         "            hi = fmax(hi, ohi);\n          }\n        }\n";
    o << "        const unsigned long prev = __shfl_up(child, 1, 64);\n";
    o << "        const bool head = lane == 0u || prev != child;\n";
    o << "        if (active && head) {\n          atomicMin(it.child_lb + child, lo);\n"
         "          atomicMax(it.child_ub + child, hi);\n        }\n";
    o << "      }\n";
  }
  o << "    }\n";
  o << "  }\n";
  if (aggregate) {
    o << "  if (wave == 0u && run_item != 0xffffffffu) {\n"
         "    flush(items, run_item, run_child, run_lo, run_hi, lane);\n  }\n";
  }
  o << "}\n";
  return o.str();
}

struct CompiledItems::Module
{
  int device = 0;
  std::vector<Instr> code;
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;

  ~Module()
  {
    if (module != nullptr) {
      int prev = 0;
      (void)hipGetDevice(&prev);
      (void)hipSetDevice(device);
      (void)hipModuleUnload(module);
      (void)hipSetDevice(prev);
    }
  }
};

void CompiledItems::launch(const void* items, const std::uint32_t* tile_item, std::uint32_t n_tiles,
                           hipStream_t stream) const
{
  if (n_tiles == 0) {
    return;
  }
  void* args[] = {&items, &tile_item, &n_tiles};
  std::uint32_t const grid = std::min<std::uint32_t>(n_tiles, n_cus_ * wgs_per_cu_);
  hip_check(hipModuleLaunchKernel(mod_->fn, grid, 1, 1, 64 * waves_, 1, 1, 0, stream, args, nullptr),
            "hipModuleLaunchKernel(compiled items)");
}

std::shared_ptr<CompiledItems> compile_items(const tape::Tape& t, int device, bool aggregate, bool compile)
{
  static std::mutex mu;
  static std::map<std::tuple<int, bool, std::uint64_t, std::size_t>, std::weak_ptr<const CompiledItems::Module>> cache;
  std::uint64_t h = fnv1a(t.code.data(), t.code.size() * sizeof(Instr));
  h = fnv1a(&t.n_vars, sizeof t.n_vars, h);
  h = fnv1a(t.group_begin.data(), t.group_begin.size() * sizeof(std::uint32_t), h);
  h = fnv1a(&t.objective_group, sizeof t.objective_group, h);
  auto const key = std::make_tuple(device, aggregate, h, t.code.size());
  std::shared_ptr<CompiledItems> ci(new CompiledItems());
  ci->aggregate_ = aggregate;
  ci->waves_ = std::max<std::uint32_t>(t.n_groups, 1);
  int cus = 0;
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute(CUs)");
  ci->n_cus_ = static_cast<std::uint32_t>(std::max(cus, 1));
  // resident workgroups per CU: as many as LDS and 32 waves per CU allow, so
  // that each SIMD has several waves to switch between while one waits on
  // its FP64 and LDS latencies
  std::uint64_t const lds = std::uint64_t {t.n_vars} * 64 * 16 + std::uint64_t {ci->waves_} * 64 * 4 + 128 * 8;
  ci->wgs_per_cu_ = static_cast<std::uint32_t>(
      std::clamp<std::uint64_t>(std::min<std::uint64_t>(kLdsBytes / lds, 32 / ci->waves_), 1, 32));
  std::lock_guard<std::mutex> lock(mu);
  if (auto it = cache.find(key); it != cache.end()) {
    if (auto sp = it->second.lock();
        sp && sp->code.size() == t.code.size()
        && std::memcmp(sp->code.data(), t.code.data(), t.code.size() * sizeof(Instr)) == 0) {
      ci->mod_ = sp;
      return ci;
    }
  }
  if (!compile) {
    return nullptr;
  }
  auto const t0 = std::chrono::steady_clock::now();
  std::vector<char> const code = compile_code_object(items_source(t, aggregate), false);
  auto m = std::make_shared<CompiledItems::Module>();
  m->device = device;
  m->code = t.code;
  hip_check(hipModuleLoadData(&m->module, code.data()), "hipModuleLoadData(compiled items)");
  hip_check(hipModuleGetFunction(&m->fn, m->module, kItemsKernel), "hipModuleGetFunction(compiled items)");
  ci->mod_ = m;
  ci->compile_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  cache[key] = m;
  return ci;
}

/// A loaded code object; unloaded when the last handle goes.
struct CompiledBoxes::Module
{
  int device = 0;
  std::vector<Instr> code;
  bool persistent = false;
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;

  ~Module()
  {
    if (module != nullptr) {
      int prev = 0;
      (void)hipGetDevice(&prev);
      (void)hipSetDevice(device);
      (void)hipModuleUnload(module);
      (void)hipSetDevice(prev);
    }
  }
};

void CompiledBoxes::launch(const double* lb, const double* ub, std::uint64_t ld, std::uint64_t n,
                           std::uint8_t* feasible, double* obj_lb, double* obj_ub, double* roots,
                           hipStream_t stream) const
{
  if (n == 0) {
    return;
  }
  if (n > kMaxBoxes) {
    throw ShapeMismatch("compiled box kernels take at most 2^29 boxes per call");
  }
  unsigned long ld_ = ld, n_ = n;
  void* args[] = {&lb, &ub, &ld_, &n_, &feasible, &obj_lb, &obj_ub, &roots};
  std::uint64_t const tiles = (n + 63) / 64;
  std::uint64_t const blocks = persistent_ ? std::min<std::uint64_t>(tiles, std::uint64_t(n_cus_) * wgs_per_cu_) : tiles;
  hip_check(hipModuleLaunchKernel(mod_->fn, static_cast<unsigned>(blocks), 1, 1, 64 * waves_, 1, 1, 0, stream, args,
                                  nullptr),
            "hipModuleLaunchKernel(compiled boxes)");
}

std::shared_ptr<CompiledBoxes> compile_boxes(const tape::Tape& t, int device, const Shape& shape)
{
  static std::mutex mu;
  static std::map<std::tuple<int, bool, bool, bool, std::uint32_t, std::uint32_t, std::uint64_t, std::size_t>,
                  std::weak_ptr<const CompiledBoxes::Module>>
      cache;
  bool const persistent = shape.persistent;
  std::uint32_t const waves = shape_waves(t, shape);
  std::uint64_t h = fnv1a(t.code.data(), t.code.size() * sizeof(Instr));
  h = fnv1a(&t.n_vars, sizeof t.n_vars, h);
  h = fnv1a(t.group_begin.data(), t.group_begin.size() * sizeof(std::uint32_t), h);
  h = fnv1a(&t.objective_group, sizeof t.objective_group, h);
  std::uint32_t const load_batch = persistent ? 0u : shape.load_batch;
  auto const key =
      std::make_tuple(device, persistent, shape.ilp_schedule, shape.roots, waves, load_batch, h, t.code.size());

  std::shared_ptr<CompiledBoxes> cb(new CompiledBoxes());
  cb->groups_ = std::max<std::uint32_t>(t.n_groups, 1);
  cb->waves_ = waves;
  cb->persistent_ = persistent;
  cb->ilp_schedule_ = shape.ilp_schedule;
  cb->roots_ = shape.roots;
  cb->load_batch_ = load_batch;
  int cus = 0;
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute(CUs)");
  cb->n_cus_ = static_cast<std::uint32_t>(std::max(cus, 1));

  std::lock_guard<std::mutex> lock(mu);
  if (auto it = cache.find(key); it != cache.end()) {
    if (auto sp = it->second.lock();
        sp && sp->code.size() == t.code.size()
        && std::memcmp(sp->code.data(), t.code.data(), t.code.size() * sizeof(Instr)) == 0) {
      cb->mod_ = sp;
      return cb;
    }
  }
  auto const t0 = std::chrono::steady_clock::now();
  std::vector<char> const code = compile_code_object(boxes_source(t, shape), shape.ilp_schedule);
  auto m = std::make_shared<CompiledBoxes::Module>();
  m->device = device;
  m->code = t.code;
  m->persistent = persistent;
  hip_check(hipModuleLoadData(&m->module, code.data()), "hipModuleLoadData(compiled boxes)");
  hip_check(hipModuleGetFunction(&m->fn, m->module, kBoxesKernel), "hipModuleGetFunction(compiled boxes)");
  cb->mod_ = m;
  cb->code_bytes_ = code.size();
  cb->compile_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  cache[key] = m;
  return cb;
}

std::vector<char> compile_code_object(const std::string& src, bool ilp_schedule)
{
  std::vector<const char*> opts(std::begin(kOptions), std::end(kOptions));
  if (ilp_schedule) {
    opts.insert(opts.end(), std::begin(kIlpOptions), std::end(kIlpOptions));
  }
  hiprtcProgram prog = nullptr;
  rtc_check(hiprtcCreateProgram(&prog, src.c_str(), "cmi_boxes.hip", 0, nullptr, nullptr), "hiprtcCreateProgram");
  hiprtcResult const rc = hiprtcCompileProgram(prog, static_cast<int>(opts.size()), opts.data());
  if (rc != HIPRTC_SUCCESS) {
    std::size_t n = 0;
    (void)hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n > 0) {
      (void)hiprtcGetProgramLog(prog, log.data());
    }
    (void)hiprtcDestroyProgram(&prog);
    throw DeviceError(static_cast<int>(rc), "compiling the tape kernel failed: " + log);
  }
  std::size_t size = 0;
  rtc_check(hiprtcGetCodeSize(prog, &size), "hiprtcGetCodeSize");
  std::vector<char> code(size);
  rtc_check(hiprtcGetCode(prog, code.data()), "hiprtcGetCode");
  (void)hiprtcDestroyProgram(&prog);
  return code;
}

}  // namespace cmi::jit
