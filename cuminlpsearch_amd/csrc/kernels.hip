// Device evaluation of DAG tapes on gfx950 (MI355X, CDNA4, wave64).
//
// Each lane evaluates the tape for one element (a box, a region, or a sample
// point); a wave covers 64 elements. The tape word stream is uniform across the
// wave; it arrives in 64-word chunks through VGPRs, one chunk ahead, and each
// word is read out of its lane into SGPRs. Per-lane slot values live
// in LDS, slot-major ([slot][lane]), so every operand read or result write is
// one conflict-free ds_read_b128/ds_write_b128 (16 B per lane, 1 KiB per wave);
// the accumulator that carries sums lives in VGPRs. For explicit boxes the
// functions of one problem are split into groups, one wave per group, so the
// waves of a workgroup evaluate the same 64 boxes together and share the
// variable slots.
//
// Outward rounding: the interval path runs inside the round-toward-+infinity
// region of interval_ops.hpp, on intervals held as (-lb, ub), one asm
// instruction per bound. The point (sampler/enumerator) interpreter never
// enters the region and is ordinary round-to-nearest code.
//
// The op semantics follow the reference's interval and point functors
// (backend/graph/ops.cuh:24-355, kernels.cuh:177-340) for the QCQP op set.
// Multiplication takes the min/max of the four directed endpoint products
// (equal in value to the IEEE-1788 case table for finite operands). A zero
// bound in any output is reported as +0.

#include <hip/hip_runtime.h>

#include <algorithm>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "cmi/errors.hpp"
#include "cmi/tape.hpp"
#include "device.hpp"
#include "interval_ops.hpp"

namespace cmi::dev
{

using tape::Instr;
using tape::TOp;

namespace
{

using namespace cmi_iv;

__device__ __forceinline__ double2 to_d2(Bnd b) { return make_double2(b.lb, b.ub); }

// ---------------------------------------------------------------------------
// Region decode and sampling (same arithmetic as region::decode::slot_bounds
// and the reference's subregion_sampler.cuh:53-120), round to nearest
// ---------------------------------------------------------------------------

struct Bounds
{
  double lo, hi;
};

__device__ __forceinline__ Bounds slot_bounds(std::uint8_t kind, Bounds parent, std::uint32_t part,
                                              std::uint32_t fan_out)
{
  // SlotKind: 0 Continuous, 1 IntegerPartition, 2 IntegerEnumerate, 3 BinaryEnumerate
  if (kind >= 2) {
    double const lo = ceil(parent.lo);
    double v = lo + static_cast<double>(part);
    if (v > parent.hi) {
      v = parent.hi;
    }
    return {v, v};
  }
  double const width = (parent.hi - parent.lo) / static_cast<double>(fan_out);
  double const step_lo = width * static_cast<double>(part);
  Bounds out;
  out.lo = parent.lo + step_lo;
  if (part + 1 == fan_out) {
    out.hi = parent.hi;
  } else {
    double const step_hi = width * static_cast<double>(part + 1);
    out.hi = parent.lo + step_hi;
  }
  return out;
}

__device__ __forceinline__ std::uint64_t sample_hash(std::uint64_t vid, std::uint64_t i, std::uint64_t region,
                                                     std::uint64_t salt)
{
  std::uint64_t x = vid * 0x9e3779b97f4a7c15ULL;
  x ^= i * 0xbf58476d1ce4e5b9ULL;
  x ^= region * 0x94d049bb133111ebULL;
  x ^= salt * 0xd6e8feb86659fd93ULL;
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

__device__ __forceinline__ double draw(bool discrete, double lb, double ub, std::uint64_t vid, std::uint64_t i,
                                       std::uint64_t region, std::uint64_t salt)
{
  std::uint64_t const x = sample_hash(vid, i, region, salt);
  if (!discrete) {
    double const u = static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0);
    return __builtin_fma(u, ub - lb, lb);
  }
  double const lo = ceil(lb);
  double const hi = floor(ub);
  if (hi < lo) {
    return lo;
  }
  std::uint64_t const count = static_cast<std::uint64_t>(hi - lo) + 1;
  return lo + static_cast<double>(x % count);
}

// ---------------------------------------------------------------------------
// Tape interpreters
// ---------------------------------------------------------------------------

// Tape words are streamed through VGPRs in chunks of 64: lane j of a wave
// loads word chunk*64 + j with one coalesced 16-byte load, the next chunk is
// in flight while the current one runs, and each word is taken out of its lane
// with v_readlane into SGPRs (the pc is wave-uniform, so the interpreter's
// dispatch stays scalar). A tape larger than the scalar data cache would
// otherwise wait out an L2 round trip on every word. Every device tape
// carries one trailing Nop word, so reading one word past the last
// instruction stays in bounds; lanes past that word load nothing.
typedef std::uint32_t u32x4 __attribute__((ext_vector_type(4)));

/// A load through the global address space: the compiler then issues a
/// global (not flat) load, which only the vector-memory counter tracks, so LDS
/// traffic does not have to wait for it.
template<typename T>
__device__ __forceinline__ T gload(const T* p)
{
  using GlobalPtr = const __attribute__((address_space(1))) T*;
  return *(GlobalPtr)(p);
}

/// Lane j's part of the 64-word tape chunk starting at word c0 = i - j;
/// lanes past the last readable word repeat it.
__device__ __forceinline__ u32x4 tape_chunk(const u32x4* words, std::uint32_t i, std::uint32_t limit)
{
  return gload(words + (i < limit ? i : limit));
}

/// Word `slot` of a chunk held one word per lane, as wave-uniform SGPRs.
__device__ __forceinline__ u32x4 lane_word(u32x4 c, std::uint32_t slot)
{
  u32x4 w;
  w.x = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c.x), static_cast<int>(slot)));
  w.y = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c.y), static_cast<int>(slot)));
  w.z = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c.z), static_cast<int>(slot)));
  w.w = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c.w), static_cast<int>(slot)));
  return w;
}

struct Word
{
  std::uint32_t op, dst, a, b;
  double imm;
};

__device__ __forceinline__ Word decode_word(u32x4 w)
{
  Word out;
  out.op = w.x & 0xffu;
  out.dst = w.x >> 16;
  out.a = w.y & 0xffffu;
  out.b = w.y >> 16;
  out.imm = __builtin_bit_cast(double, (static_cast<std::uint64_t>(w.w) << 32) | w.z);
  return out;
}

struct IntervalResult
{
  Iv obj;
  bool feasible;
};

__device__ __forceinline__ Iv load_iv(const double2* S, std::uint32_t slot, unsigned lane)
{
  double2 const v = S[slot * kTile + lane];
  return {v.x, v.y};
}

__device__ __forceinline__ void store_iv(double2* S, std::uint32_t slot, unsigned lane, Iv v)
{
  S[slot * kTile + lane] = make_double2(v.nl, v.ub);
}

// One word of the extended op set (interval): the result, and `ok` cleared
// when it is empty.
__device__ __forceinline__ Iv ext_interval(std::uint32_t op, Iv a, const double2* S, const Word& ins, unsigned lane,
                                           int t, bool& ok)
{
  switch (static_cast<TOp>(op)) {
    case TOp::Div: return iv_div(a, load_iv(S, ins.b, lane), t, ok);
    case TOp::DivImm: return iv_div_imm(a, ins.imm, t, ok);
    case TOp::ImmDiv: return iv_imm_div(ins.imm, a, t, ok);
    case TOp::PowN: return iv_pown(a, static_cast<int>(ins.imm), t, ok);
    case TOp::Sqrt: return iv_sqrt(a, t, ok);
    case TOp::Exp: return iv_exp(a, t);
    case TOp::Log: return iv_log(a, t, ok);
    case TOp::Sin: return iv_sincos(a, 0, t);
    case TOp::Cos: return iv_sincos(a, 1, t);
    case TOp::Tanh: return iv_tanh(a, t);
    case TOp::Abs: return iv_abs(a);
    case TOp::Min: return iv_min(a, load_iv(S, ins.b, lane));
    case TOp::MinImm: return iv_min(a, iv_const(ins.imm, t));
    case TOp::Max: return iv_max(a, load_iv(S, ins.b, lane));
    case TOp::MaxImm: return iv_max(a, iv_const(ins.imm, t));
    default: return a;
  }
}

// Interval tape code[0, n_instr) over LDS slots S[slot * kTile + lane] holding
// (-lb, ub), and the accumulator register. Must run inside the round-up region
// (token t). Output instructions store [lb, ub] into roots[b*stride + e] when
// roots is non-null.
// With Ext the extended op set is interpreted too.
template<bool Ext>
__device__ __forceinline__ IntervalResult run_interval(const Instr* code, std::uint32_t n_instr, double2* S,
                                                       unsigned lane, int t, double2* roots,
                                                       std::uint64_t roots_stride, std::uint64_t e)
{
  IntervalResult res {{0.0, 0.0}, true};
  Iv acc {0.0, 0.0};
  // the tape arrives one 64-word chunk at a time, the next chunk in flight
  // while this one runs
  const u32x4* const words = reinterpret_cast<const u32x4*>(code);
  u32x4 cur = tape_chunk(words, lane, n_instr);
  for (std::uint32_t c0 = 0; c0 < n_instr; c0 += kTile) {
    u32x4 const nxt = tape_chunk(words, c0 + kTile + lane, n_instr);
    std::uint32_t const end = std::min<std::uint32_t>(n_instr - c0, kTile);
    for (std::uint32_t slot = 0; slot < end; ++slot) {
      Word const ins = decode_word(lane_word(cur, slot));
      std::uint32_t const op = ins.op;
      if (op >= tape::kTermBase) {
        // acc <- acc (set|+|-) term(a, b, imm)
        std::uint32_t const term = op & 7u;
        std::uint32_t const mode = (op - tape::kTermBase) >> 3;
        Iv const a = load_iv(S, ins.a, lane);
        Iv x;
        if (term == static_cast<std::uint32_t>(tape::Term::SSC)) {
          x = iv_mul_imm(iv_mul(a, load_iv(S, ins.b, lane), t), ins.imm, t);
        } else if (term == static_cast<std::uint32_t>(tape::Term::QC)) {
          x = iv_mul_imm(iv_sqr(a, t), ins.imm, t);
        } else if (term == static_cast<std::uint32_t>(tape::Term::SC)) {
          x = iv_mul_imm(a, ins.imm, t);
        } else if (term == static_cast<std::uint32_t>(tape::Term::SS)) {
          x = iv_mul(a, load_iv(S, ins.b, lane), t);
        } else if (term == static_cast<std::uint32_t>(tape::Term::Q)) {
          x = iv_sqr(a, t);
        } else {
          x = a;
        }
        if (mode == static_cast<std::uint32_t>(tape::AccMode::Add)) {
          acc = iv_add(acc, x, t);
        } else if (mode == static_cast<std::uint32_t>(tape::AccMode::Set)) {
          acc = x;
        } else {
          acc = iv_sub(acc, x, t);
        }
        continue;
      }
      if (op >= static_cast<std::uint32_t>(TOp::AccStore) && op < tape::kExtBase) {
        switch (static_cast<TOp>(op)) {
          case TOp::AccStore: store_iv(S, ins.dst, lane, acc); break;
          case TOp::AccAddImm: acc = iv_add_imm(acc, ins.imm, t); break;
          case TOp::AccSubImm: acc = iv_sub_imm(acc, ins.imm, t); break;
          case TOp::AccImmSub: acc = iv_imm_sub(ins.imm, acc, t); break;
          case TOp::AccScale: acc = iv_mul_imm(acc, ins.imm, t); break;
          case TOp::AccNeg: acc = iv_neg(acc); break;
          case TOp::AccCheckLE: res.feasible = res.feasible && check_le(acc, ins.imm); break;
          case TOp::AccCheckEQ: res.feasible = res.feasible && check_eq(acc, ins.imm); break;
          case TOp::AccObjective: res.obj = acc; break;
          case TOp::AccOutput:
            if (roots != nullptr) {
              roots[static_cast<std::uint64_t>(ins.b) * roots_stride + e] = to_d2(to_output(acc, t));
            }
            break;
          default: break;
        }
        continue;
      }
      Iv const a = load_iv(S, ins.a, lane);
      if constexpr (Ext) {
        if (op >= tape::kExtBase && op < tape::kExtEnd) {
          bool ok = true;
          Iv const r = ext_interval(op, a, S, ins, lane, t, ok);
          res.feasible = res.feasible && ok;
          store_iv(S, ins.dst, lane, r);
          continue;
        }
      }
      Iv r;
      switch (static_cast<TOp>(op)) {
        case TOp::LoadConst: r = iv_const(ins.imm, t); break;
        case TOp::Add: r = iv_add(a, load_iv(S, ins.b, lane), t); break;
        case TOp::AddImm: r = iv_add_imm(a, ins.imm, t); break;
        case TOp::Sub: r = iv_sub(a, load_iv(S, ins.b, lane), t); break;
        case TOp::SubImm: r = iv_sub_imm(a, ins.imm, t); break;
        case TOp::ImmSub: r = iv_imm_sub(ins.imm, a, t); break;
        case TOp::Mul: r = iv_mul(a, load_iv(S, ins.b, lane), t); break;
        case TOp::MulImm: r = iv_mul_imm(a, ins.imm, t); break;
        case TOp::Sqr: r = iv_sqr(a, t); break;
        case TOp::Neg: r = iv_neg(a); break;
        case TOp::CheckLE:
          res.feasible = res.feasible && check_le(a, ins.imm);
          continue;
        case TOp::CheckEQ:
          res.feasible = res.feasible && check_eq(a, ins.imm);
          continue;
        case TOp::Objective:
          res.obj = a;
          continue;
        case TOp::Output:
          if (roots != nullptr) {
            roots[static_cast<std::uint64_t>(ins.b) * roots_stride + e] = to_d2(to_output(a, t));
          }
          continue;
        default:
          continue;
      }
      store_iv(S, ins.dst, lane, r);
    }
    cur = nxt;
  }
  return res;
}

__device__ __forceinline__ bool almost_equal(double a, double b)
{
  double const diff = fabs(a - b);
  if (diff <= 1e-8) {
    return true;
  }
  return diff <= 1e-6 * fmax(fabs(a), fabs(b));
}

struct PointResult
{
  double obj;
  bool feasible;
};

// One word of the extended op set (round to nearest).
__device__ __forceinline__ double ext_point(std::uint32_t op, double a, const double* P, const Word& ins,
                                            unsigned lane, std::uint32_t stride = kTile)
{
  switch (static_cast<TOp>(op)) {
    case TOp::Div: return a / P[ins.b * stride + lane];
    case TOp::DivImm: return a / ins.imm;
    case TOp::ImmDiv: return ins.imm / a;
    case TOp::PowN: return pt_pown(a, static_cast<int>(ins.imm));
    case TOp::Sqrt: return __builtin_sqrt(a);
    case TOp::Exp: return pt_exp(a);
    case TOp::Log: return pt_log(a);
    case TOp::Sin: return pt_sincos(a, 0);
    case TOp::Cos: return pt_sincos(a, 1);
    case TOp::Tanh: return pt_tanh(a);
    case TOp::Abs: return __builtin_fabs(a);
    case TOp::Min: return __builtin_fmin(a, P[ins.b * stride + lane]);
    case TOp::MinImm: return __builtin_fmin(a, ins.imm);
    case TOp::Max: return __builtin_fmax(a, P[ins.b * stride + lane]);
    case TOp::MaxImm: return __builtin_fmax(a, ins.imm);
    default: return a;
  }
}

// Round-to-nearest point tape over LDS slots P[slot * kTile + lane] and the
// accumulator register.
template<bool Ext>
__device__ __forceinline__ PointResult run_point(const Instr* code, std::uint32_t n_instr, double* P, unsigned lane,
                                                 double* roots, std::uint64_t roots_stride, std::uint64_t e)
{
  PointResult res {0.0, true};
  double acc = 0.0;
  // the tape arrives one 64-word chunk at a time, the next chunk in flight
  // while this one runs
  const u32x4* const words = reinterpret_cast<const u32x4*>(code);
  u32x4 cur = tape_chunk(words, lane, n_instr);
  for (std::uint32_t c0 = 0; c0 < n_instr; c0 += kTile) {
    u32x4 const nxt = tape_chunk(words, c0 + kTile + lane, n_instr);
    std::uint32_t const end = std::min<std::uint32_t>(n_instr - c0, kTile);
    for (std::uint32_t slot = 0; slot < end; ++slot) {
      Word const ins = decode_word(lane_word(cur, slot));
      std::uint32_t const op = ins.op;
      if (op >= tape::kTermBase) {
        std::uint32_t const term = op & 7u;
        std::uint32_t const mode = (op - tape::kTermBase) >> 3;
        double const a = P[ins.a * kTile + lane];
        double x;
        if (term == static_cast<std::uint32_t>(tape::Term::SSC)) {
          x = (a * P[ins.b * kTile + lane]) * ins.imm;
        } else if (term == static_cast<std::uint32_t>(tape::Term::QC)) {
          x = (a * a) * ins.imm;
        } else if (term == static_cast<std::uint32_t>(tape::Term::SC)) {
          x = a * ins.imm;
        } else if (term == static_cast<std::uint32_t>(tape::Term::SS)) {
          x = a * P[ins.b * kTile + lane];
        } else if (term == static_cast<std::uint32_t>(tape::Term::Q)) {
          x = a * a;
        } else {
          x = a;
        }
        if (mode == static_cast<std::uint32_t>(tape::AccMode::Add)) {
          acc = acc + x;
        } else if (mode == static_cast<std::uint32_t>(tape::AccMode::Set)) {
          acc = x;
        } else {
          acc = acc - x;
        }
        continue;
      }
      if (op >= static_cast<std::uint32_t>(TOp::AccStore) && op < tape::kExtBase) {
        switch (static_cast<TOp>(op)) {
          case TOp::AccStore: P[ins.dst * kTile + lane] = acc; break;
          case TOp::AccAddImm: acc = acc + ins.imm; break;
          case TOp::AccSubImm: acc = acc - ins.imm; break;
          case TOp::AccImmSub: acc = ins.imm - acc; break;
          case TOp::AccScale: acc = acc * ins.imm; break;
          case TOp::AccNeg: acc = -acc; break;
          case TOp::AccCheckLE: res.feasible = res.feasible && (acc <= ins.imm); break;
          case TOp::AccCheckEQ: res.feasible = res.feasible && almost_equal(acc, ins.imm); break;
          case TOp::AccObjective: res.obj = acc; break;
          case TOp::AccOutput:
            if (roots != nullptr) {
              roots[static_cast<std::uint64_t>(ins.b) * roots_stride + e] = acc;
            }
            break;
          default: break;
        }
        continue;
      }
      double const a = P[ins.a * kTile + lane];
      if constexpr (Ext) {
        if (op >= tape::kExtBase && op < tape::kExtEnd) {
          P[ins.dst * kTile + lane] = ext_point(op, a, P, ins, lane);
          continue;
        }
      }
      double r;
      switch (static_cast<TOp>(op)) {
        case TOp::LoadConst: r = ins.imm; break;
        case TOp::Add: r = a + P[ins.b * kTile + lane]; break;
        case TOp::AddImm: r = a + ins.imm; break;
        case TOp::Sub: r = a - P[ins.b * kTile + lane]; break;
        case TOp::SubImm: r = a - ins.imm; break;
        case TOp::ImmSub: r = ins.imm - a; break;
        case TOp::Mul: r = a * P[ins.b * kTile + lane]; break;
        case TOp::MulImm: r = a * ins.imm; break;
        case TOp::Sqr: r = a * a; break;
        case TOp::Neg: r = -a; break;
        case TOp::CheckLE:
          res.feasible = res.feasible && (a <= ins.imm);
          continue;
        case TOp::CheckEQ:
          res.feasible = res.feasible && almost_equal(a, ins.imm);
          continue;
        case TOp::Objective:
          res.obj = a;
          continue;
        case TOp::Output:
          if (roots != nullptr) {
            roots[static_cast<std::uint64_t>(ins.b) * roots_stride + e] = a;
          }
          continue;
        default:
          continue;
      }
      P[ins.dst * kTile + lane] = r;
    }
    cur = nxt;
  }
  return res;
}

// Wave64 (value, index) min with the lowest index winning ties.
__device__ __forceinline__ void wave_argmin(double& v, std::int64_t& i)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double const ov = __shfl_xor(v, off, 64);
    std::int64_t const oi = __shfl_xor(i, off, 64);
    if (ov < v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Explicit boxes, one wave per function group: workgroup = n_groups waves over
// the same 64 boxes. Variables are loaded once per workgroup into the shared
// slots; each wave runs its group's code; wave 0 combines the groups'
// feasibility and writes the outputs.
template<bool Ext>
__global__ __launch_bounds__(kTile * kMaxGroups) void k_boxes(const Instr* __restrict__ code_g, GroupTable groups,
                                                               std::uint32_t n_vars, std::uint32_t n_slots,
                                                               const double* __restrict__ lb,
                                                               const double* __restrict__ ub, std::uint64_t ld,
                                                               std::uint64_t n_boxes,
                                                               std::uint8_t* __restrict__ feasible,
                                                               double* __restrict__ obj_lb,
                                                               double* __restrict__ obj_ub, double2* __restrict__ roots)
{
  extern __shared__ double2 S[];
  unsigned const lane = threadIdx.x % kTile;
  // the wave index is uniform across the wave; say so, so the tape pointer,
  // the word fetches and the dispatch stay scalar
  unsigned const wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);
  unsigned const n_groups = groups.n_groups;
  std::uint64_t const e = static_cast<std::uint64_t>(blockIdx.x) * kTile + lane;
  bool const active = e < n_boxes;
  std::uint64_t const ee = active ? e : n_boxes - 1;
  for (std::uint32_t v = wave; v < n_vars; v += n_groups) {
    // each box is read once per launch: stream it past the caches
    S[v * kTile + lane] = make_double2(-__builtin_nontemporal_load(lb + v * ld + ee),
                                       __builtin_nontemporal_load(ub + v * ld + ee));
  }
  __syncthreads();
  std::uint32_t const begin = groups.begin[wave];
  int const t = enter_round_up();
  IntervalResult const res = run_interval<Ext>(code_g + begin, groups.begin[wave + 1] - begin, S, lane,
                                               t, active ? roots : nullptr, n_boxes, e);
  Bnd const obj = to_output(res.obj, t);
  leave_round_up(t, obj.lb, obj.ub, res.feasible);
  // per-group results after the slots: feasibility [n_groups][64], objective [64]
  auto* feas_g = reinterpret_cast<std::uint32_t*>(S + n_slots * kTile);
  auto* obj_s = reinterpret_cast<double2*>(feas_g + kMaxGroups * kTile);
  feas_g[wave * kTile + lane] = res.feasible ? 1u : 0u;
  if (wave == groups.objective_group) {
    obj_s[lane] = to_d2(obj);
  }
  __syncthreads();
  if (wave == 0 && active) {
    bool f = true;
    for (unsigned g = 0; g < n_groups; ++g) {
      f = f && feas_g[g * kTile + lane] != 0u;
    }
    double2 const o = obj_s[lane];
    feasible[e] = f ? 1 : 0;
    obj_lb[e] = o.x;
    if (obj_ub != nullptr) {
      obj_ub[e] = o.y;
    }
  }
}

template<bool Ext>
__global__ __launch_bounds__(kTile) void k_points(const Instr* __restrict__ code_g, std::uint32_t n_instr,
                                                   std::uint32_t n_vars, const double* __restrict__ x,
                                                   std::uint64_t ld, std::uint64_t n_points,
                                                   std::uint8_t* __restrict__ feasible, double* __restrict__ obj,
                                                   double* __restrict__ roots)
{
  extern __shared__ double P[];
  unsigned const lane = threadIdx.x;
  std::uint64_t const e = static_cast<std::uint64_t>(blockIdx.x) * kTile + lane;
  bool const active = e < n_points;
  std::uint64_t const ee = active ? e : n_points - 1;
  for (std::uint32_t v = 0; v < n_vars; ++v) {
    P[v * kTile + lane] = x[v * ld + ee];
  }
  PointResult const res =
      run_point<Ext>(code_g, n_instr, P, lane, active ? roots : nullptr, n_points, e);
  if (active) {
    feasible[e] = res.feasible ? 1 : 0;
    obj[e] = res.obj;
  }
}


// The round-to-nearest tape over N points per lane: point j of lane l is in
// LDS at P[(slot * N + j) * kTile + l]. The word stream and its dispatch are
// shared by the N points, so a sampler tile of N x 64 points costs one
// dispatch per word instead of N; each point's arithmetic is that of
// run_point.
template<bool Ext, int N>
struct PointsResult
{
  double obj[N];
  bool feasible[N];
};

template<bool Ext, int N>
__device__ __forceinline__ PointsResult<Ext, N> run_points(const Instr* code, std::uint32_t n_instr, double* P,
                                                           unsigned lane)
{
  PointsResult<Ext, N> res;
  double acc[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    res.obj[j] = 0.0;
    res.feasible[j] = true;
    acc[j] = 0.0;
  }
  auto at = [&](std::uint32_t slot, int j) -> double& { return P[(slot * N + j) * kTile + lane]; };
  // each dispatch branch below runs its operation on every point, so a word
  // is decoded and branched on once
#define CMI_EACH(stmt)            \
  _Pragma("unroll") for (int j = 0; j < N; ++j) { stmt; }
  const u32x4* const words = reinterpret_cast<const u32x4*>(code);
  u32x4 cur = tape_chunk(words, lane, n_instr);
  for (std::uint32_t c0 = 0; c0 < n_instr; c0 += kTile) {
    u32x4 const nxt = tape_chunk(words, c0 + kTile + lane, n_instr);
    std::uint32_t const end = std::min<std::uint32_t>(n_instr - c0, kTile);
    for (std::uint32_t slot = 0; slot < end; ++slot) {
      Word const ins = decode_word(lane_word(cur, slot));
      std::uint32_t const op = ins.op;
      if (op >= tape::kTermBase) {
        std::uint32_t const term = op & 7u;
        std::uint32_t const mode = (op - tape::kTermBase) >> 3;
        double x[N];
        if (term == static_cast<std::uint32_t>(tape::Term::SSC)) {
          CMI_EACH(x[j] = (at(ins.a, j) * at(ins.b, j)) * ins.imm)
        } else if (term == static_cast<std::uint32_t>(tape::Term::QC)) {
          CMI_EACH(double const a = at(ins.a, j); x[j] = (a * a) * ins.imm)
        } else if (term == static_cast<std::uint32_t>(tape::Term::SC)) {
          CMI_EACH(x[j] = at(ins.a, j) * ins.imm)
        } else if (term == static_cast<std::uint32_t>(tape::Term::SS)) {
          CMI_EACH(x[j] = at(ins.a, j) * at(ins.b, j))
        } else if (term == static_cast<std::uint32_t>(tape::Term::Q)) {
          CMI_EACH(double const a = at(ins.a, j); x[j] = a * a)
        } else {
          CMI_EACH(x[j] = at(ins.a, j))
        }
        if (mode == static_cast<std::uint32_t>(tape::AccMode::Add)) {
          CMI_EACH(acc[j] = acc[j] + x[j])
        } else if (mode == static_cast<std::uint32_t>(tape::AccMode::Set)) {
          CMI_EACH(acc[j] = x[j])
        } else {
          CMI_EACH(acc[j] = acc[j] - x[j])
        }
        continue;
      }
      if (op >= static_cast<std::uint32_t>(TOp::AccStore) && op < tape::kExtBase) {
        switch (static_cast<TOp>(op)) {
          case TOp::AccStore: CMI_EACH(at(ins.dst, j) = acc[j]) break;
          case TOp::AccAddImm: CMI_EACH(acc[j] = acc[j] + ins.imm) break;
          case TOp::AccSubImm: CMI_EACH(acc[j] = acc[j] - ins.imm) break;
          case TOp::AccImmSub: CMI_EACH(acc[j] = ins.imm - acc[j]) break;
          case TOp::AccScale: CMI_EACH(acc[j] = acc[j] * ins.imm) break;
          case TOp::AccNeg: CMI_EACH(acc[j] = -acc[j]) break;
          case TOp::AccCheckLE: CMI_EACH(res.feasible[j] = res.feasible[j] && (acc[j] <= ins.imm)) break;
          case TOp::AccCheckEQ: CMI_EACH(res.feasible[j] = res.feasible[j] && almost_equal(acc[j], ins.imm)) break;
          case TOp::AccObjective: CMI_EACH(res.obj[j] = acc[j]) break;
          default: break;
        }
        continue;
      }
      if constexpr (Ext) {
        if (op >= tape::kExtBase && op < tape::kExtEnd) {
          CMI_EACH(at(ins.dst, j) = ext_point(op, at(ins.a, j), P + j * kTile, ins, lane, N * kTile))
          continue;
        }
      }
      switch (static_cast<TOp>(op)) {
        case TOp::LoadConst: CMI_EACH(at(ins.dst, j) = ins.imm) break;
        case TOp::Add: CMI_EACH(at(ins.dst, j) = at(ins.a, j) + at(ins.b, j)) break;
        case TOp::AddImm: CMI_EACH(at(ins.dst, j) = at(ins.a, j) + ins.imm) break;
        case TOp::Sub: CMI_EACH(at(ins.dst, j) = at(ins.a, j) - at(ins.b, j)) break;
        case TOp::SubImm: CMI_EACH(at(ins.dst, j) = at(ins.a, j) - ins.imm) break;
        case TOp::ImmSub: CMI_EACH(at(ins.dst, j) = ins.imm - at(ins.a, j)) break;
        case TOp::Mul: CMI_EACH(at(ins.dst, j) = at(ins.a, j) * at(ins.b, j)) break;
        case TOp::MulImm: CMI_EACH(at(ins.dst, j) = at(ins.a, j) * ins.imm) break;
        case TOp::Sqr: CMI_EACH(double const a = at(ins.a, j); at(ins.dst, j) = a * a) break;
        case TOp::Neg: CMI_EACH(at(ins.dst, j) = -at(ins.a, j)) break;
        case TOp::CheckLE: CMI_EACH(res.feasible[j] = res.feasible[j] && (at(ins.a, j) <= ins.imm)) break;
        case TOp::CheckEQ: CMI_EACH(res.feasible[j] = res.feasible[j] && almost_equal(at(ins.a, j), ins.imm)) break;
        case TOp::Objective: CMI_EACH(res.obj[j] = at(ins.a, j)) break;
        default: break;
      }
    }
    cur = nxt;
  }
#undef CMI_EACH
  return res;
}

// ---------------------------------------------------------------------------
// An item's parent box and slot data, fetched lane-parallel (one memory round
// trip for the whole wave) and handed out by index with v_readlane: the
// decode loops then run on registers and LDS only.
// ---------------------------------------------------------------------------

__device__ __forceinline__ std::uint32_t lane_u32(std::uint32_t x, std::uint32_t k)
{
  return static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), static_cast<int>(k)));
}

__device__ __forceinline__ double lane_f64(double x, std::uint32_t k)
{
  std::uint64_t const b = __builtin_bit_cast(std::uint64_t, x);
  std::uint64_t const lo = lane_u32(static_cast<std::uint32_t>(b), k);
  std::uint64_t const hi = lane_u32(static_cast<std::uint32_t>(b >> 32), k);
  return __builtin_bit_cast(double, (hi << 32) | lo);
}

/// Slot j of an item in lane j (slot_count <= kMaxSlotsPerItem = kTile): its
/// variable, width and kind, and that variable's parent bounds.
struct LaneSlots
{
  std::uint32_t var, width, kind;
  double plo, phi;
};

__device__ __forceinline__ LaneSlots lane_slots(const Item& it, unsigned lane)
{
  LaneSlots s {0u, 1u, 0u, 0.0, 0.0};
  if (lane < it.slot_count) {
    s.var = gload(it.slot_var + lane);
    s.width = gload(it.slot_width + lane);
    s.kind = gload(it.slot_kind + lane);
    s.plo = gload(it.parent + 2 * s.var);
    s.phi = gload(it.parent + 2 * s.var + 1);
  }
  return s;
}

/// Agent-scope relaxed load of a double another workgroup wrote (past this
/// CU's L1).
__device__ __forceinline__ double atomic_load_f64(const double* p)
{
  std::uint64_t const b = __hip_atomic_load(reinterpret_cast<const std::uint64_t*>(p), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_bit_cast(double, b);
}

/// Wave 0 of one slice of tile lt, after its groups' verdicts are combined
/// into `feasible` (N per lane: one per point, or the lane's one region) and
/// the slice that holds the objective group has stored the tile's objective
/// values: folds the infeasible lanes of each point j into the tile's mask j
/// and counts the slice. Returns true in the last slice of the tile to
/// finish, with `feasible` then the verdict of every group; that slice clears
/// the tile's masks and count for the next launch. Nothing waits on another
/// workgroup, so the slices may run in any order.
template<int N>
__device__ __forceinline__ bool join_slices(const Item& it, std::uint32_t lt, std::uint32_t n_slices, bool* feasible,
                                            unsigned lane)
{
  std::uint64_t bad[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    bad[j] = __ballot(!feasible[j]);
  }
  std::uint64_t* const masks = it.tile_infeasible + static_cast<std::size_t>(lt) * kSliceMasks;
  std::uint32_t ticket = 0;
  __threadfence();
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (bad[j] != 0) {
        atomicOr(reinterpret_cast<unsigned long long*>(masks + j), static_cast<unsigned long long>(bad[j]));
      }
    }
    __threadfence();
    ticket = atomicAdd(it.tile_slices_done + lt, 1u);
  }
  ticket = __shfl(ticket, 0, 64);
  if (ticket != n_slices - 1) {
    return false;
  }
  __threadfence();
#pragma unroll
  for (int j = 0; j < N; ++j) {
    std::uint64_t const mask = __hip_atomic_load(masks + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    feasible[j] = ((mask >> lane) & 1u) == 0u;
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      masks[j] = 0;
    }
    it.tile_slices_done[lt] = 0;
  }
  return true;
}

// Decode-mode item kernels. Each workgroup is one 64-element tile of one item,
// one wave per function group of the item's tape (blockDim.x / 64 waves, the
// most groups of any item in the launch; a wave past its item's groups runs
// no code). Wave 0 decodes the tile's sub-boxes or points into the shared
// variable slots; each wave then runs its group's code over them; the groups'
// verdicts and the objective meet in LDS after the lds_slots slots, and wave
// 0 writes the results. Groups share only the (read-only) variable slots, so
// the values are those of the whole tape run by one wave.
//
// N (sampler and enumerator only): points per lane. A tile is then 64 x N
// elements, point j of lane l being element tile * 64N + j * 64 + l, and
// every word of the tape is fetched and dispatched once for N points (its
// arithmetic per point is unchanged, so each point's values are those of
// N = 1).
template<int M, bool Ext, int N>
__global__ __launch_bounds__(kTile * kMaxGroups) void k_items(const Item* __restrict__ items,
                                                               const std::uint32_t* __restrict__ tile_item,
                                                               std::uint32_t lds_slots)
{
  extern __shared__ double2 lds[];
  unsigned const lane = threadIdx.x % kTile;
  unsigned const wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);
  unsigned const n_waves = blockDim.x / kTile;
  Item const& it = items[tile_item[blockIdx.x]];
  // workgroup -> (tile, slice): the slices of one tile are consecutive
  std::uint32_t const n_slices = it.n_slices > 1 ? it.n_slices : 1u;
  std::uint32_t const local = blockIdx.x - it.tile_begin;
  std::uint32_t const lt = n_slices == 1 ? local : local / n_slices;
  std::uint32_t const slice = local - lt * n_slices;
  std::uint32_t const S = it.samples;
  // this wave's function group, its code range, and the objective's group
  std::uint32_t const group = slice * n_waves + wave;
  std::uint32_t g_begin = 0, g_len = 0, objective = 0;
  if (n_slices == 1) {
    objective = it.groups.objective_group;
    if (group < it.groups.n_groups) {
      g_begin = it.groups.begin[group];
      g_len = it.groups.begin[group + 1] - g_begin;
    }
  } else {
    objective = it.slice_objective;
    if (group < it.slice_groups) {
      g_begin = gload(it.slice_begin + group);
      g_len = gload(it.slice_begin + group + 1) - g_begin;
    }
  }
  const Instr* const code = it.code + g_begin;
  std::uint32_t const n_vars = it.n_vars;
  constexpr bool kInterval = M == static_cast<int>(Mode::Bound) || M == static_cast<int>(Mode::Aggregate);
  // per-group verdicts [kMaxGroups][N][64] and the objective [N][64], after
  // the slots (N points of 8 bytes, or one interval of 16, per lane and slot)
  auto* const feas_g = reinterpret_cast<std::uint32_t*>(reinterpret_cast<unsigned char*>(lds)
                                                        + std::size_t {lds_slots} * kTile * (kInterval ? 16u : 8u * N));

  if constexpr (kInterval) {
    static_assert(N == 1, "interval items take one region per lane");
    std::uint64_t const e = static_cast<std::uint64_t>(lt) * kTile + lane;
    bool const active = e < it.n_elems;
    std::uint32_t const region = static_cast<std::uint32_t>(active ? e : it.n_elems - 1);
    double2* const Sl = lds;
    if (wave == 0) {
      LaneSlots const ls = lane_slots(it, lane);
      for (std::uint32_t v0 = 0; v0 < n_vars; v0 += kTile) {
        double lo = 0.0, hi = 0.0;
        if (v0 + lane < n_vars) {
          lo = gload(it.parent + 2 * (v0 + lane));
          hi = gload(it.parent + 2 * (v0 + lane) + 1);
        }
        std::uint32_t const cnt = std::min<std::uint32_t>(kTile, n_vars - v0);
        for (std::uint32_t k = 0; k < cnt; ++k) {
          Sl[(v0 + k) * kTile + lane] = make_double2(-lane_f64(lo, k), lane_f64(hi, k));
        }
      }
      std::uint32_t idx = region;
      for (std::uint32_t j = 0; j < it.slot_count; ++j) {
        std::uint32_t const w = lane_u32(ls.width, j);
        std::uint32_t const part = idx % w;
        idx /= w;
        std::uint32_t const vid = lane_u32(ls.var, j);
        Bounds const b = slot_bounds(static_cast<std::uint8_t>(lane_u32(ls.kind, j)),
                                     {lane_f64(ls.plo, j), lane_f64(ls.phi, j)}, part, w);
        Sl[vid * kTile + lane] = make_double2(-b.lo, b.hi);
      }
    }
    if (n_waves > 1) {
      __syncthreads();
    }
    int const t = enter_round_up();
    IntervalResult res = run_interval<Ext>(code, g_len, Sl, lane, t, nullptr, 0, 0);
    Bnd obj = to_output(res.obj, t);
    leave_round_up(t, obj.lb, obj.ub, res.feasible);
    if (n_waves > 1) {
      auto* const obj_s = reinterpret_cast<double2*>(feas_g + kMaxGroups * kTile);
      feas_g[wave * kTile + lane] = res.feasible ? 1u : 0u;
      if (group == objective) {
        obj_s[lane] = to_d2(obj);
      }
      __syncthreads();
      if (wave != 0) {
        return;
      }
      bool f = true;
      for (unsigned g = 0; g < n_waves; ++g) {
        f = f && feas_g[g * kTile + lane] != 0u;
      }
      res.feasible = f;
      double2 const o = obj_s[lane];
      obj = {o.x, o.y};
    }
    if (n_slices > 1) {
      auto* const tobj = reinterpret_cast<double2*>(it.tile_obj + static_cast<std::size_t>(lt) * kSliceObjDoubles);
      if (objective / n_waves == slice) {
        tobj[lane] = to_d2(obj);
      }
      if (!join_slices<1>(it, lt, n_slices, &res.feasible, lane)) {
        return;
      }
      obj = {atomic_load_f64(&tobj[lane].x), atomic_load_f64(&tobj[lane].y)};
    }
    if constexpr (M == static_cast<int>(Mode::Bound)) {
      if (active) {
        it.feasible[region] = res.feasible ? 1 : 0;
        it.obj_lb[region] = obj.lb;
      }
    } else {
      // Mask by feasibility with the reductions' identities (reference
      // backend/aggregate/kernels.cuh:32-47), reduce each child's run of
      // lanes with shuffles, and fold the run into the child's bound with one
      // atomic min and max per child and tile. Min and max are exact, so the
      // result does not depend on the order the atomics land in.
      bool const live = active && res.feasible;
      double lo = live ? obj.lb : INFINITY;
      double hi = live ? obj.ub : -INFINITY;
      std::uint64_t const child = static_cast<std::uint64_t>(region) / it.refine;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        double const olo = __shfl_down(lo, off, 64);
        double const ohi = __shfl_down(hi, off, 64);
        std::uint64_t const oc = __shfl_down(child, off, 64);
        if (lane + off < 64 && oc == child) {
          lo = fmin(lo, olo);
          hi = fmax(hi, ohi);
        }
      }
      std::uint64_t const prev = __shfl_up(child, 1, 64);
      bool const head = lane == 0 || prev != child;
      if (active && head) {
        atomicMin(it.child_lb + child, lo);
        atomicMax(it.child_ub + child, hi);
      }
    }
    return;
  } else {
    double* const P = reinterpret_cast<double*>(lds);
    bool const sampling = M == static_cast<int>(Mode::Sample);
    // point j of this lane: element e[j], its region and sample
    std::uint64_t e[N];
    bool active[N];
    std::uint32_t region[N], sample[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      e[j] = (static_cast<std::uint64_t>(lt) * N + j) * kTile + lane;
      active[j] = e[j] < it.n_elems;
      std::uint32_t const ee = static_cast<std::uint32_t>(active[j] ? e[j] : it.n_elems - 1);
      region[j] = sampling ? ee / S : ee;
      sample[j] = sampling ? ee - region[j] * S : 0;
    }
    auto at = [&](std::uint32_t v, int j) -> double& { return P[(v * N + j) * kTile + lane]; };
    if (wave == 0) {
      LaneSlots const ls = lane_slots(it, lane);
      std::uint32_t skind = 0;  // lane j: kind of the variable of slot j
      if (lane < it.slot_count) {
        skind = gload(it.var_kinds + ls.var);
      }
      for (std::uint32_t v0 = 0; v0 < n_vars; v0 += kTile) {
        double lo = 0.0, hi = 0.0;
        std::uint32_t vk = 0;
        if (v0 + lane < n_vars) {
          lo = gload(it.parent + 2 * (v0 + lane));
          hi = gload(it.parent + 2 * (v0 + lane) + 1);
          vk = gload(it.var_kinds + v0 + lane);
        }
        std::uint32_t const cnt = std::min<std::uint32_t>(kTile, n_vars - v0);
        for (std::uint32_t k = 0; k < cnt; ++k) {
          std::uint32_t const v = v0 + k;
          double const l = lane_f64(lo, k);
          double const h = lane_f64(hi, k);
          bool const integral = lane_u32(vk, k) != 0;
#pragma unroll
          for (int j = 0; j < N; ++j) {
            at(v, j) = sampling ? draw(integral, l, h, v, sample[j], region[j], it.salt) : l;
          }
        }
      }
      std::uint32_t idx[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        idx[j] = region[j];
      }
      for (std::uint32_t k = 0; k < it.slot_count; ++k) {
        std::uint32_t const w = lane_u32(ls.width, k);
        std::uint32_t const vid = lane_u32(ls.var, k);
        auto const kind = static_cast<std::uint8_t>(lane_u32(ls.kind, k));
        bool const integral = lane_u32(skind, k) != 0;
        Bounds const pb {lane_f64(ls.plo, k), lane_f64(ls.phi, k)};
#pragma unroll
        for (int j = 0; j < N; ++j) {
          std::uint32_t const part = idx[j] % w;
          idx[j] /= w;
          Bounds const b = slot_bounds(kind, pb, part, w);
          at(vid, j) = sampling ? draw(integral, b.lo, b.hi, vid, sample[j], region[j], it.salt) : b.lo;
        }
      }
    }
    if (n_waves > 1) {
      __syncthreads();
    }
    double obj[N];
    bool feasible[N];
    if constexpr (N == 1) {
      PointResult const r = run_point<Ext>(code, g_len, P, lane, nullptr, 0, 0);
      obj[0] = r.obj;
      feasible[0] = r.feasible;
    } else {
      PointsResult<Ext, N> const r = run_points<Ext, N>(code, g_len, P, lane);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        obj[j] = r.obj[j];
        feasible[j] = r.feasible[j];
      }
    }
    if (n_waves > 1) {
      auto* const obj_s = reinterpret_cast<double*>(feas_g + kMaxGroups * kTile * N);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        feas_g[(wave * N + j) * kTile + lane] = feasible[j] ? 1u : 0u;
        if (group == objective) {
          obj_s[j * kTile + lane] = obj[j];
        }
      }
      __syncthreads();
      if (wave != 0) {
        return;
      }
#pragma unroll
      for (int j = 0; j < N; ++j) {
        bool f = true;
        for (unsigned g = 0; g < n_waves; ++g) {
          f = f && feas_g[(g * N + j) * kTile + lane] != 0u;
        }
        feasible[j] = f;
        obj[j] = obj_s[j * kTile + lane];
      }
    }
    if (n_slices > 1) {
      double* const tobj = it.tile_obj + static_cast<std::size_t>(lt) * kSliceObjDoubles;
      if (objective / n_waves == slice) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          tobj[j * kTile + lane] = obj[j];
        }
      }
      if (!join_slices<N>(it, lt, n_slices, feasible, lane)) {
        return;
      }
#pragma unroll
      for (int j = 0; j < N; ++j) {
        obj[j] = atomic_load_f64(&tobj[j * kTile + lane]);
      }
    }
    // this lane's best point (lowest element on ties), then the tile's
    double v = INFINITY;
    std::int64_t i = INT64_MAX;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      // a point needs a finite objective (log(0) = -inf is no incumbent)
      double vj = (active[j] && feasible[j]) ? obj[j] : INFINITY;
      if (vj != vj || vj == -INFINITY) {
        vj = INFINITY;
      }
      std::int64_t const ij = active[j] ? static_cast<std::int64_t>(e[j]) : INT64_MAX;
      if (vj < v || (vj == v && ij < i)) {
        v = vj;
        i = ij;
      }
    }
    wave_argmin(v, i);
    if (lane == 0) {
      it.tile_min[lt] = v;
      it.tile_arg[lt] = i;
    }
    __threadfence();
    std::uint32_t ticket = 0;
    if (lane == 0) {
      ticket = atomicAdd(it.done, 1u);
    }
    ticket = __shfl(ticket, 0, 64);
    if (ticket != it.n_tiles - 1) {
      return;
    }
    __threadfence();
    double bv = INFINITY;
    std::int64_t bi = INT64_MAX;
    for (std::uint32_t k = lane; k < it.n_tiles; k += kTile) {
      std::uint64_t const vb = __hip_atomic_load(reinterpret_cast<std::uint64_t*>(&it.tile_min[k]),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      std::int64_t const ib = __hip_atomic_load(&it.tile_arg[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      double const vv = __builtin_bit_cast(double, vb);
      if (vv < bv || (vv == bv && ib < bi)) {
        bv = vv;
        bi = ib;
      }
    }
    wave_argmin(bv, bi);
    bool const found = bv < DBL_MAX;
    if (lane == 0) {
      *it.cand_value = found ? bv : DBL_MAX;
      *it.cand_index = found ? bi : -1;
      *it.done = 0;
    }
    std::uint32_t const best = found ? static_cast<std::uint32_t>(bi) : 0;
    std::uint32_t const b_region = sampling ? best / S : best;
    std::uint32_t const b_sample = sampling ? best - b_region * S : 0;
    LaneSlots const ls = lane_slots(it, lane);
    for (std::uint32_t v0 = 0; v0 < n_vars; v0 += kTile) {
      std::uint32_t const v = v0 + lane;
      double lo = 0.0, hi = 0.0;
      if (v < n_vars) {
        lo = gload(it.parent + 2 * v);
        hi = gload(it.parent + 2 * v + 1);
      }
      std::uint32_t bidx = b_region;
      for (std::uint32_t j = 0; j < it.slot_count; ++j) {
        std::uint32_t const w = lane_u32(ls.width, j);
        std::uint32_t const vj = lane_u32(ls.var, j);
        auto const kj = static_cast<std::uint8_t>(lane_u32(ls.kind, j));
        std::uint32_t const part = bidx % w;
        bidx /= w;
        if (vj == v) {
          Bounds const b = slot_bounds(kj, {lo, hi}, part, w);
          lo = b.lo;
          hi = b.hi;
        }
      }
      if (v >= n_vars) {
        continue;
      }
      double const x = sampling ? draw(it.var_kinds[v] != 0, lo, hi, v, b_sample, b_region, it.salt) : lo;
      it.witness[v] = found ? x : NAN;
    }
  }
}

int lds_optin_bytes()
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || optin <= 0) {
    (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  }
  return optin;
}

template<typename K>
void allow_large_lds(K kernel)
{
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            lds_optin_bytes());
}

}  // namespace

std::uint32_t max_lds_slots(bool interval)
{
  return static_cast<std::uint32_t>(lds_optin_bytes()) / (kTile * (interval ? 16u : 8u));
}

std::size_t boxes_lds_bytes(std::uint32_t n_slots)
{
  return static_cast<std::size_t>(n_slots) * kTile * 16 + kMaxGroups * kTile * 4 + kTile * 16;
}

std::size_t lds_limit_bytes() { return static_cast<std::size_t>(lds_optin_bytes()); }

namespace
{

template<int M, bool Ext>
void launch_items_as(const Item* items, const std::uint32_t* tile_item, std::uint32_t n_tiles, std::size_t bytes,
                     std::uint32_t lds_slots, std::uint32_t waves, hipStream_t stream)
{
  static bool once = (allow_large_lds(k_items<M, Ext, 1>), true);
  (void)once;
  hipLaunchKernelGGL((k_items<M, Ext, 1>), dim3(n_tiles), dim3(kTile * waves), bytes, stream, items, tile_item,
                     lds_slots);
}

template<int M, bool Ext, int N>
const void* items_kernel_fn()
{
  static bool once = (allow_large_lds(k_items<M, Ext, N>), true);
  (void)once;
  return reinterpret_cast<const void*>(&k_items<M, Ext, N>);
}

template<int M, bool Ext>
const void* point_items_kernel(std::uint32_t points_per_lane)
{
  switch (points_per_lane) {
    case 2: return items_kernel_fn<M, Ext, 2>();
    case 4: return items_kernel_fn<M, Ext, 4>();
    default: return items_kernel_fn<M, Ext, 1>();
  }
}

template<bool Ext>
const void* items_kernel_mode(Mode mode, std::uint32_t points_per_lane)
{
  switch (mode) {
    case Mode::Bound: return items_kernel_fn<0, Ext, 1>();
    case Mode::Sample: return point_items_kernel<1, Ext>(points_per_lane);
    case Mode::Enumerate: return point_items_kernel<2, Ext>(points_per_lane);
    case Mode::Aggregate: return items_kernel_fn<3, Ext, 1>();
  }
  return nullptr;
}

template<bool Ext>
void launch_items_mode(Mode mode, const Item* items, const std::uint32_t* tile_item, std::uint32_t n_tiles,
                       std::size_t bytes, std::uint32_t lds_slots, std::uint32_t waves, hipStream_t stream)
{
  switch (mode) {
    case Mode::Bound: launch_items_as<0, Ext>(items, tile_item, n_tiles, bytes, lds_slots, waves, stream); break;
    case Mode::Sample: launch_items_as<1, Ext>(items, tile_item, n_tiles, bytes, lds_slots, waves, stream); break;
    case Mode::Enumerate:
      launch_items_as<2, Ext>(items, tile_item, n_tiles, bytes, lds_slots, waves, stream);
      break;
    case Mode::Aggregate:
      launch_items_as<3, Ext>(items, tile_item, n_tiles, bytes, lds_slots, waves, stream);
      break;
  }
}

template<bool Ext>
void launch_boxes_as(const Instr* code, const GroupTable& groups, std::uint32_t n_slots, std::uint32_t n_vars,
                     const double* lb, const double* ub, std::uint64_t ld, std::uint64_t n_boxes,
                     std::uint8_t* feasible, double* obj_lb, double* obj_ub, double* roots, hipStream_t stream)
{
  static bool once = (allow_large_lds(k_boxes<Ext>), true);
  (void)once;
  std::uint64_t const blocks = (n_boxes + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_boxes<Ext>, dim3(static_cast<unsigned>(blocks)), dim3(kTile * groups.n_groups),
                     boxes_lds_bytes(n_slots), stream, code, groups, n_vars, n_slots, lb, ub, ld, n_boxes, feasible,
                     obj_lb, obj_ub, reinterpret_cast<double2*>(roots));
}

template<bool Ext>
void launch_points_as(const Instr* code, std::uint32_t n_instr, std::uint32_t n_slots, std::uint32_t n_vars,
                      const double* x, std::uint64_t ld, std::uint64_t n_points, std::uint8_t* feasible,
                      double* obj, double* roots, hipStream_t stream)
{
  static bool once = (allow_large_lds(k_points<Ext>), true);
  (void)once;
  std::uint64_t const blocks = (n_points + kTile - 1) / kTile;
  std::size_t const bytes = static_cast<std::size_t>(n_slots) * kTile * 8;
  hipLaunchKernelGGL(k_points<Ext>, dim3(static_cast<unsigned>(blocks)), dim3(kTile), bytes, stream, code, n_instr,
                     n_vars, x, ld, n_points, feasible, obj, roots);
}

}  // namespace

std::size_t items_lds_bytes(Mode mode, std::uint32_t lds_slots, std::uint32_t waves, std::uint32_t points_per_lane)
{
  bool const interval = mode == Mode::Bound || mode == Mode::Aggregate;
  std::size_t const n = interval ? 1 : std::max<std::uint32_t>(points_per_lane, 1);
  std::size_t const slots = static_cast<std::size_t>(lds_slots) * kTile * (interval ? 16 : 8) * n;
  return waves > 1 ? slots + (kMaxGroups * kTile * 4 + kTile * 16) * n : slots;
}

void launch_items(Mode mode, bool extended, const Item* items, const std::uint32_t* tile_item, std::uint32_t n_tiles,
                  std::uint32_t lds_slots, std::uint32_t waves, hipStream_t stream)
{
  if (n_tiles == 0) {
    return;
  }
  waves = std::clamp<std::uint32_t>(waves, 1, kMaxGroups);
  std::size_t const bytes = items_lds_bytes(mode, lds_slots, waves);
  if (extended) {
    launch_items_mode<true>(mode, items, tile_item, n_tiles, bytes, lds_slots, waves, stream);
  } else {
    launch_items_mode<false>(mode, items, tile_item, n_tiles, bytes, lds_slots, waves, stream);
  }
}

ItemsLaunch items_launch(Mode mode, bool extended, std::uint32_t n_tiles, std::uint32_t lds_slots,
                         std::uint32_t waves, std::uint32_t points_per_lane)
{
  waves = std::clamp<std::uint32_t>(waves, 1, kMaxGroups);
  if (points_per_lane != 2 && points_per_lane != 4) {
    points_per_lane = 1;
  }
  ItemsLaunch l;
  l.func = extended ? items_kernel_mode<true>(mode, points_per_lane) : items_kernel_mode<false>(mode, points_per_lane);
  l.grid = n_tiles;
  l.block = kTile * waves;
  l.lds_bytes = items_lds_bytes(mode, lds_slots, waves, points_per_lane);
  return l;
}

void launch_boxes(const Instr* code, bool extended, const GroupTable& groups, std::uint32_t n_slots,
                  std::uint32_t n_vars, const double* lb, const double* ub, std::uint64_t ld, std::uint64_t n_boxes,
                  std::uint8_t* feasible, double* obj_lb, double* obj_ub, double* roots, hipStream_t stream)
{
  if (n_boxes == 0) {
    return;
  }
  if (extended) {
    launch_boxes_as<true>(code, groups, n_slots, n_vars, lb, ub, ld, n_boxes, feasible, obj_lb, obj_ub, roots, stream);
  } else {
    launch_boxes_as<false>(code, groups, n_slots, n_vars, lb, ub, ld, n_boxes, feasible, obj_lb, obj_ub, roots, stream);
  }
}

void launch_points(const Instr* code, bool extended, std::uint32_t n_instr, std::uint32_t n_slots,
                   std::uint32_t n_vars, const double* x, std::uint64_t ld, std::uint64_t n_points,
                   std::uint8_t* feasible, double* obj, double* roots, hipStream_t stream)
{
  if (n_points == 0) {
    return;
  }
  if (extended) {
    launch_points_as<true>(code, n_instr, n_slots, n_vars, x, ld, n_points, feasible, obj, roots, stream);
  } else {
    launch_points_as<false>(code, n_instr, n_slots, n_vars, x, ld, n_points, feasible, obj, roots, stream);
  }
}

}  // namespace cmi::dev

// ---------------------------------------------------------------------------
// Static scratch for the entry points that allocate nothing
// ---------------------------------------------------------------------------

namespace
{

// Staging, tile partials and the done counter of a no-allocation launch. In
// host code `scratch` is only a stand-in for the symbol: its device address
// comes from hipGetSymbolAddress, per device (cmi::dev::static_scratch).
static __device__ __align__(256) unsigned char scratch[cmi::dev::kStaticScratchBytes];

__global__ void k_fill_u32(std::uint32_t* __restrict__ p, std::uint64_t n, std::uint32_t value)
{
  std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    p[i] = value;
  }
}

}  // namespace

namespace cmi::dev
{

void* static_scratch()
{
  static std::mutex mu;
  static void* address[kMaxDevices] = {};
  int device = -1;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) {
    throw DeviceError(static_cast<int>(hipErrorInvalidDevice), "static scratch: no current device in [0, " + std::to_string(kMaxDevices) + ")");
  }
  std::lock_guard<std::mutex> lock(mu);
  if (address[device] == nullptr) {
    void* p = nullptr;
    hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      throw DeviceError(static_cast<int>(e), std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    address[device] = p;
  }
  return address[device];
}

void launch_fill_u32(std::uint32_t* p, std::size_t n, std::uint32_t value, hipStream_t stream)
{
  if (n == 0) {
    return;
  }
  constexpr unsigned kBlock = 256;
  auto const blocks = static_cast<unsigned>((n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(kBlock), 0, stream, p, static_cast<std::uint64_t>(n), value);
}

}  // namespace cmi::dev
