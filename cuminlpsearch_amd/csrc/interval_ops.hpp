// Outward-rounded FP64 interval operations for gfx950, shared by the tape
// interpreter (kernels.hip) and by the kernels compiled at run time from a tape
// (jit.cpp embeds this file's text). Self-contained: no includes, builtin types
// only, so the run-time compiler needs nothing else.
//
// gfx950 has no per-instruction rounding modes for FP64. These operations run
// inside a region where the MODE register's FP64 rounding field is set to
// round-toward-+infinity, and an interval [lb, ub] is held as (nl, ub) with
// nl = -lb: the lower bound is then computed as -RD(x) = RU(-x) through VALU
// source negation, which is exact, so every interval operation is one
// instruction per bound. Every rounding FP64 instruction here is inline asm
// that takes the region's token as an operand, so the compiler can neither fold
// the negations away nor hoist an operation above the region's entry; the
// region's exit takes the values that must be final as operands, so nothing
// that feeds them sinks below it. Entry and exit are volatile asm with a memory
// clobber.
//
// Op semantics follow the reference's interval functors
// (backend/graph/ops.cuh:24-355) for the QCQP op set. Multiplication takes the
// max of the four directed endpoint products for each bound (equal in value to
// the IEEE-1788 case table for finite operands).

#ifndef CMI_INTERVAL_OPS_HPP
#define CMI_INTERVAL_OPS_HPP

namespace cmi_iv
{

/// Enter the round-toward-+infinity region; returns its token.
__device__ __forceinline__ int enter_round_up()
{
  int tok;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 1\n\t"
      "s_mov_b32 %0, 0"
      : "=s"(tok)
      :
      : "memory");
  return tok;
}

/// Leave the region (round to nearest even again).
__device__ __forceinline__ void leave_round_up(int tok)
{
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0 ; %0"
      :
      : "s"(tok)
      : "memory");
}

/// Leave the region after the two values a and b are computed.
__device__ __forceinline__ void leave_round_up(int tok, double a, double b)
{
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0 ; %0 %1 %2"
      :
      : "s"(tok), "v"(a), "v"(b)
      : "memory");
}

/// An interval [lb, ub] held as nl = -lb and ub.
struct Iv
{
  double nl, ub;
};

/// An interval as [lb, ub].
struct Bnd
{
  double lb, ub;
};

/// Box variable as (-lb, ub) from the double at byte offset `off` of the
/// columns lb and ub. The sign flip is opaque to the compiler, so it is done
/// once and not repeated at every use. With lb, ub uniform and a 32-bit byte
/// offset, the loads take the scalar-base + vector-offset form.
__device__ __forceinline__ Iv load_box(const double* lb, const double* ub, unsigned off)
{
  double const l = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(lb) + off);
  double const u = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(ub) + off);
  unsigned long long const bits = __builtin_bit_cast(unsigned long long, l);
  unsigned hi = static_cast<unsigned>(bits >> 32);
  unsigned const lo = static_cast<unsigned>(bits);
  asm("v_xor_b32 %0, 0x80000000, %0" : "+v"(hi));
  Iv r;
  r.nl = __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
  r.ub = u;
  return r;
}

// ---------------------------------------------------------------------------
// One instruction per asm statement. Each is pure (the compiler may schedule,
// interleave or merge them freely) and names the region token, so none can
// move above the region's entry. Operands in SGPRs are wave-uniform constants.
// ---------------------------------------------------------------------------

#define CMI_IV_OP2(name, text)                                                   \
  __device__ __forceinline__ double name(double a, double b, int t)              \
  {                                                                              \
    double r;                                                                    \
    asm(text " ; %3" : "=v"(r) : "v"(a), "v"(b), "s"(t));                        \
    return r;                                                                    \
  }
#define CMI_IV_OP2S(name, text)                                                  \
  __device__ __forceinline__ double name(double a, double c, int t)              \
  {                                                                              \
    double r;                                                                    \
    asm(text " ; %3" : "=v"(r) : "v"(a), "s"(c), "s"(t));                        \
    return r;                                                                    \
  }

CMI_IV_OP2(ru_add, "v_add_f64 %0, %1, %2")          // RU(a + b)
CMI_IV_OP2(ru_mul, "v_mul_f64 %0, %1, %2")          // RU(a * b)
CMI_IV_OP2(ru_mul_na, "v_mul_f64 %0, -%1, %2")      // RU(-a * b)
CMI_IV_OP2(vmax, "v_max_f64 %0, %1, %2")            // max(a, b)
CMI_IV_OP2(vmax_na, "v_max_f64 %0, -%1, %2")        // max(-a, b)
CMI_IV_OP2(vmax_abs, "v_max_f64 %0, |%1|, |%2|")    // max(|a|, |b|)
CMI_IV_OP2S(ru_add_s, "v_add_f64 %0, %1, %2")       // RU(a + c)
CMI_IV_OP2S(ru_sub_s, "v_add_f64 %0, %1, -%2")      // RU(a - c)
CMI_IV_OP2S(ru_mul_s, "v_mul_f64 %0, %1, %2")       // RU(a * c)
CMI_IV_OP2S(ru_mul_na_s, "v_mul_f64 %0, -%1, %2")   // RU(-a * c)

#undef CMI_IV_OP2
#undef CMI_IV_OP2S

__device__ __forceinline__ double vmin0(double a, int t)  // min(a, 0)
{
  double r;
  asm("v_min_f64 %0, %1, 0 ; %2" : "=v"(r) : "v"(a), "s"(t));
  return r;
}
__device__ __forceinline__ double plus0(double a, int t)  // a + 0: a, or +0 for a zero
{
  double r;
  asm("v_add_f64 %0, %1, 0 ; %2" : "=v"(r) : "v"(a), "s"(t));
  return r;
}
__device__ __forceinline__ double neg_plus0(double a, int t)  // -a + 0
{
  double r;
  asm("v_add_f64 %0, -%1, 0 ; %2" : "=v"(r) : "v"(a), "s"(t));
  return r;
}
__device__ __forceinline__ double plus0_s(double c, int t)  // c + 0 for a uniform c
{
  double r;
  asm("v_add_f64 %0, %1, 0 ; %2" : "=v"(r) : "s"(c), "s"(t));
  return r;
}
__device__ __forceinline__ double neg_plus0_s(double c, int t)  // -c + 0 for a uniform c
{
  double r;
  asm("v_add_f64 %0, -%1, 0 ; %2" : "=v"(r) : "s"(c), "s"(t));
  return r;
}

// ---------------------------------------------------------------------------
// Interval operations on (nl, ub) = (-lb, ub)
// ---------------------------------------------------------------------------

// [a] + [b]: -lb = RU(-al + -bl), ub = RU(au + bu)
__device__ __forceinline__ Iv iv_add(Iv a, Iv b, int t) { return {ru_add(a.nl, b.nl, t), ru_add(a.ub, b.ub, t)}; }

// [a] - [b]: -lb = RU(-al + bu), ub = RU(au + -bl)
__device__ __forceinline__ Iv iv_sub(Iv a, Iv b, int t) { return {ru_add(a.nl, b.ub, t), ru_add(a.ub, b.nl, t)}; }

// [a] + c: -lb = RU(-al - c), ub = RU(au + c)
__device__ __forceinline__ Iv iv_add_imm(Iv a, double c, int t) { return {ru_sub_s(a.nl, c, t), ru_add_s(a.ub, c, t)}; }

// [a] - c: -lb = RU(-al + c), ub = RU(au - c)
__device__ __forceinline__ Iv iv_sub_imm(Iv a, double c, int t) { return {ru_add_s(a.nl, c, t), ru_sub_s(a.ub, c, t)}; }

// c - [a]: -lb = RU(au - c), ub = RU(-al + c)
__device__ __forceinline__ Iv iv_imm_sub(double c, Iv a, int t) { return {ru_sub_s(a.ub, c, t), ru_add_s(a.nl, c, t)}; }

__device__ __forceinline__ Iv iv_neg(Iv a) { return {a.ub, a.nl}; }

// The point interval [c, c] (exact: -c + 0 and c + 0 are -c and c, or +0).
__device__ __forceinline__ Iv iv_const(double c, int t) { return {neg_plus0_s(c, t), plus0_s(c, t)}; }

// [x] * [y] from the four endpoint products (xl = -xnl, yl = -ynl):
// -lb = max RU(-xi*yj), ub = max RU(xi*yj).
__device__ __forceinline__ Iv iv_mul(Iv x, Iv y, int t)
{
  double const l0 = ru_mul_na(x.nl, y.nl, t);  // -(xl*yl)
  double const l1 = ru_mul(x.nl, y.ub, t);     // -(xl*yu)
  double const l2 = ru_mul(x.ub, y.nl, t);     // -(xu*yl)
  double const l3 = ru_mul_na(x.ub, y.ub, t);  // -(xu*yu)
  double const u0 = ru_mul(x.nl, y.nl, t);     // xl*yl
  double const u1 = ru_mul_na(x.nl, y.ub, t);  // xl*yu
  double const u2 = ru_mul_na(x.ub, y.nl, t);  // xu*yl
  double const u3 = ru_mul(x.ub, y.ub, t);     // xu*yu
  return {vmax(vmax(l0, l1, t), vmax(l2, l3, t), t), vmax(vmax(u0, u1, t), vmax(u2, u3, t), t)};
}

// [a] * c for c > 0: [RD(al*c), RU(au*c)]; -lb = RU(-al*c) = RU(anl*c)
__device__ __forceinline__ Iv iv_mul_pos(Iv a, double c, int t) { return {ru_mul_s(a.nl, c, t), ru_mul_s(a.ub, c, t)}; }

// [a] * c for c < 0: [RD(au*c), RU(al*c)]; -lb = RU(-au*c), ub = RU(-anl*c)
__device__ __forceinline__ Iv iv_mul_neg(Iv a, double c, int t)
{
  return {ru_mul_na_s(a.ub, c, t), ru_mul_na_s(a.nl, c, t)};
}

// [a] * c: by the sign of c, [0, 0] for c == 0.
__device__ __forceinline__ Iv iv_mul_imm(Iv a, double c, int t)
{
  if (c > 0.0) {
    return iv_mul_pos(a, c, t);
  }
  if (c < 0.0) {
    return iv_mul_neg(a, c, t);
  }
  return {0.0, 0.0};
}

// [a]^2: s = max(lb, min(ub, 0)) is the endpoint nearest zero, or 0 when the
// interval straddles zero; -lb = RU(-s*s); ub = RU(m*m), m = max(|lb|, |ub|).
__device__ __forceinline__ Iv iv_sqr(Iv a, int t)
{
  double const s = vmax_na(a.nl, vmin0(a.ub, t), t);
  double const m = vmax_abs(a.nl, a.ub, t);
  return {ru_mul_na(s, s, t), ru_mul(m, m, t)};
}

/// [lb, ub] with a zero bound reported as +0 (x + 0 is exact otherwise).
__device__ __forceinline__ Bnd to_output(Iv a, int t) { return {neg_plus0(a.nl, t), plus0(a.ub, t)}; }

/// Interval constraint checks (reference kernels.cuh:283-340):
/// LE holds while lb <= rhs, EQ while lb <= rhs <= ub.
__device__ __forceinline__ bool check_le(Iv a, double rhs) { return -a.nl <= rhs; }
__device__ __forceinline__ bool check_eq(Iv a, double rhs) { return -a.nl <= rhs && rhs <= a.ub; }

}  // namespace cmi_iv

#endif
