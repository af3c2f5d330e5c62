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

/// Leave the region after a, b and the flag f are computed.
__device__ __forceinline__ void leave_round_up(int tok, double a, double b, bool f)
{
  int const fi = f ? 1 : 0;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0 ; %0 %1 %2 %3"
      :
      : "s"(tok), "v"(a), "v"(b), "v"(fi)
      : "memory");
}

/// Leave the region after the flag f is computed.
__device__ __forceinline__ void leave_round_up(int tok, bool f)
{
  int const fi = f ? 1 : 0;
  asm volatile(
      "s_nop 1\n\t"
      "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0 ; %0 %1"
      :
      : "s"(tok), "v"(fi)
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

/// (-lb, ub) from the two doubles lb, ub. The sign flip is opaque to the
/// compiler, so it is done once and not repeated at every use.
__device__ __forceinline__ Iv box_iv(double l, double u)
{
  unsigned long long const bits = __builtin_bit_cast(unsigned long long, l);
  unsigned hi = static_cast<unsigned>(bits >> 32);
  unsigned const lo = static_cast<unsigned>(bits);
  asm("v_xor_b32 %0, 0x80000000, %0" : "+v"(hi));
  Iv r;
  r.nl = __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
  r.ub = u;
  return r;
}

/// Box variable as (-lb, ub) from the double at byte offset `off` of the
/// columns lb and ub. With lb, ub uniform and a 32-bit byte offset, the loads
/// take the scalar-base + vector-offset form. A box variable that only one
/// wave of a workgroup reads is read once per launch, so these loads are
/// non-temporal (streamed past the caches).
__device__ __forceinline__ Iv load_box(const double* lb, const double* ub, unsigned off)
{
  return box_iv(
      __builtin_nontemporal_load(reinterpret_cast<const double*>(reinterpret_cast<const char*>(lb) + off)),
      __builtin_nontemporal_load(reinterpret_cast<const double*>(reinterpret_cast<const char*>(ub) + off)));
}

/// The same for a box variable that several waves of a workgroup read: plain
/// loads, so the second wave finds the lines in L2.
__device__ __forceinline__ Iv load_box_shared(const double* lb, const double* ub, unsigned off)
{
  return box_iv(*reinterpret_cast<const double*>(reinterpret_cast<const char*>(lb) + off),
                *reinterpret_cast<const double*>(reinterpret_cast<const char*>(ub) + off));
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
//
// Each bound is one asm statement of 7 instructions (4 products, then the max
// tree max(max(p0, p1), max(p2, p3))). The compiler's hazard check assumes that
// any inline asm may write a 16-bit half of a VGPR and puts an s_nop between
// two asm statements when the second reads what the first wrote; with one
// statement per instruction that was one s_nop per 4 to 5 instructions of the
// product chains. Inside a statement nothing is inserted, and none is needed:
// a dependent full-width FP64 VALU instruction waits in hardware (the compiler
// emits such pairs back to back in its own code). The bounds stay separate
// statements so that a bound nothing reads is still deleted.
__device__ __forceinline__ double mul_nl(Iv x, Iv y, int t)
{
  double r, p0, p1, p2, p3;
  asm("v_mul_f64 %1, -%6, %7\n\t"  // -(xl*yl)
      "v_mul_f64 %2, %6, %8\n\t"   // -(xl*yu)
      "v_mul_f64 %3, %9, %7\n\t"   // -(xu*yl)
      "v_mul_f64 %4, -%9, %8\n\t"  // -(xu*yu)
      "v_max_f64 %1, %1, %2\n\t"
      "v_max_f64 %3, %3, %4\n\t"
      "v_max_f64 %0, %1, %3 ; %5"
      : "=v"(r), "=&v"(p0), "=&v"(p1), "=&v"(p2), "=&v"(p3)
      : "s"(t), "v"(x.nl), "v"(y.nl), "v"(y.ub), "v"(x.ub));
  return r;
}
__device__ __forceinline__ double mul_ub(Iv x, Iv y, int t)
{
  double r, p0, p1, p2, p3;
  asm("v_mul_f64 %1, %6, %7\n\t"   // xl*yl
      "v_mul_f64 %2, -%6, %8\n\t"  // xl*yu
      "v_mul_f64 %3, -%9, %7\n\t"  // xu*yl
      "v_mul_f64 %4, %9, %8\n\t"   // xu*yu
      "v_max_f64 %1, %1, %2\n\t"
      "v_max_f64 %3, %3, %4\n\t"
      "v_max_f64 %0, %1, %3 ; %5"
      : "=v"(r), "=&v"(p0), "=&v"(p1), "=&v"(p2), "=&v"(p3)
      : "s"(t), "v"(x.nl), "v"(y.nl), "v"(y.ub), "v"(x.ub));
  return r;
}
__device__ __forceinline__ Iv iv_mul(Iv x, Iv y, int t) { return {mul_nl(x, y, t), mul_ub(x, y, t)}; }

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
// One asm statement per bound, as for the product.
__device__ __forceinline__ Iv iv_sqr(Iv a, int t)
{
  double l, u, s, m;
  asm("v_min_f64 %1, %4, 0\n\t"    // min(ub, 0)
      "v_max_f64 %1, -%3, %1\n\t"  // max(lb, min(ub, 0))
      "v_mul_f64 %0, -%1, %1 ; %2"
      : "=v"(l), "=&v"(s)
      : "s"(t), "v"(a.nl), "v"(a.ub));
  asm("v_max_f64 %1, |%3|, |%4|\n\t"
      "v_mul_f64 %0, %1, %1 ; %2"
      : "=v"(u), "=&v"(m)
      : "s"(t), "v"(a.nl), "v"(a.ub));
  return {l, u};
}

/// [lb, ub] with a zero bound reported as +0 (x + 0 is exact otherwise). A
/// NaN bound -- no information, as from the product of [0, 0] and an interval
/// unbounded on both sides -- is reported as the infinite bound on its side.
__device__ __forceinline__ Bnd to_output(Iv a, int t)
{
  double const nl = a.nl != a.nl ? __builtin_huge_val() : a.nl;
  double const ub = a.ub != a.ub ? __builtin_huge_val() : a.ub;
  return {neg_plus0(nl, t), plus0(ub, t)};
}

/// Interval constraint checks (reference kernels.cuh:283-340): LE fails once
/// lb > rhs, EQ once lb > rhs or ub < rhs; a NaN bound never fails a check.
__device__ __forceinline__ bool check_le(Iv a, double rhs) { return !(-a.nl > rhs); }
__device__ __forceinline__ bool check_eq(Iv a, double rhs) { return !(-a.nl > rhs) && !(rhs > a.ub); }

// ===========================================================================
// Extended operation set: Div, PowN, Sqrt, Exp, Log, Sin, Cos, Tanh, Abs, Min,
// Max. Same (nl, ub) representation, same region. Empty results (IEEE 1788:
// division by [0, 0], sqrt or log of an interval below the domain) clear the
// `ok` flag, which makes the box infeasible: every value reaches a root, and an
// empty value makes that root empty.
//
// Division and square root are correctly rounded in the directed sense: an
// approximation from the hardware reciprocal (square root) seed and Newton
// steps is corrected one ulp at a time against the exact residual a - b*q
// (v - s*s), whose sign a fused multiply-add gives exactly. Exp, log, sin, cos
// and tanh are enclosures: argument reduction with directed error terms and
// Horner schemes whose every step is rounded in the direction that keeps a
// bound (the partial sums have known signs), plus a bound on the series tail.
// ===========================================================================

#define CMI_CONST static constexpr
// BEGIN GENERATED CONSTANTS (make_enclosure_constants.py)
// 1/ln 2 (nearest); only picks the reduction integer
CMI_CONST double kInvLn2 = 0x1.71547652b82fep+0;
// ln 2 to 32 bits: k * kLn2Hi is exact for |k| < 2^21
CMI_CONST double kLn2Hi = 0x1.62e42fee00000p-1;
// ln 2 - kLn2Hi (nearest)
CMI_CONST double kLn2Lo = 0x1.a39ef35793c76p-33;
// bound on |ln 2 - kLn2Hi - kLn2Lo|
CMI_CONST double kLn2Err = 0x1.cc01f97b57a08p-87;
// RU(ln 2)
CMI_CONST double kLn2Up = 0x1.62e42fefa39f0p-1;
// RD(ln 2)
CMI_CONST double kLn2Dn = 0x1.62e42fefa39efp-1;
// 2/pi (nearest); only picks the quadrant
CMI_CONST double kTwoOverPi = 0x1.45f306dc9c883p-1;
// pi/2 to 32 bits: k * kPio2A is exact for |k| < 2^21
CMI_CONST double kPio2A = 0x1.921fb54400000p+0;
// next 53 bits of pi/2
CMI_CONST double kPio2B = 0x1.0b4611a626331p-34;
// next 53 bits of pi/2
CMI_CONST double kPio2C = 0x1.1701b839a2520p-88;
// bound on |pi/2 - A - B - C|
CMI_CONST double kPio2Err = 0x1.27044533e63a1p-142;
// RU(1/j!)
CMI_CONST double kExpUp[18] = {0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555556p-3, 0x1.5555555555556p-5, 0x1.1111111111112p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01bp-13, 0x1.a01a01a01a01bp-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5dp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d0ap-33, 0x1.93974a8c07c9ep-37, 0x1.ae7f3e733b820p-41, 0x1.ae7f3e733b820p-45, 0x1.952c77030ad4bp-49};
// RD(1/j!)
CMI_CONST double kExpDn[18] = {0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c16p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c733p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e3p-26, 0x1.1eed8eff8d897p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49};
// bound on the exp series tail for r <= 0.7
CMI_CONST double kExpTail = 0x1.2e613ec06638dp-61;
// RN(1/j!)
CMI_CONST double kExpRn[14] = {0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
// RU(1/(2j+1))
CMI_CONST double kAtanhUp[12] = {0x1.0000000000000p+0, 0x1.5555555555556p-2, 0x1.999999999999ap-3, 0x1.2492492492493p-3, 0x1.c71c71c71c71dp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111112p-4, 0x1.e1e1e1e1e1e1fp-5, 0x1.af286bca1af29p-5, 0x1.8618618618619p-5, 0x1.642c8590b2165p-5};
// RD(1/(2j+1))
CMI_CONST double kAtanhDn[12] = {0x1.0000000000000p+0, 0x1.5555555555555p-2, 0x1.9999999999999p-3, 0x1.2492492492492p-3, 0x1.c71c71c71c71cp-4, 0x1.745d1745d1745p-4, 0x1.3b13b13b13b13p-4, 0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5};
// bound on the atanh series tail (sum/u) for u^2 <= 0.0295
CMI_CONST double kAtanhTail = 0x1.522e290785893p-66;
// RN(1/(2j+1))
CMI_CONST double kAtanhRn[12] = {0x1.0000000000000p+0, 0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3, 0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5};
// RU((-1)^j/(2j+1)!)
CMI_CONST double kSinUp[10] = {0x1.0000000000000p+0, -0x1.5555555555555p-3, 0x1.1111111111112p-7, -0x1.a01a01a01a01ap-13, 0x1.71de3a556c734p-19, -0x1.ae64567f544e3p-26, 0x1.6124613a86d0ap-33, -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4bp-49, -0x1.2f49b46814157p-57};
// RD((-1)^j/(2j+1)!)
CMI_CONST double kSinDn[10] = {0x1.0000000000000p+0, -0x1.5555555555556p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01bp-13, 0x1.71de3a556c733p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b820p-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814158p-57};
// RU((-1)^j/(2j)!)
CMI_CONST double kCosUp[11] = {0x1.0000000000000p+0, -0x1.0000000000000p-1, 0x1.5555555555556p-5, -0x1.6c16c16c16c16p-10, 0x1.a01a01a01a01bp-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b820p-45, -0x1.6827863b97d97p-53, 0x1.e542ba4020226p-62};
// RD((-1)^j/(2j)!)
CMI_CONST double kCosDn[11] = {0x1.0000000000000p+0, -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5dp-22, 0x1.1eed8eff8d897p-29, -0x1.93974a8c07c9ep-37, 0x1.ae7f3e733b81fp-45, -0x1.6827863b97d98p-53, 0x1.e542ba4020225p-62};
// bound on the sin(r)/r series tail for r^2 <= 0.62 (alternating, decreasing)
CMI_CONST double kSinTail = 0x1.8d31a7266d8a7p-73;
// bound on the cos series tail for r^2 <= 0.62
CMI_CONST double kCosTail = 0x1.66325d05cbfa9p-78;
// RN((-1)^j/(2j+1)!)
CMI_CONST double kSinRn[10] = {0x1.0000000000000p+0, -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13, 0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57};
// RN((-1)^j/(2j)!)
CMI_CONST double kCosRn[11] = {0x1.0000000000000p+0, -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53, 0x1.e542ba4020225p-62};
// RU(1/(j+1)!)
CMI_CONST double kExpm1Up[19] = {0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555556p-3, 0x1.5555555555556p-5, 0x1.1111111111112p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01bp-13, 0x1.a01a01a01a01bp-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5dp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d0ap-33, 0x1.93974a8c07c9ep-37, 0x1.ae7f3e733b820p-41, 0x1.ae7f3e733b820p-45, 0x1.952c77030ad4bp-49, 0x1.6827863b97d98p-53, 0x1.2f49b46814158p-57};
// RD(1/(j+1)!)
CMI_CONST double kExpm1Dn[19] = {0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c16p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c733p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e3p-26, 0x1.1eed8eff8d897p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49, 0x1.6827863b97d97p-53, 0x1.2f49b46814157p-57};
// bound on the expm1(u)/u series tail for 0 <= u <= 1
CMI_CONST double kExpm1Tail = 0x1.49fa2cbaf721ap-60;
// RN(1/(j+1)!)
CMI_CONST double kExpm1Rn[19] = {0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49, 0x1.6827863b97d97p-53, 0x1.2f49b46814157p-57};
// END GENERATED CONSTANTS

constexpr double kInfD = __builtin_huge_val();
constexpr double kMaxD = 1.7976931348623157e308;
constexpr double kTinyD = 4.9406564584124654e-324;  // 2^-1074

#define CMI_IV_FMA(name, text, ca, cb, cc)                                        \
  __device__ __forceinline__ double name(double a, double b, double c, int t)   \
  {                                                                              \
    double r;                                                                    \
    asm(text " ; %4" : "=v"(r) : ca(a), cb(b), cc(c), "s"(t));                   \
    return r;                                                                    \
  }

CMI_IV_FMA(ru_fma, "v_fma_f64 %0, %1, %2, %3", "v", "v", "v")            // RU(a*b + c)
CMI_IV_FMA(ru_fma_na, "v_fma_f64 %0, -%1, %2, %3", "v", "v", "v")        // RU(-a*b + c)
CMI_IV_FMA(ru_fma_nc, "v_fma_f64 %0, %1, %2, -%3", "v", "v", "v")        // RU(a*b - c)
CMI_IV_FMA(ru_fma_s, "v_fma_f64 %0, %1, %2, %3", "v", "v", "s")          // RU(a*b + c), uniform c
CMI_IV_FMA(ru_fma_nc_s, "v_fma_f64 %0, %1, %2, -%3", "v", "v", "s")      // RU(a*b - c), uniform c
CMI_IV_FMA(ru_fma_na_s, "v_fma_f64 %0, -%1, %2, %3", "v", "v", "s")      // RU(-a*b + c), uniform c
CMI_IV_FMA(ru_fma_na_nc_s, "v_fma_f64 %0, -%1, %2, -%3", "v", "v", "s")  // RU(-a*b - c), uniform c
CMI_IV_FMA(ru_fma_bs, "v_fma_f64 %0, %1, %2, %3", "v", "s", "v")         // RU(a*b + c), uniform b
CMI_IV_FMA(ru_fma_na_bs, "v_fma_f64 %0, -%1, %2, %3", "v", "s", "v")     // RU(-a*b + c), uniform b
CMI_IV_FMA(ru_fma_nc_bs, "v_fma_f64 %0, %1, %2, -%3", "v", "s", "v")     // RU(a*b - c), uniform b

#undef CMI_IV_FMA

#define CMI_IV_UN(name, text)                                                    \
  __device__ __forceinline__ double name(double a, int t)                        \
  {                                                                              \
    double r;                                                                    \
    asm(text " ; %2" : "=v"(r) : "v"(a), "s"(t));                                \
    return r;                                                                    \
  }

CMI_IV_UN(ru_rcp, "v_rcp_f64 %0, %1")                  // ~1/a (an approximation)
CMI_IV_UN(ru_rsq, "v_rsq_f64 %0, %1")                  // ~1/sqrt(a) (an approximation)
CMI_IV_UN(ru_minus_one, "v_add_f64 %0, %1, -1.0")      // RU(a - 1)
CMI_IV_UN(ru_plus_one, "v_add_f64 %0, %1, 1.0")        // RU(a + 1)
CMI_IV_UN(ru_neg_minus_one, "v_add_f64 %0, -%1, -1.0")  // RU(-a - 1)
CMI_IV_UN(ru_plus_two, "v_add_f64 %0, %1, 2.0")        // RU(a + 2)
CMI_IV_UN(ru_neg_minus_two, "v_add_f64 %0, -%1, -2.0")  // RU(-a - 2)
CMI_IV_UN(ru_twice, "v_add_f64 %0, %1, %1")            // 2a (exact below overflow)

#undef CMI_IV_UN

__device__ __forceinline__ double ru_add_na_s(double a, double c, int t)  // RU(-a + c), uniform c
{
  double r;
  asm("v_add_f64 %0, -%1, %2 ; %3" : "=v"(r) : "v"(a), "s"(c), "s"(t));
  return r;
}
__device__ __forceinline__ double ru_fma2(double a, double c, int t)  // RU(2a + c)
{
  double r;
  asm("v_fma_f64 %0, %1, 2.0, %2 ; %3" : "=v"(r) : "v"(a), "v"(c), "s"(t));
  return r;
}

/// Exact conversions between a double holding an integer and int.
__device__ __forceinline__ int to_int(double a, int t)
{
  int r;
  asm("v_cvt_i32_f64 %0, %1 ; %2" : "=v"(r) : "v"(a), "s"(t));
  return r;
}
__device__ __forceinline__ double to_double(int a, int t)
{
  double r;
  asm("v_cvt_f64_i32 %0, %1 ; %2" : "=v"(r) : "v"(a), "s"(t));
  return r;
}

/// 2^k for -1074 <= k <= 1023, built from its bits.
__device__ __forceinline__ double pow2i(int k)
{
  unsigned long long const bits =
      k >= -1022 ? static_cast<unsigned long long>(k + 1023) << 52 : 1ull << static_cast<unsigned>(k + 1074);
  return __builtin_bit_cast(double, bits);
}

/// RU(x * 2^k) for 1/4 <= |x| <= 4 and any k: the first factor keeps the
/// product normal (exact), only the second can round.
__device__ __forceinline__ double ru_scale(double x, int k, int t)
{
  k = k < -2100 ? -2100 : (k > 1030 ? 1030 : k);
  int const k1 = k < -1018 ? -1018 : (k > 1018 ? 1018 : k);
  int k2 = k - k1;
  k2 = k2 < -1074 ? -1074 : k2;
  return ru_mul(ru_mul(x, pow2i(k1), t), pow2i(k2), t);
}

__device__ __forceinline__ double ru_nextup(double q, int t) { return ru_add_s(q, kTinyD, t); }
__device__ __forceinline__ double ru_nextdown(double q, int t) { return -ru_add_na_s(q, kTinyD, t); }

/// RU(a / b) for finite a and finite b != 0, correctly rounded.
__device__ __forceinline__ double ru_div(double a, double b, int t)
{
  if (a == 0.0) {
    return ru_mul(a, b, t);  // the zero quotient, with its sign
  }
  double const ma = __builtin_amdgcn_frexp_mant(a);
  int const ea = __builtin_amdgcn_frexp_exp(a);
  double const mb = __builtin_amdgcn_frexp_mant(b);
  int const eb = __builtin_amdgcn_frexp_exp(b);
  double y = ru_rcp(mb, t);
  y = ru_fma(y, ru_fma_na_s(mb, y, 1.0, t), y, t);
  y = ru_fma(y, ru_fma_na_s(mb, y, 1.0, t), y, t);
  double q = ru_mul(ma, y, t);
  q = ru_fma(ru_fma_na(mb, q, ma, t), y, q, t);
  // a/b > q exactly when the residual ma - mb*q is nonzero with the sign of mb
  bool const bpos = mb > 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double const r = ru_fma_na(mb, q, ma, t);
    q = (r != 0.0 && (r > 0.0) == bpos) ? ru_nextup(q, t) : q;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double const qd = ru_nextdown(q, t);
    double const r = ru_fma_na(mb, qd, ma, t);
    q = (r != 0.0 && (r > 0.0) == bpos) ? q : qd;
  }
  return ru_scale(q, ea - eb, t);
}

/// RU(sqrt(v)) for finite v > 0, correctly rounded; `down` receives
/// RD(sqrt(v)).
__device__ __forceinline__ double ru_sqrt(double v, int t, double& down)
{
  double m = __builtin_amdgcn_frexp_mant(v);
  int e = __builtin_amdgcn_frexp_exp(v);
  if (e & 1) {
    m = ru_twice(m, t);
    e -= 1;
  }
  // m in [1/2, 2), v = m * 2^e with e even
  double const y = ru_rsq(m, t);
  double g = ru_mul(m, y, t);
  double h = ru_mul_s(y, 0.5, t);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double const r = ru_fma_na_s(g, h, 0.5, t);
    g = ru_fma(g, r, g, t);
    h = ru_fma(h, r, h, t);
  }
  g = ru_fma(ru_fma_na(g, g, m, t), h, g, t);
  // sqrt(m) > g exactly when m - g*g > 0
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    g = ru_fma_na(g, g, m, t) > 0.0 ? ru_nextup(g, t) : g;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double const gd = ru_nextdown(g, t);
    g = ru_fma_na(gd, gd, m, t) > 0.0 ? g : gd;
  }
  double const gdn = ru_fma_na(g, g, m, t) == 0.0 ? g : ru_nextdown(g, t);
  double const s = pow2i(e / 2);
  down = ru_mul(gdn, s, t);
  return ru_mul(g, s, t);
}

/// RU(n / d) for d > 0 (finite or +inf) and a bound n (+inf allowed).
__device__ __forceinline__ double ru_div_bound(double n, double d, int t)
{
  if (n == kInfD || n == -kInfD) {
    return n;
  }
  if (d == kInfD) {
    return ru_mul(n, 0.0, t);  // a zero with the sign of n
  }
  return ru_div(n, d, t);
}

/// x / d for d = [dl, du], 0 < dl.
__device__ __forceinline__ Iv div_positive(Iv x, double dl, double du, int t)
{
  return {ru_div_bound(x.nl, x.nl >= 0.0 ? dl : du, t), ru_div_bound(x.ub, x.ub >= 0.0 ? dl : du, t)};
}

// [x] / [y], IEEE 1788: empty for y = [0, 0]; [0, 0] for x = [0, 0]; a
// half-line when 0 is an endpoint of y and x keeps one sign; else the line.
__device__ __forceinline__ Iv iv_div(Iv x, Iv y, int t, bool& ok)
{
  if (y.nl < 0.0) {  // y > 0
    return div_positive(x, -y.nl, y.ub, t);
  }
  if (y.ub < 0.0) {  // y < 0: x / y = -(x / -y)
    return iv_neg(div_positive(x, -y.ub, y.nl, t));
  }
  bool const y0l = y.nl == 0.0;
  bool const y0u = y.ub == 0.0;
  if (y0l && y0u) {
    ok = false;
    return {0.0, 0.0};
  }
  if (x.nl == 0.0 && x.ub == 0.0) {
    return {0.0, 0.0};
  }
  if (x.nl < 0.0) {  // x > 0
    if (y0l) {
      return {ru_div_bound(x.nl, y.ub, t), kInfD};
    }
    if (y0u) {
      return {kInfD, ru_div_bound(x.nl, y.nl, t)};
    }
  } else if (x.ub < 0.0) {  // x < 0
    if (y0l) {
      return {kInfD, ru_div_bound(x.ub, y.ub, t)};
    }
    if (y0u) {
      return {ru_div_bound(x.ub, y.nl, t), kInfD};
    }
  }
  return {kInfD, kInfD};
}

// [a] / c and c / [a] for a constant c (a point interval).
__device__ __forceinline__ Iv iv_div_imm(Iv a, double c, int t, bool& ok) { return iv_div(a, iv_const(c, t), t, ok); }
__device__ __forceinline__ Iv iv_imm_div(double c, Iv a, int t, bool& ok) { return iv_div(iv_const(c, t), a, t, ok); }

// sqrt[x] on x's part in [0, inf); empty if x < 0.
__device__ __forceinline__ Iv iv_sqrt(Iv x, int t, bool& ok)
{
  if (!(x.ub >= 0.0)) {
    ok = false;
    return {0.0, 0.0};
  }
  double dn = 0.0;
  double ub = x.ub;
  if (ub > 0.0 && ub < kInfD) {
    ub = ru_sqrt(ub, t, dn);
  }
  double nl = 0.0;
  double const xl = -x.nl;
  if (xl > 0.0) {
    double lo = 0.0;
    if (xl < kInfD) {
      (void)ru_sqrt(xl, t, lo);
    } else {
      lo = kInfD;
    }
    nl = -lo;
  }
  return {nl, ub};
}

__device__ __forceinline__ Iv iv_abs(Iv x)
{
  if (x.nl <= 0.0) {
    return x;
  }
  if (x.ub <= 0.0) {
    return iv_neg(x);
  }
  return {0.0, fmax(x.nl, x.ub)};
}

__device__ __forceinline__ Iv iv_min(Iv x, Iv y) { return {fmax(x.nl, y.nl), fmin(x.ub, y.ub)}; }
__device__ __forceinline__ Iv iv_max(Iv x, Iv y) { return {fmin(x.nl, y.nl), fmax(x.ub, y.ub)}; }

/// RU(s * m^n) for m >= 0, n >= 1 and s = -1 when `neg`, else +1: binary
/// powering from the low bit, every square and product rounded up in the
/// direction that keeps s * (its value) a bound.
__device__ __forceinline__ double ru_pow_signed(double m, unsigned n, bool neg, int t)
{
  double u = neg ? -m : m;
  double acc = 0.0;
  bool have = false;
  for (;;) {
    if (n & 1u) {
      acc = have ? ru_mul(neg ? -acc : acc, u, t) : u;
      have = true;
    }
    n >>= 1;
    if (n == 0u) {
      break;
    }
    u = ru_mul(neg ? -u : u, u, t);
  }
  return acc;
}

// [x]^n for an integer n (IEEE 1788 pown): [1, 1] for n = 0; 1 / [x]^-n for n < 0.
__device__ __forceinline__ Iv iv_pown(Iv x, int n, int t, bool& ok)
{
  if (n == 0) {
    return {-1.0, 1.0};
  }
  if (n == 1) {
    return x;
  }
  unsigned const m = static_cast<unsigned>(n < 0 ? -n : n);
  Iv r;
  if (m & 1u) {
    r = {ru_pow_signed(fabs(x.nl), m, x.nl < 0.0, t), ru_pow_signed(fabs(x.ub), m, x.ub < 0.0, t)};
  } else {
    double const s = fabs(vmax_na(x.nl, vmin0(x.ub, t), t));  // distance of [x] from 0
    r = {ru_pow_signed(s, m, true, t), ru_pow_signed(vmax_abs(x.nl, x.ub, t), m, false, t)};
  }
  if (n < 0) {
    return iv_imm_div(1.0, r, t, ok);
  }
  return r;
}

/// Upper bound of e^x.
__device__ __forceinline__ double exp_up(double x, int t)
{
  if (x != x) {
    return x;
  }
  if (x >= 709.8) {
    return kInfD;
  }
  x = x < -1100.0 ? -1100.0 : x;
  double const kd = __builtin_floor(ru_mul_s(x, kInvLn2, t));
  // x = kd ln2 + r; r <= RU(x - kd A) - kd B + |kd| err, 0 <~ r <~ ln2
  double r = ru_fma_na_bs(kd, kLn2Hi, x, t);
  r = ru_fma_na_bs(kd, kLn2Lo, r, t);
  r = ru_fma_bs(__builtin_fabs(kd), kLn2Err, r, t);
  r = r > 0.0 ? r : 0.0;  // e^r <= e^max(r, 0)
  double p = kExpUp[17];
#pragma unroll
  for (int j = 16; j >= 0; --j) {
    p = ru_fma_s(p, r, kExpUp[j], t);
  }
  p = ru_add_s(p, kExpTail, t);
  return ru_scale(p, to_int(kd, t), t);
}

/// -RD(e^x): the nl of a lower bound of e^x.
__device__ __forceinline__ double exp_dn_neg(double x, int t)
{
  if (x != x) {
    return x;
  }
  if (x <= -1100.0) {
    return -0.0;
  }
  x = x > 710.0 ? 710.0 : x;
  double const kd = __builtin_floor(ru_mul_s(x, kInvLn2, t));
  // -r >= RU(kd A - x) + kd B + |kd| err
  double nr = ru_fma_nc_bs(kd, kLn2Hi, x, t);
  nr = ru_fma_bs(kd, kLn2Lo, nr, t);
  nr = ru_fma_bs(__builtin_fabs(kd), kLn2Err, nr, t);
  // r >= 0: Horner rounded down, held negated; r < 0 (by a few ulps): e^r >= 1 + r
  double const nrc = nr < 0.0 ? nr : 0.0;
  double m = -kExpDn[17];
#pragma unroll
  for (int j = 16; j >= 0; --j) {
    m = ru_fma_na_nc_s(m, nrc, kExpDn[j], t);  // -(p*r + c) with r = -nrc
  }
  if (nr > 0.0) {
    m = ru_minus_one(nr, t);  // -(1 + r)
  }
  return ru_scale(m, to_int(kd, t), t);
}

__device__ __forceinline__ Iv iv_exp(Iv x, int t) { return {exp_dn_neg(-x.nl, t), exp_up(x.ub, t)}; }

/// RU(u * P(u^2)) with P(w) = sum w^j/(2j+1) (atanh(u) = u P(u^2)), P bounded
/// from above (with its tail) for u >= 0 and from below for u < 0: an upper
/// bound of atanh(u) for |u| <= 0.1716.
__device__ __forceinline__ double atanh_up(double u, int t)
{
  double const wu = ru_mul(u, u, t);
  double const wd = -ru_mul_na(u, u, t);
  double p = kAtanhUp[11];
  double m = -kAtanhDn[11];
#pragma unroll
  for (int j = 10; j >= 0; --j) {
    p = ru_fma_s(p, wu, kAtanhUp[j], t);
    m = ru_fma_nc_s(m, wd, kAtanhDn[j], t);
  }
  p = ru_add_s(p, kAtanhTail, t);
  return u >= 0.0 ? ru_mul(u, p, t) : ru_mul_na(u, m, t);
}

/// v = m3 * 2^e3 with m3 in [1/sqrt2, sqrt2): returns m3 - 1 (exact), m3 and e3.
__device__ __forceinline__ double log_reduce(double v, int t, double& m3, int& e3)
{
  double m = __builtin_amdgcn_frexp_mant(v);
  int e = __builtin_amdgcn_frexp_exp(v);
  if (m < 0.70710678118654752) {
    m = ru_twice(m, t);
    e -= 1;
  }
  m3 = m;
  e3 = e;
  return ru_minus_one(m, t);
}

/// Upper bound of log(v) for finite v > 0.
__device__ __forceinline__ double log_up(double v, int t)
{
  double m3;
  int e3;
  double const num = log_reduce(v, t, m3, e3);
  // s = num / (m3 + 1) bounded above
  double const den = num >= 0.0 ? -ru_neg_minus_one(m3, t) : ru_plus_one(m3, t);
  double const a = atanh_up(ru_div(num, den, t), t);
  double const e = to_double(e3, t);
  double const el = e >= 0.0 ? ru_mul_s(e, kLn2Up, t) : ru_mul_s(e, kLn2Dn, t);
  return ru_fma2(a, el, t);
}

/// -RD(log v) for finite v > 0.
__device__ __forceinline__ double log_dn_neg(double v, int t)
{
  double m3;
  int e3;
  double const num = log_reduce(v, t, m3, e3);
  // -s_lo = -num / (m3 + 1) bounded above
  double const den = num >= 0.0 ? ru_plus_one(m3, t) : -ru_neg_minus_one(m3, t);
  double const na = atanh_up(ru_div(-num, den, t), t);  // atanh(-s) = -atanh(s)
  double const e = to_double(e3, t);
  double const el = e >= 0.0 ? ru_mul_na_s(e, kLn2Dn, t) : ru_mul_na_s(e, kLn2Up, t);
  return ru_fma2(na, el, t);
}

// log[x] on x's part in (0, inf); empty if x <= 0.
__device__ __forceinline__ Iv iv_log(Iv x, int t, bool& ok)
{
  if (!(x.ub > 0.0)) {
    ok = false;
    return {0.0, 0.0};
  }
  double const ub = x.ub == kInfD ? kInfD : log_up(x.ub, t);
  double const xl = -x.nl;
  double const nl = xl > 0.0 ? log_dn_neg(xl, t) : kInfD;
  return {nl, ub};
}

/// sin(x + shift*pi/2) at an endpoint: its enclosure (-lo, hi) and where x
/// lies among the multiples of pi/2 (x = k*pi/2 + r, r in [r_lo, r_hi]).
struct TrigEnd
{
  double nlo, hi;
  int k;
  bool r_lo_nonpos, r_hi_nonneg;
};

__device__ __forceinline__ TrigEnd trig_end(double x, int shift, int t)
{
  double const kd = __builtin_rint(ru_mul_s(x, kTwoOverPi, t));
  double const ak = __builtin_fabs(kd);
  double const t0 = ru_fma_na_bs(kd, kPio2A, x, t);  // exact
  double rh = ru_fma_na_bs(kd, kPio2B, t0, t);
  rh = ru_fma_na_bs(kd, kPio2C, rh, t);
  rh = ru_fma_bs(ak, kPio2Err, rh, t);
  double rn = ru_fma_nc_bs(kd, kPio2B, t0, t);  // -r_lo
  rn = ru_fma_bs(kd, kPio2C, rn, t);
  rn = ru_fma_bs(ak, kPio2Err, rn, t);
  // w = r^2 in [wl, wh]
  double const wh = ru_mul(vmax_abs(rn, rh, t), vmax_abs(rn, rh, t), t);
  double const s0 = fabs(vmax_na(rn, vmin0(rh, t), t));
  double const wl = -ru_mul_na(s0, s0, t);
  // S(w) = sin(r)/r and C(w) = cos(r): partial sums alternate in sign with
  // the coefficients, so each step reads the w bound that keeps a bound
  double sn = -kSinDn[9], sh = kSinUp[9];
#pragma unroll
  for (int j = 8; j >= 0; --j) {
    bool const next_pos = ((j + 1) & 1) == 0;
    sn = ru_fma_nc_s(next_pos ? wl : wh, sn, kSinDn[j], t);
    sh = ru_fma_s(next_pos ? wh : wl, sh, kSinUp[j], t);
  }
  double cn = -kCosDn[10], ch = kCosUp[10];
#pragma unroll
  for (int j = 9; j >= 0; --j) {
    bool const next_pos = ((j + 1) & 1) == 0;
    cn = ru_fma_nc_s(next_pos ? wl : wh, cn, kCosDn[j], t);
    ch = ru_fma_s(next_pos ? wh : wl, ch, kCosUp[j], t);
  }
  sn = ru_add_s(sn, kSinTail, t);
  sh = ru_add_s(sh, kSinTail, t);
  cn = ru_add_s(cn, kCosTail, t);
  ch = ru_add_s(ch, kCosTail, t);
  // sin(r) = r * S, S > 0
  double const sin_hi = rh >= 0.0 ? ru_mul(rh, sh, t) : ru_mul_na(rh, sn, t);
  double const sin_nlo = rn <= 0.0 ? ru_mul_na(rn, sn, t) : ru_mul(rn, sh, t);
  int const k = to_int(kd, t);
  int const q = (k + shift) & 3;
  TrigEnd out;
  out.nlo = q == 0 ? sin_nlo : (q == 1 ? cn : (q == 2 ? sin_hi : ch));
  out.hi = q == 0 ? sin_hi : (q == 1 ? ch : (q == 2 ? sin_nlo : cn));
  out.k = k + shift;
  out.r_lo_nonpos = rn >= 0.0;
  out.r_hi_nonneg = rh >= 0.0;
  return out;
}

// sin[x] (shift 0) or cos[x] = sin[x + pi/2] (shift 1). Whole periods,
// non-finite bounds and |x| > 2^20 give [-1, 1].
__device__ __forceinline__ Iv iv_sincos(Iv x, int shift, int t)
{
  double const a = -x.nl;
  double const b = x.ub;
  if (!(__builtin_fabs(a) <= 1048576.0) || !(__builtin_fabs(b) <= 1048576.0) || ru_add(b, x.nl, t) >= 6.28) {
    return {1.0, 1.0};
  }
  TrigEnd const ea = trig_end(a, shift, t);
  TrigEnd const eb = trig_end(b, shift, t);
  // quadrant points K*pi/2 (shifted) inside [a, b], counted conservatively
  int const lo_k = ea.r_lo_nonpos ? ea.k : ea.k + 1;
  int const hi_k = eb.r_hi_nonneg ? eb.k : eb.k - 1;
  int const first_max = lo_k + (((1 - lo_k) % 4) + 4) % 4;  // K = 1 mod 4: +1
  int const first_min = lo_k + (((3 - lo_k) % 4) + 4) % 4;  // K = 3 mod 4: -1
  double hi = first_max <= hi_k ? 1.0 : fmax(ea.hi, eb.hi);
  double nl = first_min <= hi_k ? 1.0 : fmax(ea.nlo, eb.nlo);
  return {fmin(nl, 1.0), fmin(hi, 1.0)};
}

/// Upper bound of tanh(u) for 0 <= u <= 40.
__device__ __forceinline__ double tanh_pos_up(double u, int t)
{
  double const v = ru_twice(u, t);
  double mh;  // upper bound of e^v - 1
  if (v <= 1.0) {
    double p = kExpm1Up[18];
#pragma unroll
    for (int j = 17; j >= 0; --j) {
      p = ru_fma_s(p, v, kExpm1Up[j], t);
    }
    mh = ru_mul(v, ru_add_s(p, kExpm1Tail, t), t);
  } else {
    mh = ru_minus_one(exp_up(v, t), t);
  }
  // M / (M + 2) grows with M
  double const r = ru_div(mh, -ru_neg_minus_two(mh, t), t);
  return r < 1.0 ? r : 1.0;
}

/// -RD(tanh(u)) for 0 <= u <= 40.
__device__ __forceinline__ double tanh_pos_dn_neg(double u, int t)
{
  double const v = ru_twice(u, t);
  double ml;  // lower bound of e^v - 1
  if (v <= 1.0) {
    double m = -kExpm1Dn[18];
#pragma unroll
    for (int j = 17; j >= 0; --j) {
      m = ru_fma_nc_s(m, v, kExpm1Dn[j], t);
    }
    ml = -ru_mul_na(v, -m, t);
  } else {
    ml = -ru_add_na_s(-exp_dn_neg(v, t), 1.0, t);  // RD(e^v - 1)
  }
  return ru_div(-ml, ru_plus_two(ml, t), t);
}

/// Upper bound of tanh(x).
__device__ __forceinline__ double tanh_up(double x, int t)
{
  if (x != x) {
    return x;
  }
  if (x >= 0.0) {
    return tanh_pos_up(x < 40.0 ? x : 40.0, t);
  }
  double const u = -x;
  return tanh_pos_dn_neg(u < 40.0 ? u : 40.0, t);
}

// tanh is odd: -RD(tanh(xl)) = RU(tanh(-xl)).
__device__ __forceinline__ Iv iv_tanh(Iv x, int t) { return {tanh_up(x.nl, t), tanh_up(x.ub, t)}; }

// ---------------------------------------------------------------------------
// Round-to-nearest point versions (outside the region), same reductions.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double pt_scale(double x, int k)
{
  k = k < -2100 ? -2100 : (k > 1030 ? 1030 : k);
  int const k1 = k < -1018 ? -1018 : (k > 1018 ? 1018 : k);
  int k2 = k - k1;
  k2 = k2 < -1074 ? -1074 : k2;
  return (x * pow2i(k1)) * pow2i(k2);
}

__device__ __forceinline__ double pt_exp(double x)
{
  if (x != x) {
    return x;
  }
  if (x >= 709.8) {
    return kInfD;
  }
  if (x <= -746.0) {
    return 0.0;
  }
  double const kd = __builtin_rint(x * kInvLn2);
  double r = __builtin_fma(-kd, kLn2Hi, x);
  r = __builtin_fma(-kd, kLn2Lo, r);
  double p = kExpRn[13];
#pragma unroll
  for (int j = 12; j >= 0; --j) {
    p = __builtin_fma(p, r, kExpRn[j]);
  }
  return pt_scale(p, static_cast<int>(kd));
}

__device__ __forceinline__ double pt_log(double x)
{
  if (x != x || x == kInfD) {
    return x;
  }
  if (x < 0.0) {
    return __builtin_nan("");
  }
  if (x == 0.0) {
    return -kInfD;
  }
  double m = __builtin_amdgcn_frexp_mant(x);
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.70710678118654752) {
    m = m + m;
    e -= 1;
  }
  double const s = (m - 1.0) / (m + 1.0);
  double const w = s * s;
  double p = kAtanhRn[11];
#pragma unroll
  for (int j = 10; j >= 0; --j) {
    p = __builtin_fma(p, w, kAtanhRn[j]);
  }
  double const ed = static_cast<double>(e);
  double r = (s + s) * p;
  r = __builtin_fma(ed, kLn2Lo, r);
  return __builtin_fma(ed, kLn2Hi, r);
}

__device__ __forceinline__ double pt_sincos(double x, int shift)
{
  if (!(__builtin_fabs(x) <= 1125899906842624.0)) {  // 2^50
    return __builtin_nan("");
  }
  double const kd = __builtin_rint(x * kTwoOverPi);
  double r = __builtin_fma(-kd, kPio2A, x);
  r = __builtin_fma(-kd, kPio2B, r);
  r = __builtin_fma(-kd, kPio2C, r);
  double const w = r * r;
  double s = kSinRn[9];
#pragma unroll
  for (int j = 8; j >= 0; --j) {
    s = __builtin_fma(s, w, kSinRn[j]);
  }
  double c = kCosRn[10];
#pragma unroll
  for (int j = 9; j >= 0; --j) {
    c = __builtin_fma(c, w, kCosRn[j]);
  }
  s = r * s;
  int const q = static_cast<int>(static_cast<long long>(kd) & 3) + shift;
  switch (q & 3) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
  }
}

__device__ __forceinline__ double pt_tanh(double x)
{
  if (x != x) {
    return x;
  }
  double const ax = __builtin_fabs(x);
  double r;
  if (ax > 40.0) {
    r = 1.0;
  } else {
    double const v = ax + ax;
    double mm;
    if (v <= 1.0) {
      double p = kExpm1Rn[18];
#pragma unroll
      for (int j = 17; j >= 0; --j) {
        p = __builtin_fma(p, v, kExpm1Rn[j]);
      }
      mm = v * p;
    } else {
      mm = pt_exp(v) - 1.0;
    }
    r = mm / (mm + 2.0);
  }
  return x < 0.0 ? -r : r;
}

__device__ __forceinline__ double pt_pown(double x, int n)
{
  if (n == 0) {
    return 1.0;
  }
  unsigned k = static_cast<unsigned>(n < 0 ? -n : n);
  double u = x;
  double acc = 1.0;
  bool have = false;
  for (;;) {
    if (k & 1u) {
      acc = have ? acc * u : u;
      have = true;
    }
    k >>= 1;
    if (k == 0u) {
      break;
    }
    u = u * u;
  }
  return n < 0 ? 1.0 / acc : acc;
}

}  // namespace cmi_iv

#endif
