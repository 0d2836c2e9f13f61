// PseudoQuad: double-double arithmetic shared by host and device code.
//
// The value is s + t with |t| <= ulp(s)/2 after renormalisation.  Every
// operation is an error-free transformation built on IEEE round-to-nearest
// add/mul and a fused multiply-add, written out explicitly; floating-point
// contraction is disabled inside each function so the host and the device
// execute the identical operation sequence and produce identical bits.
//
// Operations and their reference counterparts (PseudoQuad.cuh):
//   fast2sum   Dekker's Fast2Sum (|a| >= |b|)        PseudoQuad.cuh:10-15
//   two_sum    Knuth's TwoSum ("Gen2Sum")            PseudoQuad.cuh:17-25
//   two_prod   FMA TwoProd ("FMA2Mult")              PseudoQuad.cuh:27-31
//   dd_add     dd + dd, accurate ("book") version    PseudoQuad.cuh:66-75
//   dd_scale   dd * d ("scale2")                     PseudoQuad.cuh:80-86
//   dd_axpy    dd += d * d ("axpy2", cheap)          PseudoQuad.cuh:96-108
//   dd_fma     dd += d * d ("slow_axpy2", accurate)  PseudoQuad.cuh:115-118
//
// dd_axpy is not error-free when the sum cancels: its step cp = r.s - abp is
// inexact once the product a*b is not itself a double, which loses about
// 2^-53 of |c.s|.  The accumulations in this library therefore use dd_fma,
// TwoProd followed by the accurate dd + dd.
//   dd_renorm  renormalize                           PseudoQuad.cuh:137-140
//   dd_to_d    convert to double (s + t)             PseudoQuad.cuh:149-152
#pragma once

#if defined(__HIPCC__) || defined(__HIP__)
#define SKINNY_HD __host__ __device__ inline
#else
#define SKINNY_HD inline
#endif

#include <cmath>

// Contraction off inside the EFTs.  Under clang (hipcc) this is a scoped
// pragma; a gcc build of host code must pass -ffp-contract=off instead.
#if defined(__clang__)
#define SKINNY_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define SKINNY_NOCONTRACT
#endif

struct PseudoQuad {
  double s;
  double t;
};

namespace skinny {

SKINNY_HD double fma_rn(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}

SKINNY_HD PseudoQuad fast2sum(double a, double b) {
SKINNY_NOCONTRACT
  double s = a + b;
  double z = s - a;
  double t = b - z;
  return {s, t};
}

SKINNY_HD PseudoQuad two_sum(double a, double b) {
SKINNY_NOCONTRACT
  double s = a + b;
  double ap = s - b;
  double bp = s - ap;
  double da = a - ap;
  double db = b - bp;
  return {s, da + db};
}

SKINNY_HD PseudoQuad two_prod(double a, double b) {
SKINNY_NOCONTRACT
  double s = a * b;
  return {s, fma_rn(a, b, -s)};
}

SKINNY_HD PseudoQuad dd_add(PseudoQuad x, PseudoQuad y) {
SKINNY_NOCONTRACT
  PseudoQuad S = two_sum(x.s, y.s);
  PseudoQuad T = two_sum(x.t, y.t);
  double c = S.t + T.s;
  PseudoQuad V = fast2sum(S.s, c);
  double w = T.t + V.t;
  return fast2sum(V.s, w);
}

SKINNY_HD PseudoQuad dd_scale(PseudoQuad a, double b) {
SKINNY_NOCONTRACT
  PseudoQuad p = two_prod(a.s, b);
  p.t = fma_rn(a.s, 0.0, p.t);
  p.t = fma_rn(a.t, b, p.t);
  return fast2sum(p.s, p.t);
}

SKINNY_HD PseudoQuad dd_axpy(PseudoQuad c, double a, double b) {
SKINNY_NOCONTRACT
  PseudoQuad r;
  r.s = fma_rn(a, b, c.s);
  double abp = r.s - c.s;
  double cp = r.s - abp;
  double dab = fma_rn(a, b, -abp);
  double dc = c.s - cp;
  r.t = dc + dab;
  r.t = r.t + c.t;
  return r;
}

SKINNY_HD PseudoQuad dd_fma(PseudoQuad c, double a, double b) {
  return dd_add(c, two_prod(a, b));
}

// Double-double plus a third word u collecting the rounding remainders of
// dd_add's two inexact additions (c and w): v is bitwise dd_add's result and
// v + u carries the sum to the remainders' precision.  Combining partial sums
// of very different magnitude (one near 2^124 and one near 2^62) puts the
// smaller one into t, whose ulp rounds the small terms away before the large
// ones cancel; u keeps them.  HR uses it to combine partial sums.
struct DdAcc {
  PseudoQuad v;
  double u;
};

SKINNY_HD DdAcc dda_add(DdAcc x, PseudoQuad y) {
SKINNY_NOCONTRACT
  PseudoQuad S = two_sum(x.v.s, y.s);
  PseudoQuad T = two_sum(x.v.t, y.t);
  PseudoQuad Cc = two_sum(S.t, T.s);  // c of dd_add and its remainder
  PseudoQuad V = fast2sum(S.s, Cc.s);
  PseudoQuad W = two_sum(T.t, V.t);   // w of dd_add and its remainder
  DdAcc r;
  r.v = fast2sum(V.s, W.s);
  r.u = x.u + (Cc.t + W.t);
  return r;
}

SKINNY_HD DdAcc dda_add(DdAcc x, DdAcc y) {
SKINNY_NOCONTRACT
  DdAcc r = dda_add(x, y.v);
  r.u = r.u + y.u;
  return r;
}

SKINNY_HD PseudoQuad dda_value(DdAcc x) {
  return dd_add(x.v, PseudoQuad{x.u, 0.0});
}

SKINNY_HD PseudoQuad dd_renorm(PseudoQuad a) { return fast2sum(a.s, a.t); }

SKINNY_HD double dd_to_d(PseudoQuad a) {
SKINNY_NOCONTRACT
  return a.s + a.t;
}

}  // namespace skinny
