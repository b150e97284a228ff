// FP64 exp, exp2 and sincos for the period-scan kernels (host and device).
//
// Method: Cody-Waite argument reduction (ln 2 split in two parts for exp,
// pi/2 split in three parts for sin/cos) followed by Horner evaluation of the
// Taylor series with fused multiply-adds; the constants are the correctly
// rounded 1/n!.  Every operation is an explicit IEEE operation, so the host
// and device compilations of this header give identical bits.  Accuracy is
// within 2 ulp of the correctly rounded result on the argument ranges the
// scan uses (checked by the test suite against the C library).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define PSM_HD __host__ __device__ __forceinline__

namespace psm {

// A polynomial or reduction constant.  In device code it is materialised in
// scalar registers at its use: left to the optimiser, the constants of the
// per-point sincos / exp2 are hoisted out of the kernel's loops into vector
// registers that then stay live (or spilled) for the whole fit.
PSM_HD double kc(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+s"(v));
#endif
  return v;
}

PSM_HD double exp_poly(double r) {
  // sum_{n=0}^{13} r^n / n!
  double p = kc(0x1.6124613a86d09p-33);
  p = fma(p, r, kc(0x1.1eed8eff8d898p-29));
  p = fma(p, r, kc(0x1.ae64567f544e4p-26));
  p = fma(p, r, kc(0x1.27e4fb7789f5cp-22));
  p = fma(p, r, kc(0x1.71de3a556c734p-19));
  p = fma(p, r, kc(0x1.a01a01a01a01ap-16));
  p = fma(p, r, kc(0x1.a01a01a01a01ap-13));
  p = fma(p, r, kc(0x1.6c16c16c16c17p-10));
  p = fma(p, r, kc(0x1.1111111111111p-7));
  p = fma(p, r, kc(0x1.5555555555555p-5));
  p = fma(p, r, kc(0x1.5555555555555p-3));
  p = fma(p, r, kc(0x1.0000000000000p-1));
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return p;
}

PSM_HD double exp(double x) {
  if (x != x) return x;
  if (x > 709.8) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  const double k = rint(x * 0x1.71547652b82fep+0);
  double r = fma(-k, 0x1.62e42fee00000p-1, x);
  r = fma(-k, 0x1.a39ef35793c76p-33, r);
  return ldexp(exp_poly(r), (int)k);
}

PSM_HD double exp2(double y) {
  if (y != y) return y;
  if (y > 1024.0) return HUGE_VAL;
  if (y < -1075.0) return 0.0;
  const double k = rint(y);
  const double r = y - k;
  return ldexp(exp_poly(r * 0x1.62e42fefa39efp-1), (int)k);
}

PSM_HD void sincos(double x, double* s, double* c) {
  if (!(x - x == 0.0)) {  // inf or nan
    *s = x - x;
    *c = x - x;
    return;
  }
  const double k = rint(x * 0x1.45f306dc9c883p-1);
  double r = fma(-k, kc(0x1.921fb54400000p+0), x);
  r = fma(-k, kc(0x1.0b4611a600000p-34), r);
  r = fma(-k, kc(0x1.3198a2e000000p-69), r);
  const double z = r * r;
  double ps = kc(-0x1.2f49b46814157p-57);
  ps = fma(ps, z, kc(0x1.952c77030ad4ap-49));
  ps = fma(ps, z, kc(-0x1.ae7f3e733b81fp-41));
  ps = fma(ps, z, kc(0x1.6124613a86d09p-33));
  ps = fma(ps, z, kc(-0x1.ae64567f544e4p-26));
  ps = fma(ps, z, kc(0x1.71de3a556c734p-19));
  ps = fma(ps, z, kc(-0x1.a01a01a01a01ap-13));
  ps = fma(ps, z, kc(0x1.1111111111111p-7));
  ps = fma(ps, z, kc(-0x1.5555555555555p-3));
  const double sn = fma(r * z, ps, r);
  double pc = kc(-0x1.6827863b97d97p-53);
  pc = fma(pc, z, kc(0x1.ae7f3e733b81fp-45));
  pc = fma(pc, z, kc(-0x1.93974a8c07c9dp-37));
  pc = fma(pc, z, kc(0x1.1eed8eff8d898p-29));
  pc = fma(pc, z, kc(-0x1.27e4fb7789f5cp-22));
  pc = fma(pc, z, kc(0x1.a01a01a01a01ap-16));
  pc = fma(pc, z, kc(-0x1.6c16c16c16c17p-10));
  pc = fma(pc, z, kc(0x1.5555555555555p-5));
  pc = fma(pc, z, kc(-0x1.0000000000000p-1));
  const double cs = fma(z, pc, 1.0);
  const long long q = (long long)k & 3;
  if (q == 0) { *s = sn; *c = cs; }
  else if (q == 1) { *s = cs; *c = -sn; }
  else if (q == 2) { *s = -sn; *c = -cs; }
  else { *s = -cs; *c = sn; }
}

PSM_HD double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return fma(a2, b2, fma(a1, b1, a0 * b0));
}

}  // namespace psm
