/* dpsnn_exp.h -- double exp that reproduces the host glibc (2.35, x86-64)
 * result bit for bit, for use inside device code.
 *
 * Method: the table-driven algorithm of ARM optimized-routines (math/exp.c,
 * Szabolcs Nagy, 2018), which glibc >= 2.28 uses for exp:
 *   x = k ln2/128 + r,  exp(x) = 2^(k/128) * exp(r),
 *   2^(k/128) from a 128-entry table (dpsnn_exp_table.inc, computed from the
 *   definition by gen_exp_table.py), exp(r) - 1 by a degree-5 polynomial.
 * glibc on x86-64 dispatches to a build compiled with FMA contraction; the
 * fused operations below are the ones that build contracts (established by
 * comparing against the host libm over 10^8 arguments, all ranges, zero
 * mismatches).  The reference calls exp once per refractory exit,
 * DPSNN_neuron_sim.c:705. */
#ifndef DPSNN_EXP_H
#define DPSNN_EXP_H
#include <stdint.h>

#if defined(__HIPCC__)
#define DPSNN_HD __host__ __device__
#else
#define DPSNN_HD
#include <math.h>
#include <string.h>
#endif

#if defined(__HIPCC__)
__device__ __constant__ static const uint64_t dpsnn_exp_tab_d[256] = {
#include "dpsnn_exp_table.inc"
};
#endif
static const uint64_t dpsnn_exp_tab_h[256] = {
#include "dpsnn_exp_table.inc"
};

DPSNN_HD static inline uint64_t dpsnn_asu64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__double_as_longlong(x);
#else
  uint64_t u; memcpy(&u, &x, 8); return u;
#endif
}
DPSNN_HD static inline double dpsnn_asf64(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  double x; memcpy(&x, &u, 8); return x;
#endif
}

DPSNN_HD static inline double dpsnn_exp_special(double tmp, uint64_t sbits, uint64_t ki) {
  double scale, y;
  if ((ki & 0x80000000u) == 0) {
    sbits -= 1009ull << 52;
    scale = dpsnn_asf64(sbits);
    return 0x1p1009 * fma(scale, tmp, scale);
  }
  sbits += 1022ull << 52;
  scale = dpsnn_asf64(sbits);
  y = scale + scale * tmp;
  if (y < 1.0) {
    double hi, lo;
    lo = scale - y + scale * tmp;
    hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

DPSNN_HD static inline double dpsnn_exp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t *T = dpsnn_exp_tab_d;
#else
  const uint64_t *T = dpsnn_exp_tab_h;
#endif
  const double InvLn2N = 0x1.71547652b82fep0 * 128, Shift = 0x1.8p52;
  const double NegLn2hiN = -0x1.62e42fefa0000p-8, NegLn2loN = -0x1.cf79abc9e3b3ap-47;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3;
  const double C4 = 0x1.55555cf172b91p-5, C5 = 0x1.1111167a4d017p-7;
  uint32_t abstop = (uint32_t)(dpsnn_asu64(x) >> 52) & 0x7ff;
  double kd, r, r2, tail, tmp;
  uint64_t ki, idx, top, sbits;
  if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
    if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;
    if (abstop >= 0x409u) {
      if (dpsnn_asu64(x) == 0xfff0000000000000ull) return 0.0;
      if (abstop >= 0x7ffu) return 1.0 + x;
      if (dpsnn_asu64(x) >> 63) return 0x1p-767 * 0x1p-767;
      return 0x1p769 * 0x1p769;
    }
    abstop = 0;
  }
  kd = fma(InvLn2N, x, Shift);
  ki = dpsnn_asu64(kd);
  kd -= Shift;
  r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
  idx = 2 * (ki % 128);
  top = ki << 45;
  tail = dpsnn_asf64(T[idx]);
  sbits = T[idx + 1] + top;
  r2 = r * r;
  tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
  if (abstop == 0) return dpsnn_exp_special(tmp, sbits, ki);
  return fma(dpsnn_asf64(sbits), tmp, dpsnn_asf64(sbits));
}
#endif
