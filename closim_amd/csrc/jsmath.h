// jsmath.h — the JavaScript Math routines the cloth step uses, as host+device
// code with bit-reproducible results.
//
// The reference runs on V8, whose Math.log/acos/cbrt/cos are ports of the
// fdlibm / FreeBSD msun algorithms and whose Math.hypot is a max-scaled Kahan
// sum.  These are restated here from those published algorithms (fdlibm 5.3
// e_log.c, e_acos.c, k_cos.c, k_sin.c, e_rem_pio2.c medium-range reduction,
// s_cos.c; FreeBSD s_cbrt.c), using only IEEE-754 +, -, *, / and sqrt, which
// the GPU rounds exactly like the CPU.  Nothing here depends on the device
// math library, so device and host produce identical bits.
//
// Also here: double-double accumulation used for every array-wide reduction
// (dot products, norms, energy totals).  The accumulated pair is rounded once,
// which yields the correctly rounded sum independent of summation order
// (barring a tie within ~2^-100 relative), so a tree reduction on the GPU and
// a sequential loop on the CPU agree bit for bit.
//
// The log, acos, cos (kernel_cos, kernel_sin, rem_pio2) and cbrt routines and
// their coefficient tables are restated from fdlibm 5.3 and FreeBSD msun
// (s_cbrt.c), which carry this notice:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
//
//   (s_cbrt.c: optimized by Bruce D. Evans.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef JSM_HD
#define JSM_HD __host__ __device__ __forceinline__
#endif

namespace jsm {

JSM_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
JSM_HD double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
JSM_HD int32_t hi(double x) { return (int32_t)(bits(x) >> 32); }
JSM_HD uint32_t lo(double x) { return (uint32_t)bits(x); }
JSM_HD double with_hi(double x, int32_t h) {
  return from_bits(((uint64_t)(uint32_t)h << 32) | (bits(x) & 0xffffffffull));
}
JSM_HD double with_lo(double x, uint32_t l) { return from_bits((bits(x) & 0xffffffff00000000ull) | l); }
JSM_HD double make(int32_t h, uint32_t l) { return from_bits(((uint64_t)(uint32_t)h << 32) | l); }
JSM_HD double fabs_(double x) { return from_bits(bits(x) & 0x7fffffffffffffffull); }
JSM_HD bool isnan_(double x) { return x != x; }
JSM_HD bool isfinite_(double x) { return (bits(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull; }

// JS Math.max / Math.min (NaN propagates, +0 > -0).
JSM_HD double jmax(double a, double b) {
  if (isnan_(a) || isnan_(b)) return from_bits(0x7ff8000000000000ull);
  if (a == 0 && b == 0) return (bits(a) >> 63) ? b : a;
  return a > b ? a : b;
}
JSM_HD double jmin(double a, double b) {
  if (isnan_(a) || isnan_(b)) return from_bits(0x7ff8000000000000ull);
  if (a == 0 && b == 0) return (bits(a) >> 63) ? a : b;
  return a < b ? a : b;
}

// IEEE sqrt (correctly rounded on both sides).
JSM_HD double sqrt_(double x) { return __builtin_sqrt(x); }

// V8 Math.hypot(a, b, c).
JSM_HD double hypot3(double a, double b, double c) {
  const double v[3] = {a, b, c};
  double ab[3] = {0, 0, 0};
  bool nan = false;
  double mx = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (isnan_(v[i])) {
      nan = true;
    } else {
      ab[i] = fabs_(v[i]);
      if (ab[i] > mx) mx = ab[i];
    }
  }
  if (mx == __builtin_inf()) return __builtin_inf();
  if (nan) return from_bits(0x7ff8000000000000ull);
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double n = ab[i] / mx;
    const double summand = n * n - comp;
    const double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return sqrt_(sum) * mx;
}

// fdlibm e_log.c
JSM_HD double log(double x) {
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10,
               two54 = 1.80143985094819840000e+16, Lg1 = 6.666666666666735130e-01,
               Lg2 = 3.999999999940941908e-01, Lg3 = 2.857142874366239149e-01,
               Lg4 = 2.222219843214978396e-01, Lg5 = 1.818357216161805012e-01,
               Lg6 = 1.531383769920937332e-01, Lg7 = 1.479819860511658591e-01;
  double hfsq, f, s, z, R, w, t1, t2, dk;
  int32_t k, hx, i, j;
  uint32_t lx;
  hx = hi(x);
  lx = lo(x);
  k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | lx) == 0) return -__builtin_inf();
    if (hx < 0) return from_bits(0x7ff8000000000000ull);
    k -= 54;
    x *= two54;
    hx = hi(x);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  i = (hx + 0x95f64) & 0x100000;
  x = with_hi(x, hx | (i ^ 0x3ff00000));
  k += (i >> 20);
  f = x - 1.0;
  if ((0x000fffff & (2 + hx)) < 3) {
    if (f == 0.0) {
      if (k == 0) return 0.0;
      dk = (double)k;
      return dk * ln2_hi + dk * ln2_lo;
    }
    R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    dk = (double)k;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  s = f / (2.0 + f);
  dk = (double)k;
  z = s * s;
  i = hx - 0x6147a;
  w = z * z;
  j = 0x6b851 - hx;
  t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  R = t2 + t1;
  if (i > 0) {
    hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

// fdlibm e_acos.c
JSM_HD double acos(double x) {
  const double pi = 3.14159265358979311600e+00, pio2_hi = 1.57079632679489655800e+00,
               pio2_lo = 6.12323399573676603587e-17, pS0 = 1.66666666666666657415e-01,
               pS1 = -3.25565818622400915405e-01, pS2 = 2.01212532134862925881e-01,
               pS3 = -4.00555345006794114027e-02, pS4 = 7.91534994289814532176e-04,
               pS5 = 3.47933107596021167570e-05, qS1 = -2.40339491173441421878e+00,
               qS2 = 2.02094576023350569471e+00, qS3 = -6.88283971605453293030e-01,
               qS4 = 7.70381505559019352791e-02;
  double z, p, q, r, w, s, c, df;
  const int32_t hx = hi(x);
  const int32_t ix = hx & 0x7fffffff;
  if (ix >= 0x3ff00000) {
    if (((ix - 0x3ff00000) | (int32_t)lo(x)) == 0) {
      if (hx > 0) return 0.0;
      return pi + 2.0 * pio2_lo;
    }
    return from_bits(0x7ff8000000000000ull);
  }
  if (ix < 0x3fe00000) {
    if (ix <= 0x3c600000) return pio2_hi + pio2_lo;
    z = x * x;
    p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    r = p / q;
    return pio2_hi - (x - (pio2_lo - x * r));
  } else if (hx < 0) {
    z = (1.0 + x) * 0.5;
    p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    s = sqrt_(z);
    r = p / q;
    w = r * s - pio2_lo;
    return pi - 2.0 * (s + w);
  }
  z = (1.0 - x) * 0.5;
  s = sqrt_(z);
  df = with_lo(s, 0);
  c = (z - df * df) / (s + df);
  p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
  q = 1.0 + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
  r = p / q;
  w = r * s + c;
  return 2.0 * (df + w);
}

// FreeBSD msun s_cbrt.c
JSM_HD double cbrt(double x) {
  const uint32_t B1 = 715094163, B2 = 696219795;
  const double P0 = 1.87595182427177009643, P1 = -1.88497979543377169875, P2 = 1.621429720105354466140,
               P3 = -0.758397934778766047437, P4 = 0.145996192886612446982;
  int32_t hx = hi(x);
  const uint32_t low = lo(x);
  const uint32_t sign = (uint32_t)hx & 0x80000000u;
  hx ^= (int32_t)sign;
  if (hx >= 0x7ff00000) return x + x;
  double t = 0.0;
  if (hx < 0x00100000) {
    if ((hx | (int32_t)low) == 0) return x;
    t = make(0x43500000, 0);
    t *= x;
    const uint32_t high = (uint32_t)hi(t);
    t = make((int32_t)(sign | ((high & 0x7fffffffu) / 3 + B2)), 0);
  } else {
    t = make((int32_t)(sign | ((uint32_t)hx / 3 + B1)), 0);
  }
  double r = (t * t) * (t / x);
  t = t * ((P0 + r * (P1 + r * P2)) + ((r * r) * r) * (P3 + r * P4));
  t = from_bits((bits(t) + 0x80000000ull) & 0xffffffffc0000000ull);
  const double s = t * t;
  r = x / s;
  const double w = t + t;
  r = (r - t) / (w + r);
  t = t + t * r;
  return t;
}

// fdlibm k_cos.c
JSM_HD double kernel_cos(double x, double y) {
  const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
               C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
               C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  const int32_t ix = hi(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return 1.0;
  }
  const double z = x * x;
  const double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
  if (ix < 0x3FD33333) return 1.0 - (0.5 * z - (z * r - x * y));
  double qx;
  if (ix > 0x3fe90000) {
    qx = 0.28125;
  } else {
    qx = make(ix - 0x00200000, 0);
  }
  const double hz = 0.5 * z - qx;
  const double a = 1.0 - qx;
  return a - (hz - (z * r - x * y));
}

// fdlibm k_sin.c
JSM_HD double kernel_sin(double x, double y, int iy) {
  const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
               S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
               S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  const int32_t ix = hi(x) & 0x7fffffff;
  if (ix < 0x3e400000) {
    if ((int)x == 0) return x;
  }
  const double z = x * x;
  const double v = z * x;
  const double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
  if (iy == 0) return x + v * (S1 + z * r);
  return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

// fdlibm e_rem_pio2.c for |x| <= 2^19 * pi/2 (the cubic solver's arguments
// are below 2 pi); larger arguments return NaN rather than reducing.
JSM_HD int rem_pio2(double x, double* y) {
  const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
               pio2_1t = 6.07710050650619224932e-11, pio2_2 = 6.07710050630396597660e-11,
               pio2_2t = 2.02226624879595063154e-21, pio2_3 = 2.02226624871116645580e-21,
               pio2_3t = 8.47842766036889956997e-32;
  const int32_t hx = hi(x);
  const int32_t ix = hx & 0x7fffffff;
  if (ix <= 0x3fe921fb) {
    y[0] = x;
    y[1] = 0;
    return 0;
  }
  if (ix < 0x4002d97c) {
    double z;
    if (hx > 0) {
      z = x - pio2_1;
      if (ix != 0x3ff921fb) {
        y[0] = z - pio2_1t;
        y[1] = (z - y[0]) - pio2_1t;
      } else {
        z -= pio2_2;
        y[0] = z - pio2_2t;
        y[1] = (z - y[0]) - pio2_2t;
      }
      return 1;
    }
    z = x + pio2_1;
    if (ix != 0x3ff921fb) {
      y[0] = z + pio2_1t;
      y[1] = (z - y[0]) + pio2_1t;
    } else {
      z += pio2_2;
      y[0] = z + pio2_2t;
      y[1] = (z - y[0]) + pio2_2t;
    }
    return -1;
  }
  if (ix <= 0x413921fb) {
    double t = fabs_(x);
    const int n = (int)(t * invpio2 + 0.5);
    const double fn = (double)n;
    double r = t - fn * pio2_1;
    double w = fn * pio2_1t;
    const int32_t j = ix >> 20;
    y[0] = r - w;
    int32_t i = j - ((hi(y[0]) >> 20) & 0x7ff);
    if (i > 16) {
      t = r;
      w = fn * pio2_2;
      r = t - w;
      w = fn * pio2_2t - ((t - r) - w);
      y[0] = r - w;
      i = j - ((hi(y[0]) >> 20) & 0x7ff);
      if (i > 49) {
        t = r;
        w = fn * pio2_3;
        r = t - w;
        w = fn * pio2_3t - ((t - r) - w);
        y[0] = r - w;
      }
    }
    y[1] = (r - y[0]) - w;
    if (hx < 0) {
      y[0] = -y[0];
      y[1] = -y[1];
      return -n;
    }
    return n;
  }
  y[0] = y[1] = from_bits(0x7ff8000000000000ull);
  return 0;
}

// fdlibm s_cos.c
JSM_HD double cos(double x) {
  const int32_t ix = hi(x) & 0x7fffffff;
  if (ix <= 0x3fe921fb) return kernel_cos(x, 0.0);
  if (ix >= 0x7ff00000) return x - x;
  double y[2];
  const int n = rem_pio2(x, y);
  switch (n & 3) {
    case 0: return kernel_cos(y[0], y[1]);
    case 1: return -kernel_sin(y[0], y[1], 1);
    case 2: return -kernel_cos(y[0], y[1]);
    default: return kernel_sin(y[0], y[1], 1);
  }
}

// ---- double-double accumulation ------------------------------------------
struct dd {
  double hi, lo;
};

JSM_HD void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
JSM_HD void fast_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  e = b - (s - a);
}
// Exact product error via Veltkamp/Dekker splitting (identical to the FMA
// residual, which is unique).
JSM_HD void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
#if defined(__HIP_DEVICE_COMPILE__)
  e = __builtin_fma(a, b, -p);
#else
  const double split = 134217729.0;  // 2^27 + 1
  double t = split * a;
  const double ah = t - (t - a), al = a - ah;
  t = split * b;
  const double bh = t - (t - b), bl = b - bh;
  e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
#endif
}
JSM_HD dd dd_zero() { return dd{0.0, 0.0}; }
JSM_HD void dd_add(dd& acc, double x) {
  double s, e;
  two_sum(acc.hi, x, s, e);
  e += acc.lo;
  fast_two_sum(s, e, acc.hi, acc.lo);
}
JSM_HD void dd_add_prod(dd& acc, double a, double b) {
  double p, pe;
  two_prod(a, b, p, pe);
  double s, e;
  two_sum(acc.hi, p, s, e);
  e += acc.lo + pe;
  fast_two_sum(s, e, acc.hi, acc.lo);
}
JSM_HD dd dd_merge(dd a, dd b) {
  double s, e;
  two_sum(a.hi, b.hi, s, e);
  e += a.lo + b.lo;
  dd r;
  fast_two_sum(s, e, r.hi, r.lo);
  return r;
}
JSM_HD double dd_value(dd a) { return a.hi + a.lo; }

}  // namespace jsm
