/*
 * tmx_rng.h -- counter-based synthetic-input generator shared by host and device.
 *
 * Every value is a pure function of (seed, tensor_id, element index), where the
 * element index is the row-major position of the element in the tensor's full
 * dense index space (global orbital indices, not block-local ones).  Values are
 * therefore independent of tiling, block layout and device, and the host and
 * the GPU produce the same bits: only integer mixing, IEEE-exact scaling,
 * correctly rounded +,-,*,/,sqrt and explicit fused multiply-adds are used, and
 * floating-point contraction is disabled inside these functions.
 *
 * Normal deviates come from the Box-Muller transform.  ln() and cos() are
 * evaluated here with fixed polynomial schemes (atanh series for ln, Taylor
 * series on [0, pi/2) for sin/cos) so that no vendor math library enters the
 * result.
 */
#ifndef TMX_RNG_H
#define TMX_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define TMX_RNG_FN __host__ __device__ static inline
#else
#define TMX_RNG_FN static inline
#endif

#ifdef __cplusplus
#include <cmath>
#define TMX_RNG_FMA(a, b, c) std::fma((a), (b), (c))
#define TMX_RNG_SQRT(a) std::sqrt(a)
#else
#include <math.h>
#define TMX_RNG_FMA(a, b, c) fma((a), (b), (c))
#define TMX_RNG_SQRT(a) sqrt(a)
#endif

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
TMX_RNG_FN uint64_t tmx_rng_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* Stream key of one tensor. */
TMX_RNG_FN uint64_t tmx_rng_key(uint64_t seed, uint64_t tensor_id) {
  return tmx_rng_mix64(seed ^ tmx_rng_mix64(tensor_id * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
}

/* Uniform double in [0, 1) with 53 random bits; exact. */
TMX_RNG_FN double tmx_rng_uniform(uint64_t key, uint64_t counter) {
  uint64_t h = tmx_rng_mix64(key + tmx_rng_mix64(counter));
  return (double)(h >> 11) * 0x1.0p-53;
}

/* Natural log for u in (0, 1]: u = m * 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln(m) = 2*atanh(f), f = (m-1)/(m+1), |f| < 0.1716; 12 odd terms. */
TMX_RNG_FN double tmx_rng_ln(double u) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  int e = 0;
  /* manual frexp: scale u into [0.5, 1) by powers of two (exact) */
  double m = u;
  while (m < 0.5) { m *= 2.0; e -= 1; }
  if (m < 0.70710678118654752440) { m *= 2.0; e -= 1; }
  double f = (m - 1.0) / (m + 1.0);
  double f2 = f * f;
  /* sum_{k=0}^{11} f^(2k) / (2k+1) by Horner in f2 */
  double p = 1.0 / 23.0;
  p = TMX_RNG_FMA(p, f2, 1.0 / 21.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 19.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 17.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 15.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 13.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 11.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 9.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 7.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 5.0);
  p = TMX_RNG_FMA(p, f2, 1.0 / 3.0);
  p = TMX_RNG_FMA(p, f2, 1.0);
  double lnm = 2.0 * f * p;
  return TMX_RNG_FMA((double)e, 0.69314718055994530942, lnm);
}

/* cos(2*pi*t) for t in [0, 1). */
TMX_RNG_FN double tmx_rng_cos2pi(double t) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double t4 = 4.0 * t;                 /* exact */
  double qd = (double)(int)t4;         /* quadrant 0..3, exact */
  int q = (int)qd;
  double a = (t4 - qd) * 1.57079632679489661923; /* [0, pi/2) */
  double a2 = a * a;
  /* cos a = sum_{k=0}^{11} (-1)^k a^(2k)/(2k)!, Horner in a2 */
  double c = -1.0 / 1124000727777607680000.0;        /* -1/22! */
  c = TMX_RNG_FMA(c, a2, 1.0 / 2432902008176640000.0); /* 1/20! */
  c = TMX_RNG_FMA(c, a2, -1.0 / 6402373705728000.0);   /* -1/18! */
  c = TMX_RNG_FMA(c, a2, 1.0 / 20922789888000.0);      /* 1/16! */
  c = TMX_RNG_FMA(c, a2, -1.0 / 87178291200.0);        /* -1/14! */
  c = TMX_RNG_FMA(c, a2, 1.0 / 479001600.0);           /* 1/12! */
  c = TMX_RNG_FMA(c, a2, -1.0 / 3628800.0);            /* -1/10! */
  c = TMX_RNG_FMA(c, a2, 1.0 / 40320.0);               /* 1/8! */
  c = TMX_RNG_FMA(c, a2, -1.0 / 720.0);                /* -1/6! */
  c = TMX_RNG_FMA(c, a2, 1.0 / 24.0);                  /* 1/4! */
  c = TMX_RNG_FMA(c, a2, -0.5);                        /* -1/2! */
  double cosa = TMX_RNG_FMA(c, a2, 1.0);
  /* sin a = a * sum_{k=0}^{11} (-1)^k a^(2k)/(2k+1)! */
  double s = -1.0 / 25852016738884976640000.0;       /* -1/23! */
  s = TMX_RNG_FMA(s, a2, 1.0 / 51090942171709440000.0); /* 1/21! */
  s = TMX_RNG_FMA(s, a2, -1.0 / 121645100408832000.0);  /* -1/19! */
  s = TMX_RNG_FMA(s, a2, 1.0 / 355687428096000.0);      /* 1/17! */
  s = TMX_RNG_FMA(s, a2, -1.0 / 1307674368000.0);       /* -1/15! */
  s = TMX_RNG_FMA(s, a2, 1.0 / 6227020800.0);           /* 1/13! */
  s = TMX_RNG_FMA(s, a2, -1.0 / 39916800.0);            /* -1/11! */
  s = TMX_RNG_FMA(s, a2, 1.0 / 362880.0);               /* 1/9! */
  s = TMX_RNG_FMA(s, a2, -1.0 / 5040.0);                /* -1/7! */
  s = TMX_RNG_FMA(s, a2, 1.0 / 120.0);                  /* 1/5! */
  s = TMX_RNG_FMA(s, a2, -1.0 / 6.0);                   /* -1/3! */
  s = TMX_RNG_FMA(s, a2, 1.0);
  double sina = a * s;
  switch (q) {
    case 0: return cosa;
    case 1: return -sina;
    case 2: return -cosa;
    default: return sina;
  }
}

/* Standard normal deviate for element `index` of tensor `tensor_id`. */
TMX_RNG_FN double tmx_rng_normal(uint64_t seed, uint64_t tensor_id, uint64_t index) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  uint64_t key = tmx_rng_key(seed, tensor_id);
  double u1 = 1.0 - tmx_rng_uniform(key, 2 * index);     /* (0, 1] */
  double u2 = tmx_rng_uniform(key, 2 * index + 1);       /* [0, 1) */
  double r = TMX_RNG_SQRT(-2.0 * tmx_rng_ln(u1));
  return r * tmx_rng_cos2pi(u2);
}

/* Uniform deviate on [lo, hi). */
TMX_RNG_FN double tmx_rng_uniform_range(uint64_t seed, uint64_t tensor_id, uint64_t index, double lo, double hi) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  uint64_t key = tmx_rng_key(seed, tensor_id);
  double u = tmx_rng_uniform(key, index);
  return lo + (hi - lo) * u;
}

#endif /* TMX_RNG_H */
