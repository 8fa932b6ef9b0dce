// rr_math.h -- canonical double-precision elementary functions.
//
// One source, compiled twice: by hipcc for the gfx950 fitness kernels and by
// g++ for host code.  Every function is written with IEEE-754 basic
// operations (+ - * / sqrt), explicit fused multiply-add and integer bit
// manipulation only, and both sides are built with -ffp-contract=off, so a
// given input produces the same bits on the CPU and on the GPU.  That is what
// lets the GPU fitness path be compared bit-for-bit against a CPU restatement.
//
// The reference evaluates node operators with Eigen's array functions
// (rils_rols_cpp/node.cpp:41-52: a.sin(), a.cos(), a.log(), a.exp(),
// a.sqrt(), a*a).  The methods here are textbook ones:
//   exp  : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor series of
//          e^r - 1 to degree 13, scaling by 2^k built from exponent bits.
//   log  : x = 2^e m, m in [sqrt2/2, sqrt2), s = f/(2+f) with f = m-1,
//          log(1+f) = f - (hfsq - s(hfsq + R(s^2))) where R is the atanh
//          series 2z/3 + 2z^2/5 + ... (the rearrangement used by Sun's fdlibm).
//   sin/cos: reduction x = q pi/2 + r with pi/2 held as three doubles and an
//          fma-exact double-double subtraction for |x| < 2^28; a Payne-Hanek
//          reduction against a 1408-bit table of 2/pi for larger |x|; Taylor
//          polynomials of degree 17 (sin) / 18 (cos) on |r| <= pi/4.
// Accuracy against mpmath is checked by tests/test_oracle.py.
//
// The includer defines RR_HD as the function qualifiers for its compiler
// (empty for g++, "__host__ __device__" for hipcc).
#pragma once
#include <stdint.h>

#ifndef RR_HD
#error "define RR_HD before including rr_math.h"
#endif

namespace rrm {

RR_HD inline double bits2d(uint64_t u) { return __builtin_bit_cast(double, u); }
RR_HD inline uint64_t d2bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// The quiet NaN returned for invalid arguments: the x86 default NaN
// (sign bit set, printed "-nan" by printf), which is what the reference's
// host build produces for log(-1) or sqrt(-1).  Written out as bits so both
// compilers return the same pattern.
RR_HD inline double qnan() { return bits2d(0xfff8000000000000ULL); }
RR_HD inline double pinf() { return bits2d(0x7ff0000000000000ULL); }

// Operations with a constant operand k: a * b + k, a * k + c, a * k.  The same
// correctly rounded fma / product on every compiler.  In GPU code each is one
// VOP3 instruction that reads k from a scalar register pair: left to itself
// the compiler picks v_fmac_f64 (the addend in the destination register) or a
// vector operand, and copies each 64-bit constant into vector registers with
// two v_mov_b32 -- vector-unit issue slots the FP64 kernels are short of.
RR_HD inline double fma_k(double a, double b, double k) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(k));
  return d;
#else
  return __builtin_fma(a, b, k);
#endif
}
RR_HD inline double fma_kv(double a, double k, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(k), "v"(c));
  return d;
#else
  return __builtin_fma(a, k, c);
#endif
}
RR_HD inline double mul_k(double a, double k) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_mul_f64 %0, %1, %2" : "=v"(d) : "v"(a), "s"(k));
  return d;
#else
  return a * k;
#endif
}

// The canonical order of every sum over samples in the FP64 form: 1024
// partial sums, row i added to partial (i mod 1024) in increasing i; the
// partials form 16 groups of 64 (one wavefront each on the GPU), each group
// is combined by an xor butterfly at distances 32, 16, 8, 4, 2, 1, and the 16
// group sums are added in group order.  canon_sum below is the host form; the
// kernels realise the same order with 16-wave workgroups (rr_device.h).
constexpr int CANON_GROUPS = 16;
constexpr int CANON_PARTIALS = 64 * CANON_GROUPS;

template <class F>
inline double canon_sum(int64_t lo, int64_t hi, F f) {
  double acc[CANON_PARTIALS];
  for (int l = 0; l < CANON_PARTIALS; ++l) acc[l] = 0.0;
  for (int64_t i = lo; i < hi; ++i) acc[i & (CANON_PARTIALS - 1)] = acc[i & (CANON_PARTIALS - 1)] + f(i);
  double total = 0.0;
  for (int g = 0; g < CANON_GROUPS; ++g) {
    double* a = acc + 64 * g;
    for (int s = 32; s >= 1; s >>= 1) {
      double nx[64];
      for (int l = 0; l < 64; ++l) nx[l] = a[l] + a[l ^ s];
      for (int l = 0; l < 64; ++l) a[l] = nx[l];
    }
    total = (g == 0) ? a[0] : total + a[0];
  }
  return total;
}

// The order of the initial column norms of the FP64 OLS form and of its
// first Householder step's tail sum: the rows fall into chunks of NORM_CHUNK;
// within a chunk, 256 partial sums (row i added to partial i mod 256 in
// increasing i), each group of 64 combined by the xor butterfly above, the 4
// group sums added in order; the chunk sums are added in chunk order.  The
// factor-evaluation kernel computes one chunk per workgroup while it writes
// the column, so the QR kernel needs no pass over the column for its norms.
constexpr int64_t NORM_CHUNK = 4096;
constexpr int NORM_GROUPS = 4;
constexpr int NORM_PARTIALS = 64 * NORM_GROUPS;

template <class F>
inline double chunk_sum(int64_t lo, int64_t hi, F f) {
  double total = 0.0;
  for (int64_t c0 = 0; c0 < hi; c0 += NORM_CHUNK) {
    double acc[NORM_PARTIALS];
    for (int l = 0; l < NORM_PARTIALS; ++l) acc[l] = 0.0;
    const int64_t a0 = lo > c0 ? lo : c0, a1 = hi < c0 + NORM_CHUNK ? hi : c0 + NORM_CHUNK;
    for (int64_t i = a0; i < a1; ++i) acc[i & (NORM_PARTIALS - 1)] = acc[i & (NORM_PARTIALS - 1)] + f(i);
    double ct = 0.0;
    for (int g = 0; g < NORM_GROUPS; ++g) {
      double* a = acc + 64 * g;
      for (int s = 32; s >= 1; s >>= 1) {
        double nx[64];
        for (int l = 0; l < 64; ++l) nx[l] = a[l] + a[l ^ s];
        for (int l = 0; l < 64; ++l) a[l] = nx[l];
      }
      ct = (g == 0) ? a[0] : ct + a[0];
    }
    total = (c0 == 0) ? ct : total + ct;
  }
  return total;
}

// 2^k for -1022 <= k <= 1023 (exact).
RR_HD inline double pow2i(int k) { return bits2d((uint64_t)(k + 1023) << 52); }

// ---------------------------------------------------------------- exp
RR_HD inline double exp(double x) {
  if (!(x == x)) return x;                       // NaN
  if (x > 709.8) return pinf();                  // > ln(DBL_MAX) + margin
  if (x < -745.2) return 0.0;                    // < ln(min subnormal / 2)
  const double INV_LN2 = 0x1.71547652b82fep+0;
  const double LN2_HI = 0x1.62e42fefa39efp-1;
  const double LN2_LO = 0x1.abc9e3b39803fp-56;
  const double kd = __builtin_rint(mul_k(x, INV_LN2));
  double r = fma_kv(-kd, LN2_HI, x);
  r = fma_kv(-kd, LN2_LO, r);
  // P(r) = sum_{j=2}^{13} r^(j-2)/j!
  double p = fma_kv(r, 0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29);   // 1/13!, 1/12!
  p = fma_k(p, r, 0x1.ae64567f544e4p-26); // 1/11!
  p = fma_k(p, r, 0x1.27e4fb7789f5cp-22); // 1/10!
  p = fma_k(p, r, 0x1.71de3a556c734p-19); // 1/9!
  p = fma_k(p, r, 0x1.a01a01a01a01ap-16); // 1/8!
  p = fma_k(p, r, 0x1.a01a01a01a01ap-13); // 1/7!
  p = fma_k(p, r, 0x1.6c16c16c16c17p-10); // 1/6!
  p = fma_k(p, r, 0x1.1111111111111p-7);  // 1/5!
  p = fma_k(p, r, 0x1.5555555555555p-5);  // 1/4!
  p = fma_k(p, r, 0x1.5555555555555p-3);  // 1/3!
  p = fma_k(p, r, 0x1.0000000000000p-1);  // 1/2!
  const double em1 = __builtin_fma(r * r, p, r);  // e^r - 1
  const double e = 1.0 + em1;
  const int k = (int)kd;
  if (k > 1023) return (e * pow2i(1023)) * pow2i(k - 1023);
  if (k < -1022) return (e * pow2i(k + 60)) * pow2i(-60);
  return e * pow2i(k);
}

// ---------------------------------------------------------------- log
RR_HD inline double log(double x) {
  if (!(x == x)) return x;                       // NaN
  if (x < 0.0) return qnan();
  if (x == 0.0) return -pinf();
  if (x == pinf()) return x;
  int e = 0;
  if (x < 0x1p-1022) { x *= 0x1p54; e = -54; }   // subnormal input
  const uint64_t u = d2bits(x);
  e += (int)((u >> 52) & 0x7ff) - 1023;
  double m = bits2d((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1,2)
  if (m > 0x1.6a09e667f3bcdp+0) { m *= 0.5; e += 1; }                       // [sqrt2/2, sqrt2]
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  // R(z) = sum_{j=1}^{12} 2 z^j/(2j+1)
  double R = fma_kv(z, 0x1.47ae147ae147bp-4, 0x1.642c8590b2164p-4);   // 2/25, 2/23
  R = fma_k(R, z, 0x1.8618618618618p-4);  // 2/21
  R = fma_k(R, z, 0x1.af286bca1af28p-4);  // 2/19
  R = fma_k(R, z, 0x1.e1e1e1e1e1e1ep-4);  // 2/17
  R = fma_k(R, z, 0x1.1111111111111p-3);  // 2/15
  R = fma_k(R, z, 0x1.3b13b13b13b14p-3);  // 2/13
  R = fma_k(R, z, 0x1.745d1745d1746p-3);  // 2/11
  R = fma_k(R, z, 0x1.c71c71c71c71cp-3);  // 2/9
  R = fma_k(R, z, 0x1.2492492492492p-2);  // 2/7
  R = fma_k(R, z, 0x1.999999999999ap-2);  // 2/5
  R = fma_k(R, z, 0x1.5555555555555p-1);  // 2/3
  R = R * z;
  const double hfsq = 0.5 * f * f;
  const double LN2_HI32 = 0x1.62e42fee00000p-1;  // ln2 to 32 bits: e*LN2_HI32 exact
  const double LN2_LO32 = 0x1.a39ef35793c76p-33;
  const double de = (double)e;
  return mul_k(de, LN2_HI32) + (f - (hfsq - (s * (hfsq + R) + mul_k(de, LN2_LO32))));
}

// ---------------------------------------------------------------- sin / cos
// 2/pi as a binary fraction, 64 bits per word, most significant first.
RR_HD inline uint64_t two_over_pi_word(int i) {
  switch (i) {
    case 0: return 0xa2f9836e4e441529ULL;  case 1: return 0xfc2757d1f534ddc0ULL;
    case 2: return 0xdb6295993c439041ULL;  case 3: return 0xfe5163abdebbc561ULL;
    case 4: return 0xb7246e3a424dd2e0ULL;  case 5: return 0x06492eea09d1921cULL;
    case 6: return 0xfe1deb1cb129a73eULL;  case 7: return 0xe88235f52ebb4484ULL;
    case 8: return 0xe99c7026b45f7e41ULL;  case 9: return 0x3991d639835339f4ULL;
    case 10: return 0x9c845f8bbdf9283bULL; case 11: return 0x1ff897ffde05980fULL;
    case 12: return 0xef2f118b5a0a6d1fULL; case 13: return 0x6d367ecf27cb09b7ULL;
    case 14: return 0x4f463f669e5fea2dULL; case 15: return 0x7527bac7ebe5f17bULL;
    case 16: return 0x3d0739f78a5292eaULL; case 17: return 0x6bfb5fb11f8d5d08ULL;
    case 18: return 0x56033046fc7b6babULL; case 19: return 0xf0cfbc209af4361dULL;
    case 20: return 0xa9e391615ee61b08ULL; default: return 0x6599855f14a06840ULL;
  }
}

// bits [pos, pos+64) of the 256-bit little-endian number r3:r2:r1:r0
// (selects instead of array indexing so that no private memory is used)
RR_HD inline uint64_t limb4(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3, int i) {
  return i == 0 ? r0 : i == 1 ? r1 : i == 2 ? r2 : i == 3 ? r3 : 0;
}
RR_HD inline uint64_t bits256(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3, int pos) {
  const int limb = pos >> 6, off = pos & 63;
  uint64_t v = limb4(r0, r1, r2, r3, limb) >> off;
  if (off != 0) v |= limb4(r0, r1, r2, r3, limb + 1) << (64 - off);
  return v;
}

// Payne-Hanek: |x| >= 2^28, finite.  Returns x = q pi/2 + (rh + rl).
RR_HD inline void reduce_large(double ax, double* rh, double* rl, int* q) {
  const uint64_t u = d2bits(ax);
  const int E = (int)((u >> 52) & 0x7ff) - 1075;            // ax = M 2^E
  const uint64_t M = (u & 0x000fffffffffffffULL) | 0x0010000000000000ULL;
  const int s = (E - 1 > 1) ? E - 1 : 1;                     // first needed bit of 2/pi (1-based)
  const int o = s - 1, wi = o >> 6, sh = o & 63;
  const uint64_t w0 = two_over_pi_word(wi), w1 = two_over_pi_word(wi + 1);
  const uint64_t w2 = two_over_pi_word(wi + 2), w3 = two_over_pi_word(wi + 3);
  uint64_t W2, W1, W0;
  if (sh == 0) { W2 = w0; W1 = w1; W0 = w2; }
  else {
    W2 = (w0 << sh) | (w1 >> (64 - sh));
    W1 = (w1 << sh) | (w2 >> (64 - sh));
    W0 = (w2 << sh) | (w3 >> (64 - sh));
  }
  unsigned __int128 t = (unsigned __int128)M * W0;
  const uint64_t r0 = (uint64_t)t;
  t = (unsigned __int128)M * W1 + (uint64_t)(t >> 64);
  const uint64_t r1 = (uint64_t)t;
  t = (unsigned __int128)M * W2 + (uint64_t)(t >> 64);
  const uint64_t r2 = (uint64_t)t;
  const uint64_t r3 = (uint64_t)(t >> 64);
  const int P = s + 191 - E;                                  // fractional bits of M*W
  int qq = (int)(bits256(r0, r1, r2, r3, P) & 3);
  uint64_t fh = bits256(r0, r1, r2, r3, P - 64), fl = bits256(r0, r1, r2, r3, P - 128);
  double sign = 1.0;
  if (fh >> 63) {                                             // fraction >= 1/2: use F - 1
    qq = (qq + 1) & 3;
    fh = ~fh; fl = ~fl;                                       // 2^128 - F = ~F + 1
    fl += 1; if (fl == 0) fh += 1;
    sign = -1.0;
  }
  // normalise the 128-bit fraction and split it into exact 53-bit pieces
  int lz = 0;
  if (fh == 0) { fh = fl; fl = 0; lz = 64; }
  if (fh == 0) { *rh = 0.0; *rl = 0.0; *q = qq; return; }
  const int c = __builtin_clzll(fh);
  if (c) { fh = (fh << c) | (fl >> (64 - c)); fl <<= c; }
  lz += c;
  const double hi = (double)(fh >> 11) * 0x1p-53;
  const double mid = (double)(((fh & 0x7ffULL) << 42) | (fl >> 22)) * 0x1p-106;
  const double scale = (lz < 960) ? pow2i(-lz) : 0.0;
  const double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
  const double ph = hi * PIO2_HI;
  double pe = __builtin_fma(hi, PIO2_HI, -ph);
  pe = pe + (hi * PIO2_LO + mid * PIO2_HI);
  const double sh2 = ph + pe;
  const double sl2 = pe - (sh2 - ph);
  *rh = sign * (sh2 * scale);
  *rl = sign * (sl2 * scale);
  *q = qq;
}

// How sin and cos reach reduce_large: inline (the default), or through a
// policy of the caller's (GPU code that inlines sin and cos keeps the rare
// large-argument path out of line: its 256-bit arithmetic would otherwise
// hold registers for the whole kernel).  The arithmetic is the same.
struct LargeInline {
  RR_HD static void run(double ax, double* rh, double* rl, int* q) { reduce_large(ax, rh, rl, q); }
};

// x = q pi/2 + (rh + rl), |rh| <~ pi/4.  x finite.
template <class L = LargeInline>
RR_HD inline void reduce(double x, double* rh, double* rl, int* q) {
  const double ax = __builtin_fabs(x);
  if (ax < 0x1p28) {
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54, P3 = -0x1.f1976b7ed8fbcp-110;
    const double kd = __builtin_rint(mul_k(x, TWO_OVER_PI));
    const double p1 = mul_k(kd, P1), e1 = fma_kv(kd, P1, -p1);   // kd*P1 = p1 + e1 exactly
    const double t = x - p1;                                      // exact (Sterbenz) for kd != 0
    const double p2 = mul_k(kd, P2), e2 = fma_kv(kd, P2, -p2);
    const double u = t - p2;                                      // two-sum: u + ue = t - p2
    const double bv = u - t;
    double ue = (t - (u - bv)) + (-p2 - bv);
    ue = ue - e1;
    ue = ue - e2;
    ue = ue - mul_k(kd, P3);
    const double h = u + ue;
    *rl = ue - (h - u);
    *rh = h;
    *q = ((int)kd) & 3;
  } else {
    int qq;
    L::run(ax, rh, rl, &qq);
    if (x < 0.0) { *rh = -*rh; *rl = -*rl; qq = (4 - qq) & 3; }
    *q = qq;
  }
}

// sin(rh + rl) for |rh| <= ~pi/4
RR_HD inline double ksin(double rh, double rl) {
  const double z = rh * rh;
  double S = fma_kv(z, 0x1.952c77030ad4ap-49, -0x1.ae7f3e733b81fp-41);   //  1/17!, -1/15!
  S = fma_k(S, z, 0x1.6124613a86d09p-33);  //  1/13!
  S = fma_k(S, z, -0x1.ae64567f544e4p-26); // -1/11!
  S = fma_k(S, z, 0x1.71de3a556c734p-19);  //  1/9!
  S = fma_k(S, z, -0x1.a01a01a01a01ap-13); // -1/7!
  S = fma_k(S, z, 0x1.1111111111111p-7);   //  1/5!
  S = fma_k(S, z, -0x1.5555555555555p-3);  // -1/3!
  const double tail = (z * rh) * S + rl * (1.0 - 0.5 * z);
  return rh + tail;
}

// cos(rh + rl) for |rh| <= ~pi/4:
//   cos r = 1 - z/2 + z^2 C(z),  C(z) = 1/4! - z/6! + z^2/8! - ... - z^7/18!
RR_HD inline double kcos(double rh, double rl) {
  const double z = rh * rh;
  double C = fma_kv(z, -0x1.6827863b97d97p-53, 0x1.ae7f3e733b81fp-45);   // -1/18!, 1/16!
  C = fma_k(C, z, -0x1.93974a8c07c9dp-37); // -1/14!
  C = fma_k(C, z, 0x1.1eed8eff8d898p-29);  //  1/12!
  C = fma_k(C, z, -0x1.27e4fb7789f5cp-22); // -1/10!
  C = fma_k(C, z, 0x1.a01a01a01a01ap-16);  //  1/8!
  C = fma_k(C, z, -0x1.6c16c16c16c17p-10); // -1/6!
  C = fma_k(C, z, 0x1.5555555555555p-5);   //  1/4!
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  const double err = (1.0 - w) - hz;              // rounding error of w, exact
  return w + (err + ((z * z) * C - rh * rl));
}

template <class L = LargeInline>
RR_HD inline double sin(double x) {
  if (!(x - x == 0.0)) return qnan();            // NaN or inf
  if (__builtin_fabs(x) < 0x1p-27) return x;
  double rh, rl; int q;
  reduce<L>(x, &rh, &rl, &q);
  // quadrant q: sin, cos, -sin, -cos.  One kernel per parity and the sign
  // apart, so lanes of a wavefront in different quadrants run each of the
  // two polynomials once (a four-way switch runs each twice).
  const double v = (q & 1) ? kcos(rh, rl) : ksin(rh, rl);
  return (q & 2) ? -v : v;
}

template <class L = LargeInline>
RR_HD inline double cos(double x) {
  if (!(x - x == 0.0)) return qnan();
  if (__builtin_fabs(x) < 0x1p-27) return 1.0;
  double rh, rl; int q;
  reduce<L>(x, &rh, &rl, &q);
  // quadrant q: cos, -sin, -cos, sin
  const double v = (q & 1) ? ksin(rh, rl) : kcos(rh, rl);
  return ((q + 1) & 2) ? -v : v;
}

// IEEE-754 correctly rounded square root on both compilers.
RR_HD inline double sqrt(double x) { return __builtin_sqrt(x); }

}  // namespace rrm
