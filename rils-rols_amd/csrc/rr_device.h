// rr_device.h -- device-side building blocks shared by the fitness kernels:
// the canonical wave reduction, the row iteration order that goes with it,
// and the bytecode interpreter for candidate programs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RR_HD __host__ __device__
#include "rr_math.h"
#include "../../include/rr_fitness.h"

namespace rr {

constexpr int WAVE = 64;

// Canonical sums over samples (the order rrm::canon_sum defines): a
// workgroup of CW = 16 waves, thread t = 64 w + l, adds rows i with
// i mod 1024 == t in increasing order starting from +0.0; each wave combines
// its 64 partials by an xor butterfly at distances 32, 16, 8, 4, 2, 1 (every
// lane ends with the same bits: a + b == b + a), and the 16 wave sums are
// added in wave order.
constexpr int CW = rrm::CANON_GROUPS;
constexpr int CP = rrm::CANON_PARTIALS;

// first row >= lo that thread t (of CP) owns
__device__ inline int64_t first_row(int64_t lo, int t) {
  const int64_t i = (lo & ~int64_t(CP - 1)) + t;
  return i < lo ? i + CP : i;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v = v + __shfl_xor(v, s);
  return v;
}

// Canonical block sums of NV values at once; every thread gets the results.
// red: LDS scratch of CW x NV doubles.  Contains two barriers.
template <int NV>
__device__ inline void block_sums(double (&v)[NV], double (*red)[NV]) {
  const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
#pragma unroll
  for (int t = 0; t < NV; ++t) v[t] = wave_sum(v[t]);
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < NV; ++t) red[wave][t] = v[t];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    double s = red[0][t];
#pragma nounroll
    for (int w = 1; w < CW; ++w) s = s + red[w][t];
    v[t] = s;
  }
  __syncthreads();
}

__device__ inline double bin_f64(uint32_t bop, double a, double b) {
  switch (bop) {
    case RR_BOP_ADD: return a + b;
    case RR_BOP_SUB: return a - b;
    case RR_BOP_MUL: return a * b;
    case RR_BOP_DIV: return a / b;
    case RR_BOP_LT: return a < b ? 1.0 : 0.0;
    case RR_BOP_GT: return a > b ? 1.0 : 0.0;
    case RR_BOP_EQ: return a == b ? 1.0 : 0.0;
    case RR_BOP_NEQ: return a != b ? 1.0 : 0.0;
    case RR_BOP_MIN: return a < b ? a : b;
    default: return a > b ? a : b;
  }
}

// Transcendentals are called, not inlined: inlining all four into the
// interpreter's switch costs ~100 extra VGPRs (occupancy 1-2 waves/SIMD);
// as calls the interpreter needs ~56.
__device__ __noinline__ double call_sin(double x) { return rrm::sin(x); }
__device__ __noinline__ double call_cos(double x) { return rrm::cos(x); }
__device__ __noinline__ double call_log(double x) { return rrm::log(x); }
__device__ __noinline__ double call_exp(double x) { return rrm::exp(x); }

// The same for the R samples of one interpreter step in one call: the R
// independent evaluations share the call and overlap in the callee.
template <int R>
struct Rows {
  double v[R];
};
#define RR_CALL_ROWS(NAME, FN)                                      \
  template <int R>                                                  \
  __device__ __noinline__ Rows<R> NAME(Rows<R> x) {                 \
    Rows<R> y;                                                      \
    _Pragma("unroll") for (int r = 0; r < R; ++r) y.v[r] = FN(x.v[r]); \
    return y;                                                       \
  }
RR_CALL_ROWS(call_sin_rows, rrm::sin)
RR_CALL_ROWS(call_cos_rows, rrm::cos)
RR_CALL_ROWS(call_log_rows, rrm::log)
RR_CALL_ROWS(call_exp_rows, rrm::exp)
#undef RR_CALL_ROWS

// Rows are padded so that a tile (CP threads x up to 4 rows, or 64 lanes x up
// to 8 rows) never reads past the end of a column: the row stride of every
// dataset and work column is a multiple of RR_ROW_PAD.
constexpr int64_t RR_ROW_PAD = 4 * CP;

// Evaluate program code[0..len) at R samples: row0, row0+ST, ...,
// row0+ST(R-1).  X is feature-major with column stride ldx.  Per-sample
// semantics: node::evaluate_inner (rils_rols_cpp/node.cpp:23-95) with
// rr_math functions.  The opcode is decoded once per R samples.
template <int R, int ST = CP>
__device__ inline void eval_f64(const uint32_t* __restrict__ code, int len, const double* __restrict__ K,
                                const double* __restrict__ X, int64_t ldx, int64_t row0, double (&acc)[R]) {
  // stack slots 0-1 in registers; deeper slots (rare with Sethi-Ullman
  // ordering) in a private array
  double st0[R], st1[R], stx[RR_MAX_STACK - 2][R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  int sp = 0;
  for (int pc = 0; pc < len; ++pc) {
    const uint32_t ins = code[pc];
    const uint32_t op = ins & 0xffu;
    const uint32_t arg = ins >> 8;
    if (op >= RR_OP_BIN) {
      const uint32_t b = op - RR_OP_BIN;
      const uint32_t bop = b >> 3, mode = b & 7u;
      double l[R], rr[R];
      switch (mode) {
        case RR_M_S:
        case RR_M_SR: {
          --sp;
          double s[R];
          if (sp == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = st0[r];
          } else if (sp == 1) {
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = st1[r];
          } else {
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = stx[sp - 2][r];
          }
#pragma unroll
          for (int r = 0; r < R; ++r) { l[r] = mode == RR_M_S ? s[r] : acc[r]; rr[r] = mode == RR_M_S ? acc[r] : s[r]; }
          break;
        }
        case RR_M_V:
#pragma unroll
          for (int r = 0; r < R; ++r) { l[r] = acc[r]; rr[r] = X[(int64_t)arg * ldx + row0 + (int64_t)ST * r]; }
          break;
        case RR_M_VR:
#pragma unroll
          for (int r = 0; r < R; ++r) { l[r] = X[(int64_t)arg * ldx + row0 + (int64_t)ST * r]; rr[r] = acc[r]; }
          break;
        case RR_M_C: {
          const double k = K[arg];
#pragma unroll
          for (int r = 0; r < R; ++r) { l[r] = acc[r]; rr[r] = k; }
          break;
        }
        default: {
          const double k = K[arg];
#pragma unroll
          for (int r = 0; r < R; ++r) { l[r] = k; rr[r] = acc[r]; }
          break;
        }
      }
      switch (bop) {   // one uniform branch per instruction, not per row
#define RR_BIN(B) case B: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = bin_f64(B, l[r], rr[r]); break;
        RR_BIN(RR_BOP_ADD) RR_BIN(RR_BOP_SUB) RR_BIN(RR_BOP_MUL) RR_BIN(RR_BOP_DIV) RR_BIN(RR_BOP_LT)
        RR_BIN(RR_BOP_GT) RR_BIN(RR_BOP_EQ) RR_BIN(RR_BOP_NEQ) RR_BIN(RR_BOP_MIN)
        default: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = bin_f64(RR_BOP_MAX, l[r], rr[r]); break;
#undef RR_BIN
      }
    } else {
      switch (op) {
        case RR_OP_LDV:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = X[(int64_t)arg * ldx + row0 + (int64_t)ST * r];
          break;
        case RR_OP_LDC: {
          const double k = K[arg];
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = k;
          break;
        }
        case RR_OP_PUSH:
          if (sp == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) st0[r] = acc[r];
          } else if (sp == 1) {
#pragma unroll
            for (int r = 0; r < R; ++r) st1[r] = acc[r];
          } else {
#pragma unroll
            for (int r = 0; r < R; ++r) stx[sp - 2][r] = acc[r];
          }
          ++sp;
          break;
        case RR_OP_SIN: {
          Rows<R> t;
#pragma unroll
          for (int r = 0; r < R; ++r) t.v[r] = acc[r];
          t = call_sin_rows<R>(t);
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = t.v[r];
          break;
        }
        case RR_OP_COS: {
          Rows<R> t;
#pragma unroll
          for (int r = 0; r < R; ++r) t.v[r] = acc[r];
          t = call_cos_rows<R>(t);
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = t.v[r];
          break;
        }
        case RR_OP_LN: {
          Rows<R> t;
#pragma unroll
          for (int r = 0; r < R; ++r) t.v[r] = acc[r];
          t = call_log_rows<R>(t);
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = t.v[r];
          break;
        }
        case RR_OP_EXP: {
          Rows<R> t;
#pragma unroll
          for (int r = 0; r < R; ++r) t.v[r] = acc[r];
          t = call_exp_rows<R>(t);
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = t.v[r];
          break;
        }
        case RR_OP_SQRT:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = rrm::sqrt(acc[r]);
          break;
        default:  // RR_OP_SQR
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = acc[r] * acc[r];
          break;
      }
    }
  }
}

// sin / cos with the large-argument reduction as a call (rrm::reduce's policy)
__device__ __noinline__ void reduce_large_call(double ax, double* rh, double* rl, int* q) {
  rrm::reduce_large(ax, rh, rl, q);
}
struct LargeCall {
  __device__ static void run(double ax, double* rh, double* rl, int* q) { reduce_large_call(ax, rh, rl, q); }
};

// The same program semantics as eval_f64 for R samples per thread (row0 +
// ST r) with the operand stack in LDS -- slot s of sample r of thread t at
// stk[(s R + r) ST + t], so each thread touches only its own words and no
// barrier is needed; stk holds the program's stack depth x R x ST doubles.
// The accumulator stays in registers.  Only it and a few addresses are live
// across an operation, so the transcendentals are inlined (two samples at a
// time; the large-argument reduction of sin and cos stays a call) and no
// registers are saved around calls.
template <int R, int ST>
__device__ inline void eval_lds_f64(const uint32_t* __restrict__ code, int len, const double* __restrict__ K,
                                    const double* __restrict__ X, int64_t ldx, int64_t row0, double (&acc)[R],
                                    double* __restrict__ stk) {
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  int sp = 0;
  for (int pc = 0; pc < len; ++pc) {
    const uint32_t ins = code[pc];
    const uint32_t op = ins & 0xffu;
    const uint32_t arg = ins >> 8;
    if (op >= RR_OP_BIN) {
      const uint32_t b = op - RR_OP_BIN;
      const uint32_t bop = b >> 3, mode = b & 7u;
      double o[R];   // the operand that is not the accumulator
      if (mode == RR_M_S || mode == RR_M_SR) {
        --sp;
#pragma unroll
        for (int r = 0; r < R; ++r) o[r] = stk[(sp * R + r) * ST + t];
      } else if (mode == RR_M_V || mode == RR_M_VR) {
        const double* xc = X + (int64_t)arg * ldx + row0;
#pragma unroll
        for (int r = 0; r < R; ++r) o[r] = xc[(int64_t)ST * r];
      } else {
        const double k = K[arg];
#pragma unroll
        for (int r = 0; r < R; ++r) o[r] = k;
      }
      // the accumulator is the left operand for SR, V and C, the right one
      // for S, VR and CR (one uniform branch each way)
      if (mode == RR_M_SR || mode == RR_M_V || mode == RR_M_C) {
        switch (bop) {
#define RR_BIN(B) case B: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = bin_f64(B, acc[r], o[r]); break;
          RR_BIN(RR_BOP_ADD) RR_BIN(RR_BOP_SUB) RR_BIN(RR_BOP_MUL) RR_BIN(RR_BOP_DIV) RR_BIN(RR_BOP_LT)
          RR_BIN(RR_BOP_GT) RR_BIN(RR_BOP_EQ) RR_BIN(RR_BOP_NEQ) RR_BIN(RR_BOP_MIN)
          default: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = bin_f64(RR_BOP_MAX, acc[r], o[r]); break;
#undef RR_BIN
        }
      } else {
        switch (bop) {
#define RR_BIN(B) case B: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = bin_f64(B, o[r], acc[r]); break;
          RR_BIN(RR_BOP_ADD) RR_BIN(RR_BOP_SUB) RR_BIN(RR_BOP_MUL) RR_BIN(RR_BOP_DIV) RR_BIN(RR_BOP_LT)
          RR_BIN(RR_BOP_GT) RR_BIN(RR_BOP_EQ) RR_BIN(RR_BOP_NEQ) RR_BIN(RR_BOP_MIN)
          default: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = bin_f64(RR_BOP_MAX, o[r], acc[r]); break;
#undef RR_BIN
        }
      }
    } else {
      switch (op) {
        case RR_OP_LDV: {
          const double* xc = X + (int64_t)arg * ldx + row0;
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = xc[(int64_t)ST * r];
          break;
        }
        case RR_OP_LDC: {
          const double k = K[arg];
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = k;
          break;
        }
        case RR_OP_PUSH:
#pragma unroll
          for (int r = 0; r < R; ++r) stk[(sp * R + r) * ST + t] = acc[r];
          ++sp;
          break;
#define RR_UN2(OPC, FN)                                          \
  case OPC:                                                      \
    _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = FN(acc[r]); \
    break;
        RR_UN2(RR_OP_SIN, rrm::sin<LargeCall>)
        RR_UN2(RR_OP_COS, rrm::cos<LargeCall>)
        RR_UN2(RR_OP_LN, rrm::log)
        RR_UN2(RR_OP_EXP, rrm::exp)
#undef RR_UN2
        case RR_OP_SQRT:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = rrm::sqrt(acc[r]);
          break;
        default:  // RR_OP_SQR
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = acc[r] * acc[r];
          break;
      }
    }
  }
}

}  // namespace rr
