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

// Canonical sum over rows [lo, n): row i goes to lane (i mod 64), each lane
// adds its rows in increasing order starting from +0.0, then the 64 partials
// are combined by an xor butterfly at distances 32, 16, 8, 4, 2, 1.  Every
// lane ends with the same bits (a + b == b + a in IEEE arithmetic).
__device__ inline int64_t first_row(int64_t lo, int lane) {
  const int64_t i = (lo & ~int64_t(63)) + lane;
  return i < lo ? i + 64 : i;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v = v + __shfl_xor(v, s);
  return v;
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

// Operand stack held in named registers; the stack pointer is wave-uniform,
// so the switch is a scalar branch and no scratch memory is used.
static_assert(RR_MAX_STACK == 8, "Stack64 is written out for 8 slots");
struct Stack64 {
  double s[RR_MAX_STACK];
  __device__ inline void set(int i, double v) {
    switch (i) {
#define RR_SET(k) case k: s[k] = v; break;
      RR_SET(0) RR_SET(1) RR_SET(2) RR_SET(3) RR_SET(4) RR_SET(5) RR_SET(6) default: s[7] = v; break;
#undef RR_SET
    }
  }
  __device__ inline double get(int i) const {
    switch (i) {
#define RR_GET(k) case k: return s[k];
      RR_GET(0) RR_GET(1) RR_GET(2) RR_GET(3) RR_GET(4) RR_GET(5) RR_GET(6) default: return s[7];
#undef RR_GET
    }
  }
};

// Evaluate program code[0..len) at sample `row`.  X is feature-major with
// column stride ldx.  Per-sample semantics: node::evaluate_inner
// (rils_rols_cpp/node.cpp:23-95) with rr_math functions.
__device__ inline double eval_f64(const uint32_t* __restrict__ code, int len, const double* __restrict__ K,
                                  const double* __restrict__ X, int64_t ldx, int64_t row) {
  double acc = 0.0;
  Stack64 st;
  int sp = 0;
  for (int pc = 0; pc < len; ++pc) {
    const uint32_t ins = code[pc];
    const uint32_t op = ins & 0xffu;
    const uint32_t arg = ins >> 8;
    if (op >= RR_OP_BIN) {
      const uint32_t b = op - RR_OP_BIN;
      const uint32_t bop = b >> 3, mode = b & 7u;
      double l, r;
      switch (mode) {
        case RR_M_S: --sp; l = st.get(sp); r = acc; break;
        case RR_M_SR: --sp; l = acc; r = st.get(sp); break;
        case RR_M_V: l = acc; r = X[(int64_t)arg * ldx + row]; break;
        case RR_M_VR: l = X[(int64_t)arg * ldx + row]; r = acc; break;
        case RR_M_C: l = acc; r = K[arg]; break;
        default: l = K[arg]; r = acc; break;
      }
      acc = bin_f64(bop, l, r);
    } else {
      switch (op) {
        case RR_OP_LDV: acc = X[(int64_t)arg * ldx + row]; break;
        case RR_OP_LDC: acc = K[arg]; break;
        case RR_OP_PUSH: st.set(sp, acc); ++sp; break;
        case RR_OP_SIN: acc = call_sin(acc); break;
        case RR_OP_COS: acc = call_cos(acc); break;
        case RR_OP_LN: acc = call_log(acc); break;
        case RR_OP_EXP: acc = call_exp(acc); break;
        case RR_OP_SQRT: acc = rrm::sqrt(acc); break;
        default: acc = acc * acc; break;  // RR_OP_SQR
      }
    }
  }
  return acc;
}

}  // namespace rr
