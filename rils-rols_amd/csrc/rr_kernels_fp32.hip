// rr_kernels_fp32.hip -- FP32 precision form of the candidate-fitness path.
//
// One wavefront owns one candidate (like the FP64 form) and makes two passes
// over the samples, each evaluating the candidate's factor programs in FP32
// for R rows per lane:
//   pass 1  accumulates the Gram matrix of [factors | 1] and A^T y in FP64
//           (a product of two FP32 numbers is exact in FP64, so only the
//           accumulation rounds);
//   solve   FP64 pivoted Cholesky of the Gram matrix on the wave, with the
//           pivot rule of ColPivHouseholderQR (largest remaining column
//           norm, first index on ties);
//   pass 2  residual r = y - A c in FP64, A^T r for one refinement step, SSR,
//           and the error terms of the bound below;
//   refine  c += G^-1 A^T r, SSR -= dc . A^T r.
//
// Every FP32 value carries a running error bound e >= |v32 - v| where v is
// the FP64 value of the same expression (u = 2^-23, twice the FP32 unit
// roundoff, absorbs the rounding of the bound arithmetic and the FP64 path's
// own error).  A value is "uncertain" (e = +inf) when FP32 and FP64 can
// disagree on finiteness (overflow, a domain boundary inside the error bound);
// a NaN with e == 0 is a "certain NaN": FP64 produces NaN at that sample too.
//
// Certification, per candidate:
//   * any uncertain value                                   -> RR_F_NEEDS64
//   * a certain-NaN factor column: the FP64 QR reaches that column's
//     reflector (no rank cut before it, guaranteed when the finite columns
//     are well conditioned, or when it is factor 0 and the threshold is NaN)
//     and then every coefficient is NaN                      -> sentinel, exact
//   * Schur complement of a pivot below TAU_C of its column, or n < k
//                                                            -> RR_F_NEEDS64
//   * otherwise |sqrt(SSR32) - sqrt(SSR64)| <= D with
//       D = ||dA |c|||  +  sum_j (dc_j + |refinement_j|) ||dA_j||
//           + sum_j 1e-12 ||A_j|| (1 + |c_j|)          (OLS-tree thresholds)
//     where dA is bounded elementwise by the running errors and
//       dc = 2 |G^-1| (|dA|^T |r| + |A|^T |dA| |c|)     (first-order LS
//     perturbation, doubled); a coefficient whose |c| vs 1e-12 or |c - 1|
//     vs 1e-12 decision lies inside dc                    -> RR_F_NEEDS64.
//   The bound on SSR gives bounds on 1-R2 and RMSE, and rr_fit.err is the
//   relative bound on the fitness value (1+e0)(1+e1)(1+size*penalty).
#include "rr_device.h"
#include "rr_kernels_fp32.h"

namespace rr {

namespace {

constexpr float U32 = 0x1p-23f;
constexpr float FINF = __builtin_huge_valf();
constexpr double TAU_C = 1e-4;       // smallest accepted Schur complement / column norm^2
constexpr double TAU_S = 1e-20;      // smallest accepted Schur complement / (max norm)^2
constexpr double EPS12 = 1e-12;
constexpr int MAXF = 8;              // factors per candidate in this form (k <= 9)

struct VE { float v, e; };

__device__ inline bool is_certain_nan(float v, float e) { return v != v && e == 0.f; }
__device__ inline float fix_e(float e) { return (e == e) ? e : FINF; }

// Transcendentals: HIP's FP32 functions, with the error bounds assumed (and
// checked on the GPU against the FP64 functions by the test suite): sinf/cosf
// 2^-21 absolute, expf/logf 4 ulp relative, sqrtf 1 ulp relative.
__device__ __noinline__ VE f_sin(VE a) {
  if (is_certain_nan(a.v, a.e)) return {a.v, 0.f};
  const float v = sinf(a.v);
  return {v, fix_e(fminf(a.e, 2.f) + 8.f * 0x1p-24f)};
}
__device__ __noinline__ VE f_cos(VE a) {
  if (is_certain_nan(a.v, a.e)) return {a.v, 0.f};
  const float v = cosf(a.v);
  return {v, fix_e(fminf(a.e, 2.f) + 8.f * 0x1p-24f)};
}
__device__ __noinline__ VE f_exp(VE a) {
  if (is_certain_nan(a.v, a.e)) return {a.v, 0.f};
  const float v = expf(a.v);
  if (!(v < FINF)) return {v, FINF};
  // |e^(a+d) - e^a| <= e^a (e^|d| - 1), and e^x - 1 <= x (1 + x) for 0 <= x <= 1
  const float g = (a.e <= 1.f) ? a.e * (1.f + a.e) : expf(a.e) - 1.f;
  return {v, fix_e(v * g * (1.f + 4.f * U32) + v * 4.f * U32 + 0x1p-126f)};
}
__device__ __noinline__ VE f_log(VE a) {
  if (is_certain_nan(a.v, a.e)) return {a.v, 0.f};
  const float lo = a.v - a.e * (1.f + U32);
  if (lo > 0.f) {
    const float v = logf(a.v);
    return {v, fix_e(a.e / lo * (1.f + 2.f * U32) + fabsf(v) * 4.f * U32)};
  }
  if (a.v + a.e * (1.f + U32) < 0.f) return {__builtin_nanf(""), 0.f};   // certainly negative
  return {logf(a.v), FINF};
}
__device__ inline VE f_sqrt(VE a) {
  if (is_certain_nan(a.v, a.e)) return {a.v, 0.f};
  const float lo = a.v - a.e * (1.f + U32);
  if (lo >= 0.f) {
    const float v = sqrtf(a.v);
    const float den = v + sqrtf(lo);
    const float e = den > 0.f ? a.e / den * (1.f + 2.f * U32) : sqrtf(a.e);
    return {v, fix_e(e + v * U32)};
  }
  if (a.v + a.e * (1.f + U32) < 0.f) return {__builtin_nanf(""), 0.f};
  return {sqrtf(a.v), FINF};
}

__device__ inline VE f_bin(uint32_t bop, VE a, VE b) {
  const bool an = is_certain_nan(a.v, a.e), bn = is_certain_nan(b.v, b.e);
  switch (bop) {
    case RR_BOP_ADD: case RR_BOP_SUB: case RR_BOP_MUL: case RR_BOP_DIV: {
      float v;
      if (bop == RR_BOP_ADD) v = a.v + b.v;
      else if (bop == RR_BOP_SUB) v = a.v - b.v;
      else if (bop == RR_BOP_MUL) v = a.v * b.v;
      else v = a.v / b.v;
      if (an || bn) return {__builtin_nanf(""), 0.f};   // NaN op anything = NaN in both precisions
      if (!(fabsf(v) < FINF)) return {v, FINF};
      float e;
      if (bop <= RR_BOP_SUB) e = a.e + b.e;
      else if (bop == RR_BOP_MUL) e = fabsf(a.v) * b.e + fabsf(b.v) * a.e + a.e * b.e;
      else {
        const float bl = fabsf(b.v) - b.e * (1.f + U32);
        if (!(bl > 0.f)) return {v, FINF};
        e = (fabsf(a.v) * b.e + fabsf(b.v) * a.e) / (fabsf(b.v) * bl);
      }
      return {v, fix_e(e * (1.f + 4.f * U32) + fabsf(v) * U32)};
    }
    case RR_BOP_LT: case RR_BOP_GT: case RR_BOP_EQ: case RR_BOP_NEQ: {
      float v;
      if (bop == RR_BOP_LT) v = a.v < b.v ? 1.f : 0.f;
      else if (bop == RR_BOP_GT) v = a.v > b.v ? 1.f : 0.f;
      else if (bop == RR_BOP_EQ) v = a.v == b.v ? 1.f : 0.f;
      else v = a.v != b.v ? 1.f : 0.f;
      if (an || bn) return {v, 0.f};   // a comparison with NaN is false (!= true) in both precisions
      const bool sep = fabsf(a.v - b.v) > (a.e + b.e) * (1.f + 4.f * U32) && bop != RR_BOP_EQ && bop != RR_BOP_NEQ;
      const bool exact = a.e == 0.f && b.e == 0.f;
      return {v, (sep || exact) ? 0.f : 1.f};
    }
    default: {   // MIN / MAX: a<b ? a : b  /  a>b ? a : b
      const bool take_a = (bop == RR_BOP_MIN) ? (a.v < b.v) : (a.v > b.v);
      const VE r = take_a ? a : b;
      if (a.e == FINF || b.e == FINF) return {r.v, FINF};
      if (an || bn) return (a.e == 0.f && b.e == 0.f) ? r : VE{r.v, FINF};
      const bool sep = fabsf(a.v - b.v) > (a.e + b.e) * (1.f + 4.f * U32);
      return {r.v, sep ? r.e : fix_e(fmaxf(a.e, b.e) + fabsf(a.v - b.v))};
    }
  }
}

// FP32 evaluation of one program at R samples (rows row0 + 64 r).
template <int R>
__device__ inline void eval_f32(const uint32_t* __restrict__ code, int len, const double* __restrict__ K,
                                const float* __restrict__ X, int64_t ldx, int64_t row0, VE (&acc)[R]) {
  // slots 0-1 in registers; deeper slots (rare with Sethi-Ullman ordering)
  // in a private array
  VE st0[R], st1[R], stx[RR_MAX_STACK - 2][R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = {0.f, 0.f};
  int sp = 0;
  for (int pc = 0; pc < len; ++pc) {
    const uint32_t ins = code[pc];
    const uint32_t op = ins & 0xffu;
    const uint32_t arg = ins >> 8;
    if (op >= RR_OP_BIN) {
      const uint32_t b = op - RR_OP_BIN;
      const uint32_t bop = b >> 3, mode = b & 7u;
      VE l[R], rr[R];
      if (mode <= RR_M_SR) {
        --sp;
        VE s[R];
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
      } else if (mode <= RR_M_VR) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float x = X[(int64_t)arg * ldx + row0 + 64 * r];
          const VE xv{x, fabsf(x) * U32};
          l[r] = mode == RR_M_V ? acc[r] : xv;
          rr[r] = mode == RR_M_V ? xv : acc[r];
        }
      } else {
        const double kd = K[arg];
        const float kf = (float)kd;
        const VE kv{kf, (fabsf(kf) < FINF) ? fabsf(kf) * U32 : FINF};
#pragma unroll
        for (int r = 0; r < R; ++r) { l[r] = mode == RR_M_C ? acc[r] : kv; rr[r] = mode == RR_M_C ? kv : acc[r]; }
      }
      switch (bop) {   // one uniform branch per instruction, not per row
#define RR_BIN(B) case B: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = f_bin(B, l[r], rr[r]); break;
        RR_BIN(RR_BOP_ADD) RR_BIN(RR_BOP_SUB) RR_BIN(RR_BOP_MUL) RR_BIN(RR_BOP_DIV) RR_BIN(RR_BOP_LT)
        RR_BIN(RR_BOP_GT) RR_BIN(RR_BOP_EQ) RR_BIN(RR_BOP_NEQ) RR_BIN(RR_BOP_MIN)
        default: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = f_bin(RR_BOP_MAX, l[r], rr[r]); break;
#undef RR_BIN
      }
    } else {
      switch (op) {
        case RR_OP_LDV:
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const float x = X[(int64_t)arg * ldx + row0 + 64 * r];
            acc[r] = {x, fabsf(x) * U32};
          }
          break;
        case RR_OP_LDC: {
          const float kf = (float)K[arg];
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = {kf, (fabsf(kf) < FINF) ? fabsf(kf) * U32 : FINF};
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
        case RR_OP_SIN:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = f_sin(acc[r]);
          break;
        case RR_OP_COS:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = f_cos(acc[r]);
          break;
        case RR_OP_LN:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = f_log(acc[r]);
          break;
        case RR_OP_EXP:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = f_exp(acc[r]);
          break;
        case RR_OP_SQRT:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = f_sqrt(acc[r]);
          break;
        default: {  // SQR
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const VE a = acc[r];
            if (is_certain_nan(a.v, a.e)) { acc[r] = {a.v, 0.f}; continue; }
            const float v = a.v * a.v;
            if (!(v < FINF)) { acc[r] = {v, FINF}; continue; }
            acc[r] = {v, fix_e((2.f * fabsf(a.v) * a.e + a.e * a.e) * (1.f + 4.f * U32) + v * U32)};
          }
          break;
        }
      }
    }
  }
}

struct Fp32Args {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const float* X32;
  const double* y;
  int64_t n, ldx;
  double sst, ysum;
  double* coefs;
  rr_fit* fit;
  const int32_t* order;   // candidates of this launch (indices into the batch)
};

// reasons a candidate is left to the FP64 form (in rr_fit.rank, negated)
enum { NR_UNCERTAIN = 1, NR_SHORT = 2, NR_NAN_COND = 3, NR_COND = 4, NR_THRESHOLD = 5, NR_BOUND = 6, NR_WIDE = 7 };

__device__ inline void write_needs64(rr_fit* f, int reason) {
  rr_fit o;
  o.e0 = 0; o.e1 = 0; o.err = __builtin_huge_val(); o.rank = -reason; o.flags = RR_F_NEEDS64;
  *f = o;
}

// LDS scratch of the per-candidate solve
template <int KB>
struct Solve {
  double M[KB][KB];    // Gram matrix, then Cholesky factor (lower)
  double Ginv[KB][KB];
  double g[KB], c[KB], w[KB];
  int perm[KB];
};

// Pivoted Cholesky of the k x k matrix in S.M (ColPivHouseholderQR pivot
// rule); returns false when a pivot fails the conditioning checks.  Only
// columns with use[j] take part (the others are left out entirely).
template <int KB>
__device__ inline bool chol(Solve<KB>& S, int k, double mx2) {
  double diag0[KB];
  for (int j = 0; j < k; ++j) { S.perm[j] = j; diag0[j] = S.M[j][j]; }
  for (int p = 0; p < k; ++p) {
    int bi = p;
    double best = S.M[S.perm[p]][S.perm[p]];
    for (int j = p + 1; j < k; ++j) { const double v = S.M[S.perm[j]][S.perm[j]]; if (v > best) { best = v; bi = j; } }
    const int t = S.perm[p]; S.perm[p] = S.perm[bi]; S.perm[bi] = t;
    const int q = S.perm[p];
    if (!(best > TAU_C * diag0[q]) || !(best > TAU_S * mx2)) return false;
    const double l = __builtin_sqrt(best);
    S.M[q][q] = l;
    for (int j = p + 1; j < k; ++j) { const int a = S.perm[j]; S.M[a][q] = S.M[a][q] / l; }
    for (int j = p + 1; j < k; ++j) {
      const int a = S.perm[j];
      for (int i = p + 1; i <= j; ++i) { const int b = S.perm[i]; S.M[a][b] -= S.M[a][q] * S.M[b][q]; }
      for (int i = p + 1; i <= j; ++i) { const int b = S.perm[i]; S.M[b][a] = S.M[a][b]; }
    }
  }
  return true;
}

// x = G^-1 b using the pivoted factor (lower L in S.M over permuted indices)
template <int KB>
__device__ inline void chol_solve(const Solve<KB>& S, int k, const double* b, double* x) {
  double z[KB];
  for (int p = 0; p < k; ++p) {            // L z = P^T b
    const int q = S.perm[p];
    double s = b[q];
    for (int i = 0; i < p; ++i) s -= S.M[q][S.perm[i]] * z[i];
    z[p] = s / S.M[q][q];
  }
  for (int p = k - 1; p >= 0; --p) {       // L^T w = z
    const int q = S.perm[p];
    double s = z[p];
    for (int i = p + 1; i < k; ++i) s -= S.M[S.perm[i]][q] * z[i];
    z[p] = s / S.M[q][q];
  }
  for (int p = 0; p < k; ++p) x[S.perm[p]] = z[p];
}

template <int KB, int R>
__global__ __launch_bounds__(64) void k_ols_fp32(Fp32Args a) {
  constexpr int NF = KB - 1;
  __shared__ Solve<KB> S;
  const int cand = a.order[blockIdx.x];
  const int lane = (int)threadIdx.x;
  const int pbeg = a.cand_start[cand];
  const int nf = a.cand_start[cand + 1] - pbeg;
  const int k = nf + 1;
  const int64_t n = a.n, ldx = a.ldx;
  double* const co = a.coefs + (int64_t)pbeg + cand;
  if (n < k) { if (lane == 0) write_needs64(a.fit + cand, NR_SHORT); return; }

  const uint32_t* codes[NF];
  int lens[NF];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    if (j < nf) {
      codes[j] = a.code + a.prog_start[pbeg + j];
      lens[j] = a.prog_start[pbeg + j + 1] - a.prog_start[pbeg + j];
    } else {
      codes[j] = a.code;
      lens[j] = 0;
    }
  }

  // ---- pass 1: Gram of the factors, factor sums, A^T y; uncertain / NaN columns
  double G[NF * (NF + 1) / 2], S1[NF], SY[NF];
#pragma unroll
  for (int j = 0; j < NF * (NF + 1) / 2; ++j) G[j] = 0.0;
#pragma unroll
  for (int j = 0; j < NF; ++j) { S1[j] = 0.0; SY[j] = 0.0; }
  uint32_t unc = 0, nanc = 0;
  bool nan0_tail = false;
  for (int64_t base = 0; base < n; base += 64 * R) {
    const int64_t row0 = base + lane;
    double v[NF][R];
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      if (j < nf) {
        VE f[R];
        eval_f32<R>(codes[j], lens[j], a.K, a.X32, ldx, row0, f);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const bool ok = row0 + 64 * r < n;
          if (ok && f[r].e == FINF) unc |= 1u << j;
          if (ok && is_certain_nan(f[r].v, f[r].e)) {
            nanc |= 1u << j;
            if (j == 0 && row0 + 64 * r >= 1) nan0_tail = true;
          }
          v[j][r] = (double)f[r].v;
        }
      }
    }
    // A certain NaN of factor 0 at a row >= 1 fixes the FP64 outcome: its
    // column norm is NaN, so Eigen's threshold is NaN (no rank cut), factor 0
    // pivots first, its tail norm is NaN and so is its reflector; every
    // coefficient is then NaN.  The rest of the pass is not needed.
    if (__any((int)nan0_tail)) {
      for (int i = lane; i < k; i += WAVE) co[i] = rrm::qnan();
      if (lane == 0) {
        rr_fit o;
        o.e0 = 1000; o.e1 = 1000; o.err = 0.0; o.rank = k; o.flags = RR_F_SENTINEL;
        a.fit[cand] = o;
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = row0 + 64 * r;
      if (i < n) {
        const double yi = a.y[i];
        int g = 0;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
#pragma unroll
          for (int l = 0; l <= j; ++l) {
            if (j < nf) G[g] = G[g] + v[j][r] * v[l][r];
            ++g;
          }
          if (j < nf) { S1[j] = S1[j] + v[j][r]; SY[j] = SY[j] + v[j][r] * yi; }
        }
      }
    }
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) { unc |= __shfl_xor(unc, s); nanc |= __shfl_xor(nanc, s); }
  if (unc) { if (lane == 0) write_needs64(a.fit + cand, NR_UNCERTAIN); return; }
#pragma unroll
  for (int j = 0; j < NF * (NF + 1) / 2; ++j) G[j] = wave_sum(G[j]);
#pragma unroll
  for (int j = 0; j < NF; ++j) { S1[j] = wave_sum(S1[j]); SY[j] = wave_sum(SY[j]); }

  // full Gram of [factors | 1] in LDS (every lane holds the same values)
  if (lane == 0) {
    int g = 0;
#pragma unroll
    for (int j = 0; j < NF; ++j) {
#pragma unroll
      for (int l = 0; l <= j; ++l) { if (j < nf) { S.M[j][l] = G[g]; S.M[l][j] = G[g]; } ++g; }
      if (j < nf) { S.M[j][nf] = S1[j]; S.M[nf][j] = S1[j]; S.g[j] = SY[j]; }
    }
    S.M[nf][nf] = (double)n;
    S.g[nf] = a.ysum;
  }
  __syncthreads();

  if (nanc) {
    // The FP64 QR pivots the first NaN factor column j0 at step j0 (a NaN norm
    // never wins the pivot search and stays in place); a cut before that step
    // needs a near-dependent finite column.  With factor 0 NaN the threshold
    // is NaN and no cut can happen at all.
    bool certain = (nanc & 1u) != 0;
    if (!certain && lane == 0) {
      // Gram of the finite columns only
      int idx[KB];
      int m = 0;
      for (int j = 0; j < k; ++j) if (j == nf || !((nanc >> j) & 1u)) idx[m++] = j;
      double T[KB][KB];
      for (int i = 0; i < m; ++i) for (int j = 0; j < m; ++j) T[i][j] = S.M[idx[i]][idx[j]];
      for (int i = 0; i < m; ++i) for (int j = 0; j < m; ++j) S.M[i][j] = T[i][j];
      double mx2 = 0.0;
      for (int i = 0; i < m; ++i) mx2 = mx2 < S.M[i][i] ? S.M[i][i] : mx2;
      S.w[0] = chol<KB>(S, m, mx2) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (!certain) certain = S.w[0] == 1.0;
    if (!certain) { if (lane == 0) write_needs64(a.fit + cand, NR_NAN_COND); return; }
    for (int i = lane; i < k; i += WAVE) co[i] = rrm::qnan();
    if (lane == 0) {
      rr_fit o;
      o.e0 = 1000; o.e1 = 1000; o.err = 0.0; o.rank = k; o.flags = RR_F_SENTINEL;
      a.fit[cand] = o;
    }
    return;
  }

  // ---- solve G c = g
  if (lane == 0) {
    double mx2 = 0.0;
    for (int j = 0; j < k; ++j) mx2 = mx2 < S.M[j][j] ? S.M[j][j] : mx2;
    double Gs[KB][KB];
    for (int i = 0; i < k; ++i) for (int j = 0; j < k; ++j) Gs[i][j] = S.M[i][j];
    const bool ok = chol<KB>(S, k, mx2);
    S.w[0] = ok ? 1.0 : 0.0;
    if (ok) {
      chol_solve<KB>(S, k, S.g, S.c);
      double e[KB], col[KB];
      for (int j = 0; j < k; ++j) {          // explicit inverse for the error bound
        for (int i = 0; i < k; ++i) e[i] = (i == j) ? 1.0 : 0.0;
        chol_solve<KB>(S, k, e, col);
        for (int i = 0; i < k; ++i) S.Ginv[i][j] = col[i];
      }
    }
    for (int i = 0; i < k; ++i) for (int j = 0; j < k; ++j) S.M[i][j] = Gs[i][j];
  }
  __syncthreads();
  if (S.w[0] != 1.0) { if (lane == 0) write_needs64(a.fit + cand, NR_COND); return; }
  double c[KB];
#pragma unroll
  for (int j = 0; j < KB; ++j) c[j] = (j < k) ? S.c[j] : 0.0;
  const double cfree = S.c[nf];

  // ---- pass 2: residual, A^T r, SSR and the bound terms
  double ssr = 0.0, atr_free = 0.0, d2 = 0.0, aad_free = 0.0;
  double ATR[NF], BER[NF], AAD[NF], BE2[NF];
#pragma unroll
  for (int j = 0; j < NF; ++j) { ATR[j] = 0.0; BER[j] = 0.0; AAD[j] = 0.0; BE2[j] = 0.0; }
  for (int64_t base = 0; base < n; base += 64 * R) {
    const int64_t row0 = base + lane;
    double v[NF][R];
    float e[NF][R];
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      if (j < nf) {
        VE f[R];
        eval_f32<R>(codes[j], lens[j], a.K, a.X32, ldx, row0, f);
#pragma unroll
        for (int r = 0; r < R; ++r) { v[j][r] = (double)f[r].v; e[j][r] = f[r].e; }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = row0 + 64 * r;
      if (i < n) {
        double yp = cfree, dl = 0.0;
#pragma unroll
        for (int j = 0; j < NF; ++j)
          if (j < nf) { yp = yp + c[j] * v[j][r]; dl = dl + __builtin_fabs(c[j]) * (double)e[j][r]; }
        const double res = a.y[i] - yp;
        ssr = ssr + res * res;
        atr_free = atr_free + res;
        d2 = d2 + dl * dl;
        aad_free = aad_free + dl;
        const double ar = __builtin_fabs(res);
#pragma unroll
        for (int j = 0; j < NF; ++j) {
          if (j < nf) {
            const double ej = (double)e[j][r];
            ATR[j] = ATR[j] + v[j][r] * res;
            BER[j] = BER[j] + ej * ar;
            AAD[j] = AAD[j] + __builtin_fabs(v[j][r]) * dl;
            BE2[j] = BE2[j] + ej * ej;
          }
        }
      }
    }
  }
  ssr = wave_sum(ssr); atr_free = wave_sum(atr_free); d2 = wave_sum(d2); aad_free = wave_sum(aad_free);
#pragma unroll
  for (int j = 0; j < NF; ++j) { ATR[j] = wave_sum(ATR[j]); BER[j] = wave_sum(BER[j]); AAD[j] = wave_sum(AAD[j]); BE2[j] = wave_sum(BE2[j]); }

  if (lane == 0) {
    double atr[KB], vv[KB], dc[KB], cr[KB], be[KB];
    for (int j = 0; j < k; ++j) {
      atr[j] = (j == nf) ? atr_free : ATR[j];
      vv[j] = (j == nf) ? aad_free : BER[j] + AAD[j];
      be[j] = (j == nf) ? 0.0 : __builtin_sqrt(BE2[j]);
    }
    // refinement step
    double ref[KB];
    for (int i = 0; i < k; ++i) { double s = 0.0; for (int j = 0; j < k; ++j) s += S.Ginv[i][j] * atr[j]; ref[i] = s; }
    double dssr = 0.0;
    for (int i = 0; i < k; ++i) { cr[i] = c[i] + ref[i]; dssr += ref[i] * atr[i]; }
    double ssr2 = ssr - dssr;
    if (ssr2 < 0.0) ssr2 = 0.0;
    // coefficient bound and the bound on sqrt(SSR)
    double D = __builtin_sqrt(d2);
    bool ambiguous = false;
    for (int i = 0; i < k; ++i) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += __builtin_fabs(S.Ginv[i][j]) * vv[j];
      dc[i] = 2.0 * s + 1e-12 * __builtin_fabs(cr[i]) + 2.0 * __builtin_fabs(ref[i]);
      D += (dc[i] + __builtin_fabs(ref[i])) * be[i] + EPS12 * __builtin_sqrt(S.M[i][i]) * (1.0 + __builtin_fabs(cr[i]));
      const double ac = __builtin_fabs(cr[i]);
      if (!(ac - dc[i] >= EPS12 || ac + dc[i] < EPS12)) ambiguous = true;
      const double a1 = __builtin_fabs(cr[i] - 1.0);
      if (i != nf && !(a1 - dc[i] >= EPS12 || a1 + dc[i] < EPS12)) ambiguous = true;
    }
    for (int i = 0; i < k; ++i) co[i] = cr[i];
    rr_fit o;
    const double rs = __builtin_sqrt(ssr2);
    const double slack = 1e-10;
    const double hi = (rs + D) * (rs + D) * (1.0 + slack) + 1e-300;
    const double lo = (rs > D) ? (rs - D) * (rs - D) * (1.0 - slack) : 0.0;
    const double r2 = 1 - ssr2 / a.sst;
    o.e0 = 1 - r2;
    o.e1 = __builtin_sqrt(ssr2 / (double)n);
    const double e0h = hi / a.sst, e0l = lo / a.sst;
    const double e1h = __builtin_sqrt(hi / (double)n), e1l = __builtin_sqrt(lo / (double)n);
    const double fv = (1 + o.e0) * (1 + o.e1);
    const double up = (1 + e0h) * (1 + e1h) / fv - 1.0, dn = 1.0 - (1 + e0l) * (1 + e1l) / fv;
    o.err = (up > dn ? up : dn) + 1e-12;
    o.rank = k;
    o.flags = 0;
    int reason = ambiguous ? NR_THRESHOLD : 0;
    if (r2 != r2 || o.e1 != o.e1 || !(o.err < 1.0)) reason = NR_BOUND;
    if (reason) { o.flags = RR_F_NEEDS64; o.err = __builtin_huge_val(); o.rank = -reason; }
    a.fit[cand] = o;
  }
}

// raw scoring of one tree in FP32 with the same certification
template <int R>
__global__ __launch_bounds__(64) void k_raw_fp32(Fp32Args a) {
  const int cand = a.order[blockIdx.x];
  const int lane = (int)threadIdx.x;
  const int p = a.cand_start[cand];
  const uint32_t* code = a.code + a.prog_start[p];
  const int len = a.prog_start[p + 1] - a.prog_start[p];
  double ssr = 0.0, d2 = 0.0;
  uint32_t unc = 0, nanc = 0;
  for (int64_t base = 0; base < a.n; base += 64 * R) {
    const int64_t row0 = base + lane;
    VE f[R];
    eval_f32<R>(code, len, a.K, a.X32, a.ldx, row0, f);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = row0 + 64 * r;
      if (i < a.n) {
        if (f[r].e == FINF) unc = 1;
        if (is_certain_nan(f[r].v, f[r].e)) nanc = 1;
        const double d = a.y[i] - (double)f[r].v;
        ssr = ssr + d * d;
        d2 = d2 + (double)f[r].e * (double)f[r].e;
      }
    }
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) { unc |= __shfl_xor(unc, s); nanc |= __shfl_xor(nanc, s); }
  ssr = wave_sum(ssr);
  d2 = wave_sum(d2);
  if (lane != 0) return;
  rr_fit o;
  o.rank = 0;
  if (unc) { write_needs64(a.fit + cand, NR_UNCERTAIN); return; }
  if (nanc) { o.e0 = 1000; o.e1 = 1000; o.err = 0.0; o.flags = RR_F_SENTINEL; a.fit[cand] = o; return; }
  const double r2 = 1 - ssr / a.sst;
  o.e0 = 1 - r2;
  o.e1 = __builtin_sqrt(ssr / (double)a.n);
  const double rs = __builtin_sqrt(ssr), D = __builtin_sqrt(d2);
  const double hi = (rs + D) * (rs + D) * (1.0 + 1e-10), lo = rs > D ? (rs - D) * (rs - D) * (1.0 - 1e-10) : 0.0;
  const double fv = (1 + o.e0) * (1 + o.e1);
  const double up = (1 + hi / a.sst) * (1 + __builtin_sqrt(hi / (double)a.n)) / fv - 1.0;
  const double dn = 1.0 - (1 + lo / a.sst) * (1 + __builtin_sqrt(lo / (double)a.n)) / fv;
  o.err = (up > dn ? up : dn) + 1e-12;
  o.flags = (r2 != r2 || o.e1 != o.e1 || !(o.err < 1.0)) ? RR_F_NEEDS64 : 0;
  if (o.flags) o.err = __builtin_huge_val();
  a.fit[cand] = o;
}

__global__ void k_to_fp32(const double* src, float* dst, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (float)src[i];
}

__global__ void k_needs64(const int32_t* order, int count, rr_fit* fit) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < count) write_needs64(fit + order[i], NR_WIDE);
}

}  // namespace

hipError_t launch_to_fp32(const double* src, float* dst, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_fp32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_fp32(const Fp32Launch& L, const int32_t* order, const int* bucket_count, hipStream_t s) {
  Fp32Args a;
  a.cand_start = L.cand_start; a.prog_start = L.prog_start; a.code = L.code; a.K = L.K; a.X32 = L.X32;
  a.y = L.y; a.n = L.n; a.ldx = L.ldx; a.sst = L.sst; a.ysum = L.ysum; a.coefs = L.coefs; a.fit = L.fit;
  // order lists candidates grouped by bucket: k<=2, k<=5, k<=9, larger (no FP32 form)
  int off = 0;
  if (L.raw) {
    a.order = order;
    if (L.n_cands > 0) hipLaunchKernelGGL(k_raw_fp32<4>, dim3(L.n_cands), dim3(64), 0, s, a);
    return hipGetLastError();
  }
  if (bucket_count[0] > 0) { a.order = order + off; hipLaunchKernelGGL((k_ols_fp32<2, 4>), dim3(bucket_count[0]), dim3(64), 0, s, a); }
  off += bucket_count[0];
  if (bucket_count[1] > 0) { a.order = order + off; hipLaunchKernelGGL((k_ols_fp32<5, 2>), dim3(bucket_count[1]), dim3(64), 0, s, a); }
  off += bucket_count[1];
  if (bucket_count[2] > 0) { a.order = order + off; hipLaunchKernelGGL((k_ols_fp32<9, 1>), dim3(bucket_count[2]), dim3(64), 0, s, a); }
  off += bucket_count[2];
  if (bucket_count[3] > 0)
    hipLaunchKernelGGL(k_needs64, dim3((bucket_count[3] + 255) / 256), dim3(256), 0, s, order + off, bucket_count[3], L.fit);
  return hipGetLastError();
}

}  // namespace rr
