// rr_kernels_fp32.hip -- FP32 precision form of the candidate-fitness path.
//
// On this hardware plain FP32 and FP64 vector arithmetic issue at the same
// rate; FP32 is cheaper only through the hardware transcendental, reciprocal
// and square-root instructions and through moving half the bytes.  The form
// is built around that: ONE pass over the samples, split into row chunks --
// k_acc_fp32 gives each (candidate, chunk) a 4-wave workgroup that evaluates
// the factor programs in FP32 with the hardware instructions for R rows per
// lane, while every value carries a running bound e >= |v32 - v64| on its
// distance from the FP64 form's value of the same expression, and k_fin_fp32
// (one wave per candidate) adds the chunk partials in chunk order.  The pass
// accumulates, per candidate,
//     G  = A^T A   (factors; FP64 -- a product of two FP32 numbers is exact)
//     S1 = column sums, SY = A^T y, Bd = sum_i e_ij^2       (FP64)
// and then, on one lane:
//   solve   FP64 pivoted Cholesky of the Gram matrix of [factors | 1] (the
//           pivot rule of ColPivHouseholderQR: largest remaining column, first
//           index on ties), one refinement step on the normal equations;
//   SSR     yy - 2 c.g + c^T G c (its rounding counted in the bound);
//   bound   |sqrt(SSR32) - sqrt(SSR64)| <= D,
//             D = sum_j |c_j| sqrt(Bd_j)                    (||dA |c|||)
//               + sum_j dc_j sqrt(Bd_j)                     (coefficient shift)
//               + sum_j 1e-12 sqrt(G_jj) (1 + |c_j|)        (OLS-tree thresholds)
//               + 2 sqrt(dcr^T G dcr)                       (solve residue)
//           dc = 2 |G^-1| (|dA|^T |r| + |A|^T |dA| |c|) + ...   (first-order
//           least-squares perturbation, doubled), with the products bounded by
//           Cauchy-Schwarz from G, Bd and y.y:
//             (|dA|^T|r|)_j    <= sqrt(Bd_j) (sqrt(yy) + sum_l |c_l| sqrt(G_ll))
//             (|A|^T|dA||c|)_j <= sqrt(G_jj) sum_l |c_l| sqrt(Bd_l)
// Candidates that cannot be certified are flagged RR_F_NEEDS64 (reason in
// rr_fit.rank, negated) and re-scored by the FP64 form: a value whose
// finiteness FP32 and FP64 can disagree on (overflow, a domain boundary
// inside the bound); a rank decision (Schur complement below tau_c of its
// column); a coefficient whose |c| < 1e-12 or |c - 1| < 1e-12 decision lies
// inside its bound; more than 8 factors; n < k.
//
// NaN: +, -, *, / and the unary operators give NaN for a NaN argument, so a
// domain violation that is certain (sqrt or ln of an argument certainly
// negative) makes the FP64 value NaN at that sample; the row state carries
// that.  The QR outcome with certain-NaN columns is decided as in the FP64
// form (see k_fin_fp32).  MIN and MAX can absorb a NaN: under them such a row
// becomes uncertain; comparisons give 0 or 1 either way (bound 1).
//
// Operator bounds (u = 2^-23, twice the FP32 unit roundoff; T = 2^-125 covers
// a result that underflows, flushed or not):
//   + -   e = ea + eb + |v| u
//   *     e = (|a| + ea) eb + |b| ea + |v| u + T
//   /     e = (|v| eb + ea) / (|b| - eb) (1 + 8u) + 2 |v| u + T
//                                              (needs |b| > eb, |b| < 2^100)
//   sqr   e = (2|a| + ea) ea + |v| u + T
//   sqrt  e = ea / (sqrt(a) + sqrt(a - ea)) (1 + 4u) + 2 |v| u   (a > ea or exact 0)
//   ln    e = ea / (a - ea) (1 + 4u) + H_LOG (1 + |v|)            (a > ea)
//   exp   e = v (ea (1 + ea) + 2|a| u + H_EXP) + 2^-124           (ea <= 1)
//   sin, cos  e = min(2|a| u + ea, 2) + H_TRIG   (|a| < 1600; beyond: v = 0, e = 1)
// The 2|a|u terms are the rounding of the argument scaling (x / 2pi, x log2 e)
// the hardware instructions need.  H_* bound the instructions' own errors
// against the FP64 functions; the GPU test suite measures them.  The bound
// arithmetic itself rounds (relative 3 * 2^-24 per operation at most), which
// the final D absorbs through a factor 1 + 2^-8 (programs of up to ~20000
// operations).
#include "rr_device.h"
#include "rr_kernels_fp32.h"

namespace rr {

namespace {

constexpr float U32 = 0x1p-23f;
constexpr float FINF = __builtin_huge_valf();
constexpr float TINY = 0x1p-125f;
constexpr float H_EXP = 0x1p-20f;     // relative
constexpr float H_LOG = 0x1p-20f;     // absolute, scaled by (1 + |v|)
constexpr float H_TRIG = 0x1p-19f;    // absolute
constexpr float TRIG_RANGE = 1600.f;  // hardware sin/cos take |x / 2pi| <= 256
// smallest accepted Schur complement / column norm^2: TAU_C_REL times the
// largest relative RMS error of a column (so that the FP32 perturbation of
// every pivot stays a few percent of it and the first-order bound applies),
// and never below TAU_C_MIN, far above the FP64 QR's own rank cut (~1e-31)
constexpr double TAU_C_REL = 64.0;
constexpr double TAU_C_MIN = 1e-9;
constexpr double TAU_S = 1e-20;       // smallest accepted Schur complement / (max norm)^2
constexpr double EPS12 = 1e-12;
constexpr double D_ROUND = 1.0 + 0x1p-8;
constexpr float INV_2PI = 0.15915494309189535f;
constexpr float LOG2E = 1.4426950408889634f;
constexpr float LN2 = 0.6931471805599453f;

// row state bits
constexpr uint32_t S_NAN = 1u;   // the FP64 value is NaN at this sample
constexpr uint32_t S_UNC = 2u;   // FP32 cannot vouch for this sample

struct VE { float v, e; };

__device__ inline float hw_exp(float x) { return __builtin_amdgcn_exp2f(x * LOG2E); }
__device__ inline float hw_log(float x) { return __builtin_amdgcn_logf(x) * LN2; }
__device__ inline float hw_sin(float x) { return __builtin_amdgcn_sinf(x * INV_2PI); }
__device__ inline float hw_cos(float x) { return __builtin_amdgcn_cosf(x * INV_2PI); }
__device__ inline float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline VE op_bin(uint32_t bop, VE a, VE b, uint32_t& st) {
  switch (bop) {
    case RR_BOP_ADD: { const float v = a.v + b.v; return {v, fmaf(fabsf(v), U32, a.e + b.e)}; }
    case RR_BOP_SUB: { const float v = a.v - b.v; return {v, fmaf(fabsf(v), U32, a.e + b.e)}; }
    case RR_BOP_MUL: {
      const float v = a.v * b.v;
      return {v, fmaf(fabsf(a.v) + a.e, b.e, fmaf(fabsf(b.v), a.e, fmaf(fabsf(v), U32, TINY)))};
    }
    case RR_BOP_DIV: {
      // |a/b - a'/b'| <= (|a| eb + |b| ea) / (|b| (|b| - eb)) = (|a/b| eb + ea) / (|b| - eb);
      // this form cannot overflow where the bound itself does not
      const float v = a.v * hw_rcp(b.v);
      const float ab = fabsf(b.v);
      const float bl = ab - b.e * (1.f + U32);
      const float e = fmaf(fmaf(fabsf(v), b.e, a.e) * hw_rcp(bl), 1.f + 8.f * U32, fmaf(fabsf(v), 2.f * U32, TINY));
      return {v, (bl > 0.f && ab < 0x1p100f) ? e : FINF};
    }
    case RR_BOP_MIN: case RR_BOP_MAX: {
      if (st & S_NAN) st = (st & ~S_NAN) | S_UNC;   // min(NaN, b) = b, min(a, NaN) = NaN
      const bool ta = (bop == RR_BOP_MIN) ? (a.v < b.v) : (a.v > b.v);
      const VE r = ta ? a : b;
      const float d = fabsf(a.v - b.v);
      const bool sep = d > (a.e + b.e) * (1.f + 4.f * U32);
      return {r.v, sep ? r.e : fmaxf(a.e, b.e) + d};
    }
    default: {   // comparisons: 0 or 1 in both precisions, whatever the arguments
      st &= ~S_NAN;
      float v;
      if (bop == RR_BOP_LT) v = a.v < b.v ? 1.f : 0.f;
      else if (bop == RR_BOP_GT) v = a.v > b.v ? 1.f : 0.f;
      else if (bop == RR_BOP_EQ) v = a.v == b.v ? 1.f : 0.f;
      else v = a.v != b.v ? 1.f : 0.f;
      const bool sep = fabsf(a.v - b.v) > (a.e + b.e) * (1.f + 4.f * U32);
      const bool exact = a.e == 0.f && b.e == 0.f && !(st & S_UNC);
      return {v, (sep || exact) && !(st & S_UNC) ? 0.f : 1.f};
    }
  }
}

__device__ inline VE op_un(uint32_t op, VE a, uint32_t& st) {
  switch (op) {
    case RR_OP_SIN: case RR_OP_COS: {
      const float aa = fabsf(a.v);
      const bool in = aa < TRIG_RANGE;
      const float v = in ? ((op == RR_OP_SIN) ? hw_sin(a.v) : hw_cos(a.v)) : 0.f;
      const float e = in ? fminf(fmaf(aa, 2.f * U32, a.e), 2.f) + H_TRIG : 1.f;
      return {v, a.e < FINF ? e : FINF};
    }
    case RR_OP_EXP: {
      const float v = hw_exp(a.v);
      const float e = fmaf(v, fmaf(a.e, 1.f + a.e, fmaf(fabsf(a.v), 2.f * U32, H_EXP)), 0x1p-124f);
      return {v, a.e <= 1.f ? e : FINF};
    }
    case RR_OP_LN: {
      const float slack = a.e * (1.f + U32);
      const float lo = a.v - slack;
      if (a.e < FINF && a.v + slack < 0.f) st |= S_NAN;   // argument certainly negative
      const float v = hw_log(a.v);
      const float e = fmaf(a.e * hw_rcp(lo), 1.f + 4.f * U32, H_LOG * (1.f + fabsf(v)));
      return {v, lo > 0.f ? e : FINF};
    }
    case RR_OP_SQRT: {
      const float slack = a.e * (1.f + U32);
      const float lo = a.v - slack;
      if (a.e < FINF && a.v + slack < 0.f) st |= S_NAN;
      const float v = hw_sqrt(a.v);
      const float den = v + hw_sqrt(fmaxf(lo, 0.f));
      const float e = fmaf(a.e * hw_rcp(den), 1.f + 4.f * U32, 2.f * v * U32);
      return {v, lo > 0.f ? e : ((a.v == 0.f && a.e == 0.f) ? 0.f : FINF)};
    }
    default: {  // SQR
      const float v = a.v * a.v;
      return {v, fmaf(2.f * fabsf(a.v) + a.e, a.e, fmaf(v, U32, TINY))};
    }
  }
}

// A constant of the program as an FP32 value with its bound (a nonzero
// constant below 2^-100 in magnitude is left uncertain: its conversion can
// underflow).
__device__ inline VE const_ve(double kd) {
  const float kf = (float)kd;
  const double ad = __builtin_fabs(kd);
  const bool ok = ad == 0.0 || ad >= 0x1p-100;
  return {kf, ok ? fabsf(kf) * U32 : FINF};
}

// FP32 evaluation of one program at R samples (rows row0 + 64 r); st[r]
// receives the row state.  X holds the FP32 copy of the features (tiny
// nonzero and non-finite entries stored as NaN, see k_to_fp32).
template <int R>
__device__ inline void eval_f32(const uint32_t* __restrict__ code, int len, const double* __restrict__ K,
                                const float* __restrict__ X, int64_t ldx, int64_t row0, VE (&acc)[R],
                                uint32_t (&st)[R]) {
  VE s0[R], s1[R], sx[RR_MAX_STACK - 2][R];
  uint32_t t0[R], t1[R], tx[RR_MAX_STACK - 2][R];
#pragma unroll
  for (int r = 0; r < R; ++r) { acc[r] = {0.f, 0.f}; st[r] = 0; }
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
        uint32_t t[R];
        if (sp == 0) {
#pragma unroll
          for (int r = 0; r < R; ++r) { s[r] = s0[r]; t[r] = t0[r]; }
        } else if (sp == 1) {
#pragma unroll
          for (int r = 0; r < R; ++r) { s[r] = s1[r]; t[r] = t1[r]; }
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) { s[r] = sx[sp - 2][r]; t[r] = tx[sp - 2][r]; }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          l[r] = mode == RR_M_S ? s[r] : acc[r];
          rr[r] = mode == RR_M_S ? acc[r] : s[r];
          st[r] |= t[r];
        }
      } else if (mode <= RR_M_VR) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float x = X[(int64_t)arg * ldx + row0 + 64 * r];
          const VE xv{x, fabsf(x) * U32};
          l[r] = mode == RR_M_V ? acc[r] : xv;
          rr[r] = mode == RR_M_V ? xv : acc[r];
        }
      } else {
        const VE kv = const_ve(K[arg]);
#pragma unroll
        for (int r = 0; r < R; ++r) { l[r] = mode == RR_M_C ? acc[r] : kv; rr[r] = mode == RR_M_C ? kv : acc[r]; }
      }
      switch (bop) {   // one uniform branch per instruction, not per row
#define RR_BIN(B) case B: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = op_bin(B, l[r], rr[r], st[r]); break;
        RR_BIN(RR_BOP_ADD) RR_BIN(RR_BOP_SUB) RR_BIN(RR_BOP_MUL) RR_BIN(RR_BOP_DIV) RR_BIN(RR_BOP_LT)
        RR_BIN(RR_BOP_GT) RR_BIN(RR_BOP_EQ) RR_BIN(RR_BOP_NEQ) RR_BIN(RR_BOP_MIN)
        default: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = op_bin(RR_BOP_MAX, l[r], rr[r], st[r]); break;
#undef RR_BIN
      }
    } else {
      switch (op) {
        case RR_OP_LDV:
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const float x = X[(int64_t)arg * ldx + row0 + 64 * r];
            acc[r] = {x, fabsf(x) * U32};
            st[r] = 0;
          }
          break;
        case RR_OP_LDC: {
          const VE kv = const_ve(K[arg]);
#pragma unroll
          for (int r = 0; r < R; ++r) { acc[r] = kv; st[r] = 0; }
          break;
        }
        case RR_OP_PUSH:
          if (sp == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) { s0[r] = acc[r]; t0[r] = st[r]; }
          } else if (sp == 1) {
#pragma unroll
            for (int r = 0; r < R; ++r) { s1[r] = acc[r]; t1[r] = st[r]; }
          } else {
#pragma unroll
            for (int r = 0; r < R; ++r) { sx[sp - 2][r] = acc[r]; tx[sp - 2][r] = st[r]; }
          }
          ++sp;
          break;
#define RR_UN(O) case O: _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = op_un(O, acc[r], st[r]); break;
        RR_UN(RR_OP_SIN) RR_UN(RR_OP_COS) RR_UN(RR_OP_LN) RR_UN(RR_OP_EXP) RR_UN(RR_OP_SQRT)
#undef RR_UN
        default:
#pragma unroll
          for (int r = 0; r < R; ++r) acc[r] = op_un(RR_OP_SQR, acc[r], st[r]);
          break;
      }
    }
  }
  // final state: NaN (certain), uncertain, or certified value
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (st[r] & S_UNC) st[r] = S_UNC;
    else if (!(st[r] & S_NAN) && !(acc[r].e < FINF && acc[r].v - acc[r].v == 0.f)) st[r] = S_UNC;
  }
}

struct Fp32Args {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  // the candidate's dataset, bound at kernel start
  const float* X32;
  const double* y;
  int64_t n, ldx;
  double sst, ysum, yy;
  double* coefs;
  rr_fit* fit;
  const int32_t* order;   // candidates of this launch (indices into the batch)
  const int32_t* dead;    // per candidate: outcome written by the factor-0 pre-pass, or null
};

__device__ inline void bind_dataset(Fp32Args& a, int cand) {
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  a.X32 = D.X32; a.y = D.y; a.n = D.n; a.ldx = D.ldx; a.sst = D.sst; a.ysum = D.ysum; a.yy = D.yy;
}

// reasons a candidate is left to the FP64 form (in rr_fit.rank, negated)
enum { NR_UNCERTAIN = 1, NR_SHORT = 2, NR_NAN_COND = 3, NR_COND = 4, NR_THRESHOLD = 5, NR_BOUND = 6, NR_WIDE = 7,
       NR_R2 = 8 };
// the largest certified |R2_32 - R2_64|: a bound above it sends the candidate to FP64
constexpr double R2_TOL = 1e-9;

__device__ inline void write_needs64(rr_fit* f, int reason) {
  rr_fit o;
  o.e0 = 0; o.e1 = 0; o.err = __builtin_huge_val(); o.rank = -reason; o.flags = RR_F_NEEDS64;
  *f = o;
}

// the same with the candidate's k coefficients set to NaN (one lane)
__device__ inline void write_needs64(rr_fit* f, double* co, int k, int reason) {
  for (int i = 0; i < k; ++i) co[i] = rrm::qnan();
  write_needs64(f, reason);
}

__device__ inline void write_sentinel(rr_fit* f, int rank) {
  rr_fit o;
  o.e0 = 1000; o.e1 = 1000; o.err = 0.0; o.rank = rank; o.flags = RR_F_SENTINEL;
  *f = o;
}

// LDS scratch of the per-candidate solve
template <int KB>
struct Solve {
  double M[KB][KB];    // Gram matrix of [factors | 1]
  double L[KB][KB];    // pivoted Cholesky factor (lower, permuted indices)
  double g[KB], bd[KB];
  int perm[KB];
  int ok;
};

// Pivoted Cholesky of the m x m matrix T (pivot: largest remaining diagonal,
// first index on ties) into S.L / S.perm; false when a pivot fails the
// conditioning checks.
template <int KB>
__device__ inline bool chol(Solve<KB>& S, const double (&T)[KB][KB], int m, double tau_c) {
  double diag0[KB], mx2 = 0.0;
  for (int j = 0; j < m; ++j) {
    S.perm[j] = j;
    diag0[j] = T[j][j];
    mx2 = mx2 < T[j][j] ? T[j][j] : mx2;
    for (int i = 0; i < m; ++i) S.L[j][i] = T[j][i];
  }
  for (int p = 0; p < m; ++p) {
    int bi = p;
    double best = S.L[S.perm[p]][S.perm[p]];
    for (int j = p + 1; j < m; ++j) { const double v = S.L[S.perm[j]][S.perm[j]]; if (v > best) { best = v; bi = j; } }
    const int t = S.perm[p]; S.perm[p] = S.perm[bi]; S.perm[bi] = t;
    const int q = S.perm[p];
    if (!(best > tau_c * diag0[q]) || !(best > TAU_S * mx2)) return false;
    const double l = __builtin_sqrt(best);
    S.L[q][q] = l;
    for (int j = p + 1; j < m; ++j) { const int a = S.perm[j]; S.L[a][q] = S.L[a][q] / l; }
    for (int j = p + 1; j < m; ++j) {
      const int a = S.perm[j];
      for (int i = p + 1; i <= j; ++i) { const int b = S.perm[i]; S.L[a][b] -= S.L[a][q] * S.L[b][q]; }
      for (int i = p + 1; i <= j; ++i) { const int b = S.perm[i]; S.L[b][a] = S.L[a][b]; }
    }
  }
  return true;
}

// x = G^-1 b with the pivoted factor
template <int KB>
__device__ inline void chol_solve(const Solve<KB>& S, int k, const double* b, double* x) {
  double z[KB];
  for (int p = 0; p < k; ++p) {
    const int q = S.perm[p];
    double s = b[q];
    for (int i = 0; i < p; ++i) s -= S.L[q][S.perm[i]] * z[i];
    z[p] = s / S.L[q][q];
  }
  for (int p = k - 1; p >= 0; --p) {
    const int q = S.perm[p];
    double s = z[p];
    for (int i = p + 1; i < k; ++i) s -= S.L[S.perm[i]][q] * z[i];
    z[p] = s / S.L[q][q];
  }
  for (int p = 0; p < k; ++p) x[S.perm[p]] = z[p];
}

// Single-pass OLS fitness of the candidates with at most NF factors.
// Chunked single pass.  The samples of a candidate are cut into chunks of
// chunk_rows rows; k_acc_fp32 gives every (candidate, chunk) pair a workgroup
// of ACC_WAVES waves that evaluates the factors over the chunk and writes
// partial sums (Gram, column sums, A^T y, sum e^2) and flags to a workspace;
// k_fin_fp32 adds a candidate's chunk partials in chunk order and solves.
// Wave w of a chunk takes its tiles of 64 R rows numbered w, w + ACC_WAVES,
// ..., and partial sums are combined in wave order, so every result is fixed
// by the inputs and chunk_rows alone.
constexpr int ACC_WAVES = 4;
constexpr int ACC_STRIDE = FP32_ACC_STRIDE;   // doubles per chunk partial (>= NG + 3 NF for NF <= 8)
constexpr uint32_t CH_NAN0 = 1u << 31;   // flag word 1: factor 0 certainly NaN at a row >= 1

struct ChunkArgs {
  const int32_t* blk_cand;    // per workgroup of k_acc: candidate (batch index)
  const int32_t* blk_chunk;   // per workgroup of k_acc: chunk of that candidate
  const int32_t* chunk_base;  // per candidate: index of its first chunk in the workspace
  int64_t chunk_rows;
  double* part;               // ACC_STRIDE doubles per chunk
  uint32_t* flags;            // 2 words per chunk: uncertain columns, NaN columns (| CH_NAN0)
};

template <int NF, int R>
__global__ __launch_bounds__(64 * ACC_WAVES) void k_acc_fp32(Fp32Args a, ChunkArgs w, int blk0) {
  constexpr int NG = NF * (NF + 1) / 2;
  constexpr int NACC = NG + 3 * NF;
  static_assert(NACC <= ACC_STRIDE, "chunk partial stride");
  __shared__ double part[ACC_WAVES][NACC];
  __shared__ uint32_t part_flags[ACC_WAVES][2];
  __shared__ int abort_flag;
  const int q = blk0 + (int)blockIdx.x;
  const int cand = w.blk_cand[q];
  if (a.dead && a.dead[cand]) return;   // uniform; k_fin_fp32 skips it too
  const int64_t chunk = w.blk_chunk[q];
  bind_dataset(a, cand);
  const int lane = (int)(threadIdx.x & 63u);
  const int wave = (int)(threadIdx.x >> 6);
  const int pbeg = a.cand_start[cand];
  const int nf = a.cand_start[cand + 1] - pbeg;
  const int64_t n = a.n, ldx = a.ldx;
  const int64_t r0 = chunk * w.chunk_rows;
  const int64_t r1 = (r0 + w.chunk_rows < n) ? r0 + w.chunk_rows : n;
  if (threadIdx.x == 0) abort_flag = 0;
  __syncthreads();

  double G[NG], S1[NF], SY[NF], BD[NF];
#pragma unroll
  for (int j = 0; j < NG; ++j) G[j] = 0.0;
#pragma unroll
  for (int j = 0; j < NF; ++j) { S1[j] = 0.0; SY[j] = 0.0; BD[j] = 0.0; }
  uint32_t unc = 0, nanc = 0;
  const volatile int* abort_seen = &abort_flag;
  for (int64_t base = r0 + (int64_t)wave * 64 * R; base < r1; base += (int64_t)ACC_WAVES * 64 * R) {
    if (*abort_seen) break;
    const int64_t row0 = base + lane;
    double v[NF][R];
    float e[NF][R];
#pragma unroll
    for (int j = 0; j < NF; ++j)
#pragma unroll
      for (int r = 0; r < R; ++r) { v[j][r] = 0.0; e[j][r] = 0.f; }
    bool nan0_tail = false;
    // one interpreter instance for all factors; the results go to their
    // registers by a uniform index
#pragma nounroll
    for (int j = 0; j < nf; ++j) {
      const int ps = a.prog_start[pbeg + j];
      VE f[R];
      uint32_t st[R];
      eval_f32<R>(a.code + ps, a.prog_start[pbeg + j + 1] - ps, a.K, a.X32, ldx, row0, f, st);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int64_t i = row0 + 64 * r;
        const bool in = i < r1;
        if (in && st[r] == S_UNC) unc |= 1u << j;
        if (in && st[r] == S_NAN) { nanc |= 1u << j; if (j == 0 && i >= 1) nan0_tail = true; }
        const double fv = in ? (double)f[r].v : 0.0;
        const float fe = in ? f[r].e : 0.f;
#pragma unroll
        for (int jj = 0; jj < NF; ++jj) if (jj == j) { v[jj][r] = fv; e[jj][r] = fe; }
      }
    }
    // Factor 0 certainly NaN at a row >= 1 decides the candidate (see
    // k_fin_fp32); the chunk stops there.
    if (__any((int)nan0_tail)) {
      if (lane == 0) abort_flag = 1;
      break;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double yi = a.y[row0 + 64 * r];   // zero past n (padded column)
      int g = 0;
#pragma unroll
      for (int j = 0; j < NF; ++j) {
#pragma unroll
        for (int l = 0; l <= j; ++l) { G[g] = __builtin_fma(v[j][r], v[l][r], G[g]); ++g; }
        S1[j] = S1[j] + v[j][r];
        SY[j] = __builtin_fma(v[j][r], yi, SY[j]);
        const double ej = (double)e[j][r];
        BD[j] = __builtin_fma(ej, ej, BD[j]);
      }
    }
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) { unc |= __shfl_xor(unc, s); nanc |= __shfl_xor(nanc, s); }
#pragma unroll
  for (int j = 0; j < NG; ++j) G[j] = wave_sum(G[j]);
#pragma unroll
  for (int j = 0; j < NF; ++j) { S1[j] = wave_sum(S1[j]); SY[j] = wave_sum(SY[j]); BD[j] = wave_sum(BD[j]); }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NG; ++j) part[wave][j] = G[j];
#pragma unroll
    for (int j = 0; j < NF; ++j) { part[wave][NG + j] = S1[j]; part[wave][NG + NF + j] = SY[j]; part[wave][NG + 2 * NF + j] = BD[j]; }
    part_flags[wave][0] = unc;
    part_flags[wave][1] = nanc;
  }
  __syncthreads();
  const int64_t cid = (int64_t)w.chunk_base[cand] + chunk;
  if ((int)threadIdx.x < NACC) {
    const int j = (int)threadIdx.x;
    double t = part[0][j];
    for (int ww = 1; ww < ACC_WAVES; ++ww) t = t + part[ww][j];
    w.part[cid * ACC_STRIDE + j] = t;
  }
  if (threadIdx.x == 0) {
    uint32_t fu = 0, fn = 0;
    for (int ww = 0; ww < ACC_WAVES; ++ww) { fu |= part_flags[ww][0]; fn |= part_flags[ww][1]; }
    w.flags[2 * cid] = fu;
    w.flags[2 * cid + 1] = fn | (abort_flag ? CH_NAN0 : 0u);
  }
}

// One wave per candidate: chunk partials in chunk order, then the solve and
// the certificate on one lane.
template <int NF>
__global__ __launch_bounds__(64) void k_fin_fp32(Fp32Args a, ChunkArgs w) {
  constexpr int KB = NF + 1;
  constexpr int NG = NF * (NF + 1) / 2;
  constexpr int NACC = NG + 3 * NF;
  __shared__ Solve<KB> S;
  __shared__ double accs[NACC];
  const int cand = a.order[blockIdx.x];
  if (a.dead && a.dead[cand]) return;   // uniform (one wave)
  bind_dataset(a, cand);
  const int lane = (int)threadIdx.x;
  const int pbeg = a.cand_start[cand];
  const int nf = a.cand_start[cand + 1] - pbeg;
  const int k = nf + 1;
  const int64_t n = a.n;
  double* const co = a.coefs + (int64_t)pbeg + cand;
  const int64_t c0 = w.chunk_base[cand];
  const int64_t nch = (n + w.chunk_rows - 1) / w.chunk_rows;
  uint32_t unc = 0, nanc = 0;
  for (int64_t c = lane; c < nch; c += WAVE) { unc |= w.flags[2 * (c0 + c)]; nanc |= w.flags[2 * (c0 + c) + 1]; }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) { unc |= __shfl_xor(unc, s); nanc |= __shfl_xor(nanc, s); }
  if (lane < NACC) {
    double t = w.part[c0 * ACC_STRIDE + lane];
    for (int64_t c = 1; c < nch; ++c) t = t + w.part[(c0 + c) * ACC_STRIDE + lane];
    accs[lane] = t;
  }
  __syncthreads();
  if (lane != 0) return;
  if (n < k) { write_needs64(a.fit + cand, co, k, NR_SHORT); return; }
  // Factor 0 certainly NaN at a row >= 1 decides the FP64 outcome: its
  // column norm is NaN, so the rank threshold is NaN (no cut), factor 0
  // pivots first and its reflector is NaN -- every coefficient is NaN.
  if (nanc & CH_NAN0) {
    for (int i = 0; i < k; ++i) co[i] = rrm::qnan();
    write_sentinel(a.fit + cand, k);
    return;
  }
  nanc &= ~CH_NAN0;
  if (unc) { write_needs64(a.fit + cand, co, k, NR_UNCERTAIN); return; }
  double G[NG], S1[NF], SY[NF], BD[NF];
#pragma unroll
  for (int j = 0; j < NG; ++j) G[j] = accs[j];
#pragma unroll
  for (int j = 0; j < NF; ++j) { S1[j] = accs[NG + j]; SY[j] = accs[NG + NF + j]; BD[j] = accs[NG + 2 * NF + j]; }

  // ---- the rest runs on one lane
  {
    int g = 0;
#pragma unroll
    for (int j = 0; j < NF; ++j) {
#pragma unroll
      for (int l = 0; l <= j; ++l) { if (j < nf) { S.M[j][l] = G[g]; S.M[l][j] = G[g]; } ++g; }
      if (j < nf) { S.M[j][nf] = S1[j]; S.M[nf][j] = S1[j]; S.g[j] = SY[j]; S.bd[j] = BD[j]; }
    }
    S.M[nf][nf] = (double)n;
    S.g[nf] = a.ysum;
    S.bd[nf] = 0.0;
  }
  if (nanc) {
    // The FP64 QR pivots the first certain-NaN column at the step it sits at
    // (a NaN norm never wins the pivot search), unless a finite column falls
    // below its rank threshold first; with factor 0 NaN the threshold itself
    // is NaN and nothing is cut.  Reaching a NaN reflector makes every
    // coefficient NaN (the sentinel).  Certified when the finite columns are
    // well conditioned.
    bool certain = (nanc & 1u) != 0;
    if (!certain) {
      int idx[KB];
      int m = 0;
      double rho = 0.0;
      for (int j = 0; j < k; ++j) {
        if (j == nf || !((nanc >> j) & 1u)) {
          idx[m++] = j;
          if (j < nf) { const double r = __builtin_sqrt(S.bd[j] / S.M[j][j]); rho = (r > rho) ? r : rho; }
        }
      }
      double T[KB][KB];
      for (int i = 0; i < m; ++i) for (int j = 0; j < m; ++j) T[i][j] = S.M[idx[i]][idx[j]];
      const double tau_c = (TAU_C_REL * rho > TAU_C_MIN) ? TAU_C_REL * rho : TAU_C_MIN;
      certain = chol<KB>(S, T, m, tau_c);
    }
    if (!certain) { write_needs64(a.fit + cand, co, k, NR_NAN_COND); return; }
    for (int i = 0; i < k; ++i) co[i] = rrm::qnan();
    write_sentinel(a.fit + cand, k);
    return;
  }
  {
    double rho = 0.0;
    for (int j = 0; j < nf; ++j) { const double r = __builtin_sqrt(S.bd[j] / S.M[j][j]); rho = (r > rho) ? r : rho; }
    const double tau_c = (TAU_C_REL * rho > TAU_C_MIN) ? TAU_C_REL * rho : TAU_C_MIN;
    if (!chol<KB>(S, S.M, k, tau_c)) { write_needs64(a.fit + cand, co, k, NR_COND); return; }
  }

  // solve, one refinement step on the normal equations
  double c[KB], r1[KB], dcr[KB];
  chol_solve<KB>(S, k, S.g, c);
  for (int i = 0; i < k; ++i) {
    double s = S.g[i];
    for (int j = 0; j < k; ++j) s -= S.M[i][j] * c[j];
    r1[i] = s;
  }
  chol_solve<KB>(S, k, r1, dcr);
  for (int i = 0; i < k; ++i) c[i] += dcr[i];

  // SSR by the normal-equation identity; its rounding and that of the
  // accumulations, bounded through Cauchy-Schwarz (|sum a b| <= sqrt(G_jj yy))
  double cg = 0.0, cGc = 0.0, sab = 0.0, sbd = 0.0, dGd = 0.0;
  double absc[KB], nrm[KB], sbdj[KB];
  for (int i = 0; i < k; ++i) {
    absc[i] = __builtin_fabs(c[i]);
    nrm[i] = __builtin_sqrt(S.M[i][i]);
    sbdj[i] = __builtin_sqrt(S.bd[i] * (1.0 + (double)(n / 64 + 64) * 0x1p-52));
    cg += c[i] * S.g[i];
    double s = 0.0, sd = 0.0;
    for (int j = 0; j < k; ++j) { s += S.M[i][j] * c[j]; sd += S.M[i][j] * dcr[j]; }
    cGc += c[i] * s;
    dGd += dcr[i] * sd;
    sab += absc[i] * nrm[i];
    sbd += absc[i] * sbdj[i];
  }
  const double ny = __builtin_sqrt(a.yy);
  const double mag = (ny + sab) * (ny + sab);
  const double acc_err = 2.0 * ((double)(n / 64 + 128 + 8 * k) * 0x1p-52) * mag;
  double ssr = a.yy - 2.0 * cg + cGc;
  if (ssr < 0.0) ssr = 0.0;

  // coefficient bound dc and the bound D on |sqrt(SSR32) - sqrt(SSR64)|
  double Ginv[KB][KB];
  {
    double ev[KB], col[KB];
    for (int j = 0; j < k; ++j) {
      for (int i = 0; i < k; ++i) ev[i] = (i == j) ? 1.0 : 0.0;
      chol_solve<KB>(S, k, ev, col);
      for (int i = 0; i < k; ++i) Ginv[i][j] = col[i];
    }
  }
  double vv[KB];
  for (int j = 0; j < k; ++j) vv[j] = sbdj[j] * (ny + sab) + nrm[j] * sbd;
  double D = sbd + 2.0 * __builtin_sqrt(__builtin_fabs(dGd));
  bool ambiguous = false;
  for (int i = 0; i < k; ++i) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += __builtin_fabs(Ginv[i][j]) * vv[j];
    const double dc = 2.0 * s + 1e-12 * absc[i] + 2.0 * __builtin_fabs(dcr[i]);
    D += dc * sbdj[i] + EPS12 * nrm[i] * (1.0 + absc[i]);
    if (!(absc[i] - dc >= EPS12 || absc[i] + dc < EPS12)) ambiguous = true;
    const double a1 = __builtin_fabs(c[i] - 1.0);
    if (i != nf && !(a1 - dc >= EPS12 || a1 + dc < EPS12)) ambiguous = true;
  }
  D *= D_ROUND;

  const double rhi = __builtin_sqrt(ssr + acc_err) + D;
  const double rlo = __builtin_sqrt(ssr > acc_err ? ssr - acc_err : 0.0) - D;
  const double hi = rhi * rhi * (1.0 + 1e-12), lo = rlo > 0.0 ? rlo * rlo * (1.0 - 1e-12) : 0.0;
  rr_fit o;
  const double r2 = 1 - ssr / a.sst;
  o.e0 = 1 - r2;
  o.e1 = __builtin_sqrt(ssr / (double)n);
  const double fv = (1 + o.e0) * (1 + o.e1);
  const double up = (1 + hi / a.sst) * (1 + __builtin_sqrt(hi / (double)n)) / fv - 1.0;
  const double dn = 1.0 - (1 + lo / a.sst) * (1 + __builtin_sqrt(lo / (double)n)) / fv;
  o.err = (up > dn ? up : dn) + 1e-12;
  o.rank = k;
  o.flags = 0;
  int reason = ambiguous ? NR_THRESHOLD : 0;
  // |R2_32 - R2_64| = |SSR32 - SSR64| / SST <= max(hi - ssr, ssr - lo) / SST
  if (!reason && !((hi - ssr) <= R2_TOL * a.sst && (ssr - lo) <= R2_TOL * a.sst)) reason = NR_R2;
  if (r2 != r2 || o.e1 != o.e1 || !(o.err < 1.0)) reason = NR_BOUND;
  if (reason) { write_needs64(a.fit + cand, co, k, reason); return; }
  for (int i = 0; i < k; ++i) co[i] = c[i];
  a.fit[cand] = o;
}

// Raw scoring (a tree as given), chunked like the OLS form: partial SSR and
// sum e^2 per chunk, then the bound per tree.
template <int R>
__global__ __launch_bounds__(64 * ACC_WAVES) void k_rawacc_fp32(Fp32Args a, ChunkArgs w, int blk0) {
  __shared__ double part[ACC_WAVES][2];
  __shared__ uint32_t part_flags[ACC_WAVES][2];
  __shared__ int abort_flag;
  const int q = blk0 + (int)blockIdx.x;
  const int cand = w.blk_cand[q];
  const int64_t chunk = w.blk_chunk[q];
  bind_dataset(a, cand);
  const int lane = (int)(threadIdx.x & 63u);
  const int wave = (int)(threadIdx.x >> 6);
  const int p = a.cand_start[cand];
  const uint32_t* code = a.code + a.prog_start[p];
  const int len = a.prog_start[p + 1] - a.prog_start[p];
  const int64_t r0 = chunk * w.chunk_rows;
  const int64_t r1 = (r0 + w.chunk_rows < a.n) ? r0 + w.chunk_rows : a.n;
  if (threadIdx.x == 0) abort_flag = 0;
  __syncthreads();
  double ssr = 0.0, d2 = 0.0;
  uint32_t unc = 0, nanc = 0;
  const volatile int* abort_seen = &abort_flag;
  for (int64_t base = r0 + (int64_t)wave * 64 * R; base < r1; base += (int64_t)ACC_WAVES * 64 * R) {
    if (*abort_seen) break;
    const int64_t row0 = base + lane;
    VE f[R];
    uint32_t st[R];
    eval_f32<R>(code, len, a.K, a.X32, a.ldx, row0, f, st);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = row0 + 64 * r;
      if (i < r1) {
        if (st[r] == S_NAN) nanc = 1;
        else if (st[r] == S_UNC) unc = 1;
        else {
          const double d = a.y[i] - (double)f[r].v;
          ssr = __builtin_fma(d, d, ssr);
          d2 = __builtin_fma((double)f[r].e, (double)f[r].e, d2);
        }
      }
    }
    if (__any((int)nanc)) {   // a certain NaN prediction: the fitness is the sentinel
      if (lane == 0) abort_flag = 1;
      break;
    }
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) { unc |= __shfl_xor(unc, s); nanc |= __shfl_xor(nanc, s); }
  ssr = wave_sum(ssr);
  d2 = wave_sum(d2);
  if (lane == 0) { part[wave][0] = ssr; part[wave][1] = d2; part_flags[wave][0] = unc; part_flags[wave][1] = nanc; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int64_t cid = (int64_t)w.chunk_base[cand] + chunk;
  double ts = part[0][0], td = part[0][1];
  uint32_t fu = part_flags[0][0], fn = part_flags[0][1];
  for (int ww = 1; ww < ACC_WAVES; ++ww) { ts = ts + part[ww][0]; td = td + part[ww][1]; fu |= part_flags[ww][0]; fn |= part_flags[ww][1]; }
  w.part[cid * ACC_STRIDE] = ts;
  w.part[cid * ACC_STRIDE + 1] = td;
  w.flags[2 * cid] = fu;
  w.flags[2 * cid + 1] = fn | (abort_flag ? CH_NAN0 : 0u);
}

__global__ __launch_bounds__(64) void k_rawfin_fp32(Fp32Args a, ChunkArgs w) {
  const int cand = a.order[blockIdx.x];
  bind_dataset(a, cand);
  const int lane = (int)threadIdx.x;
  const int64_t c0 = w.chunk_base[cand];
  const int64_t nch = (a.n + w.chunk_rows - 1) / w.chunk_rows;
  uint32_t unc = 0, nanc = 0;
  for (int64_t c = lane; c < nch; c += WAVE) { unc |= w.flags[2 * (c0 + c)]; nanc |= w.flags[2 * (c0 + c) + 1]; }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) { unc |= __shfl_xor(unc, s); nanc |= __shfl_xor(nanc, s); }
  if (lane != 0) return;
  if (nanc) { write_sentinel(a.fit + cand, 0); return; }
  if (unc) { write_needs64(a.fit + cand, NR_UNCERTAIN); return; }
  double ssr = w.part[c0 * ACC_STRIDE], d2 = w.part[c0 * ACC_STRIDE + 1];
  for (int64_t c = 1; c < nch; ++c) { ssr = ssr + w.part[(c0 + c) * ACC_STRIDE]; d2 = d2 + w.part[(c0 + c) * ACC_STRIDE + 1]; }
  rr_fit o;
  o.rank = 0;
  const double r2 = 1 - ssr / a.sst;
  o.e0 = 1 - r2;
  o.e1 = __builtin_sqrt(ssr / (double)a.n);
  // ||y - p32|| carries its accumulation rounding; ||p32 - p64|| <= sqrt(d2)
  const double acc = (double)(a.n / 64 + 128) * 0x1p-52;
  const double rs = __builtin_sqrt(ssr);
  const double D = __builtin_sqrt(d2 * (1.0 + acc)) * D_ROUND + rs * acc;
  const double hi = (rs + D) * (rs + D) * (1.0 + 1e-12), lo = rs > D ? (rs - D) * (rs - D) * (1.0 - 1e-12) : 0.0;
  const double fv = (1 + o.e0) * (1 + o.e1);
  const double up = (1 + hi / a.sst) * (1 + __builtin_sqrt(hi / (double)a.n)) / fv - 1.0;
  const double dn = 1.0 - (1 + lo / a.sst) * (1 + __builtin_sqrt(lo / (double)a.n)) / fv;
  o.err = (up > dn ? up : dn) + 1e-12;
  const bool r2_wide = !((hi - ssr) <= R2_TOL * a.sst && (ssr - lo) <= R2_TOL * a.sst);
  o.flags = (r2 != r2 || o.e1 != o.e1 || !(o.err < 1.0) || r2_wide) ? RR_F_NEEDS64 : 0;
  if (o.flags) { o.err = __builtin_huge_val(); o.rank = r2_wide && r2 == r2 && o.e1 == o.e1 ? -NR_R2 : -NR_BOUND; }
  a.fit[cand] = o;
}

// FP32 prediction with per-sample bounds: out_e = +inf where FP32 cannot
// vouch for the sample, out_v = NaN and out_e = 0 where the FP64 value is NaN.
__global__ __launch_bounds__(256) void k_predict_fp32(const uint32_t* code, int len, const double* K, const float* X,
                                                      int64_t n, int64_t ldx, float* out_v, float* out_e) {
  const int64_t base = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63u);
  const int64_t row0 = base + (threadIdx.x & 63u);
  VE f[1];
  uint32_t st[1];
  eval_f32<1>(code, len, K, X, ldx, row0, f, st);
  if (row0 < n) {
    out_v[row0] = st[0] == S_NAN ? __builtin_nanf("") : f[0].v;
    out_e[row0] = st[0] == S_NAN ? 0.f : (st[0] == S_UNC ? FINF : f[0].e);
  }
}

// FP32 copy of a column block.  Entries FP32 cannot carry with relative
// error 2^-24 (non-finite, beyond the FP32 range, nonzero below 2^-100) are
// stored as NaN, which the evaluator reports as uncertain.
__global__ void k_to_fp32(const double* src, float* dst, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const double x = src[i];
    const double ax = __builtin_fabs(x);
    const bool ok = ax == 0.0 || (ax >= 0x1p-100 && ax <= 0x1p127);
    dst[i] = ok ? (float)x : __builtin_nanf("");
  }
}

__global__ void k_needs64(const int32_t* order, int count, const int32_t* cand_start, const int32_t* dead,
                          double* coefs, rr_fit* fit) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < count) {
    const int c = order[i];
    if (dead && dead[c]) return;
    const int p = cand_start[c];
    write_needs64(fit + c, coefs + p + c, cand_start[c + 1] - p + 1, NR_WIDE);
  }
}

}  // namespace

hipError_t launch_to_fp32(const double* src, float* dst, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_fp32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_predict_fp32(const uint32_t* code, int len, const double* K, const float* X, int64_t n, int64_t ldx,
                               float* out_v, float* out_e, hipStream_t s) {
  // rows up to the padded ldx are readable, so whole blocks of 256 are safe
  // as long as ldx is a multiple of 256
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_predict_fp32, dim3((unsigned)blocks), dim3(256), 0, s, code, len, K, X, n, ldx, out_v, out_e);
  return hipGetLastError();
}

hipError_t launch_fp32(const Fp32Launch& L, const int32_t* order, const int* bucket_count, const int* bucket_blocks,
                       const hipStream_t* streams) {
  Fp32Args a;
  a.cand_start = L.cand_start; a.prog_start = L.prog_start; a.code = L.code; a.K = L.K;
  a.ds = L.ds; a.cand_ds = L.cand_ds; a.coefs = L.coefs; a.fit = L.fit;
  a.dead = L.raw ? nullptr : L.dead;
  ChunkArgs w;
  w.blk_cand = L.blk_cand; w.blk_chunk = L.blk_chunk; w.chunk_base = L.chunk_base; w.chunk_rows = L.chunk_rows;
  w.part = L.part; w.flags = L.flags;
  const bool r2 = L.rows == 2;
  const dim3 acc_blk(64 * ACC_WAVES);
  if (L.raw) {
    a.order = order;
    if (L.n_cands > 0) {
      if (r2) hipLaunchKernelGGL(k_rawacc_fp32<2>, dim3(bucket_blocks[0]), acc_blk, 0, streams[0], a, w, 0);
      else hipLaunchKernelGGL(k_rawacc_fp32<1>, dim3(bucket_blocks[0]), acc_blk, 0, streams[0], a, w, 0);
      hipLaunchKernelGGL(k_rawfin_fp32, dim3(L.n_cands), dim3(64), 0, streams[0], a, w);
    }
    return hipGetLastError();
  }
  int off[FP32_BUCKETS], boff[FP32_BUCKETS];
  for (int b = 0, acc = 0, bacc = 0; b < FP32_BUCKETS; ++b) {
    off[b] = acc; acc += bucket_count[b];
    boff[b] = bacc; bacc += bucket_blocks[b];
  }
#define RR_OLS32(B, NF)                                                                                    \
  if (bucket_count[B] > 0) {                                                                               \
    a.order = order + off[B];                                                                              \
    if (r2) hipLaunchKernelGGL((k_acc_fp32<NF, 2>), dim3(bucket_blocks[B]), acc_blk, 0, streams[B], a, w, boff[B]); \
    else hipLaunchKernelGGL((k_acc_fp32<NF, 1>), dim3(bucket_blocks[B]), acc_blk, 0, streams[B], a, w, boff[B]);   \
    hipLaunchKernelGGL((k_fin_fp32<NF>), dim3(bucket_count[B]), dim3(64), 0, streams[B], a, w);              \
  }
  // the wide buckets first: few candidates, long per-candidate runs
  RR_OLS32(3, 8)
  RR_OLS32(2, 4)
  RR_OLS32(1, 2)
  RR_OLS32(0, 1)
#undef RR_OLS32
  if (bucket_count[4] > 0)
    hipLaunchKernelGGL(k_needs64, dim3((bucket_count[4] + 255) / 256), dim3(256), 0, streams[4], order + off[4],
                       bucket_count[4], L.cand_start, a.dead, L.coefs, L.fit);
  return hipGetLastError();
}

}  // namespace rr
