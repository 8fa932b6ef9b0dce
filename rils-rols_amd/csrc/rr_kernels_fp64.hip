// rr_kernels_fp64.hip -- the plain FP64 form of the candidate-fitness path.
//
// One workgroup of 16 wavefronts owns one candidate and walks all n samples
// in tiles of 1024*R rows: thread t handles rows t, t+1024, t+2048, ..., R of
// them per tile, in increasing order.  Every sum over samples therefore
// follows the canonical order of rr_math.h / rr_device.h whatever R is, and
// the column-pivoted Householder QR
// follows Eigen's ColPivHouseholderQR (eigen/Eigen/src/QR/
// ColPivHouseholderQR.h:482-607, Householder.h:67-135) operation for
// operation, so the results are the same bits as the CPU oracle's.
//
// Per candidate, k_eval_fp64 first evaluates each factor program
// (tune_constants, rils_rols_cpp.cpp:477-482) into a kept column of a work
// slab in HBM, spread over many small workgroups (row chunks), so that a
// long program does not hold one compute unit for the whole column.  A
// factor that is a bare variable keeps no copy: its column is the dataset's.
// Then k_ols_fp64
//   1. reads the factor columns once for their norms;
//   2. runs the pivoted QR on the slab: per step one pass computing the
//      Householder dot products and one pass applying the reflector, which
//      also accumulates the tail norms the next step and the LAWN-176 norm
//      recomputation need; the reflectors are applied to y as they are made
//      (Q^T b of _solve_impl).  The first step reads every column from where
//      its values already are (kept column, dataset column, the dataset's
//      column of ones for the free term, y) and its update pass writes the
//      working copies; the essential part of a reflector is recomputed
//      (v / (c0 - beta)) where it is used rather than stored, since nothing
//      reads it after its step; the last step, with only y left, needs no
//      update pass.  The factorisation stops at the first step under Eigen's
//      rank threshold, after which nothing changes the solution, and at the
//      first reflector whose beta is NaN, after which the outcome is fixed
//      (see nan_exit below);
//   3. back-substitutes on the leading r x r block and scatters the
//      coefficients through the column permutation (zeros for the rest);
// Then k_score_fp64, one workgroup per (candidate, 4096-row chunk),
//   4. reads the kept factor values (or the dataset's columns, which the QR
//      does not touch) to form the prediction of the OLS tree that tune_constants
//      builds (rils_rols_cpp.cpp:488-517): left-deep sum of
//      c*f (or f when |c-1| < 1e-12), skipping |c| < 1e-12, free term last,
//      and forms the chunk's sum of squared residuals in the chunked order
//      (rrm::chunk_sum); k_fit_fp64 adds a candidate's chunk sums in order
//      into 1-R2 and RMSE (fitness, 520-541).  A kept NaN coefficient makes
//      every prediction NaN, so k_ols_fp64 writes that case as the sentinel
//      and it is not scored.
// A tree scored as given (RAW mode) takes the same two kernels, evaluating
// the tree in place of the OLS prediction.
#include "rr_device.h"
#include "rr_kernels.h"

namespace rr {

namespace {

constexpr double EPS12 = 1e-12;                 // node.h:13-14
constexpr double DBL_EPS = 2.220446049250313e-16;
constexpr double DBL_MIN_ = 2.2250738585072014e-308;

struct Fp64Args {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  // the candidate's dataset, bound at kernel start
  const double* X;
  const double* y;
  const double* one;   // ldx ones: the free term's column
  int64_t n, ldx;
  double sst;
  double* slab;
  const int64_t* slab_off;
  double* coefs;
  rr_fit* fit;
  int cand_base;
  const int32_t* order;   // launched position -> candidate, or null (identity)
  const int32_t* dead;    // per candidate: outcome written by the factor-0 pre-pass, or null
  const double* norm_part;   // per evaluation work item: its chunk's sums of squares (rows >= 0, >= 1)
  const int32_t* prog_item;  // per program: its first evaluation work item (bare variables: unused)
  const double* xnorm;       // the candidate's dataset: per feature, the same two sums over its column
  int32_t* info;             // per candidate: -1 (fitness written here) or rank | steps << 16 (to be scored)
};

__device__ inline void write_fit(rr_fit* f, double ssr, double sst, int64_t n, int rank, int steps = 0) {
  const double r2 = 1 - ssr / sst;
  const double rmse = __builtin_sqrt(ssr / (double)n);
  rr_fit o;
  if (r2 != r2 || rmse != rmse) {
    o.e0 = 1000; o.e1 = 1000; o.flags = RR_F_SENTINEL;
  } else {
    o.e0 = 1 - r2; o.e1 = rmse; o.flags = 0;
  }
  o.err = 0.0;
  o.rank = rank;
  o.flags |= steps << 16;
  *f = o;
}

// Bind the dataset candidate `cand` is scored against (uniform: scalar loads).
template <class A>
__device__ inline void bind_dataset(A& a, int cand) {
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  a.X = D.X; a.y = D.y; a.one = D.one; a.n = D.n; a.ldx = D.ldx; a.sst = D.sst;
}
__device__ inline void bind_xnorm(Fp64Args& a, int cand) { a.xnorm = a.ds[a.cand_ds ? a.cand_ds[cand] : 0].xnorm; }

// Variable index of a program that is one LDV, else -1.
__device__ inline int bare_variable(const uint32_t* code, int len) {
  return (len == 1 && (code[0] & 0xffu) == RR_OP_LDV) ? (int)(code[0] >> 8) : -1;
}

__device__ inline int64_t tile_start(int64_t lo, int R) { return (lo / (CP * R)) * (CP * R); }

// One workgroup of CW = 16 waves per candidate; thread t takes the rows
// i = t (mod 1024), R of them per tile of 1024 R rows, so every sum follows
// the canonical order whatever R is.
template <int MAXK, int R>
__global__ __launch_bounds__(CP) void k_ols_fp64(Fp64Args a) {
  constexpr int G = (R == 1) ? 8 : (R == 2) ? 4 : 2;   // columns per QR pass
  __shared__ double normU[MAXK], normD[MAXK], tsq[MAXK], cc[MAXK];
  __shared__ int phys[MAXK];
  __shared__ int ovar[MAXK];   // per factor: its bare variable, or -1 (kept column)
  __shared__ double redG[CW][G], redS[CW][2 * G];
  __shared__ int abort_flag;

  const int cand = a.order ? a.order[a.cand_base + (int)blockIdx.x] : a.cand_base + (int)blockIdx.x;
  if (a.dead && a.dead[cand]) return;   // uniform: the whole workgroup leaves
  bind_dataset(a, cand);
  bind_xnorm(a, cand);
  const int tid = (int)threadIdx.x;
  const int pbeg = a.cand_start[cand];
  const int nf = a.cand_start[cand + 1] - pbeg;
  const int k = nf + 1;
  const int64_t n = a.n, ldx = a.ldx;
  // slab columns 0..k-1: A (factors, free term), k: Q^T y, k+1..k+nf: the
  // factor values as evaluated (kept for the first QR step and the
  // prediction of step 4)
  double* const S = a.slab + a.slab_off[blockIdx.x];
  const double* const X = a.X;
  double* const co = a.coefs + (int64_t)pbeg + cand;
  if (tid == 0) abort_flag = 0;
  __syncthreads();

  // ---- 1. norms of the factor columns and their sums over rows >= 1, in the
  // chunked order (rrm::chunk_sum): k_eval_fp64 left one pair of sums per
  // 4096-row chunk, added here in chunk order; a bare variable's come with
  // the dataset.
  // Factor 0 NaN at a row >= 1 (its rows >= 1 sum is NaN exactly then) fixes
  // the outcome: its norm is NaN, so the rank threshold is NaN (no cut),
  // factor 0 pivots first (nothing compares greater than NaN), its tail norm
  // and hence its reflector are NaN, and with the threshold NaN no later
  // column stops the factorisation: every coefficient is NaN, rank = min(n, k).
  if (tid == 0) {
    const int64_t nch = (n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
    for (int j = 0; j < nf; ++j) {
      const uint32_t* code = a.code + a.prog_start[pbeg + j];
      const int len = a.prog_start[pbeg + j + 1] - a.prog_start[pbeg + j];
      const int var = bare_variable(code, len);
      ovar[j] = var;
      double s0, s1;
      if (var >= 0) {
        s0 = a.xnorm[2 * var];
        s1 = a.xnorm[2 * var + 1];
      } else {
        const double* pp = a.norm_part + 2 * (int64_t)a.prog_item[pbeg + j];
        s0 = pp[0];
        s1 = pp[1];
        for (int64_t c = 1; c < nch; ++c) { s0 = s0 + pp[2 * c]; s1 = s1 + pp[2 * c + 1]; }
      }
      normD[j] = __builtin_sqrt(s0);
      tsq[j] = s1;
      if (j == 0 && s1 != s1) abort_flag = 1;
    }
  }
  __syncthreads();
  if (abort_flag) {
    const int size0 = (int)((n < (int64_t)k) ? n : (int64_t)k);
    for (int i = tid; i < k; i += CP) co[i] = (i < size0) ? rrm::qnan() : 0.0;
    if (tid == 0) { write_fit(a.fit + cand, rrm::qnan(), a.sst, n, size0, 0); a.info[cand] = -1; }
    return;
  }
  double* const cy = S + (int64_t)k * ldx;
  // where the values of physical column c are before the first QR step
  auto origin = [&](int c) -> const double* {
    if (c == nf) return a.one;
    const int v = __builtin_amdgcn_readfirstlane(ovar[c]);   // uniform: keep it scalar
    return v >= 0 ? X + (int64_t)v * ldx : S + (int64_t)(k + 1 + c) * ldx;
  };
  // row 0 of every working column; rows >= 1 are written by the first
  // update pass (the single-row and top-row updates work on row p in place)
  if (tid == 0) {
    for (int c = 0; c < k; ++c) S[(int64_t)c * ldx] = origin(c)[0];
    cy[0] = a.y[0];
    normD[nf] = __builtin_sqrt((double)n);
    tsq[nf] = (double)(n - 1);
  }
  for (int j = tid; j < k; j += CP) phys[j] = j;
  __syncthreads();
  for (int j = tid; j < k; j += CP) normU[j] = normD[j];
  __syncthreads();

  // ---- 2. column-pivoted Householder QR
  const int size = (int)((n < (int64_t)k) ? n : (int64_t)k);
  double mx = normU[0];
  for (int j = 1; j < k; ++j) { const double v = normU[j]; mx = (mx < v) ? v : mx; }
  const double thr = ((mx * DBL_EPS) * (mx * DBL_EPS)) / (double)n;
  const double dd_thr = __builtin_sqrt(DBL_EPS);
  int nonzero = size;
  int nan_exit = -1;   // step whose reflector had a NaN beta
  int steps = 0;       // steps that made passes over the samples

  for (int p = 0; p < size; ++p) {
    int bi = p;
    double best = normU[p];
    for (int j = p + 1; j < k; ++j) { const double v = normU[j]; if (v > best) { best = v; bi = j; } }
    if (best * best < thr * (double)(n - p)) { nonzero = p; break; }
    __syncthreads();
    if (tid == 0 && bi != p) {
      int ti = phys[p]; phys[p] = phys[bi]; phys[bi] = ti;
      double t;
      t = normU[p]; normU[p] = normU[bi]; normU[bi] = t;
      t = normD[p]; normD[p] = normD[bi]; normD[bi] = t;
      t = tsq[p]; tsq[p] = tsq[bi]; tsq[bi] = t;
    }
    __syncthreads();

    const int pc = __builtin_amdgcn_readfirstlane(phys[p]);   // uniform: keep it scalar
    double* const v = S + (int64_t)pc * ldx;
    const double* const vr = (p == 0) ? origin(pc) : v;   // rows > p of the pivot column
    const double c0 = v[p];
    const double tailSq = (n - p == 1) ? 0.0 : tsq[p];
    double tau, beta;
    const bool zero_tail = tailSq <= DBL_MIN_;
    if (zero_tail) {
      tau = 0.0;
      beta = c0;
    } else {
      beta = __builtin_sqrt(c0 * c0 + tailSq);
      if (c0 >= 0.0) beta = -beta;
      tau = (beta - c0) / beta;
    }
    if (!zero_tail && beta != beta && n - p > 1) {
      // A NaN reflector turns every remaining column and Q^T y into NaN, so
      // R(p,p) = NaN and every coefficient up to the rank is NaN.  The norm
      // downdate leaves columns with norm 0 at 0 and makes all others NaN;
      // later steps pivot in place, and the first column at 0 stops the
      // factorisation if the threshold is positive.
      nan_exit = p;
      int r = size;
      if (thr > 0.0)
        for (int j = p + 1; j < size; ++j) if (normU[j] == 0.0) { r = j; break; }
      nonzero = r;
      break;
    }
    const double den = c0 - beta;
    const bool apply = tau != 0.0;         // applyHouseholderOnTheLeft does nothing for tau == 0
    __syncthreads();   // every thread has read c0 before row p is overwritten
    if (tid == 0) v[p] = beta;

    const int nrem = k - 1 - p;            // remaining A columns: positions p+1..k-1
    const int ntarget = nrem + 1;          // plus the y column (last target)
    ++steps;
    if (n - p == 1) {
      // applyHouseholderOnTheLeft with a single row: scale by (1 - tau)
      if (tid == 0) {
        for (int t = 0; t < nrem; ++t) { double* col = S + (int64_t)phys[p + 1 + t] * ldx; col[p] *= (1.0 - tau); }
        cy[p] *= (1.0 - tau);
      }
      __syncthreads();
      for (int t = 0; t < nrem; ++t) if (tid == 0) tsq[p + 1 + t] = 0.0;
      // the LAWN-176 recomputation below needs sums over rows p+1.. : empty
      for (int t = 0; t < nrem; ++t) {
        const int j = p + 1 + t;
        if (normU[j] != 0.0) {
          const double apj = S[(int64_t)phys[j] * ldx + p];
          double temp = __builtin_fabs(apj) / normU[j];
          temp = (1.0 + temp) * (1.0 - temp);
          temp = temp < 0.0 ? 0.0 : temp;
          const double ratio = normU[j] / normD[j];
          const double temp2 = temp * (ratio * ratio);
          __syncthreads();
          if (tid == 0) {
            if (temp2 <= dd_thr) { normD[j] = 0.0; normU[j] = 0.0; }
            else normU[j] *= __builtin_sqrt(temp);
          }
          __syncthreads();
        }
      }
      continue;
    }

    const int64_t t0 = tile_start(p + 1, R);
    for (int g0 = 0; g0 < ntarget; g0 += G) {
      const int gn = (ntarget - g0 < G) ? (ntarget - g0) : G;
      double* cols[G];          // working columns (written)
      const double* rcols[G];   // where rows > p are read from
#pragma unroll
      for (int t = 0; t < G; ++t) {
        const int tg = g0 + t;
        const int c = (t < gn && tg < nrem) ? __builtin_amdgcn_readfirstlane(phys[p + 1 + tg]) : -1;
        cols[t] = c >= 0 ? S + (int64_t)c * ldx : cy;
        rcols[t] = p > 0 ? cols[t] : (c >= 0 ? origin(c) : a.y);
      }
      double tdot[G];
#pragma unroll
      for (int t = 0; t < G; ++t) tdot[t] = 0.0;
      if (apply) {
        // Householder dot products  t = ess . a[p+1..] (+ a[p])
        double acc[G];
#pragma unroll
        for (int t = 0; t < G; ++t) acc[t] = 0.0;
        for (int64_t base = t0; base < n; base += CP * R) {
          double e[R];
          bool ok[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int64_t i = base + tid + (int64_t)CP * r;
            ok[r] = i > p && i < n;
            e[r] = vr[i] / den;   // the essential part of the reflector
          }
#pragma unroll
          for (int t = 0; t < G; ++t) {
            if (t < gn) {
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const double x = rcols[t][base + tid + (int64_t)CP * r];
                if (ok[r]) acc[t] = acc[t] + e[r] * x;
              }
            }
          }
        }
        block_sums<G>(acc, redG);
#pragma unroll
        for (int t = 0; t < G; ++t) {
          if (t < gn) {
            const double ap = cols[t][p];
            tdot[t] = acc[t] + ap;
          }
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll
          for (int t = 0; t < G; ++t) if (t < gn) cols[t][p] = cols[t][p] - tau * tdot[t];
        }
      }
      // apply the reflector and accumulate tail norms (rows >= p+1 and >= p+2);
      // with only y left nothing reads its rows > p again
      if (nrem == 0) { __syncthreads(); continue; }
      double s12[2 * G];
#pragma unroll
      for (int t = 0; t < 2 * G; ++t) s12[t] = 0.0;
      for (int64_t base = t0; base < n; base += CP * R) {
        double te[R];
        bool ok[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int64_t i = base + tid + (int64_t)CP * r;
          ok[r] = i > p && i < n;
          te[r] = apply ? tau * (vr[i] / den) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < G; ++t) {
          if (t < gn) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int64_t i = base + tid + (int64_t)CP * r;
              double x = rcols[t][i];
              if (ok[r]) {
                if (apply) x = x - te[r] * tdot[t];
                if (apply || p == 0) cols[t][i] = x;
                s12[t] = s12[t] + x * x;
                if (i >= p + 2) s12[G + t] = s12[G + t] + x * x;
              }
            }
          }
        }
      }
      block_sums<2 * G>(s12, redS);
      if (tid == 0) {
#pragma unroll
        for (int t = 0; t < G; ++t) {
          const int tg = g0 + t;
          if (t < gn && tg < nrem) { cc[p + 1 + tg] = s12[t]; tsq[p + 1 + tg] = s12[G + t]; }
        }
      }
      __syncthreads();
    }

    // LAWN-176 norm downdate (cc[] holds the recomputed tail norms^2)
    for (int j = p + 1; j < k; ++j) {
      const double nu = normU[j];
      if (nu != 0.0) {
        const double apj = S[(int64_t)phys[j] * ldx + p];
        double temp = __builtin_fabs(apj) / nu;
        temp = (1.0 + temp) * (1.0 - temp);
        temp = temp < 0.0 ? 0.0 : temp;
        const double ratio = nu / normD[j];
        const double temp2 = temp * (ratio * ratio);
        const double rec = __builtin_sqrt(cc[j]);
        __syncthreads();
        if (tid == 0) {
          if (temp2 <= dd_thr) { normD[j] = rec; normU[j] = rec; }
          else normU[j] = nu * __builtin_sqrt(temp);
        }
      }
    }
    __syncthreads();
  }

  const int r = nonzero;
  if (nan_exit >= 0) {
    for (int i = tid; i < k; i += CP) co[phys[i]] = (i < r) ? rrm::qnan() : 0.0;
    if (tid == 0) { write_fit(a.fit + cand, rrm::qnan(), a.sst, n, r, steps); a.info[cand] = -1; }
    return;
  }

  // ---- 3. back substitution on the leading nonzero x nonzero block
  for (int i = tid; i < r; i += CP) cc[i] = cy[i];
  __syncthreads();
  for (int i = r - 1; i >= 0; --i) {
    const double* coli = S + (int64_t)phys[i] * ldx;
    const double ci = cc[i] / coli[i];
    __syncthreads();
    for (int l = tid; l < i; l += CP) cc[l] = cc[l] - ci * coli[l];
    if (tid == 0) cc[i] = ci;
    __syncthreads();
  }
  bool any_nan = false;
  for (int i = 0; i < k; ++i) {
    const double cv = (i < r) ? cc[i] : 0.0;
    if (cv != cv) any_nan = true;
  }
  for (int i = tid; i < k; i += CP) {
    double cv = (i < r) ? cc[i] : 0.0;
    if (cv != cv) cv = rrm::qnan();
    co[phys[i]] = cv;
  }
  __syncthreads();
  if (tid == 0) {
    if (any_nan) {   // a kept NaN coefficient: every prediction is NaN
      write_fit(a.fit + cand, rrm::qnan(), a.sst, n, r, steps);
      a.info[cand] = -1;
    } else {
      a.info[cand] = r | (steps << 16);   // scored by k_score_fp64 / k_fit_fp64
    }
  }
}

// Factor evaluation for k_ols_fp64: one workgroup of EV_THREADS threads per
// (candidate, non-bare factor, chunk of rrm::NORM_CHUNK rows).  Work item b
// is the b-th chunk in pair order; pair_item[p] is the first item of pair p.
// Besides the values, each work item leaves its chunk's two sums of squares
// (rows >= 0 and rows >= 1) in the chunked order for the QR's norms.
struct EvalArgs {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* slab;
  const int64_t* slab_off;   // per launched position
  const int32_t* order;      // launched position -> candidate
  const int32_t* dead;
  const int32_t* pair_pos;   // per pair: launched position
  const int32_t* pair_fac;   // per pair: factor index
  const int32_t* pair_item;  // per pair: first work item
  int n_pairs;
  int items_per_pair;        // > 0: every pair has this many work items (one dataset)
  double* norm_part;         // per work item: its chunk's sums of squares (rows >= 0, rows >= 1)
};
constexpr int EV_THREADS = 256;
static_assert(EVAL_CHUNK_ROWS == rrm::NORM_CHUNK, "evaluation work items are the chunks of the norm order");

template <int R>
__global__ __launch_bounds__(EV_THREADS) void k_eval_fp64(EvalArgs a) {
  static_assert(EV_THREADS == rrm::NORM_PARTIALS, "one thread per partial of the chunked order");
  __shared__ double red[EV_THREADS / WAVE][2];
  const int item = (int)blockIdx.x;
  int lo;
  if (a.items_per_pair > 0) {
    lo = item / a.items_per_pair;
  } else {   // last pair whose first item <= item (uniform)
    lo = 0;
    int hi = a.n_pairs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.pair_item[mid] <= item) lo = mid; else hi = mid - 1;
    }
  }
  const int pos = a.pair_pos[lo], j = a.pair_fac[lo];
  const int cand = a.order[pos];
  if (a.dead && a.dead[cand]) return;
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const int pbeg = a.cand_start[cand];
  const int k = a.cand_start[cand + 1] - pbeg + 1;
  const uint32_t* code = a.code + a.prog_start[pbeg + j];
  const int len = a.prog_start[pbeg + j + 1] - a.prog_start[pbeg + j];
  double* const keep = a.slab + a.slab_off[pos] + (int64_t)(k + 1 + j) * D.ldx;
  const int64_t r0 = (int64_t)(item - a.pair_item[lo]) * rrm::NORM_CHUNK;
  const int64_t r1 = (r0 + rrm::NORM_CHUNK < D.n) ? r0 + rrm::NORM_CHUNK : D.n;
  // thread t adds the squares of rows r0 + t (mod 256) in increasing order:
  // the chunk's partials of rrm::chunk_sum.  base is a multiple of
  // EV_THREADS * R (which divides 4096) and below n, so every row touched is
  // below ldx.
  double s[2] = {0.0, 0.0};   // rows >= 0, rows >= 1
  for (int64_t base = r0; base < r1; base += (int64_t)EV_THREADS * R) {
    const int64_t row0 = base + (int64_t)threadIdx.x;
    double v[R];
    eval_f64<R, EV_THREADS>(code, len, a.K, D.X, D.ldx, row0, v);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = row0 + (int64_t)EV_THREADS * r;
      keep[i] = v[r];
      if (i < r1) {
        s[0] = s[0] + v[r] * v[r];
        if (i >= 1) s[1] = s[1] + v[r] * v[r];
      }
    }
  }
  s[0] = wave_sum(s[0]);
  s[1] = wave_sum(s[1]);
  const int wave = (int)(threadIdx.x / WAVE);
  if ((threadIdx.x & (WAVE - 1)) == 0) { red[wave][0] = s[0]; red[wave][1] = s[1]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = red[0][0], t1 = red[0][1];
    for (int w = 1; w < EV_THREADS / WAVE; ++w) { t0 = t0 + red[w][0]; t1 = t1 + red[w][1]; }
    a.norm_part[2 * (int64_t)item] = t0;
    a.norm_part[2 * (int64_t)item + 1] = t1;
  }
}

// Per feature of a dataset, the sums of squares of its column over rows >= 0
// and rows >= 1 in the chunked order (one workgroup per feature, the chunks
// in order): the norms the FP64 OLS form uses for a bare-variable factor.
__global__ __launch_bounds__(EV_THREADS) void k_xnorm_fp64(const double* X, int64_t n, int64_t ldx, double* out) {
  __shared__ double red[EV_THREADS / WAVE][2];
  const double* x = X + (int64_t)blockIdx.x * ldx;
  const int wave = (int)(threadIdx.x / WAVE);
  double tot0 = 0.0, tot1 = 0.0;
  for (int64_t r0 = 0; r0 < n; r0 += rrm::NORM_CHUNK) {
    const int64_t r1 = (r0 + rrm::NORM_CHUNK < n) ? r0 + rrm::NORM_CHUNK : n;
    double s[2] = {0.0, 0.0};
    for (int64_t i = r0 + (int64_t)threadIdx.x; i < r1; i += EV_THREADS) {
      const double v = x[i];
      s[0] = s[0] + v * v;
      if (i >= 1) s[1] = s[1] + v * v;
    }
    s[0] = wave_sum(s[0]);
    s[1] = wave_sum(s[1]);
    if ((threadIdx.x & (WAVE - 1)) == 0) { red[wave][0] = s[0]; red[wave][1] = s[1]; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double t0 = red[0][0], t1 = red[0][1];
      for (int w = 1; w < EV_THREADS / WAVE; ++w) { t0 = t0 + red[w][0]; t1 = t1 + red[w][1]; }
      tot0 = (r0 == 0) ? t0 : tot0 + t0;
      tot1 = (r0 == 0) ? t1 : tot1 + t1;
    }
    __syncthreads();   // red is reused by the next chunk
  }
  if (threadIdx.x == 0) { out[2 * blockIdx.x] = tot0; out[2 * blockIdx.x + 1] = tot1; }
}

// The factor-0 pre-pass: one workgroup of NAN0_ROWS threads per candidate,
// one row each, so it costs one evaluation of latency.
struct Nan0Args {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* coefs;
  rr_fit* fit;
  int32_t* dead;
};

__global__ __launch_bounds__(NAN0_ROWS) void k_nan0_fp64(Nan0Args a) {
  __shared__ int found;
  const int cand = (int)blockIdx.x;
  const int tid = (int)threadIdx.x;
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const int64_t n = D.n;
  const int pbeg = a.cand_start[cand];
  const int nf = a.cand_start[cand + 1] - pbeg;
  if (tid == 0) found = 0;
  __syncthreads();
  if (nf > 0) {   // row tid < NAN0_ROWS <= ldx (a multiple of 4096)
    const uint32_t* code = a.code + a.prog_start[pbeg];
    const int len = a.prog_start[pbeg + 1] - a.prog_start[pbeg];
    double v[1];
    eval_f64<1>(code, len, a.K, D.X, D.ldx, (int64_t)tid, v);
    if (__any((int)(tid >= 1 && (int64_t)tid < n && v[0] != v[0])) && (tid & 63) == 0) found = 1;
  }
  __syncthreads();
  if (tid == 0) {
    a.dead[cand] = found;
    if (found) {
      // the outcome k_ols_fp64 writes when it sees the same NaN
      const int k = nf + 1;
      const int size0 = (int)((n < (int64_t)k) ? n : (int64_t)k);
      double* co = a.coefs + (int64_t)pbeg + cand;
      for (int i = 0; i < k; ++i) co[i] = (i < size0) ? rrm::qnan() : 0.0;
      write_fit(a.fit + cand, rrm::qnan(), D.sst, n, size0, 0);
    }
  }
}

// Scoring in the chunked order: work item b is chunk (b - pos_item[p]) of
// launched position p (pos_item[p] <= b < pos_item[p+1]); items_per_pos > 0
// when every position has that many (one dataset).
struct ScoreArgs {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  const double* coefs;
  const double* slab;
  const int64_t* slab_off;   // per launched position
  const int32_t* order;      // launched position -> candidate, or null (identity)
  const int32_t* dead;
  const int32_t* info;
  const int32_t* pos_item;
  int n_pos;
  int items_per_pos;
  double* part;
  rr_fit* fit;
};

__device__ inline int score_pos(const ScoreArgs& a, int item) {
  if (a.items_per_pos > 0) return item / a.items_per_pos;
  int lo = 0, hi = a.n_pos - 1;   // last position whose first item <= item (uniform)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.pos_item[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// RAW: the tree (the candidate's one program) over the chunk; OLS: the OLS
// tree's prediction from the coefficients and the factor columns, with the
// arithmetic of tune_constants' tree (rils_rols_cpp.cpp:488-517).
template <bool RAW, int R>
__global__ __launch_bounds__(EV_THREADS) void k_score_fp64(ScoreArgs a) {
  __shared__ double red[EV_THREADS / WAVE];
  const int item = (int)blockIdx.x;
  const int pos = score_pos(a, item);
  const int cand = a.order ? a.order[pos] : pos;
  if (a.dead && a.dead[cand]) return;   // uniform
  if (!RAW && a.info[cand] < 0) return;
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const int64_t n = D.n, ldx = D.ldx;
  const int64_t first = a.items_per_pos > 0 ? (int64_t)pos * a.items_per_pos : a.pos_item[pos];
  const int64_t r0 = (item - first) * rrm::NORM_CHUNK;
  const int64_t r1 = (r0 + rrm::NORM_CHUNK < n) ? r0 + rrm::NORM_CHUNK : n;
  const int pbeg = a.cand_start[cand];
  double s = 0.0;   // thread t: rows r0 + t (mod 256) in increasing order
  if (RAW) {
    const uint32_t* code = a.code + a.prog_start[pbeg];
    const int len = a.prog_start[pbeg + 1] - a.prog_start[pbeg];
    for (int64_t base = r0; base < r1; base += (int64_t)EV_THREADS * R) {   // rows < ldx
      const int64_t row0 = base + (int64_t)threadIdx.x;
      double v[R];
      eval_f64<R, EV_THREADS>(code, len, a.K, D.X, ldx, row0, v);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int64_t i = row0 + (int64_t)EV_THREADS * r;
        if (i < r1) { const double d = D.y[i] - v[r]; s = s + d * d; }
      }
    }
  } else {
    const int nf = a.cand_start[cand + 1] - pbeg;
    const int k = nf + 1;
    const double* co = a.coefs + (int64_t)pbeg + cand;
    const double* S = a.slab + a.slab_off[pos];
    const double cf_free = co[nf];
    for (int64_t i = r0 + (int64_t)threadIdx.x; i < r1; i += EV_THREADS) {
      double yp = 0.0;
      bool any = false;
      for (int j = 0; j < nf; ++j) {
        const double cj = co[j];
        if (__builtin_fabs(cj) < EPS12) continue;
        const uint32_t* code = a.code + a.prog_start[pbeg + j];
        const int var = bare_variable(code, a.prog_start[pbeg + j + 1] - a.prog_start[pbeg + j]);
        const double f = var >= 0 ? D.X[(int64_t)var * ldx + i] : S[(int64_t)(k + 1 + j) * ldx + i];
        const double term = (__builtin_fabs(cj - 1) < EPS12) ? f : cj * f;
        yp = any ? yp + term : term;
        any = true;
      }
      if (!(__builtin_fabs(cf_free) < EPS12)) {
        const double term = cf_free * 1.0;
        yp = any ? yp + term : term;
      }
      const double d = D.y[i] - yp;
      s = s + d * d;
    }
  }
  s = wave_sum(s);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < EV_THREADS / WAVE; ++w) t = t + red[w];
    a.part[item] = t;
  }
}

// One thread per launched position: the chunk sums in chunk order, then the
// fitness (rils_rols_cpp.cpp:520-541).
template <bool RAW>
__global__ __launch_bounds__(256) void k_fit_fp64(ScoreArgs a) {
  const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pos >= a.n_pos) return;
  const int cand = a.order ? a.order[pos] : pos;
  if (a.dead && a.dead[cand]) return;
  int rank = 0, steps = 0;
  if (!RAW) {
    const int32_t info = a.info[cand];
    if (info < 0) return;
    rank = info & 0xffff;
    steps = info >> 16;
  }
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const int64_t first = a.items_per_pos > 0 ? (int64_t)pos * a.items_per_pos : a.pos_item[pos];
  const int64_t nch = (D.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
  double ssr = a.part[first];
  for (int64_t c = 1; c < nch; ++c) ssr = ssr + a.part[first + c];
  write_fit(a.fit + cand, ssr, D.sst, D.n, rank, steps);
}

// node::evaluate_all over all samples; many waves, any order (no reduction)
__global__ __launch_bounds__(256) void k_predict_fp64(const uint32_t* code, int len, const double* K, const double* X,
                                                      int64_t n, int64_t ldx, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[1];
  eval_f64<1>(code, len, K, X, ldx, i < n ? i : 0, v);
  if (i < n) out[i] = v[0];
}

// out4: mean of y, sum (y - mean)^2, sum y (canonical sums), sum y^2 (any
// order: used by the FP32 form only)
__global__ __launch_bounds__(CP) void k_stats_fp64(const double* y, int64_t n, double* out4) {
  __shared__ double red2[CW][2];
  __shared__ double red1[CW][1];
  const int tid = (int)threadIdx.x;
  double s[2] = {0.0, 0.0};
  for (int64_t i = tid; i < n; i += CP) { s[0] = s[0] + y[i]; s[1] = __builtin_fma(y[i], y[i], s[1]); }
  block_sums<2>(s, red2);
  const double ybar = s[0] / (double)n;
  double q[1] = {0.0};
  for (int64_t i = tid; i < n; i += CP) q[0] = q[0] + (y[i] - ybar) * (y[i] - ybar);
  block_sums<1>(q, red1);
  if (tid == 0) { out4[0] = ybar; out4[1] = q[0]; out4[2] = s[0]; out4[3] = s[1]; }
}

}  // namespace

hipError_t launch_ols_fp64(const OlsLaunch& L, hipStream_t s) {
  Fp64Args a;
  a.cand_start = L.cand_start; a.prog_start = L.prog_start; a.code = L.code; a.K = L.K;
  a.ds = L.ds; a.cand_ds = L.cand_ds;
  a.slab = L.slab; a.slab_off = L.slab_off; a.coefs = L.coefs; a.fit = L.fit; a.cand_base = L.cand_base;
  a.order = L.order;
  a.dead = L.dead;
  a.norm_part = L.norm_part;
  a.prog_item = L.prog_item;
  a.info = L.info;
  a.xnorm = nullptr;   // bound per candidate
  const dim3 grid(L.n_cands), block(CP);
  const int R = L.rows;
  if (L.max_cols <= 16) {
    if (R == 1) hipLaunchKernelGGL((k_ols_fp64<16, 1>), grid, block, 0, s, a);
    else if (R == 2) hipLaunchKernelGGL((k_ols_fp64<16, 2>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_ols_fp64<16, 4>), grid, block, 0, s, a);
  } else if (L.max_cols <= 64) {
    hipLaunchKernelGGL((k_ols_fp64<64, 2>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_ols_fp64<RR_MAX_COLS, 2>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_eval_fp64(const EvalLaunch& L, hipStream_t s) {
  if (L.n_items <= 0) return hipSuccess;
  EvalArgs a{L.cand_start, L.prog_start, L.code, L.K, L.ds, L.cand_ds, L.slab, L.slab_off, L.order, L.dead,
             L.pair_pos, L.pair_fac, L.pair_item, L.n_pairs, L.items_per_pair, L.norm_part};
  const dim3 grid((unsigned)L.n_items), block(EV_THREADS);
  if (L.rows == 1) hipLaunchKernelGGL(k_eval_fp64<1>, grid, block, 0, s, a);
  else if (L.rows == 4) hipLaunchKernelGGL(k_eval_fp64<4>, grid, block, 0, s, a);
  else hipLaunchKernelGGL(k_eval_fp64<2>, grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_xnorm_fp64(const double* X, int64_t n, int64_t ldx, int32_t d, double* out, hipStream_t s) {
  if (d <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_xnorm_fp64, dim3((unsigned)d), dim3(EV_THREADS), 0, s, X, n, ldx, out);
  return hipGetLastError();
}

hipError_t launch_nan0_fp64(const Nan0Launch& L, hipStream_t s) {
  if (L.n_cands <= 0) return hipSuccess;
  Nan0Args a{L.cand_start, L.prog_start, L.code, L.K, L.ds, L.cand_ds, L.coefs, L.fit, L.dead};
  hipLaunchKernelGGL(k_nan0_fp64, dim3((unsigned)L.n_cands), dim3(NAN0_ROWS), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_score_fp64(const ScoreLaunch& L, hipStream_t s) {
  if (L.n_pos <= 0) return hipSuccess;
  ScoreArgs a{L.cand_start, L.prog_start, L.code, L.K, L.ds, L.cand_ds, L.coefs, L.slab, L.slab_off, L.order,
              L.dead, L.info, L.pos_item, L.n_pos, L.items_per_pos, L.part, L.fit};
  const dim3 grid((unsigned)L.n_items), block(EV_THREADS), fgrid((unsigned)((L.n_pos + 255) / 256));
  if (L.raw) {
    if (L.rows == 1) hipLaunchKernelGGL((k_score_fp64<true, 1>), grid, block, 0, s, a);
    else if (L.rows == 4) hipLaunchKernelGGL((k_score_fp64<true, 4>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_score_fp64<true, 2>), grid, block, 0, s, a);
    hipLaunchKernelGGL(k_fit_fp64<true>, fgrid, dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_score_fp64<false, 1>), grid, block, 0, s, a);
    hipLaunchKernelGGL(k_fit_fp64<false>, fgrid, dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_predict_fp64(const uint32_t* code, int len, const double* K, const double* X, int64_t n,
                               int64_t ldx, double* out, hipStream_t s) {
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_predict_fp64, dim3((unsigned)blocks), dim3(256), 0, s, code, len, K, X, n, ldx, out);
  return hipGetLastError();
}

__global__ void k_fill_f64(double* p, int64_t count, double value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

hipError_t launch_fill_f64(double* p, int64_t count, double value, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, s, p, count, value);
  return hipGetLastError();
}

hipError_t launch_stats_fp64(const double* y, int64_t n, double* out4, hipStream_t s) {
  hipLaunchKernelGGL(k_stats_fp64, dim3(1), dim3(CP), 0, s, y, n, out4);
  return hipGetLastError();
}

}  // namespace rr
