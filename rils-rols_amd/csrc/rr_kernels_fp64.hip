// rr_kernels_fp64.hip -- the plain FP64 form of the candidate-fitness path
// (tune_constants + fitness, rils_rols_cpp.cpp:445-541), bit for bit the
// oracle's results.
//
// Every sum over samples follows an order fixed in advance (the chunked order
// of rr_math.h: 4096-row chunks, 256 partials each, xor butterflies, chunk
// sums in chunk order; the chunked QR's own order below), so each pass splits
// into independent 4096-row work items while the results stay bit-identical
// to the CPU oracle's, which sums the same way.  One call over a batch runs:
//   k_nan0_fp64   one wave per candidate: factor 0 on rows 0..63; a
//                 NaN at a row >= 1 fixes the outcome (the sentinel), which
//                 is written here, and later kernels skip the candidate;
//   k_alive_fp64, k_eval_fp64
//                 each unique factor program (equal programs over one dataset
//                 shared by the candidates that use them) that a candidate
//                 still to be scored uses, evaluated (rils_rols_cpp.cpp:
//                 477-482) per (factor, 4096-row chunk) work item into a kept
//                 column, with the chunk's sums of squares for the norms; a
//                 factor that is a bare variable keeps nothing (its column is
//                 the dataset's);
//   k_ts_init, k_ts_chunk<NC>, k_ts_solve (candidates of at most 9 columns)
//                 the column-pivoted QR (Eigen's ColPivHouseholderQR,
//                 ColPivHouseholderQR.h:482-607) carried out on the
//                 triangular factor of the data: per 4096-row chunk an
//                 unpivoted Householder QR of [finite columns | y] held in
//                 registers, then one wave per candidate folds the chunk
//                 factors together in chunk order and runs Eigen's pivot
//                 choice, rank cut, reflectors and back substitution on the
//                 small factor;
//   k_qr_* (candidates of more than 9 columns)
//                 the same factorisation step by step over the columns in
//                 HBM, every candidate of the launch in lockstep (described
//                 below);
//   k_score_fp64 + k_fit_fp64
//                 per (candidate, chunk) the OLS tree's prediction from the
//                 coefficients and the kept columns (rils_rols_cpp.cpp:
//                 488-517: left-deep sum of c*f, or f when |c-1| < 1e-12,
//                 skipping |c| < 1e-12, free term last) and the chunk's sum of
//                 squared residuals; per candidate the chunk sums in order
//                 into 1-R2 and RMSE (fitness, 520-541).
// A tree scored as given (RAW mode) takes only the two scoring kernels, the
// tree evaluated in place of the OLS prediction.
#include "rr_device.h"
#include "rr_kernels.h"

namespace rr {

namespace {

constexpr double EPS12 = 1e-12;                 // node.h:13-14
constexpr double DBL_EPS = 2.220446049250313e-16;
constexpr double DBL_MIN_ = 2.2250738585072014e-308;

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

// Variable index of a program that is one LDV, else -1.
__device__ inline int bare_variable(const uint32_t* code, int len) {
  return (len == 1 && (code[0] & 0xffu) == RR_OP_LDV) ? (int)(code[0] >> 8) : -1;
}

// ---------------------------------------------------------------- the QR
// The batch's candidates (launched positions 0..n_pos-1, sorted by column
// count, most first) factor in lockstep: step p of every candidate with more
// than p columns runs as
//   k_qr_begin   one thread per candidate: pivot choice, rank threshold,
//                reflector (beta, tau), R(p,p);
//   k_qr_dot     one workgroup per (candidate, 4096-row chunk): the chunk's
//                Householder dot products (v / (c0 - beta)) . x for every
//                target column and y;
//   k_qr_mid     one thread per candidate: dot products over the chunks in
//                order, top-row update of the targets;
//   k_qr_update  one workgroup per (candidate, chunk): x -= tau v t for rows
//                > p, and the chunk's tail sums of squares of the targets;
//   k_qr_end     one thread per candidate: tail sums over the chunks in
//                order, LAWN-176 norm downdate.
// Every sum is in the chunked order (rrm::chunk_sum), so a pass spreads over
// as many workgroups as the candidate has chunks.  State between kernels
// lives in global memory (QrState and the per-column arrays).
constexpr int QR_THREADS = 256;
static_assert(QR_THREADS == rrm::NORM_PARTIALS, "one thread per partial of the chunked order");
constexpr int QR_G = 4;   // target columns per pass over a chunk's rows

enum : int32_t { QF_DONE = 1, QF_PASS = 2, QF_APPLY = 4, QF_NOUPD = 8 };

struct QrArgs {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* coefs;
  rr_fit* fit;
  const int32_t* dead;
  int32_t* info;
  const int32_t* order;      // launched position -> candidate
  double* slab;
  const int64_t* slab_off;   // per launched position
  const int32_t* pos_item;   // per launched position: first chunk work item
  int items_per_pos;         // > 0: every position has that many (one dataset)
  int n_pos;                 // positions taking part in this launch
  const double* norm_part;   // k_eval_fp64's chunk sums of squares
  const int32_t* prog_item;  // per non-bare program: its factor's first evaluation work item
  const int64_t* prog_kept;  // per non-bare program: slab offset of its factor's kept column
  QrState* st;
  int32_t* phys;             // per position: ks entries
  double* cold;              // normU, normD, tsq, cc, tdot: 5 blocks of (positions x ks)
  int ks;
  int64_t col_block;         // positions x ks of the whole launch chunk
  double* part;              // per chunk work item: 2 ks doubles
};

struct Cand {   // a candidate's data as one thread or workgroup sees it (uniform)
  int cand, pbeg, nf, k;
  int64_t n, ldx;
  const DsDesc* D;
  double* S;
};

__device__ inline Cand cand_of(const QrArgs& a, int pos) {
  Cand c;
  c.cand = a.order[pos];
  c.D = a.ds + (a.cand_ds ? a.cand_ds[c.cand] : 0);
  c.pbeg = a.cand_start[c.cand];
  c.nf = a.cand_start[c.cand + 1] - c.pbeg;
  c.k = c.nf + 1;
  c.n = c.D->n;
  c.ldx = c.D->ldx;
  c.S = a.slab + a.slab_off[pos];
  return c;
}

// where the values of physical column col are before the first reflector:
// the free term's ones, a bare variable's X column, or the kept column
__device__ inline const double* origin_col(const QrArgs& a, const Cand& c, int col) {
  if (col == c.nf) return c.D->one;
  const int var = bare_variable(a.code + a.prog_start[c.pbeg + col],
                                a.prog_start[c.pbeg + col + 1] - a.prog_start[c.pbeg + col]);
  return var >= 0 ? c.D->X + (int64_t)var * c.ldx : a.slab + a.prog_kept[c.pbeg + col];
}

__device__ inline double* cold(const QrArgs& a, int which, int pos, int j) {
  return a.cold + which * a.col_block + (int64_t)pos * a.ks + j;
}
#define NU(j) (*cold(a, 0, pos, (j)))
#define ND(j) (*cold(a, 1, pos, (j)))
#define TS(j) (*cold(a, 2, pos, (j)))
#define CCS(j) (*cold(a, 3, pos, (j)))
#define TD(j) (*cold(a, 4, pos, (j)))
#define PHYS(j) (a.phys[(int64_t)pos * a.ks + (j)])

__device__ inline int chunk_pos(const QrArgs& a, int item, int n_pos, int64_t* first) {
  int pos;
  if (a.items_per_pos > 0) {
    pos = item / a.items_per_pos;
    *first = (int64_t)pos * a.items_per_pos;
  } else {
    int lo = 0, hi = n_pos - 1;   // last position whose first item <= item (uniform)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.pos_item[mid] <= item) lo = mid; else hi = mid - 1;
    }
    pos = lo;
    *first = a.pos_item[pos];
  }
  return pos;
}

__device__ inline int64_t first_item(const QrArgs& a, int pos) {
  return a.items_per_pos > 0 ? (int64_t)pos * a.items_per_pos : a.pos_item[pos];
}

// chunk sum of NV per-thread partials, in the chunked order; thread 0 gets them
template <int NV>
__device__ inline void chunk_sums(double (&v)[NV], double (*red)[NV]) {
  const int wave = (int)(threadIdx.x / WAVE);
#pragma unroll
  for (int t = 0; t < NV; ++t) v[t] = wave_sum(v[t]);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
#pragma unroll
    for (int t = 0; t < NV; ++t) red[wave][t] = v[t];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      double s = red[0][t];
      for (int w = 1; w < QR_THREADS / WAVE; ++w) s = s + red[w][t];
      v[t] = s;
    }
  }
  __syncthreads();   // red is reused
}

// ---- initial state: norms from the evaluation's chunk sums, row 0 of the
// working columns.  Factor 0 NaN at a row >= 1 (its rows >= 1 sum is NaN
// exactly then) fixes the outcome: its norm is NaN, so the rank threshold is
// NaN (no cut), factor 0 pivots first (nothing compares greater than NaN),
// its tail norm and hence its reflector are NaN, and with the threshold NaN
// no later column stops the factorisation: every coefficient is NaN, rank =
// min(n, k).
// Chunk sums added in chunk order, NV interleaved series at once: out[v] =
// base[v] + base[stride + v] + ... + base[(n-1) stride + v], left to right.
// The loads of eight chunks are issued before their additions.
template <int NV>
__device__ inline void ordered_sums(const double* base, int64_t stride, int64_t n, double* out) {
#pragma unroll
  for (int v = 0; v < NV; ++v) out[v] = base[v];
  int64_t q = 1;
  for (; q + 8 <= n; q += 8) {
    double x[8][NV];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) x[u][v] = base[(q + u) * stride + v];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) out[v] = out[v] + x[u][v];
  }
  for (; q < n; ++q)
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = out[v] + base[q * stride + v];
}

__device__ inline void qr_init_pos(const QrArgs& a, int pos) {
  QrState& s = a.st[pos];
  s.flags = QF_DONE;
  s.nan_exit = -2;   // nothing for k_qr_solve
  const int cand0 = a.order[pos];
  if (a.dead && a.dead[cand0]) return;
  const Cand c = cand_of(a, pos);
  const int64_t nch = (c.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
  double* const co = a.coefs + (int64_t)c.pbeg + c.cand;
  for (int j = 0; j < c.nf; ++j) {
    const int var = bare_variable(a.code + a.prog_start[c.pbeg + j],
                                  a.prog_start[c.pbeg + j + 1] - a.prog_start[c.pbeg + j]);
    double s0, s1;
    if (var >= 0) {
      s0 = c.D->xnorm[2 * var];
      s1 = c.D->xnorm[2 * var + 1];
    } else {
      double s01[2];
      ordered_sums<2>(a.norm_part + 2 * (int64_t)a.prog_item[c.pbeg + j], 2, nch, s01);
      s0 = s01[0];
      s1 = s01[1];
    }
    if (j == 0 && s1 != s1) {
      const int size0 = (int)((c.n < (int64_t)c.k) ? c.n : (int64_t)c.k);
      for (int i = 0; i < c.k; ++i) co[i] = (i < size0) ? rrm::qnan() : 0.0;
      write_fit(a.fit + c.cand, rrm::qnan(), c.D->sst, c.n, size0, 0);
      a.info[c.cand] = -1;
      return;
    }
    ND(j) = __builtin_sqrt(s0);
    TS(j) = s1;
  }
  ND(c.nf) = __builtin_sqrt((double)c.n);
  TS(c.nf) = (double)(c.n - 1);
  for (int j = 0; j < c.k; ++j) { PHYS(j) = j; NU(j) = ND(j); }
  // row 0 of every working column; rows >= 1 are written by the first
  // update pass (the single-row and top-row updates work on row p in place)
  for (int col = 0; col < c.k; ++col) c.S[(int64_t)col * c.ldx] = origin_col(a, c, col)[0];
  c.S[(int64_t)c.k * c.ldx] = c.D->y[0];
  double mx = NU(0);
  for (int j = 1; j < c.k; ++j) { const double v = NU(j); mx = (mx < v) ? v : mx; }
  s.thr = ((mx * DBL_EPS) * (mx * DBL_EPS)) / (double)c.n;
  s.k = c.k;
  s.size = (int)((c.n < (int64_t)c.k) ? c.n : (int64_t)c.k);
  s.p = 0;
  s.nonzero = s.size;
  s.nan_exit = -1;
  s.steps = 0;
  s.flags = 0;
}

// the norm downdate of column j after step p (each column on its own)
__device__ inline void lawn_downdate_col(const QrArgs& a, int pos, const Cand& c, int p, int j, bool recompute_empty) {
  const double dd_thr = __builtin_sqrt(DBL_EPS);
  const double nu = NU(j);
  if (nu != 0.0) {
    const double apj = c.S[(int64_t)PHYS(j) * c.ldx + p];
    double temp = __builtin_fabs(apj) / nu;
    temp = (1.0 + temp) * (1.0 - temp);
    temp = temp < 0.0 ? 0.0 : temp;
    const double ratio = nu / ND(j);
    const double temp2 = temp * (ratio * ratio);
    const double rec = recompute_empty ? 0.0 : __builtin_sqrt(CCS(j));
    if (temp2 <= dd_thr) { ND(j) = rec; NU(j) = rec; }
    else NU(j) = nu * __builtin_sqrt(temp);
  }
}
__device__ inline void lawn_downdate(const QrArgs& a, int pos, const Cand& c, int p, bool recompute_empty) {
  for (int j = p + 1; j < c.k; ++j) lawn_downdate_col(a, pos, c, p, j, recompute_empty);
}

__device__ inline void qr_begin_pos(const QrArgs& a, int pos) {
  QrState& s = a.st[pos];
  if (s.flags & QF_DONE) return;
  const int p = s.p;
  if (p >= s.size) { s.flags = QF_DONE; return; }
  const Cand c = cand_of(a, pos);
  int bi = p;
  double best = NU(p);
  for (int j = p + 1; j < c.k; ++j) { const double v = NU(j); if (v > best) { best = v; bi = j; } }
  if (best * best < s.thr * (double)(c.n - p)) { s.nonzero = p; s.flags = QF_DONE; return; }
  if (bi != p) {
    const int ti = PHYS(p); PHYS(p) = PHYS(bi); PHYS(bi) = ti;
    double t;
    t = NU(p); NU(p) = NU(bi); NU(bi) = t;
    t = ND(p); ND(p) = ND(bi); ND(bi) = t;
    t = TS(p); TS(p) = TS(bi); TS(bi) = t;
  }
  double* const v = c.S + (int64_t)PHYS(p) * c.ldx;
  const double c0 = v[p];
  const double tailSq = (c.n - p == 1) ? 0.0 : TS(p);
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
  if (!zero_tail && beta != beta && c.n - p > 1) {
    // A NaN reflector turns every remaining column and Q^T y into NaN, so
    // R(p,p) = NaN and every coefficient up to the rank is NaN.  The norm
    // downdate leaves columns with norm 0 at 0 and makes all others NaN;
    // later steps pivot in place, and the first column at 0 stops the
    // factorisation if the threshold is positive.
    s.nan_exit = p;
    int r = s.size;
    if (s.thr > 0.0)
      for (int j = p + 1; j < s.size; ++j) if (NU(j) == 0.0) { r = j; break; }
    s.nonzero = r;
    s.flags = QF_DONE;
    return;
  }
  const double den = c0 - beta;
  v[p] = beta;
  const int nrem = c.k - 1 - p;   // remaining A columns: positions p+1..k-1
  s.steps = s.steps + 1;
  if (c.n - p == 1) {
    // applyHouseholderOnTheLeft with a single row: scale by (1 - tau)
    for (int t = 0; t < nrem; ++t) c.S[(int64_t)PHYS(p + 1 + t) * c.ldx + p] *= (1.0 - tau);
    c.S[(int64_t)c.k * c.ldx + p] *= (1.0 - tau);
    for (int t = 0; t < nrem; ++t) TS(p + 1 + t) = 0.0;
    lawn_downdate(a, pos, c, p, true);   // the recomputation's sums over rows p+1.. are empty
    s.p = p + 1;
    s.flags = 0;
    return;
  }
  s.tau = tau;
  s.den = den;
  s.flags = QF_PASS | (tau != 0.0 ? QF_APPLY : 0) | (nrem == 0 ? QF_NOUPD : 0);
}

// target t (0..ntarget-1) of step p: physical column PHYS(p+1+t), or y last
__device__ inline double* target_col(const QrArgs& a, int pos, const Cand& c, int p, int t, int nrem) {
  return t < nrem ? c.S + (int64_t)PHYS(p + 1 + t) * c.ldx : c.S + (int64_t)c.k * c.ldx;
}
__device__ inline const double* target_src(const QrArgs& a, int pos, const Cand& c, int p, int t, int nrem) {
  if (p > 0) return target_col(a, pos, c, p, t, nrem);
  return t < nrem ? origin_col(a, c, PHYS(1 + t)) : c.D->y;
}

// Row passes over one chunk [r0, r1) for GN target columns, QR_DOT_R or
// QR_UPD_R rows per thread per step: every load of a step is issued before
// its arithmetic.
// Thread t still takes rows r0 + t (mod QR_THREADS) in increasing order (the
// partials of rrm::chunk_sum); rows at or beyond r1 load row r0 instead and
// are not used.  The columns are device memory: global address space.
// (2 and 4 rows per step measure the same: the passes stream at the same
// rate; 2 needs fewer registers)
constexpr int QR_R = 4;       // rows per thread per step of the OLS scoring
constexpr int QR_DOT_R = 2;   // ... of the dot pass
constexpr int QR_UPD_R = 2;   // ... of the update pass
typedef const __attribute__((address_space(1))) double* gcptr;
typedef __attribute__((address_space(1))) double* gptr;

// acc[t] += e_i * src[t][i] over rows i > p, e_i = vr[i] / den
template <int GN>
__device__ inline void dot_rows(const double* vr_, const double* const* src_, double den, int p, int64_t r0, int64_t r1,
                                double* acc) {
  const gcptr vr = (gcptr)vr_;
  gcptr src[GN];
#pragma unroll
  for (int t = 0; t < GN; ++t) src[t] = (gcptr)src_[t];
  for (int64_t base = r0; base < r1; base += (int64_t)QR_THREADS * QR_DOT_R) {
    double v[QR_DOT_R], x[QR_DOT_R][GN];
#pragma unroll
    for (int r = 0; r < QR_DOT_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      const int64_t ii = i < r1 ? i : r0;
      v[r] = vr[ii];
#pragma unroll
      for (int t = 0; t < GN; ++t) x[r][t] = src[t][ii];
    }
#pragma unroll
    for (int r = 0; r < QR_DOT_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      if (i < r1 && i > p) {
        const double e = v[r] / den;   // the essential part of the reflector
#pragma unroll
        for (int t = 0; t < GN; ++t) acc[t] = acc[t] + e * x[r][t];
      }
    }
  }
}

// rows i > p: x = src[t][i] (- tau * (vr[i] / den) * td[t] when applying),
// stored to dst[t][i] (rows up to wlim) when applying or at p == 0;
// s12[t] += x^2, and s12[QR_G + t] += x^2 for rows >= p + 2
template <int GN>
__device__ inline void update_rows(const double* vr_, const double* const* src_, double* const* dst_, const double* td,
                                   double den, double tau, bool apply, int p, int64_t wlim, int64_t r0, int64_t r1,
                                   double* s12) {
  const gcptr vr = (gcptr)vr_;
  gcptr src[GN];
  gptr dst[GN];
#pragma unroll
  for (int t = 0; t < GN; ++t) { src[t] = (gcptr)src_[t]; dst[t] = (gptr)dst_[t]; }
  const bool store = apply || p == 0;
  for (int64_t base = r0; base < r1; base += (int64_t)QR_THREADS * QR_UPD_R) {
    double v[QR_UPD_R], x[QR_UPD_R][GN];
#pragma unroll
    for (int r = 0; r < QR_UPD_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      const int64_t ii = i < r1 ? i : r0;
      v[r] = apply ? vr[ii] : 0.0;
#pragma unroll
      for (int t = 0; t < GN; ++t) x[r][t] = src[t][ii];
    }
#pragma unroll
    for (int r = 0; r < QR_UPD_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      if (i < r1 && i > p) {
        const double te = apply ? tau * (v[r] / den) : 0.0;
#pragma unroll
        for (int t = 0; t < GN; ++t) {
          double y = x[r][t];
          if (apply) y = y - te * td[t];
          if (store && i <= wlim) dst[t][i] = y;
          s12[t] = s12[t] + y * y;
          if (i >= p + 2) s12[QR_G + t] = s12[QR_G + t] + y * y;
        }
      }
    }
  }
}

// The last step's dot product (only y left) when the second-to-last step
// stored row p of its two targets only: rows i > p of the pivot column and
// of y are recomputed from that step's sources with its update's operations
// (x = src - tau' (vr'_i / den') td'), then acc += (x_pivot / den) x_y.
__device__ inline void dot_last_rows(const double* vrp_, const double* s0_, const double* s1_, const QrState& s,
                                     int p, int64_t r0, int64_t r1, double* acc) {
  const gcptr vrp = (gcptr)vrp_, s0 = (gcptr)s0_, s1 = (gcptr)s1_;
  const bool papply = s.papply != 0;
  for (int64_t base = r0; base < r1; base += (int64_t)QR_THREADS * QR_DOT_R) {
    double v[QR_DOT_R], x0[QR_DOT_R], x1[QR_DOT_R];
#pragma unroll
    for (int r = 0; r < QR_DOT_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      const int64_t ii = i < r1 ? i : r0;
      v[r] = papply ? vrp[ii] : 0.0;
      x0[r] = s0[ii];
      x1[r] = s1[ii];
    }
#pragma unroll
    for (int r = 0; r < QR_DOT_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      if (i < r1 && i > p) {
        double xp = x0[r], xy = x1[r];
        if (papply) {
          const double te = s.ptau * (v[r] / s.pden);
          xp = xp - te * s.ptd0;
          xy = xy - te * s.ptd1;
        }
        const double e = xp / s.den;   // the essential part of the reflector
        acc[0] = acc[0] + e * xy;
      }
    }
  }
}

// Candidates with three or more columns: the third-to-last step (q) stores
// row q+1 only of its three targets; the second-to-last step's passes
// recompute their inputs from step q's sources (x = src - tau_q (vr_q,i /
// den_q) td_q, the update's own operations) and store rows up to its p+1;
// the last step recomputes through both reflectors.
struct StepQ {   // step q's reflector and sources, for one candidate
  gcptr vr, sP, sT, sy;   // pivot of step q; the sources of physical columns P, T and of y
  double tau, den, tdP, tdT, tdy;
  bool apply;
};
__device__ inline StepQ step_q(const QrArgs& a, int pos, const Cand& c, const QrState& s, int q, int P, int T) {
  StepQ r;
  r.vr = (gcptr)((q == 0) ? origin_col(a, c, PHYS(0)) : c.S + (int64_t)PHYS(q) * c.ldx);
  r.sP = (gcptr)((q == 0) ? origin_col(a, c, P) : c.S + (int64_t)P * c.ldx);
  r.sT = (gcptr)((q == 0) ? origin_col(a, c, T) : c.S + (int64_t)T * c.ldx);
  r.sy = (gcptr)((q == 0) ? c.D->y : c.S + (int64_t)c.k * c.ldx);
  r.tau = s.qtau;
  r.den = s.qden;
  r.apply = s.qapply != 0;
  r.tdP = (P == s.qphys0) ? s.qtd0 : s.qtd1;
  r.tdT = (T == s.qphys0) ? s.qtd0 : s.qtd1;
  r.tdy = s.qtd2;
  return r;
}

// rows [r0, r1) beyond row p of x^(q+1) for P, T and y, QR_DOT_R rows per
// thread per step (thread t: rows r0 + t mod QR_THREADS, increasing), handed
// to f(i, xP, xT, xy) in that order
template <class F>
__device__ inline void rows_q1(const StepQ& Q, int p, int64_t r0, int64_t r1, F&& f) {
  for (int64_t base = r0; base < r1; base += (int64_t)QR_THREADS * QR_DOT_R) {
    double v[QR_DOT_R], xp[QR_DOT_R], xt[QR_DOT_R], xy[QR_DOT_R];
#pragma unroll
    for (int r = 0; r < QR_DOT_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      const int64_t ii = i < r1 ? i : r0;
      v[r] = Q.apply ? Q.vr[ii] : 0.0;
      xp[r] = Q.sP[ii];
      xt[r] = Q.sT[ii];
      xy[r] = Q.sy[ii];
    }
#pragma unroll
    for (int r = 0; r < QR_DOT_R; ++r) {
      const int64_t i = base + (int64_t)threadIdx.x + (int64_t)QR_THREADS * r;
      if (i < r1 && i > p) {
        double P = xp[r], T = xt[r], Y = xy[r];
        if (Q.apply) {
          const double te = Q.tau * (v[r] / Q.den);
          P = P - te * Q.tdP;
          T = T - te * Q.tdT;
          Y = Y - te * Q.tdy;
        }
        f(i, P, T, Y);
      }
    }
  }
}

__device__ inline void qr_dot_item(const QrArgs& a, int pos, int64_t first, int item, double (*red)[QR_G],
                                   double (*red1)[1]) {
  const QrState& s = a.st[pos];
  if ((s.flags & (QF_PASS | QF_APPLY)) != (QF_PASS | QF_APPLY)) return;   // uniform
  const Cand c = cand_of(a, pos);
  const int p = s.p, nrem = c.k - 1 - p, ntarget = nrem + 1;
  const double den = s.den;
  const double* const vr = (p == 0) ? origin_col(a, c, PHYS(0)) : c.S + (int64_t)PHYS(p) * c.ldx;
  const int64_t r0 = (item - first) * rrm::NORM_CHUNK;
  const int64_t r1 = (r0 + rrm::NORM_CHUNK < c.n) ? r0 + rrm::NORM_CHUNK : c.n;
  double* const out = a.part + (int64_t)item * 2 * a.ks;
  if (nrem == 1 && p >= 1) {   // the second-to-last step, from step p-1's sources
    const StepQ Q = step_q(a, pos, c, s, p - 1, PHYS(p), PHYS(p + 1));
    double acc[QR_G] = {0.0, 0.0, 0.0, 0.0};
    rows_q1(Q, p, r0, r1, [&](int64_t, double xP, double xT, double xy) {
      const double e = xP / den;   // the essential part of the reflector
      acc[0] = acc[0] + e * xT;
      acc[1] = acc[1] + e * xy;
    });
    chunk_sums<QR_G>(acc, red);
    if (threadIdx.x == 0) { out[0] = acc[0]; out[1] = acc[1]; }
    return;
  }
  if (nrem == 0 && p >= 2) {   // the last step, through the two reflectors before it
    const StepQ Q = step_q(a, pos, c, s, p - 2, PHYS(p - 1), PHYS(p));
    const bool papply = s.papply != 0;
    double acc[1] = {0.0};
    rows_q1(Q, p, r0, r1, [&](int64_t, double xP, double xT, double xy) {
      if (papply) {
        const double te = s.ptau * (xP / s.pden);
        xT = xT - te * s.ptd0;
        xy = xy - te * s.ptd1;
      }
      const double e = xT / den;
      acc[0] = acc[0] + e * xy;
    });
    chunk_sums<1>(acc, red1);
    if (threadIdx.x == 0) out[0] = acc[0];
    return;
  }
  if (nrem == 0 && p >= 1) {   // the last step of a two-column candidate: from step 0's sources
    double acc[1] = {0.0};
    const double* const vrp = (p == 1) ? origin_col(a, c, PHYS(0)) : c.S + (int64_t)PHYS(p - 1) * c.ldx;
    dot_last_rows(vrp, target_src(a, pos, c, p - 1, 0, 1), target_src(a, pos, c, p - 1, 1, 1), s, p, r0, r1, acc);
    chunk_sums<1>(acc, red1);
    if (threadIdx.x == 0) out[0] = acc[0];
    return;
  }
  for (int g0 = 0; g0 < ntarget; g0 += QR_G) {
    const int gn = (ntarget - g0 < QR_G) ? ntarget - g0 : QR_G;
    const double* src[QR_G];
#pragma unroll
    for (int t = 0; t < QR_G; ++t) src[t] = target_src(a, pos, c, p, g0 + (t < gn ? t : 0), nrem);
    double acc[QR_G];
#pragma unroll
    for (int t = 0; t < QR_G; ++t) acc[t] = 0.0;
    switch (gn) {   // uniform
      case 1: dot_rows<1>(vr, src, den, p, r0, r1, acc); break;
      case 2: dot_rows<2>(vr, src, den, p, r0, r1, acc); break;
      case 3: dot_rows<3>(vr, src, den, p, r0, r1, acc); break;
      default: dot_rows<QR_G>(vr, src, den, p, r0, r1, acc); break;
    }
    chunk_sums<QR_G>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int t = 0; t < QR_G; ++t)   // constant indices: acc stays in registers
        if (t < gn) out[g0 + t] = acc[t];
    }
  }
}

__device__ inline void qr_mid_pos(const QrArgs& a, int pos) {
  QrState& s = a.st[pos];
  if (!(s.flags & QF_PASS)) return;
  const Cand c = cand_of(a, pos);
  const int p = s.p, nrem = c.k - 1 - p, ntarget = nrem + 1;
  if (s.flags & QF_APPLY) {
    const int64_t first = first_item(a, pos), nch = (c.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
    for (int t = 0; t < ntarget; ++t) {
      double sum;
      ordered_sums<1>(a.part + first * 2 * a.ks + t, 2 * a.ks, nch, &sum);
      double* const col = target_col(a, pos, c, p, t, nrem);
      const double td = sum + col[p];
      TD(t) = td;
      col[p] = col[p] - s.tau * td;
    }
  } else {
    for (int t = 0; t < ntarget; ++t) TD(t) = 0.0;
  }
  if (nrem == 2) {   // the third-to-last step: what the next two steps recompute from
    s.qapply = (s.flags & QF_APPLY) ? 1 : 0;
    s.qtau = s.tau;
    s.qden = s.den;
    s.qtd0 = TD(0);
    s.qtd1 = TD(1);
    s.qtd2 = TD(2);
    s.qphys0 = PHYS(p + 1);
    s.qphys1 = PHYS(p + 2);
  }
  if (nrem == 1) {   // the second-to-last step: what the last step's dot recomputes from
    s.papply = (s.flags & QF_APPLY) ? 1 : 0;
    s.ptau = s.tau;
    s.pden = s.den;
    s.ptd0 = TD(0);
    s.ptd1 = TD(1);
  }
  if (s.flags & QF_NOUPD) {   // only y was left: rows below p of Q^T y are never read
    s.p = p + 1;
    s.flags = 0;
  }
}

__device__ inline void qr_update_item(const QrArgs& a, int pos, int64_t first, int item, double (*red)[2 * QR_G]) {
  const QrState& s = a.st[pos];
  if (!(s.flags & QF_PASS) || (s.flags & QF_NOUPD)) return;   // uniform
  const Cand c = cand_of(a, pos);
  const int p = s.p, nrem = c.k - 1 - p, ntarget = nrem + 1;
  const bool apply = (s.flags & QF_APPLY) != 0;
  const double den = s.den, tau = s.tau;
  // the second- and third-to-last steps store row p + 1 only (the next
  // pivot row and row p + 1 of the other targets); the steps after them
  // recompute the rows below from the sources they read
  const int64_t wlim = (nrem == 1 || nrem == 2) ? (int64_t)p + 1 : INT64_MAX;
  const double* const vr = (p == 0) ? origin_col(a, c, PHYS(0)) : c.S + (int64_t)PHYS(p) * c.ldx;
  const int64_t r0 = (item - first) * rrm::NORM_CHUNK;
  const int64_t r1 = (r0 + rrm::NORM_CHUNK < c.n) ? r0 + rrm::NORM_CHUNK : c.n;
  double* const out = a.part + (int64_t)item * 2 * a.ks;
  if (nrem == 1 && p >= 1) {   // the second-to-last step, from step p-1's sources; rows up to p+1 stored always
    const StepQ Q = step_q(a, pos, c, s, p - 1, PHYS(p), PHYS(p + 1));
    const gptr dT = (gptr)(c.S + (int64_t)PHYS(p + 1) * c.ldx), dy = (gptr)(c.S + (int64_t)c.k * c.ldx);
    const double td0 = TD(0), td1 = TD(1);
    double s12[2 * QR_G] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    rows_q1(Q, p, r0, r1, [&](int64_t i, double xP, double xT, double xy) {
      const double te = apply ? tau * (xP / den) : 0.0;
      if (apply) {
        xT = xT - te * td0;
        xy = xy - te * td1;
      }
      if (i <= wlim) { dT[i] = xT; dy[i] = xy; }
      s12[0] = s12[0] + xT * xT;
      if (i >= p + 2) s12[QR_G] = s12[QR_G] + xT * xT;
    });
    chunk_sums<2 * QR_G>(s12, red);
    if (threadIdx.x == 0) { out[0] = s12[0]; out[1] = s12[QR_G]; }
    return;
  }
  for (int g0 = 0; g0 < ntarget; g0 += QR_G) {
    const int gn = (ntarget - g0 < QR_G) ? ntarget - g0 : QR_G;
    const double* src[QR_G];
    double* dst[QR_G];
    double td[QR_G];
#pragma unroll
    for (int t = 0; t < QR_G; ++t) {
      const int tg = g0 + (t < gn ? t : 0);
      src[t] = target_src(a, pos, c, p, tg, nrem);
      dst[t] = target_col(a, pos, c, p, tg, nrem);
      td[t] = TD(tg);
    }
    double s12[2 * QR_G];   // rows >= p+1, rows >= p+2
#pragma unroll
    for (int t = 0; t < 2 * QR_G; ++t) s12[t] = 0.0;
    switch (gn) {   // uniform
      case 1: update_rows<1>(vr, src, dst, td, den, tau, apply, p, wlim, r0, r1, s12); break;
      case 2: update_rows<2>(vr, src, dst, td, den, tau, apply, p, wlim, r0, r1, s12); break;
      case 3: update_rows<3>(vr, src, dst, td, den, tau, apply, p, wlim, r0, r1, s12); break;
      default: update_rows<QR_G>(vr, src, dst, td, den, tau, apply, p, wlim, r0, r1, s12); break;
    }
    chunk_sums<2 * QR_G>(s12, red);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int t = 0; t < QR_G; ++t)   // constant indices: s12 stays in registers
        if (t < gn && g0 + t < nrem) { out[2 * (g0 + t)] = s12[t]; out[2 * (g0 + t) + 1] = s12[QR_G + t]; }
    }
  }
}

__device__ inline void qr_end_pos(const QrArgs& a, int pos) {
  QrState& s = a.st[pos];
  if (!(s.flags & QF_PASS)) return;
  const Cand c = cand_of(a, pos);
  const int p = s.p, nrem = c.k - 1 - p;
  const int64_t first = first_item(a, pos), nch = (c.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
  for (int t = 0; t < nrem; ++t) {
    double s12[2];
    ordered_sums<2>(a.part + first * 2 * a.ks + 2 * t, 2 * a.ks, nch, s12);
    CCS(p + 1 + t) = s12[0];
    TS(p + 1 + t) = s12[1];
  }
  lawn_downdate(a, pos, c, p, false);
  s.p = p + 1;
  s.flags = 0;
}

// k_qr_mid's and k_qr_end's work with a thread per target column, for a
// workgroup that owns its position (k_qr_all1): each column's arithmetic is
// that of qr_mid_pos / qr_end_pos, and the columns do not touch each other's
// values; the position's bookkeeping follows on thread 0 after a barrier.
// Every thread of the workgroup calls them (the early return is uniform).
__device__ inline void qr_mid_block(const QrArgs& a, int pos) {
  QrState& s = a.st[pos];
  if (!(s.flags & QF_PASS)) return;
  const Cand c = cand_of(a, pos);
  const int p = s.p, nrem = c.k - 1 - p, ntarget = nrem + 1;
  const bool apply = (s.flags & QF_APPLY) != 0;
  const int64_t first = first_item(a, pos), nch = (c.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
  for (int t = (int)threadIdx.x; t < ntarget; t += (int)blockDim.x) {
    if (apply) {
      double sum;
      ordered_sums<1>(a.part + first * 2 * a.ks + t, 2 * a.ks, nch, &sum);
      double* const col = target_col(a, pos, c, p, t, nrem);
      const double td = sum + col[p];
      TD(t) = td;
      col[p] = col[p] - s.tau * td;
    } else {
      TD(t) = 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (nrem == 2) {
    s.qapply = apply ? 1 : 0;
    s.qtau = s.tau;
    s.qden = s.den;
    s.qtd0 = TD(0);
    s.qtd1 = TD(1);
    s.qtd2 = TD(2);
    s.qphys0 = PHYS(p + 1);
    s.qphys1 = PHYS(p + 2);
  }
  if (nrem == 1) {
    s.papply = apply ? 1 : 0;
    s.ptau = s.tau;
    s.pden = s.den;
    s.ptd0 = TD(0);
    s.ptd1 = TD(1);
  }
  if (s.flags & QF_NOUPD) {
    s.p = p + 1;
    s.flags = 0;
  }
}
__device__ inline void qr_end_block(const QrArgs& a, int pos) {
  QrState& s = a.st[pos];
  if (!(s.flags & QF_PASS)) return;
  const Cand c = cand_of(a, pos);
  const int p = s.p, nrem = c.k - 1 - p;
  const int64_t first = first_item(a, pos), nch = (c.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
  for (int t = (int)threadIdx.x; t < nrem; t += (int)blockDim.x) {
    double s12[2];
    ordered_sums<2>(a.part + first * 2 * a.ks + 2 * t, 2 * a.ks, nch, s12);
    CCS(p + 1 + t) = s12[0];
    TS(p + 1 + t) = s12[1];
    lawn_downdate_col(a, pos, c, p, p + 1 + t, false);   // column p + 1 + t: its own values only
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s.p = p + 1;
    s.flags = 0;
  }
}

// back substitution on the leading r x r block, coefficients through the
// column permutation (zeros for the rest); the outcome for the scoring
__device__ inline void qr_solve_pos(const QrArgs& a, int pos) {
  const QrState& s = a.st[pos];
  if (s.nan_exit == -2) return;   // skipped, or decided by factor 0
  const Cand c = cand_of(a, pos);
  double* const co = a.coefs + (int64_t)c.pbeg + c.cand;
  const int r = s.nonzero;
  if (s.nan_exit >= 0) {
    for (int i = 0; i < c.k; ++i) co[PHYS(i)] = (i < r) ? rrm::qnan() : 0.0;
    write_fit(a.fit + c.cand, rrm::qnan(), c.D->sst, c.n, r, s.steps);
    a.info[c.cand] = -1;
    return;
  }
  const double* const cy = c.S + (int64_t)c.k * c.ldx;
  for (int i = 0; i < r; ++i) TD(i) = cy[i];
  for (int i = r - 1; i >= 0; --i) {
    const double* coli = c.S + (int64_t)PHYS(i) * c.ldx;
    const double ci = TD(i) / coli[i];
    for (int l = 0; l < i; ++l) TD(l) = TD(l) - ci * coli[l];
    TD(i) = ci;
  }
  bool any_nan = false;
  for (int i = 0; i < c.k; ++i) {
    double cv = (i < r) ? TD(i) : 0.0;
    if (cv != cv) { any_nan = true; cv = rrm::qnan(); }
    co[PHYS(i)] = cv;
  }
  if (any_nan) {   // a kept NaN coefficient: every prediction is NaN
    write_fit(a.fit + c.cand, rrm::qnan(), c.D->sst, c.n, r, s.steps);
    a.info[c.cand] = -1;
  } else {
    a.info[c.cand] = r | (s.steps << 16);   // scored by k_score_fp64 / k_fit_fp64
  }
}
#undef NU
#undef ND
#undef TS
#undef CCS
#undef TD
#undef PHYS

__global__ __launch_bounds__(256) void k_qr_init(QrArgs a) {
  const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pos < a.n_pos) qr_init_pos(a, pos);
}
__global__ __launch_bounds__(256) void k_qr_begin(QrArgs a) {
  const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pos < a.n_pos) qr_begin_pos(a, pos);
}
__global__ __launch_bounds__(QR_THREADS) void k_qr_dot(QrArgs a) {
  __shared__ double red[QR_THREADS / WAVE][QR_G];
  __shared__ double red1[QR_THREADS / WAVE][1];
  const int item = (int)blockIdx.x;
  int64_t first;
  const int pos = chunk_pos(a, item, a.n_pos, &first);
  qr_dot_item(a, pos, first, item, red, red1);
}
__global__ __launch_bounds__(256) void k_qr_mid(QrArgs a) {
  const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pos < a.n_pos) qr_mid_pos(a, pos);
}
__global__ __launch_bounds__(QR_THREADS) void k_qr_update(QrArgs a) {
  __shared__ double red[QR_THREADS / WAVE][2 * QR_G];
  const int item = (int)blockIdx.x;
  int64_t first;
  const int pos = chunk_pos(a, item, a.n_pos, &first);
  qr_update_item(a, pos, first, item, red);
}
__global__ __launch_bounds__(256) void k_qr_end(QrArgs a) {
  const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pos < a.n_pos) qr_end_pos(a, pos);
}
__global__ __launch_bounds__(256) void k_qr_solve(QrArgs a) {
  const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pos < a.n_pos) qr_solve_pos(a, pos);
}

// The whole step-synchronous QR of positions that have one chunk each (n <=
// rrm::NORM_CHUNK): one workgroup per position runs k_qr_init's work, then
// for each of its k steps (a position with k columns takes part in steps 0 ..
// k-1) the work of k_qr_begin, k_qr_dot, k_qr_mid, k_qr_update and k_qr_end
// on its one chunk work item, then k_qr_solve's: the same device code in the
// same order, one launch instead of five per step (the mid and end parts with
// a thread per target column: qr_mid_block, qr_end_block).  The position's state,
// per-column arrays and chunk sums live in LDS for the launch (nothing reads
// them afterwards): the serial per-position parts, on thread 0, then wait on
// LDS rather than on global memory.  The barriers between the parts make each
// part's writes visible to the next.  Dynamic LDS: 7 ks doubles + ks ints.
__global__ __launch_bounds__(QR_THREADS) void k_qr_all1(QrArgs a) {
  __shared__ double red[QR_THREADS / WAVE][QR_G];
  __shared__ double red1[QR_THREADS / WAVE][1];
  __shared__ double red2[QR_THREADS / WAVE][2 * QR_G];
  __shared__ QrState lst;
  extern __shared__ double qr_lds[];
  const int pos = (int)blockIdx.x;
  const int64_t first = first_item(a, pos);
  const int item = (int)first;
  const int64_t ks = a.ks;
  // the same addressing as the global arrays, one position and one item deep
  // (bases moved back by integer arithmetic: the parts add the offsets again)
  auto rebase = [](const void* p, int64_t bytes) { return (uintptr_t)p - (uintptr_t)bytes; };
  QrArgs b = a;
  b.st = (QrState*)rebase(&lst, pos * (int64_t)sizeof(QrState));
  b.cold = (double*)rebase(qr_lds, pos * ks * 8);
  b.col_block = ks;
  b.part = (double*)rebase(qr_lds + 5 * ks, first * 2 * ks * 8);
  b.phys = (int32_t*)rebase(qr_lds + 7 * ks, pos * ks * 4);
  const bool t0 = threadIdx.x == 0;
  if (t0) qr_init_pos(b, pos);
  __syncthreads();
  const int k = a.cand_start[a.order[pos] + 1] - a.cand_start[a.order[pos]] + 1;
  for (int st = 0; st < k; ++st) {
    if (t0) qr_begin_pos(b, pos);
    __syncthreads();
    if (lst.flags & QF_DONE) break;   // uniform; every later part returns at once
    qr_dot_item(b, pos, first, item, red, red1);
    __syncthreads();
    qr_mid_block(b, pos);
    __syncthreads();
    qr_update_item(b, pos, first, item, red2);
    __syncthreads();
    qr_end_block(b, pos);
    __syncthreads();
  }
  if (t0) qr_solve_pos(b, pos);
}

// Factor evaluation for the QR: one workgroup of EV_THREADS threads per
// (unique factor, chunk of rrm::NORM_CHUNK rows).  A unique factor is one
// program over one dataset, shared by every candidate of the launch that
// uses it.  Work item b is the b-th chunk in factor order; fac_item[f] is the
// first item of factor f.  Besides the values, each work item leaves its
// chunk's two sums of squares (rows >= 0 and rows >= 1) in the chunked order
// for the QR's norms.
struct EvalArgs {
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* slab;
  const int32_t* fac_prog;   // per factor: its program
  const int32_t* fac_cand;   // per factor: a candidate using it (for the dataset)
  const int32_t* fac_item;   // per factor: first work item
  const int64_t* fac_kept;   // per factor: slab offset of its kept column
  const int32_t* alive;      // per factor: 0 when no live candidate uses it
  int n_facs;
  int items_per_fac;         // > 0: every factor has this many work items (one dataset)
  double* norm_part;         // per work item: its chunk's sums of squares (rows >= 0, rows >= 1)
};
constexpr int EV_THREADS = 256;
static_assert(EVAL_CHUNK_ROWS == rrm::NORM_CHUNK, "evaluation work items are the chunks of the norm order");

// FORM 0: eval_f64 with its stack in registers and the transcendentals
// called; eight waves per SIMD (at two rows per thread the interpreter then
// fits in 64 VGPRs without new spills).  FORM 1: eval_lds_f64, the operand
// stack in dynamic LDS (the launch's deepest program x R x EV_THREADS
// doubles) and the transcendentals inlined.
template <int R, int FORM>
__global__ __launch_bounds__(EV_THREADS, FORM == 0 ? 8 : R <= 2 ? 6 : 4) void k_eval_fp64(EvalArgs a) {
  static_assert(EV_THREADS == rrm::NORM_PARTIALS, "one thread per partial of the chunked order");
  __shared__ double red[EV_THREADS / WAVE][2];
  extern __shared__ double ev_stk[];
  const int item = (int)blockIdx.x;
  int lo;
  if (a.items_per_fac > 0) {
    lo = item / a.items_per_fac;
  } else {   // last factor whose first item <= item (uniform)
    lo = 0;
    int hi = a.n_facs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.fac_item[mid] <= item) lo = mid; else hi = mid - 1;
    }
  }
  if (!a.alive[lo]) return;
  const int cand = a.fac_cand[lo], prog = a.fac_prog[lo];
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const uint32_t* code = a.code + a.prog_start[prog];
  const int len = a.prog_start[prog + 1] - a.prog_start[prog];
  double* const keep = a.slab + a.fac_kept[lo];
  const int64_t r0 = (int64_t)(item - a.fac_item[lo]) * rrm::NORM_CHUNK;
  const int64_t r1 = (r0 + rrm::NORM_CHUNK < D.n) ? r0 + rrm::NORM_CHUNK : D.n;
  // thread t adds the squares of rows r0 + t (mod 256) in increasing order:
  // the chunk's partials of rrm::chunk_sum.  base is a multiple of
  // EV_THREADS * R (which divides 4096) and below n, so every row touched is
  // below ldx.
  double s[2] = {0.0, 0.0};   // rows >= 0, rows >= 1
  for (int64_t base = r0; base < r1; base += (int64_t)EV_THREADS * R) {
    const int64_t row0 = base + (int64_t)threadIdx.x;
    double v[R];
    if constexpr (FORM == 1) eval_lds_f64<R, EV_THREADS>(code, len, a.K, D.X, D.ldx, row0, v, ev_stk);
    else eval_f64<R, EV_THREADS>(code, len, a.K, D.X, D.ldx, row0, v);
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

// One thread per unique factor: whether a candidate still to be scored uses it.
__global__ __launch_bounds__(256) void k_alive_fp64(const int32_t* ustart, const int32_t* users, const int32_t* dead,
                                                    int32_t n, int32_t* alive) {
  const int u = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (u >= n) return;
  int32_t any = 0;
  for (int32_t i = ustart[u]; i < ustart[u + 1] && !any; ++i) any = dead[users[i]] == 0;
  alive[u] = any;
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

// The factor-0 pre-pass: one workgroup of NAN0_ROWS threads (one wave) per candidate,
// one row each, so it costs one evaluation of latency (the LDS-stack
// interpreter: transcendentals inlined, no calls).
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
  __shared__ double stk[RR_MAX_STACK * NAN0_ROWS];   // eval_lds_f64's operand stack
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
    eval_lds_f64<1, NAN0_ROWS>(code, len, a.K, D.X, D.ldx, (int64_t)tid, v, stk);
    if (__any((int)(tid >= 1 && (int64_t)tid < n && v[0] != v[0])) && (tid & 63) == 0) found = 1;
  }
  __syncthreads();
  if (tid == 0) {
    a.dead[cand] = found;
    if (found) {
      // the outcome k_ts_init / k_qr_init write when they see the same NaN
      const int k = nf + 1;
      const int size0 = (int)((n < (int64_t)k) ? n : (int64_t)k);
      double* co = a.coefs + (int64_t)pbeg + cand;
      for (int i = 0; i < k; ++i) co[i] = (i < size0) ? rrm::qnan() : 0.0;
      write_fit(a.fit + cand, rrm::qnan(), D.sst, n, size0, 0);
      a.fit[cand].flags |= RR_F_EARLY;
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
  const int64_t* prog_kept;  // per non-bare program: slab offset of its factor's kept column
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
__global__ __launch_bounds__(EV_THREADS, 8) void k_score_fp64(ScoreArgs a) {   // as k_eval_fp64
  __shared__ double red[EV_THREADS / WAVE];
  // OLS: per factor its column and coefficient; kind 0 skipped (|c| < 1e-12),
  // 1 the column itself (|c - 1| < 1e-12), 2 c times the column
  __shared__ const double* fcol[RAW ? 1 : RR_MAX_COLS];
  __shared__ double fcoef[RAW ? 1 : RR_MAX_COLS];
  __shared__ int8_t fkind[RAW ? 1 : RR_MAX_COLS];
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
    const double* co = a.coefs + (int64_t)pbeg + cand;
    const double cf_free = co[nf];
    for (int j = (int)threadIdx.x; j < nf; j += EV_THREADS) {
      const double cj = co[j];
      const uint32_t* code = a.code + a.prog_start[pbeg + j];
      const int var = bare_variable(code, a.prog_start[pbeg + j + 1] - a.prog_start[pbeg + j]);
      fcol[j] = var >= 0 ? D.X + (int64_t)var * ldx : a.slab + a.prog_kept[pbeg + j];
      fcoef[j] = cj;
      fkind[j] = (__builtin_fabs(cj) < EPS12) ? 0 : (__builtin_fabs(cj - 1) < EPS12) ? 1 : 2;
    }
    __syncthreads();
    const gcptr yv = (gcptr)D.y;
    // thread t: rows r0 + t (mod 256) in increasing order, QR_R at a time;
    // rows at or beyond r1 read row r0 and are not added
    for (int64_t base = r0; base < r1; base += (int64_t)EV_THREADS * QR_R) {
      int64_t ii[QR_R];
#pragma unroll
      for (int r = 0; r < QR_R; ++r) {
        const int64_t i = base + (int64_t)threadIdx.x + (int64_t)EV_THREADS * r;
        ii[r] = i < r1 ? i : r0;
      }
      double yy[QR_R];   // issued with the first factor's loads: one wait per pass
#pragma unroll
      for (int r = 0; r < QR_R; ++r) yy[r] = yv[ii[r]];
      double yp[QR_R], f[QR_R];
      bool any = false;   // the same for every row
      for (int j = 0; j < nf; ++j) {
        const int kind = fkind[j];
        if (kind == 0) continue;
        const gcptr col = (gcptr)fcol[j];
        const double cj = fcoef[j];
#pragma unroll
        for (int r = 0; r < QR_R; ++r) f[r] = col[ii[r]];
#pragma unroll
        for (int r = 0; r < QR_R; ++r) {
          const double term = kind == 1 ? f[r] : cj * f[r];
          yp[r] = any ? yp[r] + term : term;
        }
        any = true;
      }
      if (!(__builtin_fabs(cf_free) < EPS12)) {
        const double term = cf_free * 1.0;
#pragma unroll
        for (int r = 0; r < QR_R; ++r) yp[r] = any ? yp[r] + term : term;
        any = true;
      }
#pragma unroll
      for (int r = 0; r < QR_R; ++r) {
        if (base + (int64_t)threadIdx.x + (int64_t)EV_THREADS * r < r1) {
          const double d = yy[r] - (any ? yp[r] : 0.0);
          s = s + d * d;
        }
      }
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
  double ssr;
  ordered_sums<1>(a.part + first, 1, nch, &ssr);
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


// ---------------------------------------------------------------- chunked QR
// Candidates with at most TS_MAX_COLS columns: the column-pivoted QR carried
// out on the triangular factor of the data (the canonical algorithm and its
// order are described with the oracle's ts_qr_solve).  State per position in
// TsState; the chunk factors in rpart.
struct TsArgs {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* coefs;
  rr_fit* fit;
  const int32_t* dead;
  int32_t* info;
  const int32_t* order;
  const double* slab;
  const double* norm_part;
  const int32_t* prog_item;
  const int64_t* prog_kept;
  const int64_t* rp_off;
  TsState* st;
  double* rpart;
  int32_t* ts_cnt;
  int64_t* ts_desc;
  int64_t list_base[3];
  int pos0, pos1, bucket;
};

// Eigen's makeHouseholder (Householder.h:67-96) from the pivot c0 and the sum
// of squares of the tail; den = c0 - beta divides the tail into the essential part
__device__ inline void ts_householder(double c0, double tail_sq, double& beta, double& tau, double& den) {
  if (tail_sq <= DBL_MIN_) {
    tau = 0.0; beta = c0; den = 1.0;
  } else {
    double b = __builtin_sqrt(c0 * c0 + tail_sq);
    if (c0 >= 0.0) b = -b;
    beta = b;
    den = c0 - b;
    tau = (b - c0) / b;
  }
}

// where physical column col of a candidate lives: the free term's ones, a
// bare variable's X column, or the kept column of its factor
__device__ inline const double* ts_col(const TsArgs& a, const DsDesc& D, int pbeg, int nf, int col) {
  if (col == nf) return D.one;
  const int var = bare_variable(a.code + a.prog_start[pbeg + col], a.prog_start[pbeg + col + 1] - a.prog_start[pbeg + col]);
  return var >= 0 ? D.X + (int64_t)var * D.ldx : a.slab + a.prog_kept[pbeg + col];
}

// Factor j's column norm (from the evaluation's chunk sums, or the
// dataset's feature norms for a bare variable) and where its column is.
__device__ inline void ts_factor(const TsArgs& a, const DsDesc& D, int pbeg, int nf, int j, double* nd,
                                 const double** col) {
  const int var = bare_variable(a.code + a.prog_start[pbeg + j], a.prog_start[pbeg + j + 1] - a.prog_start[pbeg + j]);
  double s0;
  if (var >= 0) {
    s0 = D.xnorm[2 * var];
  } else {
    const int64_t nch_eval = (D.n + rrm::NORM_CHUNK - 1) / rrm::NORM_CHUNK;
    double s01[2];
    ordered_sums<2>(a.norm_part + 2 * (int64_t)a.prog_item[pbeg + j], 2, nch_eval, s01);
    s0 = s01[0];
  }
  *nd = __builtin_sqrt(s0);
  *col = ts_col(a, D, pbeg, nf, j);
}

// One wave per position: a lane per factor finds its norm and column
// (ts_factor, into fnd / fcol); lane 0 then does the position's work (outcome
// of a non-finite step-0 pivot, finite columns, threshold, list slot); then
// the wave writes the chunk work items' descriptors, a lane per chunk.
__device__ inline bool ts_init_position(const TsArgs& a, int pos, int* slot_out, const double* fnd,
                                        const double* const* fcol) {
  TsState& s = a.st[pos];
  s.live = 0;
  const int cand = a.order[pos];
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const int pbeg = a.cand_start[cand];
  const int nf = a.cand_start[cand + 1] - pbeg;
  const int k = nf + 1;
  const int64_t n = D.n;
  double nd[TS_MAX_COLS];
  for (int j = 0; j < nf; ++j) nd[j] = fnd[j];
  nd[nf] = __builtin_sqrt((double)n);
  const int size = (int)((n < (int64_t)k) ? n : (int64_t)k);
  int bi = 0;
  double best = nd[0];
  for (int j = 1; j < k; ++j)
    if (nd[j] > best) { best = nd[j]; bi = j; }
  if (!(best - best == 0.0)) {
    // step 0 pivots a non-finite column (swapped to position 0): its reflector
    // is NaN, nothing is cut after it, every coefficient of the first `size`
    // positions is NaN
    double* const co = a.coefs + (int64_t)pbeg + cand;
    for (int i = 0; i < k; ++i) co[i] = 0.0;
    for (int i = 0; i < size; ++i) co[i == 0 ? bi : (i == bi ? 0 : i)] = rrm::qnan();
    write_fit(a.fit + cand, rrm::qnan(), D.sst, n, size, 0);
    a.info[cand] = -1;
    return false;
  }
  int kf = 0;
  for (int j = 0; j < k; ++j)
    if (nd[j] - nd[j] == 0.0) s.fin[kf++] = j;
  double mx = nd[0];
  for (int j = 1; j < k; ++j) mx = (mx < nd[j]) ? nd[j] : mx;
  s.thr = ((mx * DBL_EPS) * (mx * DBL_EPS)) / (double)n;
  for (int j = 0; j < k; ++j) s.nd[j] = nd[j];
  for (int f = 0; f < kf; ++f) s.col[f] = s.fin[f] == nf ? D.one : fcol[s.fin[f]];
  s.col[kf] = D.y;
  s.k = k;
  s.kf = kf;
  s.size = size;
  const int nch = (int)((n + TS_ROWS - 1) / TS_ROWS);
  s.nch = nch;
  s.n = n;
  s.live = 1;
  const int nb = ts_bucket(k);
  const int b = nb == 10 ? 0 : nb == 5 ? 1 : 2;
  *slot_out = atomicAdd(a.ts_cnt + b, nch);
  return true;
}

__global__ __launch_bounds__(WAVE) void k_ts_init(TsArgs a) {
  const int pos = a.pos0 + (int)blockIdx.x;
  // what the other lanes need of lane 0's work, through LDS
  __shared__ int live, slot, kf_s, nch_s, k_s;
  __shared__ int64_t n_s;
  __shared__ const double* col_s[TS_MAX_COLS + 1];
  __shared__ double fnd[TS_MAX_COLS];
  __shared__ const double* fcol[TS_MAX_COLS];
  const int cand = a.order[pos];
  if (a.dead && a.dead[cand]) {   // uniform
    if (threadIdx.x == 0) a.st[pos].live = 0;
    return;
  }
  {
    const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
    const int pbeg = a.cand_start[cand];
    const int nf = a.cand_start[cand + 1] - pbeg;   // < TS_MAX_COLS
    if ((int)threadIdx.x < nf) ts_factor(a, D, pbeg, nf, (int)threadIdx.x, &fnd[threadIdx.x], &fcol[threadIdx.x]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int sl = 0;
    live = ts_init_position(a, pos, &sl, fnd, fcol) ? 1 : 0;
    slot = sl;
    if (live) {
      const TsState& s = a.st[pos];
      kf_s = s.kf; nch_s = s.nch; k_s = s.k; n_s = s.n;
      for (int f = 0; f <= s.kf; ++f) col_s[f] = s.col[f];
    }
  }
  __syncthreads();
  if (!live) return;   // uniform
  // the position's chunk work items, appended to its bucket's list (the
  // order of the list does not matter: every item writes its own factor) as
  // self-contained descriptors: where its factor goes, its rows, its finite
  // column count, its columns from the chunk's first row
  const int kf = kf_s, nch = nch_s;
  const int64_t n = n_s;
  const int nb = ts_bucket(k_s);
  const int b = nb == 10 ? 0 : nb == 5 ? 1 : 2;
  const int dw = ts_desc_words(nb);
  int64_t* const base = a.ts_desc + a.list_base[b] + (int64_t)slot * dw;
  double* const out = a.rpart + a.rp_off[pos];
  for (int c = (int)threadIdx.x; c < nch; c += WAVE) {
    int64_t* const dst = base + (int64_t)c * dw;
    const int64_t r0 = (int64_t)c * TS_ROWS;
    dst[0] = (int64_t)(out + (int64_t)c * nb * (nb - 1));
    dst[1] = (n - r0 < TS_ROWS) ? n - r0 : TS_ROWS;
    dst[2] = kf;
    for (int f = 0; f <= kf; ++f) dst[3 + f] = (int64_t)(col_s[f] + r0);
  }
}

// A double moved between lanes by a DPP lane pattern (both halves).
template <int CTRL>
__device__ inline double ts_dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double ts_readlane(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a wave by the xor butterfly of strides 1, 2, 4,
// ..., 32 (lane l adds lane l ^ s; a + b and b + a are the same bits, so every
// lane of a combined group holds the same value): strides 1 and 2 by quad
// permutations; then the groups of 4 and 8 are uniform, so stride 4 is the
// mirror within 8 lanes and stride 8 the mirror within 16; strides 16 and 32
// from lanes 0, 16, 32, 48.  Every thread gets the result.  All 64 lanes must
// be active.
__device__ inline double ts_wave_sum(double v) {
  v = v + ts_dpp<0xB1>(v);    // quad_perm [1, 0, 3, 2]: lane ^ 1
  v = v + ts_dpp<0x4E>(v);    // quad_perm [2, 3, 0, 1]: lane ^ 2
  v = v + ts_dpp<0x141>(v);   // row_half_mirror: lane 7 - l of its 8
  v = v + ts_dpp<0x140>(v);   // row_mirror: lane 15 - l of its 16
  return (ts_readlane(v, 0) + ts_readlane(v, 16)) + (ts_readlane(v, 32) + ts_readlane(v, 48));
}

// One chunk work item: unpivoted Householder QR of the chunk's [finite
// columns | y] (m rows, kf finite columns), held in registers: thread t holds
// rows t + THR u, u = 0 .. RPT-1.  Every sum is in the oracle's order for THR
// threads: per thread its rows in order, groups of 64 by ts_wave_sum, the
// group sums in order.  FULL: m == TS_ROWS, so only rows 0..j of step j
// (thread t <= j, u = 0) need a condition.  The group sums go to LDS
// (double-buffered by step parity, buf); with up to 4 groups every thread adds
// them up itself (one barrier per step), with more one lane per column of
// wave 0 adds them and the others read its totals (two barriers).
template <int NC, int THR, bool FULL>
__device__ inline void ts_chunk_item(const int64_t* cur, double (*prow)[NC], double (*red)[THR / WAVE][NC],
                                     double (*totl)[NC], int& buf) {
  constexpr int RPT = (int)(TS_ROWS / THR);
  constexpr int NG = THR / WAVE;
  const int t = (int)threadIdx.x;
  const int lane = t & (WAVE - 1), wave = t / WAVE;
  double* const out = (double*)cur[0];
  const int64_t m = cur[1];
  const int kf = (int)cur[2];
  const int nc = kf + 1;
  double v[NC][RPT];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double* src = c < nc ? (const double*)cur[3 + c] : nullptr;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      const int64_t row = t + (int64_t)THR * u;
      v[c][u] = (src != nullptr && (FULL || row < m)) ? src[row] : 0.0;
    }
  }
  const int steps = (int)((m < (int64_t)kf) ? m : (int64_t)kf);
#pragma unroll
  for (int j = 0; j < NC - 1; ++j) {
    if (j >= steps) continue;   // uniform
    if (t == j) {
#pragma unroll
      for (int c = 0; c < NC; ++c) prow[buf][c] = v[c][0];
    }
    double ps[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) ps[c] = 0.0;
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      const int64_t row = t + (int64_t)THR * u;
      const bool in = FULL ? (u > 0 || t > j) : (row > j && row < m);
      if (in) {
        const double x = v[j][u];
        ps[j] = __builtin_fma(x, x, ps[j]);
#pragma unroll
        for (int c = j + 1; c < NC; ++c)
          if (c < nc) ps[c] = __builtin_fma(x, v[c][u], ps[c]);
      }
    }
#pragma unroll
    for (int c = j; c < NC; ++c)
      if (c < nc) ps[c] = ts_wave_sum(ps[c]);
    if (lane == 0) {
#pragma unroll
      for (int c = j; c < NC; ++c)
        if (c < nc) red[buf][wave][c] = ps[c];
    }
    __syncthreads();
    double tot[NC];
    if constexpr (NG > 4) {
      // many group sums: lane c of wave 0 adds column c's in group order and
      // every thread reads the totals, instead of every thread adding them all
      if (wave == 0 && lane >= j && lane < nc) {
        double acc = red[buf][0][lane];
#pragma unroll
        for (int g = 1; g < NG; ++g) acc = acc + red[buf][g][lane];
        totl[buf][lane] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int c = j; c < NC; ++c)
        if (c < nc) tot[c] = totl[buf][c];
    } else {
#pragma unroll
      for (int c = j; c < NC; ++c) {
        if (c >= nc) continue;
        double acc = red[buf][0][c];
#pragma unroll
        for (int g = 1; g < NG; ++g) acc = acc + red[buf][g][c];
        tot[c] = acc;
      }
    }
    double beta, tau, den;
    ts_householder(prow[buf][j], tot[j], beta, tau, den);
    if (tau != 0.0) {
#pragma unroll
      for (int c = j + 1; c < NC; ++c) {
        if (c >= nc) continue;
        const double w = prow[buf][c] + tot[c] / den;
        const double g = (tau * w) / den;
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const int64_t row = t + (int64_t)THR * u;
          if (u == 0 && t == j) v[c][u] = v[c][u] - tau * w;
          else if (FULL ? (u > 0 || t > j) : (row > j && row < m)) v[c][u] = __builtin_fma(-v[j][u], g, v[c][u]);
        }
      }
    }
    if (t == j) v[j][0] = beta;
    buf ^= 1;   // the next step writes the other half; this one is read until its barrier
  }
  if (t < kf) {   // row t of the chunk's factor (rows past m are zero)
    double* dst = out + (int64_t)t * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (c < nc) dst[c] = ((int64_t)t < m && c >= t) ? v[c][0] : 0.0;
  }
}

// Chunk work items of the live positions, one bucket (positions with NC
// columns with y) at a time: workgroups of ts_threads(NC) threads take items
// blockIdx.x, blockIdx.x + gridDim.x, ... of the bucket's list (written by
// k_ts_init) until the count is reached; the next item's descriptor is loaded
// while the current one computes (up to 5 columns: the wide bucket has no
// scalar registers to spare).
template <int NC>
__global__ __launch_bounds__(ts_threads(NC)) void k_ts_chunk(TsArgs a) {
  constexpr int THR = ts_threads(NC);
  constexpr int DW = ts_desc_words(NC);
  constexpr bool PREFETCH = NC <= 5;
  __shared__ double prow[2][NC];
  __shared__ double red[2][THR / WAVE][NC];
  __shared__ double totl[2][NC];
  const int cnt = a.ts_cnt[a.bucket];   // uniform
  const int64_t* const desc = a.ts_desc + a.list_base[a.bucket];
  int buf = 0;   // LDS half of the next step
  int item = (int)blockIdx.x;
  if (item >= cnt) return;
  int64_t cur[DW];
#pragma unroll
  for (int w = 0; w < DW; ++w) cur[w] = desc[(int64_t)item * DW + w];
  for (;;) {
    const int next = item + (int)gridDim.x;
    int64_t nxt[DW];
#pragma unroll
    for (int w = 0; w < DW; ++w) nxt[w] = (PREFETCH && next < cnt) ? desc[(int64_t)next * DW + w] : 0;
    if (cur[1] == TS_ROWS) ts_chunk_item<NC, THR, true>(cur, prow, red, totl, buf);
    else ts_chunk_item<NC, THR, false>(cur, prow, red, totl, buf);
    if (next >= cnt) break;
    item = next;
#pragma unroll
    for (int w = 0; w < DW; ++w) cur[w] = PREFETCH ? nxt[w] : desc[(int64_t)item * DW + w];
  }
}

constexpr int TS_RW = TS_MAX_COLS * (TS_MAX_COLS + 1);   // words of the largest factor (bucket 10)

// Fold count - 1 factors into R, in order: the ones at src + i * stride, i =
// 1 .. count - 1 (R already holds the one at src).  Each fold replaces R by
// the triangular factor of [R; T] (T's column j has rows 0..j): per column j
// a Householder reflector from R(j, j) and T's column j, applied to the
// columns right of j, sums in row order as in the oracle.  One wave, no
// memory between the steps: lane c (< nc) holds column c of R (r[i] =
// R(i, c), i < kf) and of T in registers, and the values of column j that
// every lane needs are broadcast from lane j by readlane.  The next factor is
// loaded while the current one is folded in.
__device__ inline void ts_load_cols(const double* p, int kf, int nc, int NC, int l, double (&dst)[TS_MAX_COLS]) {
#pragma unroll
  for (int i = 0; i < TS_MAX_COLS; ++i) dst[i] = (i < kf && l < nc) ? p[i * NC + l] : 0.0;
}
__device__ inline void ts_fold_chain(double (&r)[TS_MAX_COLS], const double* src, int64_t stride, int count,
                                     int kf, int nc, int NC, int l) {
  double nx[TS_MAX_COLS];
  if (count > 1) ts_load_cols(src + stride, kf, nc, NC, l, nx);
  for (int f = 1; f < count; ++f) {
    double t[TS_MAX_COLS];
#pragma unroll
    for (int i = 0; i < TS_MAX_COLS; ++i) t[i] = nx[i];
    if (f + 1 < count) ts_load_cols(src + (int64_t)(f + 1) * stride, kf, nc, NC, l, nx);
#pragma unroll
    for (int j = 0; j < TS_MAX_COLS; ++j) {   // [R; T] -> R
      if (j >= kf) continue;   // uniform
      double tj[TS_MAX_COLS];   // T's column j, rows 0..j
#pragma unroll
      for (int i = 0; i <= j; ++i) tj[i] = ts_readlane(t[i], j);
      double sxx = 0.0;
#pragma unroll
      for (int i = 0; i <= j; ++i) sxx = __builtin_fma(tj[i], tj[i], sxx);
      double beta, tau, den;
      ts_householder(ts_readlane(r[j], j), sxx, beta, tau, den);
      if (tau != 0.0 && l > j && l < nc) {
        double sx = 0.0;
#pragma unroll
        for (int i = 0; i <= j; ++i) sx = __builtin_fma(tj[i], t[i], sx);
        const double w = r[j] + sx / den;
        const double g = (tau * w) / den;
        r[j] = r[j] - tau * w;
#pragma unroll
        for (int i = 0; i <= j; ++i) t[i] = __builtin_fma(-tj[i], g, t[i]);
      }
      if (l == j) r[j] = beta;
    }
  }
}

// The first level of the combination: one wave per (position, group of
// TS_GROUP consecutive chunks) folds the group's chunk factors in chunk order
// and leaves the result in the slot of the group's first chunk.
__global__ __launch_bounds__(WAVE) void k_ts_group(TsArgs a) {
  const int pos = a.pos0 + (int)blockIdx.x;
  const TsState& s = a.st[pos];
  const int first = (int)blockIdx.y * TS_GROUP;
  if (!s.live || first >= s.nch) return;   // uniform
  const int count = min(TS_GROUP, s.nch - first);
  if (count < 2) return;                    // a group of one chunk is its own factor
  const int l = (int)threadIdx.x;
  const int kf = s.kf, nc = kf + 1;
  const int NC = ts_bucket(s.k);
  const int W = NC * (NC - 1);   // words per chunk factor
  double* const src = a.rpart + a.rp_off[pos] + (int64_t)first * W;
  double r[TS_MAX_COLS];
  ts_load_cols(src, kf, nc, NC, l, r);
  ts_fold_chain(r, src, W, count, kf, nc, NC, l);
  if (l < nc) {
#pragma unroll
    for (int i = 0; i < TS_MAX_COLS; ++i)
      if (i < kf) src[i * NC + l] = r[i];
  }
}

// One wave per position: the group factors (k_ts_group) combined in group
// order, Eigen's pivoted steps on the result, the solve.  The factor R (kf
// rows, nc columns with y) lives in LDS; in every step lane c carries column
// c, its sums in row order as in the oracle.
__global__ __launch_bounds__(WAVE) void k_ts_solve(TsArgs a) {
  const int pos = a.pos0 + (int)blockIdx.x;
  const TsState& s = a.st[pos];
  if (!s.live) return;   // uniform
  __shared__ double R[TS_RW];
  __shared__ int perm[TS_MAX_COLS], fidx[TS_MAX_COLS];
  __shared__ double nu[TS_MAX_COLS];
  const int l = (int)threadIdx.x;
  const int cand = a.order[pos];
  const DsDesc& D = a.ds[a.cand_ds ? a.cand_ds[cand] : 0];
  const int pbeg = a.cand_start[cand];
  const int k = s.k, kf = s.kf, nc = kf + 1, size = s.size;
  const int NC = ts_bucket(k);
  const int W = NC * (NC - 1);   // words per chunk factor
  const int64_t n = D.n;
  const double* const src = a.rpart + a.rp_off[pos];
  const int groups = (s.nch + TS_GROUP - 1) / TS_GROUP;
  {
    double r[TS_MAX_COLS];
    ts_load_cols(src, kf, nc, NC, l, r);
    ts_fold_chain(r, src, (int64_t)TS_GROUP * W, groups, kf, nc, NC, l);
    if (l < nc) {
#pragma unroll
      for (int i = 0; i < TS_MAX_COLS; ++i)
        if (i < kf) R[i * NC + l] = r[i];
    }
  }
  if (l < k) { perm[l] = l; fidx[l] = -1; nu[l] = s.nd[l]; }
  __syncthreads();
  if (l < kf) fidx[s.fin[l]] = l;
  __syncthreads();
  int nonzero = size, steps = 0;
  bool nan_out = false;
  for (int p = 0; p < size; ++p) {   // every lane follows the same (uniform) decisions
    if (fidx[perm[p]] < 0) { nan_out = true; break; }   // a NaN-norm column reached before the cut
    int bi = p;
    double best = nu[p];
    for (int j = p + 1; j < k; ++j)
      if (nu[j] > best) { best = nu[j]; bi = j; }
    if (best * best < s.thr * (double)(n - p)) { nonzero = p; break; }
    __syncthreads();
    if (l == 0) {
      const int ti = perm[p]; perm[p] = perm[bi]; perm[bi] = ti;
      const double tv = nu[p]; nu[p] = nu[bi]; nu[bi] = tv;
    }
    __syncthreads();
    const int f = fidx[perm[p]];
    double sxx = 0.0;
    for (int i = p + 1; i < kf; ++i) sxx = __builtin_fma(R[i * NC + f], R[i * NC + f], sxx);
    double beta, tau, den;
    ts_householder(R[p * NC + f], sxx, beta, tau, den);
    const int q = p + 1 + l;   // lane l: position q (k: y)
    if (q <= k) {
      const int g = q < k ? fidx[perm[q]] : kf;
      if (g >= 0) {
        if (tau != 0.0) {
          double sx = 0.0;
          for (int i = p + 1; i < kf; ++i) sx = __builtin_fma(R[i * NC + f], R[i * NC + g], sx);
          const double w = R[p * NC + g] + sx / den;
          const double gg = (tau * w) / den;
          R[p * NC + g] = R[p * NC + g] - tau * w;
          for (int i = p + 1; i < kf; ++i) R[i * NC + g] = __builtin_fma(-R[i * NC + f], gg, R[i * NC + g]);
        }
        if (q < k) {
          double sq = 0.0;
          for (int i = p + 1; i < kf; ++i) sq = __builtin_fma(R[i * NC + g], R[i * NC + g], sq);
          nu[q] = __builtin_sqrt(sq);
        }
      }
    }
    __syncthreads();
    if (l == 0) R[p * NC + f] = beta;
    __syncthreads();
    steps = p + 1;
  }
  if (l != 0) return;
  double* const co = a.coefs + (int64_t)pbeg + cand;
  if (nan_out) {
    for (int i = 0; i < k; ++i) co[perm[i]] = (i < size) ? rrm::qnan() : 0.0;
    write_fit(a.fit + cand, rrm::qnan(), D.sst, n, size, steps);
    a.info[cand] = -1;
    return;
  }
  const int r = nonzero;
  double c[TS_MAX_COLS];
  for (int i = 0; i < r; ++i) c[i] = R[i * NC + kf];
  for (int i = r - 1; i >= 0; --i) {
    const int col = fidx[perm[i]];
    c[i] = c[i] / R[i * NC + col];
    for (int q = 0; q < i; ++q) c[q] = __builtin_fma(-c[i], R[q * NC + col], c[q]);
  }
  bool any_nan = false;
  for (int i = 0; i < k; ++i) {
    double cv = (i < r) ? c[i] : 0.0;
    if (cv != cv) { any_nan = true; cv = rrm::qnan(); }
    co[perm[i]] = cv;
  }
  if (any_nan) {   // a NaN coefficient: every prediction is NaN
    write_fit(a.fit + cand, rrm::qnan(), D.sst, n, r, steps);
    a.info[cand] = -1;
  } else {
    a.info[cand] = r | (steps << 16);   // scored by k_score_fp64 / k_fit_fp64
  }
}

__global__ void k_ts_zero(int32_t* cnt) {
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
}

}  // namespace

static QrArgs qr_args(const QrLaunch& L, int n_pos) {
  QrArgs a;
  a.cand_start = L.cand_start; a.prog_start = L.prog_start; a.code = L.code; a.ds = L.ds; a.cand_ds = L.cand_ds;
  a.coefs = L.coefs; a.fit = L.fit; a.dead = L.dead; a.info = L.info;
  a.order = L.order; a.slab = L.slab; a.slab_off = L.slab_off;
  a.pos_item = L.pos_item; a.items_per_pos = L.items_per_pos; a.n_pos = n_pos;
  a.norm_part = L.norm_part; a.prog_item = L.prog_item; a.prog_kept = L.prog_kept;
  a.st = L.st; a.phys = L.phys; a.cold = L.cold; a.ks = L.ks;
  a.col_block = (int64_t)L.n_pos * L.ks;
  a.part = L.part;
  return a;
}

// The per-position kernels are latency-bound serial loops: one wave per
// workgroup spreads a launch over more compute units.
constexpr int POS_THREADS = 64;

hipError_t launch_qr_init(const QrLaunch& L, hipStream_t s) {
  if (L.n_pos <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_qr_init, dim3((unsigned)((L.n_pos + POS_THREADS - 1) / POS_THREADS)), dim3(POS_THREADS), 0, s,
                     qr_args(L, L.n_pos));
  return hipGetLastError();
}

hipError_t launch_qr_step(const QrLaunch& L, int n_act, int n_items, hipStream_t s) {
  if (n_act <= 0) return hipSuccess;
  const QrArgs a = qr_args(L, n_act);
  const dim3 pgrid((unsigned)((n_act + POS_THREADS - 1) / POS_THREADS)), pblock(POS_THREADS), cgrid((unsigned)n_items),
      cblock(QR_THREADS);
  hipLaunchKernelGGL(k_qr_begin, pgrid, pblock, 0, s, a);
  hipLaunchKernelGGL(k_qr_dot, cgrid, cblock, 0, s, a);
  hipLaunchKernelGGL(k_qr_mid, pgrid, pblock, 0, s, a);
  hipLaunchKernelGGL(k_qr_update, cgrid, cblock, 0, s, a);
  hipLaunchKernelGGL(k_qr_end, pgrid, pblock, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_qr_all1(const QrLaunch& L, hipStream_t s) {
  if (L.n_pos <= 0) return hipSuccess;
  const size_t lds = (size_t)L.ks * (7 * sizeof(double) + sizeof(int32_t));
  hipLaunchKernelGGL(k_qr_all1, dim3((unsigned)L.n_pos), dim3(QR_THREADS), lds, s, qr_args(L, L.n_pos));
  return hipGetLastError();
}

hipError_t launch_qr_solve(const QrLaunch& L, hipStream_t s) {
  if (L.n_pos <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_qr_solve, dim3((unsigned)((L.n_pos + POS_THREADS - 1) / POS_THREADS)), dim3(POS_THREADS), 0, s,
                     qr_args(L, L.n_pos));
  return hipGetLastError();
}

hipError_t launch_eval_fp64(const EvalLaunch& L, hipStream_t s) {
  if (L.n_items <= 0) return hipSuccess;
  EvalArgs a{L.prog_start, L.code,     L.K,        L.ds,       L.cand_ds, L.slab,          L.fac_prog,
             L.fac_cand,   L.fac_item, L.fac_kept, L.alive, L.n_facs,  L.items_per_fac, L.norm_part};
  const dim3 grid((unsigned)L.n_items), block(EV_THREADS);
  // form 1's LDS stack: depth x rows x EV_THREADS doubles, within the 64 KiB
  // a workgroup may ask for; deeper launches take form 0
  const size_t stk = (size_t)(L.depth > 0 ? L.depth : 1) * (size_t)L.rows * EV_THREADS * sizeof(double);
  if (L.form == 1 && stk <= 65536 && (L.rows == 1 || L.rows == 2 || L.rows == 4)) {
    if (L.rows == 1) hipLaunchKernelGGL((k_eval_fp64<1, 1>), grid, block, stk, s, a);
    else if (L.rows == 4) hipLaunchKernelGGL((k_eval_fp64<4, 1>), grid, block, stk, s, a);
    else hipLaunchKernelGGL((k_eval_fp64<2, 1>), grid, block, stk, s, a);
  } else {
    if (L.rows == 1) hipLaunchKernelGGL((k_eval_fp64<1, 0>), grid, block, 0, s, a);
    else if (L.rows == 4) hipLaunchKernelGGL((k_eval_fp64<4, 0>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_eval_fp64<2, 0>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_alive_fp64(const int32_t* ustart, const int32_t* users, const int32_t* dead, int32_t n,
                             int32_t* alive, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_alive_fp64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ustart, users, dead, n, alive);
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
  ScoreArgs a{L.cand_start, L.prog_start, L.code, L.K, L.ds, L.cand_ds, L.coefs, L.slab, L.prog_kept, L.order,
              L.dead, L.info, L.pos_item, L.n_pos, L.items_per_pos, L.part, L.fit};
  const dim3 grid((unsigned)L.n_items), block(EV_THREADS), fgrid((unsigned)((L.n_pos + POS_THREADS - 1) / POS_THREADS));
  if (L.raw) {
    if (L.rows == 1) hipLaunchKernelGGL((k_score_fp64<true, 1>), grid, block, 0, s, a);
    else if (L.rows == 4) hipLaunchKernelGGL((k_score_fp64<true, 4>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_score_fp64<true, 2>), grid, block, 0, s, a);
    hipLaunchKernelGGL(k_fit_fp64<true>, fgrid, dim3(POS_THREADS), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_score_fp64<false, 1>), grid, block, 0, s, a);
    hipLaunchKernelGGL(k_fit_fp64<false>, fgrid, dim3(POS_THREADS), 0, s, a);
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


static TsArgs ts_args(const TsLaunch& L, int pos0, int pos1, int bucket) {
  return TsArgs{L.cand_start, L.prog_start, L.code, L.ds, L.cand_ds, L.coefs, L.fit, L.dead, L.info, L.order,
                L.slab, L.norm_part, L.prog_item, L.prog_kept, L.rp_off, L.st, L.rpart, L.cnt, L.desc,
                {L.list_base[0], L.list_base[1], L.list_base[2]}, pos0, pos1, bucket};
}

hipError_t launch_ts_init(const TsLaunch& L, int pos0, int pos1, hipStream_t s) {
  if (pos1 <= pos0) return hipSuccess;
  hipLaunchKernelGGL(k_ts_zero, dim3(1), dim3(WAVE), 0, s, L.cnt);
  hipLaunchKernelGGL(k_ts_init, dim3((unsigned)(pos1 - pos0)), dim3(WAVE), 0, s, ts_args(L, pos0, pos1, 0));
  return hipGetLastError();
}

hipError_t launch_ts_chunk(const TsLaunch& L, int bucket, int max_items, hipStream_t s) {
  if (max_items <= 0) return hipSuccess;
  const TsArgs a = ts_args(L, 0, 0, bucket);
  const dim3 grid((unsigned)(max_items < TS_GRID ? max_items : TS_GRID));
  if (bucket == 2) hipLaunchKernelGGL(k_ts_chunk<3>, grid, dim3(ts_threads(3)), 0, s, a);
  else if (bucket == 1) hipLaunchKernelGGL(k_ts_chunk<5>, grid, dim3(ts_threads(5)), 0, s, a);
  else hipLaunchKernelGGL(k_ts_chunk<10>, grid, dim3(ts_threads(10)), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ts_solve(const TsLaunch& L, int pos0, int pos1, int64_t max_chunks, hipStream_t s) {
  if (pos1 <= pos0) return hipSuccess;
  const TsArgs a = ts_args(L, pos0, pos1, 0);
  const int64_t groups = (max_chunks + TS_GROUP - 1) / TS_GROUP;
  if (groups > 65535) return hipErrorInvalidValue;
  if (max_chunks > 1)
    hipLaunchKernelGGL(k_ts_group, dim3((unsigned)(pos1 - pos0), (unsigned)groups), dim3(WAVE), 0, s, a);
  hipLaunchKernelGGL(k_ts_solve, dim3((unsigned)(pos1 - pos0)), dim3(WAVE), 0, s, a);
  return hipGetLastError();
}

}  // namespace rr
