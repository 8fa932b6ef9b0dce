// rr_kernels.h -- launch descriptors of the fitness kernels (host side of
// the kernels, used by the C-ABI implementation).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rr_fitness.h"

namespace rr {

// One dataset as the kernels see it.  A launch carries a table of these and,
// per candidate, the index of the one it is scored against (cand_ds; null
// means entry 0 for every candidate), so candidates of independent runs over
// different datasets can share one launch.
struct DsDesc {
  const double* X;    // feature-major FP64, column stride ldx
  const float* X32;   // FP32 copy, same layout
  const double* y;
  int64_t n, ldx;
  double sst, ysum, yy;   // sum (y - mean)^2 (canonical order), sum y, sum y^2
  const double* one;      // ldx ones: the column of the OLS free term
  const double* xnorm;    // per feature: sums of squares over rows >= 0 and >= 1 (chunked order)
};

// Step-synchronous column-pivoted QR over the positions 0..n_pos-1 of a
// launch (sorted by column count, most first; pointers already offset to the
// launch's first position): launch_qr_init, then per step p launch_qr_step
// over the first n_act positions (those with more than p columns) and their
// chunk work items, then launch_qr_solve.  ks: the most columns of any
// position; phys: n_pos x ks ints; cold: 5 x n_pos x ks doubles; part: 2 ks
// doubles per chunk work item.
struct QrState {
  int32_t k, size, p, nonzero, nan_exit, steps, flags, papply;
  int32_t qphys0, qphys1, qapply, pad;   // the third-to-last step: its two A targets, reflector applied
  double thr, tau, den;
  double ptau, pden, ptd0, ptd1;          // the second-to-last step's reflector and dot products
  double qtau, qden, qtd0, qtd1, qtd2;    // the third-to-last step's (targets qphys0, qphys1, y)
};
struct QrLaunch {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* coefs;
  rr_fit* fit;
  const int32_t* dead;
  int32_t* info;              // per candidate: -1 (fitness written) or rank | steps << 16 (to be scored)
  const int32_t* order;       // launched position -> candidate
  double* slab;
  const int64_t* slab_off;    // per launched position
  const int32_t* pos_item;    // per launched position: first chunk work item
  int items_per_pos;          // > 0 when every position has that many chunk items
  int n_pos;
  const double* norm_part;    // the evaluation's chunk sums (EvalLaunch::norm_part)
  const int32_t* prog_item;   // per non-bare program: the first evaluation work item of its factor
  const int64_t* prog_kept;   // per non-bare program: slab offset of its factor's kept column
  QrState* st;
  int32_t* phys;
  double* cold;
  int ks;
  double* part;
};

// Chunked column-pivoted QR of the FP64 OLS form, for candidates with at most
// TS_MAX_COLS columns (free term included):
//   launch_ts_init   one thread per position: the initial norms (from the
//                    evaluation's chunk sums), the outcome when step 0 would
//                    pivot a non-finite column, the finite columns, the rank
//                    threshold; the chunk work items (position, chunk) of the
//                    live positions appended to their column bucket's list;
//   launch_ts_chunk  per bucket, a grid of workgroups over the bucket's list:
//                    unpivoted Householder QR of a chunk of TS_ROWS rows of
//                    [finite columns | y], its triangular factor to rpart;
//   launch_ts_solve  the chunk factors combined in two levels -- one wave
//                    per (position, group of TS_GROUP consecutive chunks)
//                    folds the group's factors in chunk order into the slot
//                    of its first chunk, then one wave per position folds
//                    the group factors in group order -- and on the result
//                    the pivoted steps and the solve; coefficients, info (as
//                    the step-synchronous QR).  max_chunks: the most chunks
//                    of any position in the range.
// Positions pos0..pos1-1 of the arrays (pointers not offset).  Column buckets
// 0, 1, 2: 10, 5, 3 columns with y (ts_bucket).  Position p's chunk factors
// sit at rpart + rp_off[p] + chunk * W, W = nc * (nc - 1) words for its
// bucket's nc.  desc: the chunk work items, ts_desc_words(nc) words each
// (factor address, rows, finite column count, the nc column addresses at the
// chunk's first row), bucket b's from word list_base[b] on, room for as many
// as its positions have chunks; cnt: 3 counters.
constexpr int TS_MAX_COLS = 9;
constexpr int64_t TS_ROWS = 4096;
constexpr int TS_GRID = 2048;   // workgroups of a chunk launch at most
constexpr int TS_GROUP = 16;    // chunks per group of the two-level combination
struct TsState {
  int32_t k, kf, size, live;            // live: 1 when the chunk and solve kernels have work
  int32_t fin[TS_MAX_COLS];             // the finite columns, in column order
  int32_t nch, pad;
  int64_t n;
  double thr;
  double nd[TS_MAX_COLS];               // initial column norms
  const double* col[TS_MAX_COLS + 1];   // the finite columns' data, then y
};
struct TsLaunch {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* coefs;
  rr_fit* fit;
  const int32_t* dead;
  int32_t* info;
  const int32_t* order;       // position -> candidate
  const double* slab;
  const double* norm_part;
  const int32_t* prog_item;
  const int64_t* prog_kept;
  const int64_t* rp_off;      // per position: offset of its chunk factors in rpart
  TsState* st;
  double* rpart;
  int32_t* cnt;
  int64_t* desc;
  int64_t list_base[3];
};
hipError_t launch_ts_init(const TsLaunch& L, int pos0, int pos1, hipStream_t s);
// bucket 0, 1, 2; max_items: the bucket's chunk work items if every position
// were live
hipError_t launch_ts_chunk(const TsLaunch& L, int bucket, int max_items, hipStream_t s);
hipError_t launch_ts_solve(const TsLaunch& L, int pos0, int pos1, int64_t max_chunks, hipStream_t s);
// the column count with y of the bucket of a candidate with k columns (free term included, k <= TS_MAX_COLS)
__host__ __device__ inline int ts_bucket(int k) { return k <= 2 ? 3 : k <= 4 ? 5 : 10; }
// words of a chunk work item's descriptor in the bucket of nc columns with y
__host__ __device__ constexpr int ts_desc_words(int nc) { return 3 + nc; }
// threads of the bucket's chunk sums (the order of the oracle's ts_dot): 256
// threads of 16 rows for 3 and 5 columns, 1024 of 4 rows for 10
__host__ __device__ constexpr int ts_threads(int nc) { return nc <= 5 ? 256 : 1024; }

// Scoring of the FP64 form: one workgroup per (launched position, chunk of
// EVAL_CHUNK_ROWS rows) computes the chunk's sum of squared residuals --
// RAW: of the candidate's one program; OLS: of the OLS tree from the
// coefficients and the factor columns (kept slab columns, at slab +
// prog_kept[program], or X), for the candidates whose info is >= 0 -- and
// then one thread per position adds the
// chunk sums in order and writes the fitness.  Position p owns work items
// pos_item[p] .. pos_item[p+1]-1 (items_per_pos > 0: p * items_per_pos ..).
struct ScoreLaunch {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  const double* coefs;       // OLS
  const double* slab;        // OLS
  const int64_t* prog_kept;  // OLS, per non-bare program
  const int32_t* order;      // launched position -> candidate (null: identity)
  const int32_t* dead;       // or null
  const int32_t* info;       // OLS: written by the QR (k_ts_init / k_ts_solve, or k_qr_init / k_qr_solve)
  const int32_t* pos_item;
  int n_pos;
  int n_items;
  int items_per_pos;
  int raw;
  int rows;                  // RAW: rows per thread per pass of the evaluation (1, 2, 4)
  double* part;              // per work item
  rr_fit* fit;
};

// Factor-0 pre-pass of the OLS form: one workgroup per candidate evaluates
// factor 0 in FP64 on rows 0..NAN0_ROWS-1.  A NaN at a row >= 1 fixes the outcome (every
// coefficient NaN, the sentinel fitness; the FP64 kernel reaches the same
// result when it sees it), which is written here; dead[c] = 1 then, else 0.
struct Nan0Launch {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* coefs;
  rr_fit* fit;
  int32_t* dead;
  int n_cands;
};
constexpr int NAN0_ROWS = 64;

// Factor evaluation of the FP64 OLS form into the kept slab columns: one
// workgroup per (unique factor, chunk of EVAL_CHUNK_ROWS rows).  A unique
// factor is a program over a dataset that one or more candidates use;
// factor f is program fac_prog[f] over the dataset of candidate fac_cand[f],
// kept at slab + fac_kept[f], and owns work items fac_item[f] ..
// fac_item[f+1]-1, n_items in all.  Work item b also leaves the sums of
// squares of its chunk in norm_part[2b], norm_part[2b+1].  Factors with
// alive[f] == 0 (every user candidate settled by the factor-0 pre-pass) are
// skipped.  Must run before the QR launches over the same candidates.
struct EvalLaunch {
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;
  const int32_t* cand_ds;
  double* slab;
  const int32_t* fac_prog;
  const int32_t* fac_cand;
  const int32_t* fac_item;
  const int64_t* fac_kept;
  const int32_t* alive;
  int n_facs;
  int n_items;
  int items_per_fac;         // > 0 when every factor has that many items (factor = item / items_per_fac)
  double* norm_part;         // 2 doubles per work item
  int rows;                  // rows per thread per pass: 1, 2 or 4
  int form;                  // 0: operand stack in registers, transcendentals called; 1: stack in LDS, inlined
  int depth;                 // the deepest operand stack of the launch's factor programs
};
constexpr int64_t EVAL_CHUNK_ROWS = 4096;   // the chunk of the norms' order (rrm::NORM_CHUNK)

hipError_t launch_qr_init(const QrLaunch& L, hipStream_t s);
hipError_t launch_qr_step(const QrLaunch& L, int n_act, int n_items, hipStream_t s);
hipError_t launch_qr_solve(const QrLaunch& L, hipStream_t s);
// The same QR (init, every step, solve) in one launch when every position has
// one chunk work item (n <= EVAL_CHUNK_ROWS).
hipError_t launch_qr_all1(const QrLaunch& L, hipStream_t s);
hipError_t launch_eval_fp64(const EvalLaunch& L, hipStream_t s);
// alive[u] = 1 if some candidate users[ustart[u] .. ustart[u+1]-1] has
// dead == 0, else 0, for u < n
hipError_t launch_alive_fp64(const int32_t* ustart, const int32_t* users, const int32_t* dead, int32_t n,
                             int32_t* alive, hipStream_t s);
// out[2f], out[2f+1]: sums of squares of feature f's column over rows >= 0
// and rows >= 1 in the chunked order (rrm::chunk_sum)
hipError_t launch_xnorm_fp64(const double* X, int64_t n, int64_t ldx, int32_t d, double* out, hipStream_t s);
hipError_t launch_nan0_fp64(const Nan0Launch& L, hipStream_t s);
hipError_t launch_score_fp64(const ScoreLaunch& L, hipStream_t s);
hipError_t launch_predict_fp64(const uint32_t* code, int len, const double* K, const double* X, int64_t n,
                               int64_t ldx, double* out, hipStream_t s);
hipError_t launch_fill_f64(double* p, int64_t count, double value, hipStream_t s);
// out4 = {mean(y), sum((y - mean)^2), sum(y), sum(y^2)}
hipError_t launch_stats_fp64(const double* y, int64_t n, double* out4, hipStream_t s);

}  // namespace rr
