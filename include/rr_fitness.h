/* rr_fitness.h -- C ABI of the MI355X candidate-fitness library (librr_hip.so).
 *
 * The reference has no internal seam here: its hot path is a private chain of
 * member functions in one translation unit,
 *     local_search (rils_rols_cpp/rils_rols_cpp.cpp:601-643)
 *       -> tune_constants (445-518) -> fitness (520-541) -> R2/RMSE (40-49)
 *       -> node::evaluate_all (rils_rols_cpp/node.cpp:5-95)
 *       -> Eigen ColPivHouseholderQR::solve (rils_rols_cpp.cpp:484).
 * This header is the boundary that replaces that chain with one batched call
 * per local-search sweep.  The host side (tree rewrites, factor extraction,
 * OLS-tree construction, acceptance) stays in C++ above it; everything that
 * touches the n samples happens on the GPU below it.
 *
 * Rules: plain C types only; caller-owned host buffers; library-owned device
 * memory keyed by handles; every function returns 0 on success and a negative
 * RR_E_* code on failure, with a message from rr_last_error() (per thread).
 * A context is used by one host thread at a time.
 */
#ifndef RR_FITNESS_H
#define RR_FITNESS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes */
#define RR_OK 0
#define RR_E_ARG (-1)     /* invalid argument (sizes, handles, program text) */
#define RR_E_HIP (-2)     /* HIP runtime error (no device, launch failure, ...) */
#define RR_E_NOMEM (-3)   /* device allocation failed */
#define RR_E_LIMIT (-4)   /* a candidate exceeds a compiled limit (columns, stack) */

/* ---- candidate programs ------------------------------------------------
 * Each expression tree is compiled on the host to postfix bytecode for an
 * accumulator machine with a small operand stack.  One 32-bit instruction:
 *     bits 0..7   opcode
 *     bits 8..31  operand: feature index (VAR forms) or constant-pool index
 * A binary opcode is RR_OP_BIN + 8*bop + mode, where bop is one of RR_BOP_*
 * and mode says where the two operands come from (acc = accumulator):
 *     RR_M_S  : left = pop(), right = acc        RR_M_SR : left = acc, right = pop()
 *     RR_M_V  : left = acc, right = X[arg]       RR_M_VR : left = X[arg], right = acc
 *     RR_M_C  : left = acc, right = K[arg]       RR_M_CR : left = K[arg], right = acc
 * Operators and their per-sample semantics are those of node::evaluate_inner
 * (rils_rols_cpp/node.cpp:23-95).  POW is not representable: the reference
 * never generates it (setup_nodes, rils_rols_cpp.cpp:121-148). */
#define RR_OP_LDV 1u    /* acc = X[arg]       (node.cpp:30-32) */
#define RR_OP_LDC 2u    /* acc = K[arg]       (node.cpp:25-29) */
#define RR_OP_PUSH 3u   /* push(acc) */
#define RR_OP_SIN 4u    /* node.cpp:41-52 */
#define RR_OP_COS 5u
#define RR_OP_LN 6u
#define RR_OP_EXP 7u
#define RR_OP_SQRT 8u
#define RR_OP_SQR 9u
#define RR_OP_BIN 16u

#define RR_BOP_ADD 0u   /* node.cpp:33-40 */
#define RR_BOP_SUB 1u
#define RR_BOP_MUL 2u
#define RR_BOP_DIV 3u
#define RR_BOP_LT 4u    /* node.cpp:55-78: 1.0 / 0.0 */
#define RR_BOP_GT 5u
#define RR_BOP_EQ 6u
#define RR_BOP_NEQ 7u
#define RR_BOP_MIN 8u   /* node.cpp:79-90 */
#define RR_BOP_MAX 9u

#define RR_M_S 0u
#define RR_M_SR 1u
#define RR_M_V 2u
#define RR_M_VR 3u
#define RR_M_C 4u
#define RR_M_CR 5u

#define RR_MAX_STACK 8      /* operand-stack depth a program may use */
#define RR_MAX_COLS 512     /* OLS columns per candidate (factors + free term) */

/* A batch of candidates.  Candidate c owns programs
 * [cand_start[c], cand_start[c+1]); program p owns instructions
 * [prog_start[p], prog_start[p+1]) of `code`; operands of LDC and *_C forms
 * index `consts`. */
typedef struct rr_batch {
  int32_t n_cands;
  const int32_t* cand_start;  /* n_cands + 1 */
  const int32_t* prog_start;  /* cand_start[n_cands] + 1 */
  const uint32_t* code;
  const double* consts;
  int32_t n_consts;
} rr_batch;

/* mode of rr_fitness_batch */
#define RR_MODE_OLS 0   /* tune_constants + fitness (rils_rols_cpp.cpp:445-541):
                           programs are the additive factors of the rewritten
                           candidate; the free term is implicit and last. */
#define RR_MODE_RAW 1   /* fitness only (rils_rols_cpp.cpp:520-541): exactly
                           one program per candidate, the whole tree. */

/* precision of rr_fitness_batch */
#define RR_FP64 0        /* plain form: bit-for-bit the canonical FP64 algorithm */
#define RR_FP32_REFINED 1 /* single pass: FP32 evaluation with per-value error
                             bounds, Gram sums of the FP32 values in FP64 (exact
                             products), FP64 solve with one refinement step, and
                             a certified bound rr_fit.err; a result is certified
                             only when its R2 provably lies within 1e-9 of the
                             FP64 form's; candidates flagged
                             RR_F_NEEDS64 must be re-scored in FP64 before they
                             decide anything */
#define RR_FP32_CHECKED 2 /* OLS mode only: RR_FP32_REFINED with the FP64 form
                             in the same call.  Candidates with a factor that
                             reads no variable or with more than 8 factors go
                             to RR_FP64 directly (concurrently with the FP32
                             kernels), and those the FP32 form flags
                             RR_F_NEEDS64 are re-scored in RR_FP64; their
                             results carry RR_F_FP64 and are bit-for-bit those
                             of RR_FP64.  No result is left RR_F_NEEDS64. */

/* flags of rr_fit */
#define RR_F_SENTINEL 1  /* R2 or RMSE was NaN: fitness is (1000, 1000, 1000) */
#define RR_F_NEEDS64 2   /* FP32 form could not certify this candidate: fitness
                            and coefficients (NaN) must come from RR_FP64 */
#define RR_F_FP64 4      /* RR_FP32_CHECKED: this result came from the FP64 form */
#define RR_F_EARLY 8     /* OLS mode: decided by the factor-0 pre-pass (factor 0
                            NaN at one of the first 64 samples, sample 0 aside):
                            the sentinel, no other pass over the samples */
/* FP64 OLS form: bits 16..31 of flags hold the number of Householder steps
 * that ran passes over the samples (for traffic accounting) */
#define RR_F_STEPS(flags) (((uint32_t)(flags)) >> 16)

typedef struct rr_fit {
  double e0;      /* 1 - R2 (rils_rols_cpp.cpp:538) */
  double e1;      /* RMSE */
  double err;     /* FP32 form: bound on |fitness_value error| relative; 0 in FP64 */
  int32_t rank;   /* nonzero pivots of the column-pivoted QR (OLS mode); FP32
                     form: -reason when flagged RR_F_NEEDS64 (1 uncertain
                     value, 2 fewer samples than columns, 3 NaN column with
                     ill-conditioned finite columns, 4 ill-conditioned, 5 a
                     coefficient threshold inside the bound, 6 bound too wide,
                     7 more than 8 factors, 8 the bound on |R2 - R2_FP64|
                     above 1e-9) */
  int32_t flags;  /* RR_F_* */
} rr_fit;

/* ---- context ----------------------------------------------------------- */
typedef struct rr_ctx rr_ctx;

/* Create a context on HIP device `device` (its own non-blocking stream). */
int rr_create(int device, rr_ctx** out);
int rr_destroy(rr_ctx* ctx);
/* Message of the last error on this thread ("" if none). */
const char* rr_last_error(void);

/* ---- datasets -----------------------------------------------------------
 * Replaces the feature-major vector<ArrayXd> X and ArrayXd y that fit_inner
 * builds (rils_rols_cpp.cpp:675-715, 772-795).  X_fm holds d columns of n
 * samples (column j at X_fm + j*n), already sampled/shuffled by the caller.
 * The data stay resident in HBM until rr_dataset_free. */
int rr_dataset_upload(rr_ctx* ctx, const double* X_fm, const double* y, int64_t n, int32_t d, int32_t* handle);
int rr_dataset_free(rr_ctx* ctx, int32_t handle);
/* mean(y) and sum((y - mean)^2) in the canonical order (R2, rils_rols_cpp.cpp:40-45) */
int rr_dataset_stats(rr_ctx* ctx, int32_t handle, double* ybar, double* sst);

/* ---- the hot path -------------------------------------------------------
 * One call scores every candidate of `batch` against dataset `handle`.
 * OLS mode: coefs_out receives k_c = (#programs of c) + 1 coefficients per
 * candidate, at offset cand_start[c] + c, in factor order with the free term
 * last (the vector `coefs` of rils_rols_cpp.cpp:484); NaN coefficients are
 * written as the x86 default NaN.  fit_out[c] receives 1-R2 and RMSE of the
 * OLS tree that tune_constants builds from those coefficients (the size
 * component of the fitness tuple is the caller's).  RAW mode: coefs_out may
 * be NULL. */
int rr_fitness_batch(rr_ctx* ctx, int32_t handle, const rr_batch* batch, int32_t mode, int32_t precision,
                     double* coefs_out, rr_fit* fit_out);

/* Validate a batch as rr_fitness_batch does before it launches anything:
 * structure (monotone offsets, one program per candidate in RAW mode, at most
 * RR_MAX_COLS OLS columns), and every program against the bytecode rules
 * above (known opcodes, operands inside [0, n_features) and [0, n_consts),
 * an accumulator where an operator needs one, stack depth within
 * RR_MAX_STACK, and a balanced stack with a value at the end).  Needs no
 * context or device.  No counterpart in the reference, whose trees are
 * objects rather than code. */
int rr_check_batch(const rr_batch* batch, int32_t mode, int32_t n_features);

/* The same for candidates of several runs at once (continuous batching of
 * independent searches): candidate c is scored against dataset handles[c].
 * Results per candidate are bit-for-bit those of rr_fitness_batch on its own
 * dataset (every candidate's sums run on one wave in a fixed order).  No
 * counterpart in the reference, which scores one candidate at a time; it
 * serves the request loop that keeps many regression runs resident. */
int rr_fitness_multi(rr_ctx* ctx, const int32_t* handles, const rr_batch* batch, int32_t mode, int32_t precision,
                     double* coefs_out, rr_fit* fit_out);

/* Pipelined form of rr_fitness_batch(RR_FP64) for callers with independent
 * batches: rr_fitness_submit checks, plans and launches the batch and returns
 * without waiting for it; rr_fitness_wait completes the oldest pending call,
 * writing its coefs_out / fit_out.  At most two calls are pending per context
 * (a third submit fails with RR_E_ARG), so the host can check and plan the
 * next batch while the current one runs, its kernels queued behind on the
 * context's stream.  The batch arrays may be reused once submit returns; the
 * output arrays must stay valid until the matching wait.  Any other call on
 * the context (fitness, predict, dataset upload or free, autotune) completes
 * the pending calls first; every submit is still retired by its own wait,
 * which then returns at once.  With option "phase_timing" or "graph" set,
 * submit runs the call to completion the same way.  Results are
 * bit-for-bit those of rr_fitness_batch.  No counterpart in the reference,
 * whose fitness calls are synchronous member functions. */
int rr_fitness_submit(rr_ctx* ctx, int32_t handle, const rr_batch* batch, int32_t mode, double* coefs_out,
                      rr_fit* fit_out);
/* The same for candidates of several runs (as rr_fitness_multi): the serving
 * loop submits the next merged launch while the current one runs. */
int rr_fitness_submit_multi(rr_ctx* ctx, const int32_t* handles, const rr_batch* batch, int32_t mode,
                            double* coefs_out, rr_fit* fit_out);
int rr_fitness_wait(rr_ctx* ctx);

/* The plain FP64 form of rr_fitness_batch (RR_MODE_OLS or RR_MODE_RAW, one
 * dataset) with no device or pinned-host allocation: no counterpart in the
 * reference; it is for callers that cannot allocate (code that must not call
 * hipMalloc, such as serving loops or harnesses that forbid it).  Every work
 * buffer is carved from one fixed device array of RR_NOALLOC_SCRATCH_BYTES
 * per GPU, resident from the library's first use on that GPU until the
 * process ends; the FP64 work columns run in slices that fit it, and a batch
 * that still needs more fails with RR_E_LIMIT before anything is launched
 * that depends on the missing memory.  Host-to-device copies come straight
 * from the caller's buffers (synchronously).  Every call on one GPU shares
 * that GPU's array and leaves its contents behind: calls that overlap on one
 * GPU (from any context or thread) race and must be serialized by the
 * caller.  Results are bit-for-bit those of rr_fitness_batch(RR_FP64). */
#define RR_NOALLOC_SCRATCH_MB 256
#define RR_NOALLOC_SCRATCH_BYTES (RR_NOALLOC_SCRATCH_MB * 1048576u)
int rr_fitness_batch_noalloc(rr_ctx* ctx, int32_t handle, const rr_batch* batch, int32_t mode, double* coefs_out,
                             rr_fit* fit_out);

/* Evaluate one program over all samples of the dataset (node::evaluate_all,
 * used by rils_rols::predict, rils_rols_cpp.cpp:730-750). */
int rr_predict(rr_ctx* ctx, int32_t handle, const uint32_t* code, int32_t len, const double* consts,
               int32_t n_consts, double* out);

/* FP32 evaluation of one program with a per-sample bound: |out_v[i] - v64[i]|
 * <= out_e[i], where v64 is what rr_predict returns.  out_e[i] = +inf where
 * FP32 cannot vouch for the sample (overflow, a domain edge inside the
 * bound); out_v[i] = NaN with out_e[i] = 0 where v64[i] is certainly NaN.
 * No counterpart in the reference: it exposes the evaluator of the
 * RR_FP32_REFINED precision form, whose bounds the certificates rest on. */
int rr_predict_fp32(rr_ctx* ctx, int32_t handle, const uint32_t* code, int32_t len, const double* consts,
                    int32_t n_consts, float* out_v, float* out_e);

/* ---- tuning ---------------------------------------------------------------
 * Named integer options: "slab_budget_mb" (device memory the FP64 OLS path
 * may use for its per-candidate work columns, default 32768),
 * "fp64_eval_rows" (samples per thread per pass of the FP64 evaluation of
 * factors and trees: 1, 2 or 4; default 2), "fp64_eval_form" (the FP64
 * factor interpreter: 0 operand stack in registers and transcendentals
 * called, 1 operand stack in LDS and transcendentals inlined; default 1), "fp32_rows" (the same for the FP32 kernels: 1 or 2; default 2),
 * "fp32_chunk_rows" (samples per workgroup of the FP32 pass, a multiple of
 * 512; default 4096: a candidate's samples are split into chunks that run as
 * independent workgroups, and the chunk sums are added in chunk order),
 * "fp32_streams" (1: the FP32 factor-count buckets run on concurrent streams;
 * default 0, back to back on the context's stream), "host_threads" (host
 * threads a call may use to check a large batch, 1..64; default 4),
 * "phase_timing" (1: record rr_last_phase_ms; default 0), "graph" (1: the
 * plain FP64 form's launches of a call are captured into a HIP graph and
 * replayed through an executable graph kept per launch sequence, up to 16,
 * updated in place with the call's arguments; same bits; default 0).
 * rr_get_option also reads "graph_made" and "graph_updated" (executable
 * graphs instantiated / reused so far).
 * Unknown names return RR_E_ARG. */
int rr_set_option(rr_ctx* ctx, const char* name, int64_t value);

/* Autotuner: time `batch` (mode, precision) under every schedule of that
 * precision -- FP64: samples per thread of the evaluation 1, 2, 4 x
 * fp64_eval_form 0, 1; FP32:
 * samples per
 * lane 1, 2 x chunk rows 2048..16384 x bucket streams off/on -- taking the
 * best of `reps` runs each, and keep the fastest as the context's options
 * (read them back with rr_get_option).  FP64 results are the same bits under
 * every schedule; FP32 results stay within their certified bounds.  No
 * counterpart in the reference. */
int rr_autotune(rr_ctx* ctx, int32_t handle, const rr_batch* batch, int32_t mode, int32_t precision, int32_t reps,
                double* best_ms);
int rr_get_option(rr_ctx* ctx, const char* name, int64_t* value);

/* Kernel time of the last rr_fitness_batch / rr_predict call on the context's
 * stream, in milliseconds (HIP events around the launches). */
int rr_last_kernel_ms(rr_ctx* ctx, double* ms);

/* The same split by phase (HIP events around each phase's launches on the
 * stream it runs on), for roofline accounting of the dominant kernel; no
 * counterpart in the reference.  Recorded only while option "phase_timing"
 * is 1 (default 0); a phase that did not run reads 0.  With the FP64 form on
 * a side stream (RR_FP32_CHECKED) phases overlap and their sum can exceed
 * rr_last_kernel_ms. */
#define RR_PHASE_PREPASS 0  /* k_nan0_fp64: factor 0 on the first 64 samples */
#define RR_PHASE_EVAL 1     /* k_alive_fp64 + k_eval_fp64: the FP64 factor columns */
#define RR_PHASE_QR 2       /* k_ts_* / k_qr_*: the column-pivoted QR and the solve */
#define RR_PHASE_SCORE 3    /* k_score_fp64 + k_fit_fp64: prediction, SSR, R2, RMSE */
#define RR_PHASE_FP32 4     /* the FP32 form's kernels (k_acc_fp32 + k_fin_fp32, or the raw pair) */
/* host wall-clock parts of the last rr_fitness_* call (recorded whatever
 * "phase_timing" is): */
#define RR_PHASE_HOST_CHECK 5  /* validating the batch (and hashing its programs) */
#define RR_PHASE_HOST_PLAN 6   /* forms, work lists, staging and uploads, up to the first kernel */
#define RR_PHASE_HOST_TAIL 7   /* from the first kernel to results in host memory, less the kernels'
                                  own span (launch gaps, result copies, synchronisation) */
#define RR_PHASES 8
int rr_last_phase_ms(rr_ctx* ctx, int32_t phase, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* RR_FITNESS_H */
