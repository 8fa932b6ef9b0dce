// rr_kernels_fp32.h -- launch descriptor of the FP32 precision form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rr_fitness.h"
#include "rr_kernels.h"

namespace rr {

struct Fp32Launch {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const DsDesc* ds;         // device table of datasets
  const int32_t* cand_ds;   // per candidate (batch index), or null
  double* coefs;
  rr_fit* fit;
  const int32_t* dead;   // per candidate: outcome written by the factor-0 pre-pass (OLS), or null
  int n_cands;
  int raw;            // 1: score the tree as given, 0: OLS
  int rows;           // samples per lane per tile: 1 or 2
  // chunked pass: workgroups of k_acc (in bucket order), per-candidate first
  // chunk, chunk size, and the workspace of chunk partials and flags
  const int32_t* blk_cand;
  const int32_t* blk_chunk;
  const int32_t* chunk_base;
  int64_t chunk_rows;
  double* part;       // FP32_ACC_STRIDE doubles per chunk
  uint32_t* flags;    // 2 words per chunk
};

constexpr int FP32_ACC_STRIDE = 64;

// Number of factor buckets of the FP32 OLS kernels: candidates with at most
// 1, 2, 4 and 8 factors; a fifth group (more factors) is flagged
// RR_F_NEEDS64 without evaluation.
constexpr int FP32_BUCKETS = 5;
constexpr int FP32_BUCKET_FACTORS[FP32_BUCKETS - 1] = {1, 2, 4, 8};

// `order` lists the candidates grouped by bucket, bucket_count[b] of each,
// and bucket_blocks[b] chunk workgroups of each (blk_cand/blk_chunk in the
// same order); bucket b is launched on streams[b] (the caller orders the
// streams).  RAW mode uses order as is and bucket 0's counts, on streams[0].
hipError_t launch_fp32(const Fp32Launch& L, const int32_t* order, const int* bucket_count, const int* bucket_blocks,
                       const hipStream_t* streams);
hipError_t launch_to_fp32(const double* src, float* dst, int64_t count, hipStream_t s);
// FP32 prediction with per-sample bounds (ldx a multiple of 256)
hipError_t launch_predict_fp32(const uint32_t* code, int len, const double* K, const float* X, int64_t n, int64_t ldx,
                               float* out_v, float* out_e, hipStream_t s);

}  // namespace rr
