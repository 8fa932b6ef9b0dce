// rr_kernels_fp32.h -- launch descriptor of the FP32 precision form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rr_fitness.h"

namespace rr {

struct Fp32Launch {
  const int32_t* cand_start;
  const int32_t* prog_start;
  const uint32_t* code;
  const double* K;
  const float* X32;   // FP32 copy of X, same layout
  const float* y32;
  const double* X;    // FP64 originals (bound computation)
  const double* y;
  int64_t n, ldx;
  double ybar, sst, ysum;
  double* coefs;
  rr_fit* fit;
  int n_cands;
  int raw;            // 1: score the tree as given, 0: OLS
  int block_waves;    // waves (candidates) per workgroup
};

// `order` lists the candidates grouped by column count: bucket_count[0]
// with k <= 2, [1] with k <= 5, [2] with k <= 9, [3] larger (flagged
// RR_F_NEEDS64: no FP32 form).  RAW mode uses order as is.
hipError_t launch_fp32(const Fp32Launch& L, const int32_t* order, const int* bucket_count, hipStream_t s);
hipError_t launch_to_fp32(const double* src, float* dst, int64_t count, hipStream_t s);

}  // namespace rr
