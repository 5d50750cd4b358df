// gemm.hpp -- batched dense products used by the eigensolver back-transform,
// the propagator and the spectral density (gfx950).
//
//   C[b](m,n) = alpha * sum_k A[b](m,k) B[b](k,n) + beta * C[b](m,n)
//
// Element (m,k) of A lives at A[m*a_m + k*a_k] (one of the two strides is 1),
// likewise B (k,n) at B[k*b_k + n*b_n] and C (m,n) at C[m*c_m + n*c_n].
//   CIR_PREC_FP32   FP32 MFMA (v_mfma_f32_32x32x2f32), FP32 accumulation
//   CIR_PREC_BF16X3 split-BF16 x 3 MFMA (v_mfma_f32_32x32x16_bf16):
//                   a_hi b_hi + a_hi b_lo + a_lo b_hi, FP32 accumulation
//   CIR_PREC_FP64   FP64 MFMA (v_mfma_f64_16x16x4f64), 64x64 tiles
#pragma once

#include "common.hpp"

namespace cir {

struct GemmArgs {
    int M = 0, N = 0, K = 0, batch = 1;
    const void* A = nullptr;
    int64_t a_m = 0, a_k = 0, a_b = 0;
    const void* B = nullptr;
    int64_t b_k = 0, b_n = 0, b_b = 0;
    void* C = nullptr;
    int64_t c_m = 0, c_n = 0, c_b = 0;
    double alpha = 1.0, beta = 0.0;
};

int gemm(int prec, const GemmArgs& g, hipStream_t s);

}  // namespace cir
