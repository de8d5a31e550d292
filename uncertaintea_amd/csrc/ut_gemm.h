// ut_gemm.h -- device GEMM launchers (internal to libuncertaintea_hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ut {
// C (+ z * split_stride for split z) = A[M x K] * Bt[N x K]^T over the split's K range.
// A, Bt: bfloat16 row-major; C fp32 row-major.  M, N multiples of 64, K of 32,
// lda and ldb multiples of 8.
hipError_t launch_gemm_nt(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, float* C, int ldc,
                          int M, int N, int K, int splits, long long split_stride, hipStream_t stream);
// Fused epilogues (no split-K): the product's element (m, n) v goes to
//   bias_tanh:   t = tanh(v + bias[n]); h[m ld + n] = t; hb[m ld + n] = bf16(t)
//   tanh_bwd:    g = v (1 - h[m ld + n]^2); gU[m ld + n] = bf16(g); gUf[m ld + n] = g (if set)
//   scatter_add: g[m P + pidx[n]] += v for n < np
struct GemmEpiBiasTanh {
  const float* bias;
  float* h;
  uint16_t* hb;
  int ld;
};
struct GemmEpiTanhBwd {
  const float* h;
  uint16_t* gU;
  float* gUf;
  int ld;
};
struct GemmEpiScatterAdd {
  float* g;
  int P;
  const int* pidx;
  int np;
};
hipError_t launch_gemm_nt_bias_tanh(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int M, int N, int K,
                                    const GemmEpiBiasTanh& e, hipStream_t stream);
hipError_t launch_gemm_nt_tanh_bwd(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int M, int N, int K,
                                   const GemmEpiTanhBwd& e, hipStream_t stream);
hipError_t launch_gemm_nt_scatter_add(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int M, int N, int K,
                                      const GemmEpiScatterAdd& e, hipStream_t stream);
// The same product with split-K into part (splits x M x N fp32), then summed over the
// splits in order into C (row stride ldc).
hipError_t launch_gemm_nt_splitk(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, float* C, int ldc, int M,
                                 int N, int K, int splits, float* part, hipStream_t stream);
}  // namespace ut
