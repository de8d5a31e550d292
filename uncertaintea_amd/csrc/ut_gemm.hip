// ut_gemm.hip -- BF16 x BF16 -> FP32 GEMM on the gfx950 matrix cores.
//
//   C[M x N] = A[M x K] * B[K x N],  A given row-major (lda), B given transposed as
//   Bt[N x K] row-major (ldb), C row-major fp32 (ldc).
//
// Workgroup tile 128 x 128, four wavefronts of 64 x 64 (2 x 2 MFMA 32x32 tiles,
// 64 fp32 accumulators per lane), K stepped by 32 through LDS with the next
// step's global loads in flight while the current step multiplies.  Optional
// split-K over grid.z writes one partial product per split (reduced by the caller's
// epilogue kernel).  M and N must be multiples of 128 and K of 32 (callers pad).
#include "ut_gemm.h"
#include "ut_mfma.h"

namespace ut {

namespace {
constexpr int BM = 128, BN = 128, BK = 32, LDS_STRIDE = BK + 8;  // +8 bf16 breaks bank aliasing

__global__ __launch_bounds__(256) void k_gemm_nt(const bf16_t* __restrict__ A, int lda,
                                                 const bf16_t* __restrict__ Bt, int ldb,
                                                 float* __restrict__ C, int ldc, long long split_stride,
                                                 int K, int k_per_split) {
  __shared__ __attribute__((aligned(16))) bf16_t As[BM * LDS_STRIDE];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[BN * LDS_STRIDE];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const long long m0 = (long long)blockIdx.y * BM;
  const long long n0 = (long long)blockIdx.x * BN;
  const int kb = blockIdx.z * k_per_split;
  const int ke = min(K, kb + k_per_split);

  // each thread moves two 16-byte chunks of A and two of B per K step
  int row[2], col[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int c = tid + 256 * j;
    row[j] = c >> 2;
    col[j] = (c & 3) * 8;
  }
  uint4 ra[2], rb[2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      ra[j] = *(const uint4*)(A + (m0 + row[j]) * lda + k0 + col[j]);
      rb[j] = *(const uint4*)(Bt + (n0 + row[j]) * ldb + k0 + col[j]);
    }
  };
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  if (kb < ke) gload(kb);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      *(uint4*)(As + row[j] * LDS_STRIDE + col[j]) = ra[j];
      *(uint4*)(Bs + row[j] * LDS_STRIDE + col[j]) = rb[j];
    }
    __syncthreads();
    if (k0 + BK < ke) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      bf16x8 a[2], b[2];
      const int kk = ks + 8 * (lane >> 5);
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *(const bf16x8*)(As + (64 * wm + 32 * i + (lane & 31)) * LDS_STRIDE + kk);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        b[j] = *(const bf16x8*)(Bs + (64 * wn + 32 * j + (lane & 31)) * LDS_STRIDE + kk);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
    }
  }
  float* Cz = C + (long long)blockIdx.z * split_stride;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long long m = m0 + 64 * wm + 32 * i + mfma32_row(lane, r);
        const long long n = n0 + 64 * wn + 32 * j + (lane & 31);
        Cz[m * ldc + n] = acc[i][j][r];
      }
}
}  // namespace

hipError_t launch_gemm_nt(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, float* C, int ldc,
                          int M, int N, int K, int splits, long long split_stride, hipStream_t stream) {
  if (M % BM || N % BN || K % BK || splits < 1) return hipErrorInvalidValue;
  const int ksteps = K / BK;
  const int per = (ksteps + splits - 1) / splits * BK;
  dim3 grid(N / BN, M / BM, splits);
  hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, stream, A, lda, Bt, ldb, C, ldc, split_stride, K, per);
  return hipGetLastError();
}

}  // namespace ut
