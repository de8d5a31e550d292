// ut_gemm.hip -- BF16 x BF16 -> FP32 GEMM on the gfx950 matrix cores.
//
//   C[M x N] = A[M x K] * B[K x N],  A given row-major (lda), B given transposed as
//   Bt[N x K] row-major (ldb), C row-major fp32 (ldc).
//
// Workgroup tile 64S x 64S (S = 2 for large problems, 1 when the 128-wide grid would
// leave CUs idle), four wavefronts of 32S x 32S (S x S MFMA 32x32 tiles), K stepped
// by 32 through LDS with the next step's global loads in flight while the current
// step multiplies.  Optional split-K over grid.z writes one partial product per
// split; launch_gemm_nt_splitk sums the partials in split order.  M and N must be
// multiples of 64 (128 for S = 2) and K of 32 (callers pad).
#include "ut_gemm.h"
#include "ut_mfma.h"

namespace ut {

namespace {
constexpr int BK = 32, LDS_STRIDE = BK + 8;  // +8 bf16 breaks bank aliasing

// epilogues: what happens to output element (m, n) of split z
struct EpiStore {
  float* C;
  int ldc;
  long long split_stride;
  __device__ void operator()(long long m, long long n, int z, float v) const {
    C[(long long)z * split_stride + m * ldc + n] = v;
  }
};
struct EpiBiasTanh {
  GemmEpiBiasTanh e;
  __device__ void operator()(long long m, long long n, int, float v) const {
    const float t = flow_tanh(v + e.bias[n]);
    e.h[m * e.ld + n] = t;
    e.hb[m * e.ld + n] = f2bf(t);
  }
};
struct EpiTanhBwd {
  GemmEpiTanhBwd e;
  __device__ void operator()(long long m, long long n, int, float v) const {
    const float h = e.h[m * e.ld + n];
    const float g = v * (1.0f - h * h);
    e.gU[m * e.ld + n] = f2bf(g);
    if (e.gUf) e.gUf[m * e.ld + n] = g;
  }
};
struct EpiScatterAdd {
  GemmEpiScatterAdd e;
  __device__ void operator()(long long m, long long n, int, float v) const {
    if (n < e.np) e.g[m * e.P + e.pidx[n]] += v;
  }
};

template <int S, class Epi>
__global__ __launch_bounds__(256) void k_gemm_nt(const bf16_t* __restrict__ A, int lda,
                                                 const bf16_t* __restrict__ Bt, int ldb, Epi epi,
                                                 int K, int k_per_split) {
  constexpr int BM = 64 * S, BN = 64 * S;
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

  // each thread moves S 16-byte chunks of A and S of B per K step
  int row[S], col[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int c = tid + 256 * j;
    row[j] = c >> 2;
    col[j] = (c & 3) * 8;
  }
  uint4 ra[S], rb[S];
  auto gload = [&](int k0) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
      ra[j] = *(const uint4*)(A + (m0 + row[j]) * lda + k0 + col[j]);
      rb[j] = *(const uint4*)(Bt + (n0 + row[j]) * ldb + k0 + col[j]);
    }
  };
  f32x16 acc[S][S];
#pragma unroll
  for (int i = 0; i < S; ++i)
#pragma unroll
    for (int j = 0; j < S; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  if (kb < ke) gload(kb);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < S; ++j) {
      *(uint4*)(As + row[j] * LDS_STRIDE + col[j]) = ra[j];
      *(uint4*)(Bs + row[j] * LDS_STRIDE + col[j]) = rb[j];
    }
    __syncthreads();
    if (k0 + BK < ke) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      bf16x8 a[S], b[S];
      const int kk = ks + 8 * (lane >> 5);
#pragma unroll
      for (int i = 0; i < S; ++i)
        a[i] = *(const bf16x8*)(As + (32 * S * wm + 32 * i + (lane & 31)) * LDS_STRIDE + kk);
#pragma unroll
      for (int j = 0; j < S; ++j)
        b[j] = *(const bf16x8*)(Bs + (32 * S * wn + 32 * j + (lane & 31)) * LDS_STRIDE + kk);
#pragma unroll
      for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S; ++j) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < S; ++i)
#pragma unroll
    for (int j = 0; j < S; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long long m = m0 + 32 * S * wm + 32 * i + mfma32_row(lane, r);
        const long long n = n0 + 32 * S * wn + 32 * j + (lane & 31);
        epi(m, n, (int)blockIdx.z, acc[i][j][r]);
      }
}

// C[m * ldc + n] = sum over splits z (in order) of part[z][m][n]
__global__ void k_splitk_sum(const float* __restrict__ part, int splits, int M, int N, float* __restrict__ C,
                             int ldc) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long mn = (long long)M * N;
  if (idx >= mn) return;
  float s = part[idx];
  for (int z = 1; z < splits; ++z) s += part[z * mn + idx];
  C[(idx / N) * ldc + idx % N] = s;
}
}  // namespace

template <class Epi>
static hipError_t gemm_nt(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, const Epi& epi, int M, int N, int K,
                          int splits, hipStream_t stream) {
  if (M % 64 || N % 64 || K % BK || splits < 1) return hipErrorInvalidValue;
  const int ksteps = K / BK;
  const int per = (ksteps + splits - 1) / splits * BK;
  const bool wide = M % 128 == 0 && N % 128 == 0 && (long long)(M / 128) * (N / 128) * splits >= 256;
  if (wide) {
    hipLaunchKernelGGL((k_gemm_nt<2, Epi>), dim3(N / 128, M / 128, splits), dim3(256), 0, stream, A, lda, Bt, ldb,
                       epi, K, per);
  } else {
    hipLaunchKernelGGL((k_gemm_nt<1, Epi>), dim3(N / 64, M / 64, splits), dim3(256), 0, stream, A, lda, Bt, ldb,
                       epi, K, per);
  }
  return hipGetLastError();
}

hipError_t launch_gemm_nt(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, float* C, int ldc,
                          int M, int N, int K, int splits, long long split_stride, hipStream_t stream) {
  return gemm_nt(A, lda, Bt, ldb, EpiStore{C, ldc, split_stride}, M, N, K, splits, stream);
}

hipError_t launch_gemm_nt_bias_tanh(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, int M, int N, int K,
                                    const GemmEpiBiasTanh& e, hipStream_t stream) {
  return gemm_nt(A, lda, Bt, ldb, EpiBiasTanh{e}, M, N, K, 1, stream);
}

hipError_t launch_gemm_nt_tanh_bwd(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, int M, int N, int K,
                                   const GemmEpiTanhBwd& e, hipStream_t stream) {
  return gemm_nt(A, lda, Bt, ldb, EpiTanhBwd{e}, M, N, K, 1, stream);
}

hipError_t launch_gemm_nt_scatter_add(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, int M, int N, int K,
                                      const GemmEpiScatterAdd& e, hipStream_t stream) {
  return gemm_nt(A, lda, Bt, ldb, EpiScatterAdd{e}, M, N, K, 1, stream);
}

hipError_t launch_gemm_nt_splitk(const bf16_t* A, int lda, const bf16_t* Bt, int ldb, float* C, int ldc, int M,
                                 int N, int K, int splits, float* part, hipStream_t stream) {
  const long long mn = (long long)M * N;
  hipError_t e = launch_gemm_nt(A, lda, Bt, ldb, part, N, M, N, K, splits, mn, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, stream, part, splits, M, N, C,
                     ldc);
  return hipGetLastError();
}

}  // namespace ut
