// rf_gemm.hip -- C -= A B on MFMA for the trailing updates of the LU.
//
// A (m x k), B (k x n) and C (m x n) are column-major FP32 blocks of the
// matrix being factored.  Per mode the products run on
//   kFP32  v_mfma_f32_32x32x2_f32   (FP32 inputs, the plain --magma path)
//   kBF16  v_mfma_f32_32x32x16_bf16 (inputs rounded to BF16 when staged)
//   kFP16  v_mfma_f32_32x32x16_f16  (inputs rounded to FP16 when staged)
// always with FP32 accumulation and an FP32 C.
//
// Tiling: a 256-lane workgroup owns a 128x128 tile of C; wave w computes the
// 64x64 quarter (w & 1, w >> 1) as 2x2 MFMA blocks of 32x32.  K advances in
// chunks of 32 staged through LDS, row-major in k so that each lane's MFMA
// operand (1 value of k for FP32, 8 consecutive values of k for BF16/FP16)
// is one contiguous LDS read.  Staging loads: lane (r, half) reads 16 values
// of k for row r of A (each wave instruction reads 64 consecutive rows of one
// column, 256 bytes) and 16 consecutive values of k of column r of B.
//
// MFMA 32x32 output layout (per lane l, accumulator i): column l % 32,
// row 8*(i/4) + 4*(l/32) + i%4.
#include "rf_lu.h"

namespace rf {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

constexpr int kBM = 128, kBN = 128, kBK = 32;

template <LuMode MODE>
struct Stage;

template <>
struct Stage<LuMode::kFP32> {
  using T = float;
  static __device__ __forceinline__ T cvt(float x) { return x; }
};
template <>
struct Stage<LuMode::kBF16> {
  using T = __bf16;
  static __device__ __forceinline__ T cvt(float x) { return static_cast<__bf16>(x); }
};
template <>
struct Stage<LuMode::kFP16> {
  using T = _Float16;
  static __device__ __forceinline__ T cvt(float x) { return static_cast<_Float16>(x); }
};

template <LuMode MODE>
__global__ __launch_bounds__(256) void gemm_sub_kernel(int m, int n, int k, const float* __restrict__ A,
                                                       int64_t lda, const float* __restrict__ B, int64_t ldb,
                                                       float* __restrict__ C, int64_t ldc) {
  using T = typename Stage<MODE>::T;
  constexpr int PAD = (sizeof(T) == 4) ? 1 : 8;
  __shared__ __attribute__((aligned(16))) T As[kBM][kBK + PAD];
  __shared__ __attribute__((aligned(16))) T Bs[kBN][kBK + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int64_t m0 = static_cast<int64_t>(blockIdx.x) * kBM;
  const int64_t n0 = static_cast<int64_t>(blockIdx.y) * kBN;

  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.0f;

  const int sr = tid & 127;       // staged row of A / column of B
  const int sh = (tid >> 7) * 16; // first k of the lane's half chunk
  const bool arow_ok = m0 + sr < m;
  const bool bcol_ok = n0 + sr < n;
  const float* Arow = A + m0 + sr;
  const float* Bcol = B + (n0 + sr) * ldb;

  for (int k0 = 0; k0 < k; k0 += kBK) {
    float av[16], bv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int kk = k0 + sh + j;
      av[j] = (arow_ok && kk < k) ? Arow[static_cast<int64_t>(kk) * lda] : 0.0f;
      bv[j] = (bcol_ok && kk < k) ? Bcol[kk] : 0.0f;
    }
    __syncthreads();  // previous chunk fully consumed
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      As[sr][sh + j] = Stage<MODE>::cvt(av[j]);
      Bs[sr][sh + j] = Stage<MODE>::cvt(bv[j]);
    }
    __syncthreads();
    if constexpr (MODE == LuMode::kFP32) {
#pragma unroll
      for (int kk = 0; kk < kBK; kk += 2) {
        float a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          a[t] = As[wm + 32 * t + (lane & 31)][kk + (lane >> 5)];
          b[t] = Bs[wn + 32 * t + (lane & 31)][kk + (lane >> 5)];
        }
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
          for (int tb = 0; tb < 2; ++tb)
            acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
      }
    } else {
      using V8 = typename std::conditional<MODE == LuMode::kBF16, v8bf, v8h>::type;
#pragma unroll
      for (int kk = 0; kk < kBK; kk += 16) {
        V8 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          a[t] = *reinterpret_cast<const V8*>(&As[wm + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)]);
          b[t] = *reinterpret_cast<const V8*>(&Bs[wn + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)]);
        }
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
          for (int tb = 0; tb < 2; ++tb) {
            if constexpr (MODE == LuMode::kBF16)
              acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
            else
              acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
          }
      }
    }
  }

  // C -= acc
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb) {
      const int64_t col = n0 + wn + 32 * tb + (lane & 31);
      if (col >= n) continue;
      float* cc = C + col * ldc;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t row = m0 + wm + 32 * ta + 8 * g + 4 * (lane >> 5);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (row + i < m) cc[row + i] -= acc[ta][tb][4 * g + i];
      }
    }
}

hipError_t gemm_sub(LuMode mode, int64_t m, int64_t n, int64_t k, const float* A, int64_t lda, const float* B,
                    int64_t ldb, float* C, int64_t ldc, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((m + kBM - 1) / kBM), static_cast<unsigned>((n + kBN - 1) / kBN));
  const int mi = static_cast<int>(m), ni = static_cast<int>(n), ki = static_cast<int>(k);
  switch (mode) {
    case LuMode::kFP32:
      hipLaunchKernelGGL(gemm_sub_kernel<LuMode::kFP32>, grid, dim3(256), 0, s, mi, ni, ki, A, lda, B, ldb, C, ldc);
      break;
    case LuMode::kBF16:
      hipLaunchKernelGGL(gemm_sub_kernel<LuMode::kBF16>, grid, dim3(256), 0, s, mi, ni, ki, A, lda, B, ldb, C, ldc);
      break;
    case LuMode::kFP16:
      hipLaunchKernelGGL(gemm_sub_kernel<LuMode::kFP16>, grid, dim3(256), 0, s, mi, ni, ki, A, lda, B, ldb, C, ldc);
      break;
  }
  return hipGetLastError();
}

}  // namespace rf
