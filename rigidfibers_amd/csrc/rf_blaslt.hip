// rf_blaslt.hip -- the LU's trailing updates C -= A B through hipBLASLt for the
// mixed-precision modes: a plain GEMM (BF16 or FP16 operands packed
// K-contiguous, FP32 C column-major, FP32 accumulation).  In the library's
// column-major terms the packs are K x m and K x n matrices with leading
// dimensions ldak / ldbk, so C <- -1 * op(A) B + 1 * C with op(A) = A^T.
#include <hipblaslt/hipblaslt.h>

#include <new>

#include "rf_lu.h"

namespace rf {

struct BltGemm {
  hipblasLtHandle_t handle = nullptr;
  hipblasLtMatmulDesc_t desc = nullptr;
  hipDataType type = HIP_R_16BF;
  void* work = nullptr;
  uint64_t work_bytes = 64ull << 20;
};

hipError_t blt_create(LuMode mode, void** out, int64_t* bytes) {
  *out = nullptr;
  *bytes = 0;
  if (mode == LuMode::kFP32) return hipErrorNotSupported;
  auto* g = new (std::nothrow) BltGemm();
  if (!g) return hipErrorOutOfMemory;
  g->type = mode == LuMode::kBF16 ? HIP_R_16BF : HIP_R_16F;
  const hipblasOperation_t ta = HIPBLAS_OP_T, tb = HIPBLAS_OP_N;
  if (hipblasLtCreate(&g->handle) != HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatmulDescCreate(&g->desc, HIPBLAS_COMPUTE_32F, HIP_R_32F) != HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatmulDescSetAttribute(g->desc, HIPBLASLT_MATMUL_DESC_TRANSA, &ta, sizeof(ta)) !=
          HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatmulDescSetAttribute(g->desc, HIPBLASLT_MATMUL_DESC_TRANSB, &tb, sizeof(tb)) !=
          HIPBLAS_STATUS_SUCCESS) {
    blt_destroy(g);
    return hipErrorNotSupported;
  }
  if (hipMalloc(&g->work, g->work_bytes) != hipSuccess) {
    (void)hipGetLastError();
    blt_destroy(g);
    return hipErrorOutOfMemory;
  }
  *out = g;
  *bytes = static_cast<int64_t>(g->work_bytes);
  return hipSuccess;
}

void blt_destroy(void* p) {
  auto* g = static_cast<BltGemm*>(p);
  if (!g) return;
  if (g->work) (void)hipFree(g->work);
  if (g->desc) (void)hipblasLtMatmulDescDestroy(g->desc);
  if (g->handle) (void)hipblasLtDestroy(g->handle);
  delete g;
}

// hipErrorNotSupported when the library offers no algorithm for the shape
// (the caller then runs the built-in kernel).
hipError_t blt_gemm_sub(void* p, int64_t m, int64_t n, int64_t K, const void* At, int64_t ldak, const void* Bt,
                        int64_t ldbk, float* C, int64_t ldc, hipStream_t s) {
  auto* g = static_cast<BltGemm*>(p);
  if (!g) return hipErrorNotSupported;
  if (m <= 0 || n <= 0 || K <= 0) return hipSuccess;
  hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
  hipblasLtMatmulPreference_t pref = nullptr;
  hipError_t e = hipErrorNotSupported;
  hipblasLtMatmulHeuristicResult_t res[1];
  int nres = 0;
  if (hipblasLtMatrixLayoutCreate(&la, g->type, K, m, ldak) == HIPBLAS_STATUS_SUCCESS &&
      hipblasLtMatrixLayoutCreate(&lb, g->type, K, n, ldbk) == HIPBLAS_STATUS_SUCCESS &&
      hipblasLtMatrixLayoutCreate(&lc, HIP_R_32F, m, n, ldc) == HIPBLAS_STATUS_SUCCESS &&
      hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS &&
      hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &g->work_bytes,
                                            sizeof(g->work_bytes)) == HIPBLAS_STATUS_SUCCESS &&
      hipblasLtMatmulAlgoGetHeuristic(g->handle, g->desc, la, lb, lc, lc, pref, 1, res, &nres) ==
          HIPBLAS_STATUS_SUCCESS &&
      nres > 0) {
    const float alpha = -1.0f, beta = 1.0f;
    e = hipblasLtMatmul(g->handle, g->desc, &alpha, At, la, Bt, lb, &beta, C, lc, C, lc, &res[0].algo, g->work,
                        g->work_bytes, s) == HIPBLAS_STATUS_SUCCESS
            ? hipSuccess
            : hipErrorLaunchFailure;
  }
  if (pref) (void)hipblasLtMatmulPreferenceDestroy(pref);
  if (lc) (void)hipblasLtMatrixLayoutDestroy(lc);
  if (lb) (void)hipblasLtMatrixLayoutDestroy(lb);
  if (la) (void)hipblasLtMatrixLayoutDestroy(la);
  return e;
}

}  // namespace rf
