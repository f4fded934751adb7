// Library GEMM candidates for the contraction C = A B' (row-major C [M][ldc],
// A [M][K], B [N][K]): hipBLASLt for BF16 and FP32 inputs.  (FP64 stays on
// the hand-written MFMA kernel, which runs at about 85% of the FP64 peak.)
// They join the hand-written kernels in the timing autotuner, which keeps
// whichever is fastest for the shape.  In column-major terms the product is
// C' (N x M, leading dimension ldc) = B (stored K x N, transposed) times
// A (stored K x M), the "TN" layout.  The split-BF16 product accumulates in
// FP32, small terms first: when the scan's interleaved layout is used (rows
// [Ahi | Alo] of A and [Blo | Bhi] of B, stride 2K) the two small terms are a
// single call of depth 2K, [Blo | Bhi]' [Ahi | Alo] = Blo'Ahi + Bhi'Alo, then
// Bhi'Ahi is added; otherwise three calls.
#include <hipblaslt/hipblaslt.h>

#include <map>
#include <mutex>
#include <tuple>

#include "oagpu_gemm_lib.h"

namespace oa {

namespace {

constexpr size_t kWorkspace = 64ull << 20;   // hipBLASLt workspace per device
constexpr int kMaxDev = 64;

struct LibDev {
  bool init = false;
  hipblasLtHandle_t lt = nullptr;
  void* ws = nullptr;
};

struct LtPlan {
  bool ok = false;
  hipblasLtMatmulDesc_t desc = nullptr;
  hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
  hipblasLtMatmulAlgo_t algo{};
};

using PlanKey = std::tuple<int, int, int, int, int, int, int, int>;   // device, dtype, m, n, k, lda, ldb, ldc

std::mutex g_mu;
LibDev g_dev[kMaxDev];
std::map<PlanKey, LtPlan> g_plans;

// hipBLASLt handle and workspace of the current device (nullptr if absent).
LibDev* lt_dev() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
  LibDev& d = g_dev[dev];
  if (!d.init) {
    d.init = true;
    if (hipblasLtCreate(&d.lt) != HIPBLAS_STATUS_SUCCESS) {
      d.lt = nullptr;
    } else if (hipMalloc(&d.ws, kWorkspace) != hipSuccess) {
      (void)hipblasLtDestroy(d.lt);
      d.lt = nullptr;
      d.ws = nullptr;
    }
  }
  return d.lt ? &d : nullptr;
}

// Plan for D (m x n, ld ldc, FP32) = op_T(A: k x m, ld lda) * B (k x n, ld ldb).
const LtPlan* lt_plan(LibDev* d, int dev, hipDataType in, int m, int n, int k, int lda, int ldb, int ldc) {
  const PlanKey key{dev, (int)in, m, n, k, lda, ldb, ldc};
  auto it = g_plans.find(key);
  if (it != g_plans.end()) return it->second.ok ? &it->second : nullptr;
  LtPlan& p = g_plans[key];
  hipblasOperation_t opA = HIPBLAS_OP_T, opB = HIPBLAS_OP_N;
  if (hipblasLtMatmulDescCreate(&p.desc, HIPBLAS_COMPUTE_32F, HIP_R_32F) != HIPBLAS_STATUS_SUCCESS) return nullptr;
  if (hipblasLtMatmulDescSetAttribute(p.desc, HIPBLASLT_MATMUL_DESC_TRANSA, &opA, sizeof(opA)) !=
          HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatmulDescSetAttribute(p.desc, HIPBLASLT_MATMUL_DESC_TRANSB, &opB, sizeof(opB)) !=
          HIPBLAS_STATUS_SUCCESS)
    return nullptr;
  if (hipblasLtMatrixLayoutCreate(&p.la, in, k, m, lda) != HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatrixLayoutCreate(&p.lb, in, k, n, ldb) != HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatrixLayoutCreate(&p.lc, HIP_R_32F, m, n, ldc) != HIPBLAS_STATUS_SUCCESS)
    return nullptr;
  hipblasLtMatmulPreference_t pref = nullptr;
  if (hipblasLtMatmulPreferenceCreate(&pref) != HIPBLAS_STATUS_SUCCESS) return nullptr;
  uint64_t ws = kWorkspace;
  (void)hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &ws, sizeof(ws));
  hipblasLtMatmulHeuristicResult_t res{};
  int got = 0;
  const hipblasStatus_t s =
      hipblasLtMatmulAlgoGetHeuristic(d->lt, p.desc, p.la, p.lb, p.lc, p.lc, pref, 1, &res, &got);
  (void)hipblasLtMatmulPreferenceDestroy(pref);
  if (s != HIPBLAS_STATUS_SUCCESS || got < 1 || res.workspaceSize > kWorkspace) return nullptr;
  p.algo = res.algo;
  p.ok = true;
  return &p;
}

bool lt_call(hipDataType in, const void* A, int lda, const void* B, int ldb, float beta, void* C, int m, int n,
             int k, int ldc, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  LibDev* d = lt_dev();
  if (!d) return false;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const LtPlan* p = lt_plan(d, dev, in, m, n, k, lda, ldb, ldc);
  if (!p) return false;
  const float alpha = 1.0f;
  return hipblasLtMatmul(d->lt, p->desc, &alpha, A, p->la, B, p->lb, &beta, C, p->lc, C, p->lc, &p->algo,
                         d->ws, kWorkspace, st) == HIPBLAS_STATUS_SUCCESS;
}

}  // namespace

bool lib_gemm_bf16x3(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                     int K, int lda, int ldb, int ldc, hipStream_t st) {
  const uint16_t* ah = static_cast<const uint16_t*>(Ah);
  const uint16_t* bl = static_cast<const uint16_t*>(Bl);
  const bool interleaved = static_cast<const uint16_t*>(Al) == ah + K && lda == 2 * K &&
                           static_cast<const uint16_t*>(Bh) == bl + K && ldb == 2 * K;
  if (interleaved)
    return lt_call(HIP_R_16BF, Bl, ldb, Ah, lda, 0.0f, C, N, M, 2 * K, ldc, st) &&
           lt_call(HIP_R_16BF, Bh, ldb, Ah, lda, 1.0f, C, N, M, K, ldc, st);
  return lt_call(HIP_R_16BF, Bl, ldb, Ah, lda, 0.0f, C, N, M, K, ldc, st) &&
         lt_call(HIP_R_16BF, Bh, ldb, Al, lda, 1.0f, C, N, M, K, ldc, st) &&
         lt_call(HIP_R_16BF, Bh, ldb, Ah, lda, 1.0f, C, N, M, K, ldc, st);
}

bool lib_gemm_f32(const void* Ah, const void* Bh, void* C, int M, int N, int K, int lda, int ldb, int ldc,
                  hipStream_t st) {
  return lt_call(HIP_R_32F, Bh, ldb, Ah, lda, 0.0f, C, N, M, K, ldc, st);
}

}  // namespace oa
