// Library GEMM candidates for the contraction autotuner (oagpu_gemm_lib.hip).
// Each returns false when the library cannot run the shape; nothing is then
// launched.
#pragma once
#include <hip/hip_runtime.h>

namespace oa {

// A rows have stride lda, B rows ldb (elements).
bool lib_gemm_bf16x3(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                     int K, int lda, int ldb, int ldc, hipStream_t st);
bool lib_gemm_f32(const void* Ah, const void* Bh, void* C, int M, int N, int K, int lda, int ldb, int ldc,
                  hipStream_t st);

}  // namespace oa
