// Host-side launchers of the gfx950 scan kernels.
#pragma once
#include "oagpu_dev.h"

namespace oa {

// trait side (oagpu_kernels_trait.hip)
void launch_ystats(const DevScan& d, int prec, void* Bhi, void* Blo, int* rows_out, int* ymiss,
                   int* yzero, double* yy, double* sst, double* rt, unsigned long long* hash,
                   hipStream_t st);
void launch_verify_patterns(int t, int n, const int* rows, const int* rep_of, const int* ymiss,
                            int* mismatch, hipStream_t st);
void launch_pattern(const DevScan& d, const double* Stl_pk, int G, const int* pat_rep, double* UA,
                    double* UA_invd, int* pat_pd, hipStream_t st);
void launch_trait_solve(const DevScan& d, int lb, double* zt, double* rss0, hipStream_t st);

// SNP side (oagpu_kernels_snp.hip)
// corr: also compute the missing-row corrections of every pattern (used
// when the patterns' missing rows are few).
void launch_prep(const DevScan& d, int prec, void* Ahi, void* Alo, bool corr, hipStream_t st);
void launch_snp_gram(const DevScan& d, int prec, int lb, hipStream_t st);
void launch_mask(const DevScan& d, int lb, hipStream_t st);
void launch_epilogue(const DevScan& d, int prec, int lb, hipStream_t st);
void launch_fallback(const DevScan& d, int lb, hipStream_t st);
void launch_reset_counter(unsigned long long* c, int idx, hipStream_t st);

// contraction (oagpu_gemm.hip): C[c][j] = sum_k A[c][k] B[j][k]
// A: M rows of K (padded to gemm_row_pad() rows), B: N rows of K (padded
// to 128 rows), K a multiple of 64.  The variant is chosen by timing.
void launch_gemm(int prec, const void* Ahi, const void* Alo, const void* Bhi, const void* Blo,
                 void* C, int M, int N, int K, int ldc, hipStream_t st);
const char* gemm_variant_name(int prec, int M, int N, int K);
int gemm_row_pad();

}  // namespace oa
