// Host-side launchers of the gfx950 scan kernels.
#pragma once
#include "oagpu_dev.h"

namespace oa {

// A per-chunk kernel's first step: take the chunk start from device memory
// when the launch belongs to a replayed graph (DevScan::s0_dev).
__device__ __forceinline__ void chunk_start(DevScan& d) {
  if (d.s0_dev) d.s0 = *d.s0_dev;
}

// trait side (oagpu_kernels_trait.hip)
void launch_ystats(const DevScan& d, int prec, void* Bhi, void* Blo, int* rows_out, int* ymiss,
                   int* yzero, double* yy, double* sst, double* rt, unsigned long long* hash,
                   hipStream_t st);
void launch_verify_patterns(int t, int n, const int* rows, const int* rep_of, const int* ymiss,
                            int* mismatch, hipStream_t st);
// union U of the patterns' missing rows (mark), then each pattern's rows as
// positions in U (uidx: row -> position)
void launch_union_mark(int n, int G, const int* rows, const int* pat_rep, const int* pat_cnt,
                       unsigned char* mark, hipStream_t st);
void launch_union_list(int n, int G, const int* rows, const int* pat_rep, const int* pat_cnt, const int* off,
                       const int* uidx, int* pu_list, hipStream_t st);
// XLp[i] = XLu[pu_list[i]] (l doubles per row) for the tot listed rows
void launch_pattern_rows(const int* pu_list, long long tot, int l, const double* XLu, double* XLp, hipStream_t st);
void launch_pattern(const DevScan& d, const double* Stl_pk, int G, const int* pat_rep, double* UA,
                    double* UA_invd, int* pat_pd, hipStream_t st);
// zt = U_A^-T XL'y0, ay = U_A^-1 zt (covariate-only coefficients), rss0
void launch_trait_solve(const DevScan& d, int lb, double* zt, double* ay, double* rss0, hipStream_t st);

// SNP side (oagpu_kernels_snp.hip)
// corr: also write the imputed values at the rows of U (d.Xu), read by
// launch_corr.
void launch_prep(const DevScan& d, int prec, void* Ahi, void* Alo, bool corr, hipStream_t st);
// the operand pass below the fused contraction ([E*G, G] pairs, no subset):
// means, missing flags, Sbr and compact rows, no operand
void launch_prep_pair(const DevScan& d, bool corr, hipStream_t st);
// Xu[u][s] = G[(s0 + s) * ldg + u] for u < n, s < mc (gdt DT_F32 or DT_F16)
void launch_panel_rows(const void* G, int gdt, long long ldg, int s0, int mc, int n, void* Xu, long long ldxu,
                       hipStream_t st, const int* s0_dev = nullptr);
// Dosage model: file SNP s has file_r columns of which the first n_used are
// read; sum != 0 writes one column per SNP, sum_i f[i] * G[:, s*file_r+i],
// otherwise n_used columns f[i] * G[:, s*file_r+i].  FP64 out (n rows,
// leading dimension ldo), NaN kept.
struct DosageModel {
  int file_r, n_used, sum;
  double f[16];
};
void launch_dosage(const void* G, int gdt, long long ldg, int n, int n_snp, const DosageModel& dm, double* out,
                   long long ldo, hipStream_t st);
void launch_snp_gram(const DevScan& d, int prec, int lb, hipStream_t st);
void launch_corr(const DevScan& d, int lb, hipStream_t st);
// one thread per pair, reading the corrections from d.Ccorr
void launch_epilogue(const DevScan& d, int prec, int lb, hipStream_t st);
// Few patterns: per-(SNP, pattern) solve pieces (d.pst, patsnp_width(r, l,
// lb) doubles each), then the per-pair epilogue that reads them.
void launch_patsnp(const DevScan& d, int lb, hipStream_t st);
void launch_epilogue_pst(const DevScan& d, int prec, int lb, hipStream_t st);
int patsnp_width(int r, int l, int lb);
void launch_fallback(const DevScan& d, int lb, hipStream_t st);
void launch_reset_counter(unsigned long long* c, int idx, hipStream_t st);
// zero bytes (a multiple of 4) at p with a kernel (no memset node when the
// caller captures the stream)
void launch_zero(void* p, size_t bytes, hipStream_t st);
// *p = v, by a kernel (no memset or copy node where the caller captures)
void launch_set_int(int* p, int v, hipStream_t st);
// counters [0, n), n <= 64, in one launch
void launch_reset_counters(unsigned long long* c, int n, hipStream_t st);
// r = 1: out[s] = sum over the subset's rows (d.keep) of x^2
void launch_subset_sbr(const DevScan& d, double* out, hipStream_t st);

// contraction (oagpu_gemm.hip): C[c][j] = sum_k A[c][k] B[j][k]
// A: M rows of K, row stride lda (padded to gemm_row_pad() rows), B: N rows
// of K, row stride ldb (padded to 128 rows), K a multiple of 64.  The
// split-BF16 operands are interleaved per row: A rows [hi | lo], B rows
// [lo | hi], strides 2K.  The variant is chosen by timing.
void launch_gemm(int prec, const void* Ahi, const void* Alo, const void* Bhi, const void* Blo,
                 void* C, int M, int N, int K, int lda, int ldb, int ldc, hipStream_t st);
// The default hand-written kernel of the precision (FP64 and FP32 direct,
// split-BF16 LDS-staged 128 x 128), launched without the autotuner.
void launch_gemm_fixed(int prec, const void* Ahi, const void* Alo, const void* Bhi, const void* Blo,
                       void* C, int M, int N, int K, int lda, int ldb, int ldc, hipStream_t st);
// Split-BF16 contraction that reads the genotype columns itself
// (oagpu_gemm_fused.hip), for blocks whose SNPs are [E*G, G] pairs from one
// genotype column and one interaction column E: C row 2s = (E*G_s)' B', row
// 2s+1 = G_s' B', a missing genotype replaced by col_mean of the row.  B rows
// interleaved [lo | hi] (stride ldb = 2K).  Only for blocks that
// gemm_fused_supported accepts; M is even (two rows per SNP).
struct FusedGemmArgs {
  const void* G; long long ldg; int n, K;
  const int* col_src; int s0;        // block descriptors; SNP s reads column col_src[2 (s0 + s)]
  const float* E;                    // the interaction column, K floats, zero past n
  const double* col_mean;            // [M] chunk-local
  const int* col_nan;                // [M] chunk-local: 1 if the column has a missing value
  const uint16_t* B; long long ldb;  // launch_fused_b_layout rows
  float* C; long long ldc;
  int M, N, tiles_n, total;
};
void launch_gemm_fused(const FusedGemmArgs& a, int gdt, hipStream_t st);
// The B operand the fused contraction reads (FusedGemmArgs::B): rows of B
// (interleaved [lo | hi], stride ldb = 2K) rewritten with each 32-value
// k-step's hi then lo values adjacent, so a k-step of a row is one 128-byte
// run.  Same size and stride.
void launch_fused_b_layout(const uint16_t* B, long long ldb, int K, int rows, uint16_t* Bf, hipStream_t st);
// genotypes of element type gdt at G (column stride ldg, n rows, operand
// depth K) the fused contraction takes: float32 or float16, 4-element runs
// aligned, n % 4 == 0
bool gemm_fused_supported(int gdt, const void* G, long long ldg, int n, int K);
constexpr const char* kFusedVariant = "bf16x3_fused_256x256";
const char* gemm_variant_name(int prec, int M, int N, int K);
// shape classes timed by the autotuner in this process
int gemm_tunings();
int gemm_row_pad();
// A rows read by launch_gemm_fixed: M rounded up to its 128-row tiles
int gemm_fixed_row_pad();

}  // namespace oa
