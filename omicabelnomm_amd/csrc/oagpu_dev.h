// Device-side view of one scan: every pointer a kernel reads or writes.
// Passed to kernels by value.
#pragma once
#include "oagpu_common.h"
#include "../../include/oagpu.h"

namespace oa {

// One pair whose RSS needs the explicit residual pass.
struct FallbackEntry {
  int y;        // local trait
  int s;        // SNP within the chunk
  int flags;    // bit0: b is the untouched right-hand side (failed Cholesky)
  int pad;
  double b[PMAX];  // coefficients in the order h: covariates 0..l-1, then the SNP terms
};

enum {
  CNT_HITS = 0,      // coefficients reserved in the hit arena
  CNT_EMITTED = 1,   // emitted (trait, SNP) records
  CNT_SIG = 2,       // kept coefficients with P < minPdisp
  CNT_FALLBACK = 3,  // pairs sent to the explicit residual pass (all chunks)
  CNT_ERR = 4,       // bit0: RSS<0 / var_x<0 (reference exit); bit1: fallback overflow
  CNT_FB_CHUNK = 5,  // fallback entries of the current chunk
  CNT_N = 8
};

struct DevScan {
  // shapes
  int n, n_pad, l, r, p, Q;   // Q = l*r + r*(r+1)/2 correction terms
  long long ldop;             // row stride of the contraction operands (elements)
  const unsigned char* keep;  // [n] individuals in the analysis (nullptr = all)
  int n_stat;                 // individuals in the analysis (n when keep is null)
  int t, ldt;                 // traits, stride of per-trait arrays
  int ldc;                    // row stride of the contraction output (>= t + l)
  int G;                      // missing-value patterns
  int start_h;
  int disp_cov, store_all, need_bl;
  double min_r2_store, t_star, min_p_disp;

  // covariates
  const double* XL_cm;   // [l][n_pad] column-major
  const double* XL_rm;   // [n][l] row-major
  const double* U0;      // packed chol(Stl) (l(l+1)/2)
  const double* Stl_inv_diag;  // diag(Stl^-1) (l)

  // interactions
  const double* E;       // [n_env][lde], lde = n_pad, zero rows beyond n
  long long lde;

  // traits
  const void* Y; int ydt; long long ldy;
  const int* ymiss;      // [t]
  const int* yzero;      // [t] observed values that are exactly 0
  const double* yy;      // [t] sum y0^2
  const double* sst;     // [t]
  const double* rss0;    // [t] y0'y0 - |zt|^2
  const int* ypat;       // [t] pattern id
  const double* rt;      // [l][ldt] XL'y0
  const double* zt;      // [l][ldt] U_A^-T XL'y0
  const double* ay;      // [l][ldt] U_A^-1 zt

  // patterns
  const int* pat_cnt;    // [G] missing rows
  const int* pat_rep;    // [G] representative trait; its rows are at rows + rep*n
  const int* rows;       // sorted missing rows of trait y at rows + y*n
  const double* UA;      // [l(l+1)/2][G] packed chol(Stl - C_ll)
  const double* UA_invd; // [l][G] 1/diag
  const int* pat_pd;     // [G] 1 if Stl - C_ll is positive definite
  // trait waves: the traits of a pattern with many traits fill waves of their
  // own; the others share mixed waves
  const int* wtrait;     // [n_waves][64] trait id, -1 for an idle lane
  const int* wpat;       // [n_waves] the wave's pattern, -1 if mixed
  int n_waves;

  // compact missing rows: U = union of the patterns' missing rows (sorted)
  int nU;                // |U| (0: corrections read the genotypes directly)
  const int* uidx;       // [n_pad] position of row k in U, -1 if absent
  const double* XLu;     // [nU][l] covariate rows of U
  const int* pu_off;     // [G+1] pattern g's rows are pu_list[pu_off[g]..pu_off[g+1])
  const int* pu_list;    // positions in U
  const double* XLp;     // [pu_off[G]][l] covariate rows of pu_list's rows, in its order

  // genotypes
  const void* Gm; int gdt; long long ldg;
  int g_vec;             // Gm 16-byte aligned and ldg % 4 == 0
  const int* col_src; const int* col_env;   // [m*r] for the whole loaded block

  // chunk
  int s0, mc;            // first SNP of the chunk, SNPs in the chunk
  // when set, the chunk's first SNP is read from here instead of s0: a
  // launch sequence captured once in a HIP graph is replayed for every chunk
  // of one size, the start written before each replay
  const int* s0_dev;
  double* col_mean;      // [mc*r]
  int* col_nan;          // [mc*r] 1 if the column has a missing value (written when set)
  double* Sbr;           // [mc][r(r+1)/2] packed upper
  double* Str;           // [mc][l][r]
  double* varx;          // [mc][p]
  double* Ccorr;         // [G][Q][mc]: pattern, correction term, SNP
  void* Xu;              // [nU][ldxu] imputed genotype values at the rows of U
  int xu_dt;             // element type of Xu: DT_F64, or the panel's DT_F32 / DT_F16
                         // when its values are exact there (no missing genotypes)
  double* pst;           // [mc][G][patsnp_width] per-(SNP, pattern) solve pieces
  long long ldxu;        // >= mc*r
  const void* Cbuf;      // [mc*r][ldc]

  // outputs
  oagpu_hit* hits; long long hit_cap;
  unsigned long long* counters;
  int trait_offset, snp_offset;
  int m_total;           // SNPs in the loaded block (dense stride)
  double *d_beta, *d_se, *d_t, *d_p; uint8_t* d_kept;
  double *d_r2, *d_rss; uint8_t* d_emitted; double* d_varx;

  FallbackEntry* fb; int fb_cap;
};

}  // namespace oa
