// Internal state of one oagpu_ctx (include/oagpu.h), shared by the scan
// driver (oagpu_api.hip) and the request server (oagpu_serve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "oagpu_kernels.h"

using oa::DevScan;
using oa::FallbackEntry;

struct oagpu_ctx {
  oagpu_params prm{};
  hipStream_t st = nullptr;
  std::string err;
  int n = 0, n_pad = 0, l = 0, r = 0, p = 0, Q = 0, lb = 0;
  size_t op_size = 8, c_size = 8;
  // Row stride (elements) of the contraction operands A and B.  Split-BF16
  // rows interleave the two halves: A rows [hi | lo], B rows [lo | hi].
  long long ldop = 0;
  // individual subset (request server): rows outside it take no part
  unsigned char* keep = nullptr;
  std::vector<unsigned char> keep_h;
  int n_stat = 0;
  // U is every row (the request server's shared transposed chunk)
  bool union_identity = false;

  // covariates
  bool have_cov = false;
  std::vector<double> Stl_inv_diag;
  double *XL_cm = nullptr, *XL_rm = nullptr, *U0 = nullptr, *Stl_inv_d = nullptr, *Stl_pk = nullptr;
  std::vector<double> XL_rm_h;   // [n][l]
  void *XLop_hi = nullptr, *XLop_lo = nullptr;   // covariate rows of B, in the operand format

  // interactions
  double* E = nullptr;
  long long lde = 0;
  int n_env = 0;
  // FP32 copy for the fused contraction: n_env rows, then a row of ones (the
  // factor of a column without interaction), each n_pad long (zero past n)
  float* E32 = nullptr;

  // traits
  bool have_traits = false;
  int t = 0, ldt = 0, trait_offset = 0;
  const void* Y = nullptr;
  int ydt = 0;
  long long ldy = 0;
  void* Y_own = nullptr;
  size_t Y_own_bytes = 0;
  void* Bbuf = nullptr;
  void *Bhi = nullptr, *Blo = nullptr;   // views into Bbuf
  // the fused contraction's copy of B (launch_fused_b_layout), built on
  // first use after the traits were loaded
  uint16_t* Bf = nullptr;
  size_t Bf_cap = 0;
  bool bf_ok = false;
  int* rows = nullptr;
  int *ymiss = nullptr, *yzero = nullptr, *ypat = nullptr;
  double *yy = nullptr, *sst = nullptr, *rss0 = nullptr, *rt = nullptr, *zt = nullptr, *ay = nullptr;
  unsigned long long* hash = nullptr;
  int* scratch_i = nullptr;   // rep_of / mismatch
  int G = 0, max_pat_cnt = 0;
  // Missing phenotypes: the operand pass copies the genotype values at U (the
  // union of the patterns' missing rows) into Xu and k_corr reads them there.
  bool compact_corr = false;
  int nU = 0;
  int *uidx = nullptr, *pu_off = nullptr, *pu_list = nullptr;
  double* XLu = nullptr;
  double* XLp = nullptr;   // XLu rows of pu_list, pattern by pattern
  size_t XLp_cap = 0;
  size_t uidx_cap = 0, pu_off_cap = 0, pu_list_cap = 0, XLu_cap = 0;
  int *pat_cnt = nullptr, *pat_rep = nullptr, *pat_pd = nullptr;
  double *UA = nullptr, *UA_invd = nullptr;
  size_t pat_cap = 0;
  // trait waves of the pattern-table pair kernel: 64 trait ids per wave (-1
  // for an idle lane) and the wave's pattern (-1: several patterns)
  int *wtrait = nullptr, *wpat = nullptr;
  int n_waves = 0;
  size_t wave_cap = 0;

  // genotypes: host blocks are uploaded on a second stream into one of two
  // device slots, so the next block's copy overlaps the current scan
  bool have_geno = false;
  const void* Gm = nullptr;
  int gdt = 0;
  long long ldg = 0;
  int n_src = 0, m = 0, snp_offset = 0;
  hipStream_t cst = nullptr;
  void* G_slot[2] = {nullptr, nullptr};
  size_t G_slot_cap[2] = {0, 0};
  int* d_src[2] = {nullptr, nullptr};
  int* d_env[2] = {nullptr, nullptr};
  size_t d_desc_cap[2] = {0, 0};
  std::vector<int> h_desc[2];
  hipEvent_t up_ev[2] = {nullptr, nullptr}, free_ev[2] = {nullptr, nullptr};
  bool free_valid[2] = {false, false};
  int gslot = 1;
  bool upload_pending = false;
  int *col_src = nullptr, *col_env = nullptr;
  // the block's SNPs are [E*G, G] pairs of one genotype column, all with
  // interaction column pair_env (-1: another layout)
  int pair_env = -1;
  // dosage model: the combined FP64 columns of each slot's block
  bool dos_on = false;
  oa::DosageModel dos{};
  void* G_dos[2] = {nullptr, nullptr};
  size_t G_dos_cap[2] = {0, 0};

  // a scan in flight (oagpu_scan_async .. oagpu_scan_wait)
  bool pending = false;
  bool p_timed = true;   // per-chunk timing events recorded
  bool p_dense = false;
  int p_nchunks = 0, p_t = 0, p_m = 0;
  std::vector<hipEvent_t> pev;
  hipEvent_t done_ev = nullptr;
  unsigned long long* cnt_h = nullptr;   // pinned copy of the counters
  double *pd_beta = nullptr, *pd_se = nullptr, *pd_t = nullptr, *pd_p = nullptr, *pd_r2 = nullptr,
         *pd_rss = nullptr, *pd_varx = nullptr;
  uint8_t *pd_kept = nullptr, *pd_em = nullptr;

  // chunk buffers
  size_t chunk_budget = 0;   // bytes of chunk buffers a scan may size for (set at the first scan)
  int mc_cap = 0;
  void* Abuf = nullptr;
  void *Ahi = nullptr, *Alo = nullptr, *Cbuf = nullptr;   // Ahi, Alo: views into Abuf
  double *col_mean = nullptr, *Sbr = nullptr, *Str = nullptr, *varx = nullptr, *Ccorr = nullptr;
  int* col_nan = nullptr;   // per chunk column: 1 if it holds a missing genotype
  void* Xu = nullptr;   // FP64, or FP32 below an FP64 contraction (xu_dtype)
  long long ldxu = 0;
  // few patterns per trait block: per-(SNP, pattern) solve pieces
  bool use_pst = false;
  double* pst = nullptr;
  size_t A_cap = 0, C_cap = 0, Ccorr_cap = 0, Xu_cap = 0, pst_cap = 0;

  // outputs
  oagpu_hit* hits = nullptr;
  long long hit_cap = 0;
  unsigned long long* counters = nullptr;
  FallbackEntry* fb = nullptr;
  int fb_cap = 0;
  double timings[8] = {0};
  std::string gemm_variant = "none";
  bool last_fused = false;   // the last scan's contraction read the genotypes itself
};

#define HIPCK(ctx, expr)                                                              \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      (ctx)->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr; \
      return OAGPU_ERR_HIP;                                                           \
    }                                                                                 \
  } while (0)

inline int fail(oagpu_ctx* c, int code, const std::string& msg) {
  c->err = msg;
  return code;
}

// (Re)allocate a device buffer so it holds at least `bytes`.
inline hipError_t grow(void** p, size_t* cap, size_t bytes) {
  if (*p && *cap >= bytes) return hipSuccess;
  if (*p) {
    hipError_t e = hipFree(*p);
    if (e != hipSuccess) return e;
    *p = nullptr;
  }
  *cap = 0;
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 64);
  if (e == hipSuccess) *cap = bytes;
  return e;
}

inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

namespace oa {
// Row stride of the contraction output (traits then covariates, padded).
int out_cols(const oagpu_ctx* c);
// Device view of the context's current trait block and chunk buffers.
DevScan make_dev(const oagpu_ctx* c);
// Chunk buffers of the pair side (Str, var_x, corrections, (SNP, pattern)
// table, Sbr when a subset is set) for chunks of up to mc SNPs.
int ensure_pair_buffers(oagpu_ctx* c, int mc);
// Element type of the compact missing-row copy Xu: FP64 below an FP64
// contraction; FP32 below the FP32 and split-BF16 contractions, whose own
// rounding (2^-17 and coarser) is far above the FP32 rounding of an imputed
// genotype (2^-24 relative), while it halves the corrections' gather traffic.
// Kernels of every chunk of the loaded block on c->st, mc SNPs per chunk:
// counters reset, then per chunk the operand pass, contraction, Gram
// pieces, corrections, pair solve and statistics, fallback.  d carries the
// chunk buffers, Ahi/Alo the operand rows; E: 7 timing events per chunk, or
// nullptr; fixed_gemm: the default hand-written contraction kernel, launched
// without the autotuner (no timing launches, no library handle).
int enqueue_chunks(oagpu_ctx* c, DevScan& d, long long mc, void* Ahi, void* Alo, hipEvent_t* E, bool fixed_gemm);
// Waits for the enqueued scan and returns its hits (sorted), statistics and
// dense outputs; the scan's status.
int scan_collect(oagpu_ctx* c, oagpu_dense* dense, oagpu_hit* hits, long long hit_cap, oagpu_stats* stats);
// Looks up this device's no-allocation scratch array once (oagpu_noalloc.hip);
// called by oagpu_create with the device current.
void noalloc_setup(int device);
inline int xu_dtype(const oagpu_ctx* c) { return c->prm.precision == OAGPU_FP64 ? DT_F64 : DT_F32; }
inline size_t xu_size(const oagpu_ctx* c) { return xu_dtype(c) == DT_F64 ? sizeof(double) : sizeof(float); }
}  // namespace oa
