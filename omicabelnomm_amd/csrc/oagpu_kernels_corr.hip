// Missing-row corrections C_lr, C_rr of every (pattern, SNP) from the compact
// missing-row copy written by the operand pass.
#include "oagpu_kernels.h"
#include "oagpu_pair.h"

namespace oa {

// ---------------------------------------------------------------------------
__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ double widen(_Float16 v) { return (double)(float)v; }

// Two consecutive SNPs per lane when a SNP has one column and there are at
// most 16 covariates: every covariate value read from LDS then feeds twice
// the multiply-adds without raising the register count past 5 waves per SIMD.
template <int R, int LB>
constexpr int corr_spl() { return (R == 1 && LB <= 16) ? 2 : 1; }

template <typename XT, int N> struct XVec { typedef __attribute__((ext_vector_type(N))) XT T; };
template <typename XT> struct XVec<XT, 1> { typedef XT T; };

// acc += v(lane J of this lane's 16-lane row) * x: one FP64 multiply-add
// whose first operand is broadcast within each row of 16 lanes by the DPP
// row_newbcast control (gfx90a and later), so a covariate row held one value
// per lane feeds every lane without an LDS broadcast per value.  The caller
// keeps every lane of the wave active, and v comes from an LDS read, not
// from a VALU write (the DPP read-after-VALU-write hazard).
#define OA_FMAC_BC(J)                                                                                  \
  case J:                                                                                              \
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #J " row_mask:0xf bank_mask:0xf"                     \
        : "+v"(acc) : "v"(v), "v"(x));                                                                 \
    break;
template <int J>
__device__ __forceinline__ void fmac_bcast(double& acc, double v, double x) {
  static_assert(J >= 0 && J < 16, "row of 16 lanes");
  switch (J) {
    OA_FMAC_BC(0) OA_FMAC_BC(1) OA_FMAC_BC(2) OA_FMAC_BC(3) OA_FMAC_BC(4) OA_FMAC_BC(5) OA_FMAC_BC(6)
    OA_FMAC_BC(7) OA_FMAC_BC(8) OA_FMAC_BC(9) OA_FMAC_BC(10) OA_FMAC_BC(11) OA_FMAC_BC(12) OA_FMAC_BC(13)
    OA_FMAC_BC(14) OA_FMAC_BC(15)
  }
}
#undef OA_FMAC_BC

template <int LB, int W, int J = 0>
__device__ __forceinline__ void fmac_row(double (&lr)[LB][W], double v, const double (&xv)[W]) {
  if constexpr (J < LB) {
#pragma unroll
    for (int c = 0; c < W; ++c) fmac_bcast<J>(lr[J][c], v, xv[c]);
    fmac_row<LB, W, J + 1>(lr, v, xv);
  }
}

// generate_correction_missings_Str (src/Algorithm.cpp:1260-1315) for every
// pattern: C_lr = sum over the pattern's missing rows of xl x', C_rr of x x'.
// The genotype values come from the compact u-major rows of U written by the
// operand pass.  One wave per pattern, corr_spl() consecutive SNPs per lane
// (one vector load per row): the pattern's row positions and covariate rows
// are wave-uniform, and the wave's SNPs at one row are contiguous.  Blocks are
// numbered so that the pattern groups of one SNP tile run back to back on one
// XCD and share its L2.  LB is the covariate count the kernel is compiled
// for: the exact l, or a bucket whose entries past l are zero.
template <int R, int LB, typename XT>
__global__ __launch_bounds__(256) void k_corr(DevScan d, int gq_count) {
  chunk_start(d);
  constexpr int SPL = corr_spl<R, LB>();
  constexpr int W = SPL * R;   // Xu values per lane and row
  typedef typename XVec<XT, W>::T XV;
  const int lane = threadIdx.x & 63;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int tile = (local / gq_count) * 8 + xcd;
  const int g = __builtin_amdgcn_readfirstlane((local % gq_count) * 4 + (threadIdx.x >> 6));
  if (g >= d.G || tile * 64 * SPL >= d.mc) return;
  const int s = (tile * 64 + lane) * SPL;
  // lanes past the end read the last group (the row holds ldxu >= mc * R
  // values, rounded up to 8, so the vector stays inside it) and store nothing
  const int se = s < d.mc ? s : (d.mc - 1) / SPL * SPL;
  const int l = d.l;
  constexpr int RR = R * (R + 1) / 2;
  double lr[LB][W], rr[SPL][RR];
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int c = 0; c < W; ++c) lr[j][c] = 0.0;
#pragma unroll
  for (int p = 0; p < SPL; ++p)
#pragma unroll
    for (int e = 0; e < RR; ++e) rr[p][e] = 0.0;
  // The pattern's row positions come in 64 at a time (one coalesced load,
  // then broadcast lane by lane), so each batch of CORR_BATCH rows issues its
  // genotype loads back to back instead of one dependent round trip per row.
  // The 64 rows' covariate rows are staged in LDS by the lanes (one row per
  // lane, zero past l and past the pattern's last row) and read back as
  // wave-uniform broadcasts; the multiply-adds run without branches.  The
  // rows come from XLp, the pattern's own covariate rows in list order (read
  // beside the row positions, not a gather that waits for them; the library
  // builds XLp with the union, whenever these corrections run).
  constexpr int CORR_BATCH = 8;
  constexpr int STAGE = LB > 16 ? 32 : 64;   // rows staged at once (LDS: 4 waves x STAGE x LB doubles)
  __shared__ double xls[4][STAGE][LB];
  double (*xw)[LB] = xls[threadIdx.x >> 6];
  const int beg = d.pu_off[g], end = d.pu_off[g + 1];
  for (int i0 = beg; i0 < end; i0 += STAGE) {
    const int nb = __builtin_amdgcn_readfirstlane(min(STAGE, end - i0));
    const bool in = lane < nb;
    // lanes past the pattern's end point at row nU of Xu, which is zero (the
    // position is loaded in bounds and replaced, so no load waits in a branch)
    const int pu = d.pu_list[i0 + (in ? lane : 0)];
    const int my_u = in ? pu : d.nU;
    if (lane < STAGE) {
      // every load of the lane's covariate row first, so they are in flight
      // together, then the stores; in bounds: a lane past the end reads the
      // stage's first row, an index past l the row's last value
      const double* xr = d.XLp + (size_t)(i0 + (in ? lane : 0)) * l;
      double v[LB];
#pragma unroll
      for (int j = 0; j < LB; ++j) v[j] = xr[min(j, l - 1)];
#pragma unroll
      for (int j = 0; j < LB; ++j) xw[lane][j] = (in && j < l) ? v[j] : 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    for (int b0 = 0; b0 < nb; b0 += CORR_BATCH) {
      double xv[CORR_BATCH][W];
#pragma unroll
      for (int q = 0; q < CORR_BATCH; ++q) {
        const int u = __builtin_amdgcn_readlane(my_u, (b0 + q) & 63);   // past nb: the zero row
        const XT* xr = static_cast<const XT*>(d.Xu) + (size_t)u * d.ldxu + (size_t)se * R;
        if constexpr (W == 2 || W == 4) {   // aligned: se * R and ldxu are multiples of W
          const XV v = *reinterpret_cast<const XV*>(xr);
#pragma unroll
          for (int c = 0; c < W; ++c) xv[q][c] = widen((XT)v[c]);
        } else {
#pragma unroll
          for (int c = 0; c < W; ++c) xv[q][c] = widen(xr[c]);
        }
      }
#pragma unroll
      for (int q = 0; q < CORR_BATCH; ++q) {
        const double* xl = xw[(b0 + q) & (STAGE - 1)];
        if constexpr (LB <= 16) {
          // lane j of each 16-lane row holds covariate j (clamped inside
          // the row); one DPP-broadcast multiply-add per covariate
          const double v = xl[min(lane & 15, LB - 1)];
          fmac_row<LB, W>(lr, v, xv[q]);
        } else {
#pragma unroll
          for (int j = 0; j < LB; ++j) {
            const double a = xl[j];
#pragma unroll
            for (int c = 0; c < W; ++c) lr[j][c] += a * xv[q][c];
          }
        }
#pragma unroll
        for (int p = 0; p < SPL; ++p)
#pragma unroll
          for (int b = 0; b < R; ++b)
#pragma unroll
            for (int a = 0; a <= b; ++a) rr[p][upk(a, b)] += xv[q][p * R + a] * xv[q][p * R + b];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int p = 0; p < SPL; ++p) {
    const int sp = s + p;
    if (sp < d.mc) {
#pragma unroll
      for (int j = 0; j < LB; ++j)
        if (j < l)
#pragma unroll
          for (int c = 0; c < R; ++c) d.Ccorr[((size_t)g * d.Q + j * R + c) * d.mc + sp] = lr[j][p * R + c];
#pragma unroll
      for (int e = 0; e < RR; ++e) d.Ccorr[((size_t)g * d.Q + l * R + e) * d.mc + sp] = rr[p][e];
    }
  }
}

// ---------------------------------------------------------------------------
// launcher
void launch_corr(const DevScan& d, int lb, hipStream_t st) {
  const int gq = (d.G + 3) / 4;
  dim3 block(256);
#define OA_CR(R_, LB_)                                                                                  \
  {                                                                                                     \
    const int w = 64 * corr_spl<R_, LB_>(), tiles = (d.mc + w - 1) / w;   /* SNPs per tile */          \
    dim3 grid((unsigned)(((tiles + 7) / 8) * 8 * gq));                                                  \
    if (d.xu_dt == DT_F64) hipLaunchKernelGGL((k_corr<R_, LB_, double>), grid, block, 0, st, d, gq);    \
    else if (d.xu_dt == DT_F32) hipLaunchKernelGGL((k_corr<R_, LB_, float>), grid, block, 0, st, d, gq); \
    else hipLaunchKernelGGL((k_corr<R_, LB_, _Float16>), grid, block, 0, st, d, gq);                   \
  }
  // compiled for the exact covariate count with one column per SNP and up to
  // 16 covariates, or two columns and up to 8, so no multiply-adds run on
  // zero padding
  if (d.r == 1 && d.l <= 16) {
    switch (d.l) {
      case 1: OA_CR(1, 1) break;   case 2: OA_CR(1, 2) break;   case 3: OA_CR(1, 3) break;
      case 4: OA_CR(1, 4) break;   case 5: OA_CR(1, 5) break;   case 6: OA_CR(1, 6) break;
      case 7: OA_CR(1, 7) break;   case 8: OA_CR(1, 8) break;   case 9: OA_CR(1, 9) break;
      case 10: OA_CR(1, 10) break; case 11: OA_CR(1, 11) break; case 12: OA_CR(1, 12) break;
      case 13: OA_CR(1, 13) break; case 14: OA_CR(1, 14) break; case 15: OA_CR(1, 15) break;
      default: OA_CR(1, 16) break;
    }
  } else if (d.r == 2 && d.l <= 8) {
    switch (d.l) {
      case 1: OA_CR(2, 1) break; case 2: OA_CR(2, 2) break; case 3: OA_CR(2, 3) break;
      case 4: OA_CR(2, 4) break; case 5: OA_CR(2, 5) break; case 6: OA_CR(2, 6) break;
      case 7: OA_CR(2, 7) break; default: OA_CR(2, 8) break;
    }
  } else {
    OA_DISPATCH_RLB(0, 0, OA_CR)
  }
#undef OA_CR
}

}  // namespace oa
