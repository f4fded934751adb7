// The pair solve with a shared (SNP, pattern) table: when patterns are few
// (4 G <= t), k_patsnp computes everything that depends only on (SNP,
// pattern) once and k_epilogue_pst finishes each (trait, SNP) pair from it.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "oagpu_kernels.h"
#include "oagpu_pair.h"

namespace oa {

// ---------------------------------------------------------------------------
// Few patterns (G << t): everything in the pair solve that depends only on
// (SNP, pattern) is computed once per (SNP, pattern) instead of once per
// pair.  With X = U_A^-T B_q, the Schur factor U_S, M = U_A^-1 X and
// a_y = U_A^-1 zt (per trait), the covariate coefficients are
// b_l = a_y - M b_r.  The reference's misplaced-correction RSS terms,
// 2 b_l' K b_r - b_r' H b_r, expand to 2 (KT' zt)' b_r - b_r' H2 b_r with
// KT = U_A^-T K and H2 = H + 2 sym(M' K), so a pair needs b_l only when it
// reports or re-evaluates it.  Rows j >= l of X, KT and M are zero, so a
// pair reads its table row without branches.  What every pair reads (X, KT,
// U_S, H2, the factor flag) comes first, M last.  Rows hold an even number
// of doubles (16-byte aligned).
template <int R, int LB>
struct PatSnp {
  static constexpr int RR = R * (R + 1) / 2;
  static constexpr int KR = R > 1 ? LB * R : 0;   // no misplaced terms with one column
  static constexpr int X = 0, KT = LB * R, US = LB * R + KR, H = US + RR, OK = H + RR, M = OK + 1,
                       W = (M + LB * R + 1) & ~1;
};

template <int R, int LB>
__global__ __launch_bounds__(256) void k_patsnp(DevScan d) {
  chunk_start(d);
  using T = PatSnp<R, LB>;
  constexpr int RR = T::RR;
  const int G = d.G, l = d.l;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)d.mc * G) return;
  // SNP fastest: the corrections of one pattern are contiguous in SNPs
  const int g = (int)(idx / d.mc), s = (int)(idx % d.mc);
  const int cnt = d.pat_cnt[g];
  double X[LB][R], Clr[LB][R], Crr[RR];
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      X[j][c] = (j < l) ? d.Str[((size_t)s * l + j) * R + c] : 0.0;
      Clr[j][c] = 0.0;
    }
#pragma unroll
  for (int e = 0; e < RR; ++e) Crr[e] = 0.0;
  if (cnt > 0) {
    // every column's correction is subtracted from the first SNP column
    // (src/Algorithm.cpp:1390-1396)
#pragma unroll
    for (int j = 0; j < LB; ++j)
      if (j < l) {
        double tot = 0.0;
#pragma unroll
        for (int c = 0; c < R; ++c) {
          const double v = d.Ccorr[((size_t)g * d.Q + j * R + c) * d.mc + s];
          tot += v;
          Clr[j][c] = v;
        }
        X[j][0] -= tot;
      }
#pragma unroll
    for (int e = 0; e < RR; ++e) Crr[e] = d.Ccorr[((size_t)g * d.Q + l * R + e) * d.mc + s];
  }
  // D_q: Sbr corrected only in rows a < min(l, r) (src/Algorithm.cpp:1300)
  double Sc[RR];
#pragma unroll
  for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
    for (int a = 0; a <= b2; ++a)
      Sc[upk(a, b2)] = d.Sbr[(size_t)s * RR + upk(a, b2)] - ((a < l) ? Crr[upk(a, b2)] : 0.0);
  // X = U_A^-T B_q
#pragma unroll
  for (int j = 0; j < LB; ++j)
    if (j < l) {
      const double inv = d.UA_invd[(size_t)j * G + g];
#pragma unroll
      for (int c = 0; c < R; ++c) {
        double v = X[j][c];
#pragma unroll
        for (int i = 0; i < j; ++i) v -= d.UA[(size_t)upk(i, j) * G + g] * X[i][c];
        X[j][c] = v * inv;
      }
    }
  // Schur complement and its Cholesky factor
#pragma unroll
  for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
    for (int a = 0; a <= b2; ++a) {
      double v = Sc[upk(a, b2)];
#pragma unroll
      for (int j = 0; j < LB; ++j) v -= X[j][a] * X[j][b2];
      Sc[upk(a, b2)] = v;
    }
  bool ok = d.pat_pd[g] != 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    double ajj = Sc[upk(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) ajj -= Sc[upk(k, j)] * Sc[upk(k, j)];
    if (!(ajj > 0.0)) ok = false;
    ajj = sqrt(ajj);
    Sc[upk(j, j)] = ajj;
#pragma unroll
    for (int i = j + 1; i < R; ++i) {
      double aji = Sc[upk(j, i)];
#pragma unroll
      for (int k = 0; k < j; ++k) aji -= Sc[upk(k, j)] * Sc[upk(k, i)];
      Sc[upk(j, i)] = aji / ajj;
    }
  }
  double* out = d.pst + ((size_t)s * G + g) * T::W;
  // M = U_A^-1 X
  double Mv[LB][R];
#pragma unroll
  for (int j = LB - 1; j >= 0; --j) {
#pragma unroll
    for (int c = 0; c < R; ++c) Mv[j][c] = 0.0;
    if (j < l) {
      const double inv = d.UA_invd[(size_t)j * G + g];
#pragma unroll
      for (int c = 0; c < R; ++c) {
        double v = X[j][c];
#pragma unroll
        for (int i = j + 1; i < LB; ++i)
          if (i < l) v -= d.UA[(size_t)upk(j, i) * G + g] * Mv[i][c];
        Mv[j][c] = v * inv;
      }
    }
  }
  // K (the first SNP column carries every column's correction), P = M' K,
  // then KT = U_A^-T K in place
  double Kt[LB][R], Pm[R][R];
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      double k = 0.0;
      if (R > 1 && j < l) {
        if (c == 0) {
#pragma unroll
          for (int c2 = 1; c2 < R; ++c2) k += Clr[j][c2];
        } else {
          k = -Clr[j][c];
        }
      }
      Kt[j][c] = k;
    }
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < LB; ++j) v += Mv[j][a] * Kt[j][c];
      Pm[a][c] = v;
    }
#pragma unroll
  for (int j = 0; j < LB; ++j)
    if (j < l) {
      const double inv = d.UA_invd[(size_t)j * G + g];
#pragma unroll
      for (int c = 0; c < R; ++c) {
        double v = Kt[j][c];
#pragma unroll
        for (int i = 0; i < j; ++i) v -= d.UA[(size_t)upk(i, j) * G + g] * Kt[i][c];
        Kt[j][c] = v * inv;
      }
    }
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      out[T::X + j * R + c] = X[j][c];   // zero for j >= l
      out[T::M + j * R + c] = Mv[j][c];
      if (R > 1) out[T::KT + j * R + c] = Kt[j][c];
    }
#pragma unroll
  for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
    for (int a = 0; a <= b2; ++a) {
      // diagonal stored as its reciprocal: the pair solves multiply
      out[T::US + upk(a, b2)] = (a == b2) ? 1.0 / Sc[upk(a, b2)] : Sc[upk(a, b2)];
      const double h = (a >= l) ? (a == b2 ? 1.0 : 2.0) * Crr[upk(a, b2)] : 0.0;
      out[T::H + upk(a, b2)] = h + 2.0 * (a == b2 ? Pm[a][a] : Pm[a][b2] + Pm[b2][a]);
    }
  out[T::OK] = ok ? 1.0 : 0.0;
  if (T::W > T::M + LB * R) out[T::W - 1] = 0.0;
}

// Pointers to what a wave reads at one SNP: with UNI (every lane of the wave
// has the same pattern) the (SNP, pattern) table row and var_x are
// wave-uniform and are read through the constant address space, i.e. with
// scalar loads; otherwise each lane reads its own pattern's row.
template <typename P, bool UNI>
using pst_ptr = typename std::conditional<UNI, const __attribute__((address_space(4))) P*, const P*>::type;

// The SNP terms of one pair from its table row: b_r and the RSS (the
// downdate of RSS0 plus the misplaced-correction terms).  False when the
// pattern's or the pair's factor failed; b_r and the RSS are then
// meaningless (computed anyway, so that the loads of a row are issued
// together and the caller needs no branch).
template <int R, int LB, typename RowP>
__device__ __forceinline__ bool pst_solve(RowP tb, const double* rb, const double* zt, double rss0, double* br,
                                          double& rss) {
  using T = PatSnp<R, LB>;
  const bool ok = tb[T::OK] != 0.0;
  double z[R], kz[R];
#pragma unroll
  for (int c = 0; c < R; ++c) {
    double v = rb[c], k = 0.0;
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      v -= tb[T::X + j * R + c] * zt[j];
      if (R > 1) k += tb[T::KT + j * R + c] * zt[j];
    }
    z[c] = v;
    kz[c] = k;
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {   // U_S' z = num (diagonal reciprocals in the table)
    double v = z[i];
#pragma unroll
    for (int k = 0; k < i; ++k) v -= tb[T::US + upk(k, i)] * z[k];
    z[i] = v * tb[T::US + upk(i, i)];
  }
  double zz = 0.0;
#pragma unroll
  for (int c = 0; c < R; ++c) zz += z[c] * z[c];
#pragma unroll
  for (int i = R - 1; i >= 0; --i) {   // U_S b_r = z
    double v = z[i];
#pragma unroll
    for (int k = i + 1; k < R; ++k) v -= tb[T::US + upk(i, k)] * br[k];
    br[i] = v * tb[T::US + upk(i, i)];
  }
  rss = rss0 - zz;
  if (R > 1) {   // misplaced-correction terms (zero for a pattern without missing rows)
    double q = 0.0;
#pragma unroll
    for (int c = 0; c < R; ++c) q += 2.0 * kz[c] * br[c];
#pragma unroll
    for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
      for (int a = 0; a <= b2; ++a) q -= tb[T::H + upk(a, b2)] * br[a] * br[b2];
    rss += q;
  }
  return ok;
}

// Per-trait constants of a pair thread, loaded once for its run of SNPs.
template <int LB>
struct PstTrait {
  TraitK tk;
  double rss0, guard;   // RSS0, RSS_GUARD * y0'y0
  bool yzero;
  double zt[LB];
};

template <int LB>
__device__ __forceinline__ PstTrait<LB> pst_trait(const DevScan& d, int y) {
  PstTrait<LB> tr;
  tr.tk = trait_k(d, y);
  tr.rss0 = d.rss0[y];
  tr.guard = RSS_GUARD * d.yy[y];
  tr.yzero = d.yzero[y] > 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) tr.zt[j] = (j < d.l) ? d.zt[(size_t)j * d.ldt + y] : 0.0;
  return tr;
}

// One pair, everything included: covariate coefficients when they are
// reported or the pair is re-evaluated, the hand-over to the explicit
// residual pass, the statistics and the records.
template <int R, int LB, typename CT, bool UNI>
__device__ __forceinline__ void pst_pair_full(const DevScan& d, const PstTrait<LB>& tr, int y, int g, int s) {
  using T = PatSnp<R, LB>;
  using RowP = pst_ptr<double, UNI>;
  constexpr int P = LB + R;
  const int l = d.l;
  const CT* C = reinterpret_cast<const CT*>(d.Cbuf) + y;
  double rb[R];
#pragma unroll
  for (int c = 0; c < R; ++c) rb[c] = (double)C[((size_t)s * R + c) * d.ldc];
  const RowP tb = (RowP)(d.pst + ((size_t)s * d.G + g) * T::W);
  double b[P], br[R], rss_s, rss = 0.0;
#pragma unroll
  for (int h = 0; h < P; ++h) b[h] = 0.0;
  const bool ok = pst_solve<R, LB>(tb, rb, tr.zt, tr.rss0, br, rss_s);
  if (ok) {
    rss = rss_s;
#pragma unroll
    for (int c = 0; c < R; ++c) b[LB + c] = br[c];
    if (d.need_bl || tr.yzero || !(rss >= tr.guard)) {   // reported, or the pair is re-evaluated
#pragma unroll
      for (int j = 0; j < LB; ++j)
        if (j < l) {
          double v = d.ay[(size_t)j * d.ldt + y];
#pragma unroll
          for (int c = 0; c < R; ++c) v -= tb[T::M + j * R + c] * br[c];
          b[j] = v;
        }
    }
  }
  finish_pair<R, LB>(d, tr.tk, y, s, ok, b, rb, rss, tr.yzero, tr.guard, (RowP)(d.varx + (size_t)s * d.p));
}

using KernDev = const __attribute__((address_space(4))) DevScan;

// The full pair as a call of its own, so that its registers (the explicit
// residual, the statistics, the records) do not weigh on the first pass.
// dp is the kernel's DevScan argument in the kernel-argument segment, s0
// the chunk start the kernel uses (DevScan::s0_dev); s is wave-uniform.  The trait's constants are loaded again here (rare) rather
// than handed over, which would put them in scratch memory, and the table
// row is read per lane (a wave-uniform row in scalar registers does not fit
// beside the full pair's state for the larger R, LB).
template <int R, int LB, typename CT>
__device__ __noinline__ void pst_pair_cold(KernDev* dp, int s0, int y, int g, int s) {
  DevScan d = *(const DevScan*)dp;
  d.s0 = __builtin_amdgcn_readfirstlane(s0);   // the kernel's chunk start
  s = __builtin_amdgcn_readfirstlane(s);
  pst_pair_full<R, LB, CT, false>(d, pst_trait<LB>(d, y), y, g, s);
}

// The pairs of one trait over a run of SNPs [s_beg, s_end), in one thread.
// When only the SNP terms are tested (no covariate display, no dense output,
// no store-all) a first pass solves every pair for b_r and the RSS only,
// loads the next SNP's contraction outputs while the current one is solved,
// and marks the SNPs at which some lane of the wave needs more: a failed
// factor, the explicit residual, a var_x error, or a SNP term whose |t| may
// pass the store threshold (the test b^2 > t*^2 se^2 of the statistics,
// with the same arithmetic).  Only those SNPs are then finished by the full
// pair, on every lane of the wave: for a lane that needed nothing there the
// full pair writes and counts nothing.
template <int R, int LB, typename CT, bool UNI>
__device__ __forceinline__ void pst_trait_run(const DevScan& d, KernDev* dp, int y, int g, int s_beg, int s_end) {
  using T = PatSnp<R, LB>;
  using RowP = pst_ptr<double, UNI>;
  const int l = d.l;
  const PstTrait<LB> tr = pst_trait<LB>(d, y);
  if (d.need_bl || d.store_all || d.d_beta) {
    for (int s = s_beg; s < s_end; ++s) pst_pair_cold<R, LB, CT>(dp, d.s0, y, g, s);
    return;
  }
  const CT* C = reinterpret_cast<const CT*>(d.Cbuf) + y;
  const size_t ldc = (size_t)d.ldc;
  const double tt = d.t_star * d.t_star, min_r2 = d.min_r2_store;
  unsigned long long pend = 0;   // wave-uniform: bit s - s_beg
  // b_r and the RSS of the pair at SNP s, and whether any lane of the wave
  // needs the full pair there (branch-free: every load of the row at once)
  auto quick = [&](int s, const double* rb) {
    const RowP tb = (RowP)(d.pst + ((size_t)s * d.G + g) * T::W);
    const RowP vx = (RowP)(d.varx + (size_t)s * d.p);
    double br[R], rss;
    const bool ok = pst_solve<R, LB>(tb, rb, tr.zt, tr.rss0, br, rss);
    const double R2 = 1.0 - rss * tr.tk.r2s;
    const double s2 = (rss * tr.tk.inv_nc) * tr.tk.vf;
    const bool tested = (rss > 0.0) & (R2 > min_r2);
    bool more = !ok | tr.yzero | !(rss >= tr.guard);
#pragma unroll
    for (int c = 0; c < R; ++c) {
      const double vxh = vx[l + c];
      const double se2 = s2 * vxh;
      more |= tested & ((vxh < 0.0) | ((se2 >= 0.0) & (br[c] * br[c] > tt * se2)));
    }
    if (__ballot(more)) pend |= 1ull << (s - s_beg);
  };
  // contraction outputs one group of four SNPs ahead: a group's values are
  // converted (the only wait on them) before the next group's loads are
  // issued into the same registers
  constexpr int NG = 4;
  CT cv[NG][R];
  const int s_last = s_end - 1;
#pragma unroll
  for (int k = 0; k < NG; ++k)
#pragma unroll
    for (int c = 0; c < R; ++c) cv[k][c] = C[((size_t)min(s_beg + k, s_last) * R + c) * ldc];
  for (int s0 = s_beg; s0 < s_end; s0 += NG) {
    double rb[NG][R];
#pragma unroll
    for (int k = 0; k < NG; ++k)
#pragma unroll
      for (int c = 0; c < R; ++c) rb[k][c] = (double)cv[k][c];
#pragma unroll
    for (int k = 0; k < NG; ++k)
#pragma unroll
      for (int c = 0; c < R; ++c) cv[k][c] = C[((size_t)min(s0 + NG + k, s_last) * R + c) * ldc];
#pragma unroll
    for (int k = 0; k < NG; ++k)
      if (s0 + k < s_end) quick(s0 + k, rb[k]);
  }
  while (pend) {
    const int k = __builtin_ctzll(pend);
    pend &= pend - 1;
    pst_pair_cold<R, LB, CT>(dp, d.s0, y, g, s_beg + k);
  }
}

// A wave holds the 64 traits of one trait wave (d.wtrait) and runs them over
// `run` <= 64 consecutive SNPs.  Blocks are numbered so that all trait waves of one
// SNP run are on one XCD and read the run's table rows from its L2.
template <int R, int LB, typename CT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_epilogue_pst(DevScan d, int n_wb, int run) {
  chunk_start(d);
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int wb = local % n_wb, rq = (local / n_wb) * 8 + xcd;
  const int w = wb * 4 + (int)(threadIdx.x >> 6);
  const int s_beg = rq * run;
  if (w >= d.n_waves || s_beg >= d.mc) return;
  const int s_end = min(d.mc, s_beg + run);
  const int g = __builtin_amdgcn_readfirstlane(((pst_ptr<int, true>)d.wpat)[w]);
  const int y = d.wtrait[(size_t)w * 64 + (threadIdx.x & 63)];
  if (y < 0) return;
  // d is the first kernel argument: offset 0 of the kernel-argument segment
  KernDev* dp = (KernDev*)__builtin_amdgcn_kernarg_segment_ptr();
  if (g >= 0) pst_trait_run<R, LB, CT, true>(d, dp, y, g, s_beg, s_end);
  else pst_trait_run<R, LB, CT, false>(d, dp, y, d.ypat[y], s_beg, s_end);
}

// ---------------------------------------------------------------------------
// launchers
// With one or two columns per SNP and at most 8 covariates the table
// kernels are compiled for the exact covariate count: no padded rows in the
// table, and the epilogue's registers drop far enough for 8 waves per SIMD.
static int pst_lb(int r, int l, int lb) { return (r <= 2 && l <= 8) ? l : lb; }

#define OA_PST_DISPATCH(FN)                                          \
  if (d.r <= 2 && d.l <= 8) {                                        \
    switch (d.r * 16 + d.l) {                                        \
      case 17: FN(1, 1) break; case 18: FN(1, 2) break;              \
      case 19: FN(1, 3) break; case 20: FN(1, 4) break;              \
      case 21: FN(1, 5) break; case 22: FN(1, 6) break;              \
      case 23: FN(1, 7) break; case 24: FN(1, 8) break;              \
      case 33: FN(2, 1) break; case 34: FN(2, 2) break;              \
      case 35: FN(2, 3) break; case 36: FN(2, 4) break;              \
      case 37: FN(2, 5) break; case 38: FN(2, 6) break;              \
      case 39: FN(2, 7) break; default: FN(2, 8) break;              \
    }                                                                \
  } else {                                                           \
    OA_DISPATCH_RLB(0, 0, FN)                                        \
  }

void launch_patsnp(const DevScan& d, int lb, hipStream_t st) {
  const long long nthr = (long long)d.mc * d.G;
  dim3 grid((unsigned)((nthr + 255) / 256)), block(256);
#define OA_PS(R_, LB_) { hipLaunchKernelGGL((k_patsnp<R_, LB_>), grid, block, 0, st, d); }
  OA_PST_DISPATCH(OA_PS)
#undef OA_PS
}

int patsnp_width(int r, int l, int lb) {
  const int b = pst_lb(r, l, lb);
  const int rr = r * (r + 1) / 2;
  const int ok = b * r + (r > 1 ? b * r : 0) + 2 * rr + b * r;
  return (ok + 2) & ~1;   // PatSnp<R, LB>::W
}

void launch_epilogue_pst(const DevScan& d, int prec, int lb, hipStream_t st) {
  if (d.n_waves <= 0 || d.mc <= 0) return;
  // SNP runs of up to 64, shorter when the grid would hold fewer than 16384
  // waves; OAGPU_PST_RUN (1..64) sets the run length
  const long long work = (long long)d.n_waves * d.mc;
  int run = (int)std::max(1ll, std::min(64ll, work / 16384));
  if (const char* e = getenv("OAGPU_PST_RUN")) {
    const int v = atoi(e);
    if (v >= 1 && v <= 64) run = v;
  }
  const int n_wb = (d.n_waves + 3) / 4, n_runs = (d.mc + run - 1) / run;
  dim3 grid((unsigned)((long long)n_wb * ((n_runs + 7) / 8) * 8)), block(256);
#define OA_EQ(R_, LB_)                                                                                           \
  {                                                                                                              \
    if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue_pst<R_, LB_, double>), grid, block, 0, st, d, n_wb, run); \
    else hipLaunchKernelGGL((k_epilogue_pst<R_, LB_, float>), grid, block, 0, st, d, n_wb, run);                  \
  }
  OA_PST_DISPATCH(OA_EQ)
#undef OA_EQ
}
#undef OA_PST_DISPATCH

}  // namespace oa
