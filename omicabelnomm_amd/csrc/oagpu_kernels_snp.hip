// SNP-side kernels of one chunk.  They replace, per SNP block:
//   replace_nans + replace_nans_avgs on XR .. src/Algorithm.cpp:618-622, src/Utility.cpp:103-127
//   interaction synthesis .................... src/AIOwrapper.cpp:767-840 (fused into the loads)
//   Sbr = XR_s'XR_s, build_S, potrf/potri .... src/Algorithm.cpp:646-671
//   generate_correction_missings_Str ......... src/Algorithm.cpp:1260-1315
//   per-pair build_S/correct_missings_S/sposv  src/Algorithm.cpp:714-792, 1370-1399
//   hpc_SSY (RSS) ............................ src/Algorithm.cpp:896-1024
//   hpc_statistics ........................... src/Algorithm.cpp:1034-1214
// Everything here is FP64.  The pair solve uses the block structure of
// S_pair = [[A_g, B],[B', D]]: A_g (covariates, per missing pattern) is
// factored once per pattern, each pair factors only its r x r Schur
// complement, and RSS is the downdate RSS0(trait) - |z_r|^2 of the
// covariate-only residual, plus the terms the reference's misplaced
// corrections add.  In exact arithmetic this is the reference's
// full-matrix Cholesky solve (same pivots, same failure rule).
#include "oagpu_kernels.h"
#include "oagpu_geno.h"

namespace oa {

// ---------------------------------------------------------------------------
// One thread per SNP: Str from the contraction output, and the unmasked
// var_x = diag(S_full^-1) (src/Algorithm.cpp:659-671) through the block
// inverse with U0 = chol(Stl).
template <int R, int LB, typename CT>
__global__ void k_snp_gram(DevScan d) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.mc) return;
  const int l = d.l, p = d.p;
  constexpr int RR = R * (R + 1) / 2;
  const CT* C = reinterpret_cast<const CT*>(d.Cbuf);
  double X[LB][R];
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      X[j][c] = 0.0;
      if (j < l) {
        const double v = (double)C[(size_t)(s * R + c) * d.ldc + d.t + j];
        d.Str[((size_t)s * l + j) * R + c] = v;
        X[j][c] = v;
      }
    }
  // X = U0^-T Str
#pragma unroll
  for (int j = 0; j < LB; ++j) {
    if (j < l) {
      const double inv = 1.0 / d.U0[upk(j, j)];
#pragma unroll
      for (int c = 0; c < R; ++c) {
        double v = X[j][c];
#pragma unroll
        for (int i = 0; i < j; ++i) v -= d.U0[upk(i, j)] * X[i][c];
        X[j][c] = v * inv;
      }
    }
  }
  // Schur0 = Sbr - X'X, then its Cholesky factor.
  double S[RR];
#pragma unroll
  for (int b = 0; b < R; ++b)
#pragma unroll
    for (int a = 0; a <= b; ++a) {
      double v = d.Sbr[(size_t)s * RR + upk(a, b)];
#pragma unroll
      for (int j = 0; j < LB; ++j)
        if (j < l) v -= X[j][a] * X[j][b];
      S[upk(a, b)] = v;
    }
  bool ok = true;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    double ajj = S[upk(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) ajj -= S[upk(k, j)] * S[upk(k, j)];
    if (!(ajj > 0.0)) ok = false;
    ajj = sqrt(ajj);
    S[upk(j, j)] = ajj;
#pragma unroll
    for (int i = j + 1; i < R; ++i) {
      double aji = S[upk(j, i)];
#pragma unroll
      for (int k = 0; k < j; ++k) aji -= S[upk(k, j)] * S[upk(k, i)];
      S[upk(j, i)] = aji / ajj;
    }
  }
  double* vx = d.varx + (size_t)s * p;
  if (!ok) {
    for (int h = 0; h < p; ++h) vx[h] = __longlong_as_double(0x7ff8000000000000ll);
  } else {
    // Ui = U_S^-1 (upper), Sinv = Ui Ui'
    double Ui[R][R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
#pragma unroll
      for (int i = 0; i < R; ++i) Ui[i][j] = 0.0;
      Ui[j][j] = 1.0 / S[upk(j, j)];
#pragma unroll
      for (int i = j - 1; i >= 0; --i) {
        double v = 0.0;
#pragma unroll
        for (int k = i + 1; k <= j; ++k) v += S[upk(i, k)] * Ui[k][j];
        Ui[i][j] = -v / S[upk(i, i)];
      }
    }
    double Sinv[R][R];
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
      for (int b = 0; b < R; ++b) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) v += Ui[a][k] * Ui[b][k];
        Sinv[a][b] = v;
      }
#pragma unroll
    for (int a = 0; a < R; ++a) vx[l + a] = Sinv[a][a];
    // W = U0^-1 X (in place, back substitution)
#pragma unroll
    for (int j = LB - 1; j >= 0; --j) {
      if (j < l) {
        const double inv = 1.0 / d.U0[upk(j, j)];
#pragma unroll
        for (int c = 0; c < R; ++c) {
          double v = X[j][c];
#pragma unroll
          for (int i = j + 1; i < LB; ++i)
            if (i < l) v -= d.U0[upk(j, i)] * X[i][c];
          X[j][c] = v * inv;
        }
      }
    }
#pragma unroll
    for (int h = 0; h < LB; ++h) {
      if (h < l) {
        double v = d.Stl_inv_diag[h];
#pragma unroll
        for (int a = 0; a < R; ++a)
#pragma unroll
          for (int b = 0; b < R; ++b) v += X[h][a] * Sinv[a][b] * X[h][b];
        vx[h] = v;
      }
    }
  }
  if (d.d_varx)
    for (int h = 0; h < p; ++h) d.d_varx[(size_t)(d.s0 + s) * p + h] = vx[h];
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ double widen(float v) { return (double)v; }
__device__ __forceinline__ double widen(_Float16 v) { return (double)(float)v; }

// generate_correction_missings_Str (src/Algorithm.cpp:1260-1315) for every
// pattern: C_lr = sum over the pattern's missing rows of xl x', C_rr of x x'.
// The genotype values come from the compact u-major rows of U written by the
// operand pass.  One wave per pattern, one lane per SNP: the pattern's row
// positions and covariate rows are wave-uniform, and the 64 SNPs' values at
// one row are contiguous.  Blocks are numbered so that the pattern groups of
// one 64-SNP tile run back to back on one XCD and share its L2.
// SPL consecutive SNPs per lane (one vector load per row); two when a SNP
// has one column and there are at most 16 covariates, so every covariate
// value read from LDS feeds twice the multiply-adds without raising the
// register count past 5 waves per SIMD.
template <int R, int LB>
constexpr int corr_spl() { return (R == 1 && LB <= 16) ? 2 : 1; }

template <typename XT, int N> struct XVec { typedef __attribute__((ext_vector_type(N))) XT T; };
template <typename XT> struct XVec<XT, 1> { typedef XT T; };

template <int R, int LB, typename XT>
__global__ __launch_bounds__(256) void k_corr(DevScan d, int gq_count) {
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
  // wave-uniform broadcasts; the multiply-adds run without branches.
  constexpr int CORR_BATCH = 8;
  constexpr int STAGE = LB > 16 ? 32 : 64;   // rows staged at once (LDS: 4 waves x STAGE x LB doubles)
  __shared__ double xls[4][STAGE][LB];
  double (*xw)[LB] = xls[threadIdx.x >> 6];
  const int beg = d.pu_off[g], end = d.pu_off[g + 1];
  for (int i0 = beg; i0 < end; i0 += STAGE) {
    const int nb = __builtin_amdgcn_readfirstlane(min(STAGE, end - i0));
    // lanes past the pattern's end point at row nU of Xu, which is zero
    const int my_u = lane < nb ? d.pu_list[i0 + lane] : d.nU;
    if (lane < STAGE) {
#pragma unroll
      for (int j = 0; j < LB; ++j) xw[lane][j] = (lane < nb && j < l) ? d.XLu[(size_t)my_u * l + j] : 0.0;
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
#pragma unroll
        for (int j = 0; j < LB; ++j) {
          const double a = xl[j];
#pragma unroll
          for (int c = 0; c < W; ++c) lr[j][c] += a * xv[q][c];
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
// hpc_statistics for one (trait, SNP) pair (src/Algorithm.cpp:1034-1214):
// R2, SE from the unmasked var_x scaled by (n - Ymiss)/n, t, the
// normal-approximation P, thresholds, counts, and the hit records.
__device__ __forceinline__ double pvalue_ref(double t) {
  if (t < 1.6) return 0.5 - 0.1 * (4.4 - t) * t;
  return erfc(t * 0.70710678118654752440084436210484903928);
}

// Per-trait constants of the statistics, computed once per thread.
struct TraitK {
  int ymiss;
  int n_c;       // n_stat - ymiss - p
  double vf;     // (n_c + p) / n_stat: SE scale of the unmasked var_x
  double r2s;    // R2 = 1 - rss * r2s = 1 - (rss / (n_c - 1)) / (SST / (n_stat - 1))
  double inv_nc; // sigma^2 = rss / n_c
};

__device__ __forceinline__ TraitK trait_k(const DevScan& d, int y) {
  TraitK k;
  k.ymiss = d.ymiss[y];
  k.n_c = d.n_stat - k.ymiss - d.p;
  k.vf = (double)(k.n_c + d.p) / (double)d.n_stat;
  k.r2s = (double)(d.n_stat - 1) / ((double)(k.n_c - 1) * d.sst[y]);
  k.inv_nc = 1.0 / (double)k.n_c;
  return k;
}

// Coefficient vectors use a split layout: b[0..LB) covariates (index h < l
// valid), b[LB..LB+R) the SNP terms (coefficient index h = l + c), so every
// register index is a compile-time constant.
template <int R, int LB>
__device__ void stats_emit(const DevScan& d, const TraitK& tk, int y, int s, const double* b, double rss) {
  const int l = d.l, p = d.p;
  const int ymiss = tk.ymiss;
  const double vf = tk.vf;
  const double R2 = 1.0 - rss * tk.r2s;
  const double* vx = d.varx + (size_t)s * p;
  const size_t pair = (size_t)y * d.m_total + (size_t)(d.s0 + s);
  const bool dense = d.d_beta != nullptr;
  const double qnan = __longlong_as_double(0x7ff8000000000000ll);
  if (dense) {
    d.d_r2[pair] = R2;
    d.d_rss[pair] = rss;
#pragma unroll
    for (int j = 0; j < LB + R; ++j) {
      const int h = j < LB ? j : l + (j - LB);
      if (j >= LB || j < l) {
        d.d_beta[pair * p + h] = b[j];
        d.d_se[pair * p + h] = qnan;
        d.d_t[pair * p + h] = qnan;
        d.d_p[pair * p + h] = qnan;
        d.d_kept[pair * p + h] = 0;
      }
    }
  }
  // kept[j] for slot j of the split layout
  unsigned long long mask = 0;
  int nk = 0, nsig = 0;
  bool has_snp = false;
  double sig = 0.0;
  if (rss > 0.0 && (R2 > d.min_r2_store || d.store_all)) {
    const double sg2 = rss * tk.inv_nc;
    sig = sqrt(sg2);
    const double s2 = sg2 * vf;
    const double tt = d.t_star * d.t_star;   // exact: t_star is a float
#pragma unroll
    for (int j = 0; j < LB + R; ++j) {
      const int h = j < LB ? j : l + (j - LB);
      if (j < LB && (j >= l || j < d.start_h)) continue;
      const double vxh = vx[h];
      if (vxh < 0.0) atomicOr(&d.counters[CNT_ERR], 1ull);
      // |b/se| > t_star  <=>  b^2 > t_star^2 se^2 (se^2 >= 0): the square
      // roots and the division are only needed for what is kept
      const double se2 = s2 * vxh;
      if (!(dense || d.store_all || (se2 >= 0.0 && b[j] * b[j] > tt * se2))) continue;
      const double se = sig * sqrt(vxh * vf);
      const double tv = fabs(b[j] / se);
      if (dense) {
        d.d_se[pair * p + h] = se;
        d.d_t[pair * p + h] = tv;
      }
      if (tv > d.t_star || d.store_all) {
        mask |= 1ull << j;
        ++nk;
        if (j >= LB) has_snp = true;
        const double pv = pvalue_ref(tv);
        if (d.min_p_disp > pv) ++nsig;
        if (dense) {
          d.d_p[pair * p + h] = pv;
          d.d_kept[pair * p + h] = 1;
        }
      }
    }
  }
  if (nsig) atomicAdd(&d.counters[CNT_SIG], (unsigned long long)nsig);
  if (dense) d.d_emitted[pair] = has_snp ? 1 : 0;
  if (!has_snp) return;
  atomicAdd(&d.counters[CNT_EMITTED], 1ull);
  const unsigned long long base = atomicAdd(&d.counters[CNT_HITS], (unsigned long long)nk);
  if (base + nk > (unsigned long long)d.hit_cap) return;
  unsigned long long o = base;
  // SNP terms then covariates, each descending: h descending overall
#pragma unroll
  for (int jj = 0; jj < LB + R; ++jj) {
    const int j = jj < R ? LB + R - 1 - jj : LB - 1 - (jj - R);
    if (!((mask >> j) & 1ull)) continue;
    const int h = j < LB ? j : l + (j - LB);
    const double se = sig * sqrt(vx[h] * vf);
    const double tv = fabs(b[j] / se);
    oagpu_hit hr;
    hr.trait = d.trait_offset + y;
    hr.snp = d.snp_offset + d.s0 + s;
    hr.h = h;
    hr.n_used = d.n_stat - ymiss;
    hr.b = b[j];
    hr.se = se;
    hr.t = tv;
    hr.p = pvalue_ref(tv);
    hr.r2 = R2;
    d.hits[o++] = hr;
  }
}

// RSS over rows with y0 != 0 (src/Algorithm.cpp:976-1024), one thread.
template <int R, int LB>
__device__ double explicit_rss_thread(const DevScan& d, int y, int s, const double* b) {
  int src[R], env[R];
  col_desc<R>(d, d.s0 + s, src, env);
  double mean[R];
#pragma unroll
  for (int c = 0; c < R; ++c) mean[c] = d.col_mean[(size_t)s * R + c];
  double acc = 0.0;
  for (int k = 0; k < d.n; ++k) {
    if (d.keep && !d.keep[k]) continue;
    double yv = load_elem(d.Y, d.ydt, (size_t)y * d.ldy + k);
    if (yv != yv) yv = 0.0;
    if (yv == 0.0) continue;
    double sy = 0.0;
#pragma unroll
    for (int h = 0; h < LB; ++h)
      if (h < d.l) sy += d.XL_cm[(size_t)h * d.n_pad + k] * b[h];
#pragma unroll
    for (int c = 0; c < R; ++c) {
      double v = geno_value(d, src[c], env[c], k);
      if (v != v) v = mean[c];
      sy += v * b[LB + c];
    }
    sy -= yv;
    acc += sy * sy;
  }
  return acc;
}

// Shared end of a pair: the hand-over to the explicit residual pass and the
// statistics.  b in the split layout (b[0..LB) covariates, b[LB..LB+R) SNP
// terms).
template <int R, int LB>
__device__ __forceinline__ void finish_pair(const DevScan& d, const TraitK& tk, int y, int s, bool ok, double* b,
                                            const double* rb, double rss) {
  const bool fallback = !ok || d.yzero[y] > 0 || !(rss >= RSS_GUARD * d.yy[y]);
  if (fallback) {
    if (!ok) {   // sposv left the right-hand side in place
#pragma unroll
      for (int h = 0; h < LB; ++h) b[h] = (h < d.l) ? d.rt[(size_t)h * d.ldt + y] : 0.0;
#pragma unroll
      for (int c = 0; c < R; ++c) b[LB + c] = rb[c];
    }
    const unsigned long long slot = atomicAdd(&d.counters[CNT_FB_CHUNK], 1ull);
    atomicAdd(&d.counters[CNT_FALLBACK], 1ull);
    if (slot < (unsigned long long)d.fb_cap) {
      FallbackEntry* f = d.fb + slot;
      f->y = y;
      f->s = s;
      f->flags = ok ? 0 : 1;
#pragma unroll
      for (int h = 0; h < LB + R; ++h) f->b[h] = b[h];
      return;
    }
    rss = explicit_rss_thread<R, LB>(d, y, s, b);
  }
  stats_emit<R, LB>(d, tk, y, s, b, rss);
}

// ---------------------------------------------------------------------------
// One thread per (trait, SNP) pair when patterns are many (typically one
// per trait): a wave covers 64 consecutive SNPs of one trait, so the trait's
// and its pattern's data (U_A, zt, constants) are wave-uniform and the
// pattern's corrections of the 64 SNPs are one coalesced load per term;
// only the 64 contraction outputs are scattered (one 4- or 8-byte value per
// SNP row).  Blocks hold 4 consecutive traits and are numbered so that the
// blocks of one 64-SNP tile run back to back on one XCD, whose L2 then
// serves the tile's contraction rows to all of them.
template <int R, int LB, typename CT>
__global__ __launch_bounds__(256) void k_epilogue(DevScan d, int n_q) {
  const int lane = threadIdx.x & 63;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int q = local % n_q, tile = (local / n_q) * 8 + xcd;
  const int y = __builtin_amdgcn_readfirstlane(q * 4 + (threadIdx.x >> 6));
  if (y >= d.t || tile * 64 >= d.mc) return;   // wave-uniform
  const int s = tile * 64 + lane;
  if (s >= d.mc) return;
  constexpr int RR = R * (R + 1) / 2;
  constexpr int P = LB + R;
  const int l = d.l, G = d.G;
  const int g = d.ypat[y];
  const int cnt = d.pat_cnt[g];
  const CT* C = reinterpret_cast<const CT*>(d.Cbuf);

  double rb[R];
#pragma unroll
  for (int c = 0; c < R; ++c) rb[c] = (double)C[(size_t)(s * R + c) * d.ldc + y];

  // Off-diagonal block B_q as the reference assembles it: every column's
  // correction is subtracted from the first SNP column
  // (src/Algorithm.cpp:1390-1396).
  double X[LB][R];
  double Clr[R > 1 ? LB : 1][R];
  double Crr[RR];
#pragma unroll
  for (int j = 0; j < LB; ++j) {
#pragma unroll
    for (int c = 0; c < R; ++c) {
      X[j][c] = (j < l) ? d.Str[((size_t)s * l + j) * R + c] : 0.0;
      if (R > 1) Clr[R > 1 ? j : 0][c] = 0.0;
    }
  }
#pragma unroll
  for (int e = 0; e < RR; ++e) Crr[e] = 0.0;
  if (cnt > 0) {
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      if (j < l) {
        double tot = 0.0;
#pragma unroll
        for (int c = 0; c < R; ++c) {
          const double v = d.Ccorr[((size_t)g * d.Q + j * R + c) * d.mc + s];
          tot += v;
          if (R > 1) Clr[R > 1 ? j : 0][c] = v;
        }
        X[j][0] -= tot;
      }
    }
#pragma unroll
    for (int e = 0; e < RR; ++e) Crr[e] = d.Ccorr[((size_t)g * d.Q + l * R + e) * d.mc + s];
  }
  // D_q: Sbr corrected only in rows j < min(l, r) (src/Algorithm.cpp:1300).
  double Sc[RR];
#pragma unroll
  for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
    for (int a = 0; a <= b2; ++a)
      Sc[upk(a, b2)] = d.Sbr[(size_t)s * RR + upk(a, b2)] - ((a < l) ? Crr[upk(a, b2)] : 0.0);

  // X = U_A^-T B_q
#pragma unroll
  for (int j = 0; j < LB; ++j) {
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
  }
  // Schur complement and its Cholesky factor.
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
  double zt[LB];
#pragma unroll
  for (int j = 0; j < LB; ++j) zt[j] = (j < l) ? d.zt[(size_t)j * d.ldt + y] : 0.0;

  double b[P];
#pragma unroll
  for (int h = 0; h < P; ++h) b[h] = 0.0;
  double rss = 0.0;
  if (ok) {
    double z[R];
#pragma unroll
    for (int c = 0; c < R; ++c) {
      double v = rb[c];
#pragma unroll
      for (int j = 0; j < LB; ++j) v -= X[j][c] * zt[j];
      z[c] = v;
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {   // U_S' z = num
      double v = z[i];
#pragma unroll
      for (int k = 0; k < i; ++k) v -= Sc[upk(k, i)] * z[k];
      z[i] = v / Sc[upk(i, i)];
    }
    double zz = 0.0;
#pragma unroll
    for (int c = 0; c < R; ++c) zz += z[c] * z[c];
    double br[R];
#pragma unroll
    for (int i = R - 1; i >= 0; --i) {   // U_S b_r = z
      double v = z[i];
#pragma unroll
      for (int k = i + 1; k < R; ++k) v -= Sc[upk(i, k)] * br[k];
      br[i] = v / Sc[upk(i, i)];
    }
    rss = d.rss0[y] - zz;
#pragma unroll
    for (int c = 0; c < R; ++c) b[LB + c] = br[c];
    const bool quirk = (R > 1) && cnt > 0;
    const bool need_bl = d.need_bl || quirk || d.yzero[y] > 0 || !(rss >= RSS_GUARD * d.yy[y]);
    if (need_bl) {
      // b_l = U_A^-1 (zt - X b_r)
#pragma unroll
      for (int j = LB - 1; j >= 0; --j) {
        if (j < l) {
          double v = zt[j];
#pragma unroll
          for (int c = 0; c < R; ++c) v -= X[j][c] * br[c];
#pragma unroll
          for (int i = j + 1; i < LB; ++i)
            if (i < l) v -= d.UA[(size_t)upk(j, i) * G + g] * b[i];
          b[j] = v * d.UA_invd[(size_t)j * G + g];
        }
      }
    }
    if (quirk) {
      // RSS of the reference's b on the true masked Gram matrix: add
      // b' (S_true - S_q) b, nonzero in the B block (first SNP column got
      // every correction) and in D rows >= l (never corrected).
      double q = 0.0;
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        if (j < l) {
          double rest = 0.0;
#pragma unroll
          for (int c = 1; c < R; ++c) rest += Clr[R > 1 ? j : 0][c];
          double v = rest * br[0];
#pragma unroll
          for (int c = 1; c < R; ++c) v -= Clr[R > 1 ? j : 0][c] * br[c];
          q += 2.0 * b[j] * v;
        }
      }
#pragma unroll
      for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
        for (int a = 0; a <= b2; ++a)
          if (a >= l) q -= (a == b2 ? 1.0 : 2.0) * br[a] * Crr[upk(a, b2)] * br[b2];
      rss += q;
    }
  }
  finish_pair<R, LB>(d, trait_k(d, y), y, s, ok, b, rb, rss);
}

// ---------------------------------------------------------------------------
// Few patterns (G << t): everything in the pair solve that depends only on
// (SNP, pattern) is computed once per (SNP, pattern) instead of once per
// pair.  With X = U_A^-T B_q, the Schur factor U_S, M = U_A^-1 X and
// a_y = U_A^-1 zt (per trait), the covariate coefficients are
// b_l = a_y - M b_r.  The reference's misplaced-correction RSS terms,
// 2 b_l' K b_r - b_r' H b_r, expand to 2 (KT' zt)' b_r - b_r' H2 b_r with
// KT = U_A^-T K and H2 = H + 2 sym(M' K), so a pair needs b_l only when it
// reports or re-evaluates it.  Rows j >= l of X, KT and M are zero, so a
// pair reads its table row without branches.  Rows hold an even number of
// doubles (16-byte aligned).
template <int R, int LB>
struct PatSnp {
  static constexpr int RR = R * (R + 1) / 2;
  static constexpr int KR = R > 1 ? LB * R : 0;   // no misplaced terms with one column
  static constexpr int X = 0, KT = LB * R, US = LB * R + KR, H = US + RR, M = H + RR, OK = M + LB * R,
                       W = (OK + 2) & ~1;
};

template <int R, int LB>
__global__ __launch_bounds__(256) void k_patsnp(DevScan d) {
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
  if (T::W > T::OK + 1) out[T::W - 1] = 0.0;
}

// One thread per (trait, SNP) pair, reading the (SNP, pattern) table.  A
// wave covers 64 traits of one SNP.  Every load a pair needs in the common
// case -- the trait's constants, its two C entries, its pattern's table row
// -- is issued before the first use, so a pair costs about one memory round
// trip; M and a_y are read only when b_l is needed.  Blocks are numbered so
// that all trait blocks of one group of 4 SNPs run on one XCD and read the
// SNPs' table rows from its L2.
template <int R, int LB, typename CT>
__global__ __launch_bounds__(256) void k_epilogue_pst(DevScan d, int n_tb) {
  using T = PatSnp<R, LB>;
  constexpr int RR = T::RR;
  constexpr int P = LB + R;
  const int lane = threadIdx.x & 63;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int tbk = local % n_tb, sq = (local / n_tb) * 8 + xcd;
  const int s = __builtin_amdgcn_readfirstlane(sq * 4 + (threadIdx.x >> 6));
  if (s >= d.mc) return;
  const int y = tbk * 64 + lane;
  if (y >= d.t) return;
  const int l = d.l;
  const int g = d.ypat[y];
  const TraitK tk = trait_k(d, y);
  const double rss0 = d.rss0[y];
  const double guard = RSS_GUARD * d.yy[y];
  const bool yzero = d.yzero[y] > 0;
  double zt[LB];
#pragma unroll
  for (int j = 0; j < LB; ++j) zt[j] = (j < l) ? d.zt[(size_t)j * d.ldt + y] : 0.0;
  const CT* C = reinterpret_cast<const CT*>(d.Cbuf);
  double rb[R];
#pragma unroll
  for (int c = 0; c < R; ++c) rb[c] = (double)C[(size_t)(s * R + c) * d.ldc + y];
  const double* tb = d.pst + ((size_t)s * d.G + g) * T::W;
  double X[LB][R], KT[R > 1 ? LB : 1][R], us[RR], h2[RR];
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      X[j][c] = tb[T::X + j * R + c];
      if (R > 1) KT[R > 1 ? j : 0][c] = tb[T::KT + j * R + c];
    }
#pragma unroll
  for (int e = 0; e < RR; ++e) {
    us[e] = tb[T::US + e];
    h2[e] = tb[T::H + e];
  }
  const bool ok = tb[T::OK] != 0.0;
  double b[P];
#pragma unroll
  for (int h = 0; h < P; ++h) b[h] = 0.0;
  double rss = 0.0;
  if (ok) {
    double z[R], kz[R];
#pragma unroll
    for (int c = 0; c < R; ++c) {
      double v = rb[c], k = 0.0;
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        v -= X[j][c] * zt[j];
        if (R > 1) k += KT[R > 1 ? j : 0][c] * zt[j];
      }
      z[c] = v;
      kz[c] = k;
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {   // U_S' z = num (diagonal reciprocals in the table)
      double v = z[i];
#pragma unroll
      for (int k = 0; k < i; ++k) v -= us[upk(k, i)] * z[k];
      z[i] = v * us[upk(i, i)];
    }
    double zz = 0.0;
#pragma unroll
    for (int c = 0; c < R; ++c) zz += z[c] * z[c];
    double br[R];
#pragma unroll
    for (int i = R - 1; i >= 0; --i) {   // U_S b_r = z
      double v = z[i];
#pragma unroll
      for (int k = i + 1; k < R; ++k) v -= us[upk(i, k)] * br[k];
      br[i] = v * us[upk(i, i)];
    }
    rss = rss0 - zz;
    if (R > 1) {   // misplaced-correction terms (zero for a pattern without missing rows)
      double q = 0.0;
#pragma unroll
      for (int c = 0; c < R; ++c) q += 2.0 * kz[c] * br[c];
#pragma unroll
      for (int b2 = 0; b2 < R; ++b2)
#pragma unroll
        for (int a = 0; a <= b2; ++a) q -= h2[upk(a, b2)] * br[a] * br[b2];
      rss += q;
    }
#pragma unroll
    for (int c = 0; c < R; ++c) b[LB + c] = br[c];
    if (d.need_bl || yzero || !(rss >= guard)) {   // reported, or the pair is re-evaluated
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
  finish_pair<R, LB>(d, tk, y, s, ok, b, rb, rss);
}

// Explicit residual pass for the pairs the epilogue handed over: one wave
// per pair.
template <int R, int LB>
__global__ __launch_bounds__(256) void k_fallback(DevScan d) {
  constexpr int P = LB + R;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long total = d.counters[CNT_FB_CHUNK];
  const int count = (int)(total < (unsigned long long)d.fb_cap ? total : (unsigned long long)d.fb_cap);
  for (int e = blockIdx.x * 4 + w; e < count; e += gridDim.x * 4) {
    const FallbackEntry* f = d.fb + e;
    const int y = f->y, s = f->s;
    double b[P];
#pragma unroll
    for (int h = 0; h < P; ++h) b[h] = f->b[h];
    int src[R], env[R];
    col_desc<R>(d, d.s0 + s, src, env);
    double mean[R];
#pragma unroll
    for (int c = 0; c < R; ++c) mean[c] = d.col_mean[(size_t)s * R + c];
    double acc = 0.0;
    for (int k = lane; k < d.n; k += 64) {
      if (d.keep && !d.keep[k]) continue;
      double yv = load_elem(d.Y, d.ydt, (size_t)y * d.ldy + k);
      if (yv != yv) yv = 0.0;
      if (yv == 0.0) continue;
      double sy = 0.0;
#pragma unroll
      for (int h = 0; h < LB; ++h)
        if (h < d.l) sy += d.XL_cm[(size_t)h * d.n_pad + k] * b[h];
#pragma unroll
      for (int c = 0; c < R; ++c) {
        double v = geno_value(d, src[c], env[c], k);
        if (v != v) v = mean[c];
        sy += v * b[LB + c];
      }
      sy -= yv;
      acc += sy * sy;
    }
    acc = wave_sum(acc);
    if (lane == 0) stats_emit<R, LB>(d, trait_k(d, y), y, s, b, acc);
  }
}

__global__ void k_reset_counter(unsigned long long* c, int idx) { c[idx] = 0ull; }
__global__ void k_reset_counters(unsigned long long* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// Sbr = x'x over the subset's rows for r = 1 (a request that shares the
// chunk's operand but not its individuals); x is imputed with the chunk's
// column mean.  One block per SNP.
__global__ __launch_bounds__(256) void k_subset_sbr(DevScan d, double* out) {
  __shared__ double shd[4];
  const int s = blockIdx.x;
  const int gs = d.s0 + s;
  const int src = d.col_src ? d.col_src[gs] : gs;
  const double mean = d.col_mean[s];
  double acc = 0.0;
  for (int k0 = 4 * threadIdx.x; k0 < d.n; k0 += 1024) {
    double x[4];
    geno_value4(d, src, -1, k0, x);
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (k0 + e < d.n && d.keep[k0 + e]) {
        const double v = (x[e] == x[e]) ? x[e] : mean;
        acc += v * v;
      }
  }
  acc = block_sum256(acc, shd);
  if (threadIdx.x == 0) out[s] = acc;
}

// ---------------------------------------------------------------------------
// launchers
// (R, LB) pairs with compiled kernels: R=1: LB 8/16/24/32; R=2: 8/16; R=3,4: 8.
#define OA_DISPATCH_RLB(R_, LB_, FN)          \
  if (d.r == 1 && lb == 8) { FN(1, 8); }      \
  else if (d.r == 1 && lb == 16) { FN(1, 16); } \
  else if (d.r == 1 && lb == 24) { FN(1, 24); } \
  else if (d.r == 1) { FN(1, 32); }           \
  else if (d.r == 2 && lb == 8) { FN(2, 8); } \
  else if (d.r == 2) { FN(2, 16); }           \
  else if (d.r == 3) { FN(3, 8); }            \
  else { FN(4, 8); }

void launch_snp_gram(const DevScan& d, int prec, int lb, hipStream_t st) {
  dim3 grid((d.mc + 63) / 64), block(64);
#define OA_SG(R_, LB_)                                                                     \
  if (prec == PREC_FP64) hipLaunchKernelGGL((k_snp_gram<R_, LB_, double>), grid, block, 0, st, d); \
  else hipLaunchKernelGGL((k_snp_gram<R_, LB_, float>), grid, block, 0, st, d);
  OA_DISPATCH_RLB(0, 0, OA_SG)
#undef OA_SG
}

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
  // one-column SNPs with up to 8 covariates: compiled for the exact l, so no
  // multiply-adds run on zero padding
  if (d.r == 1 && lb == 8) {
    switch (d.l) {
      case 1: OA_CR(1, 1) break;
      case 2: OA_CR(1, 2) break;
      case 3: OA_CR(1, 3) break;
      case 4: OA_CR(1, 4) break;
      case 5: OA_CR(1, 5) break;
      case 6: OA_CR(1, 6) break;
      case 7: OA_CR(1, 7) break;
      default: OA_CR(1, 8) break;
    }
  } else {
    OA_DISPATCH_RLB(0, 0, OA_CR)
  }
#undef OA_CR
}

void launch_epilogue(const DevScan& d, int prec, int lb, hipStream_t st) {
  const int n_q = (d.t + 3) / 4, tiles = (d.mc + 63) / 64;
  dim3 grid((unsigned)((long long)((tiles + 7) / 8) * 8 * n_q)), block(256);
#define OA_EP(R_, LB_)                                                                                \
  if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue<R_, LB_, double>), grid, block, 0, st, d, n_q); \
  else hipLaunchKernelGGL((k_epilogue<R_, LB_, float>), grid, block, 0, st, d, n_q);
  OA_DISPATCH_RLB(0, 0, OA_EP)
#undef OA_EP
}

void launch_patsnp(const DevScan& d, int lb, hipStream_t st) {
  const long long nthr = (long long)d.mc * d.G;
  dim3 grid((unsigned)((nthr + 255) / 256)), block(256);
#define OA_PS(R_, LB_) hipLaunchKernelGGL((k_patsnp<R_, LB_>), grid, block, 0, st, d);
  OA_DISPATCH_RLB(0, 0, OA_PS)
#undef OA_PS
}

int patsnp_width(int r, int lb) {
  const int rr = r * (r + 1) / 2;
  const int ok = lb * r + (r > 1 ? lb * r : 0) + 2 * rr + lb * r;
  return (ok + 2) & ~1;   // PatSnp<R, LB>::W
}

void launch_epilogue_pst(const DevScan& d, int prec, int lb, hipStream_t st) {
  const int n_tb = (d.t + 63) / 64, n_sq = (d.mc + 3) / 4;
  dim3 grid((unsigned)((long long)n_tb * ((n_sq + 7) / 8) * 8)), block(256);
#define OA_EQ(R_, LB_)                                                                                    \
  if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue_pst<R_, LB_, double>), grid, block, 0, st, d, n_tb); \
  else hipLaunchKernelGGL((k_epilogue_pst<R_, LB_, float>), grid, block, 0, st, d, n_tb);
  OA_DISPATCH_RLB(0, 0, OA_EQ)
#undef OA_EQ
}

void launch_fallback(const DevScan& d, int lb, hipStream_t st) {
  dim3 grid(1024), block(256);
#define OA_FB(R_, LB_) hipLaunchKernelGGL((k_fallback<R_, LB_>), grid, block, 0, st, d);
  OA_DISPATCH_RLB(0, 0, OA_FB)
#undef OA_FB
}

void launch_subset_sbr(const DevScan& d, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_subset_sbr, dim3(d.mc), dim3(256), 0, st, d, out);
}

void launch_reset_counters(unsigned long long* c, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, c, n);
}

void launch_reset_counter(unsigned long long* c, int idx, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, c, idx);
}

}  // namespace oa
