// Pair statistics shared by the pair kernels (oagpu_kernels_snp.hip,
// oagpu_kernels_pst.hip, oagpu_kernels_epi.hip): hpc_statistics, the explicit
// residual, the hand-over to the fallback pass, and the (R, LB) dispatch of
// the compiled kernels.
//
// The pair solve (all FP64) uses the block structure of
// S_pair = [[A_g, B],[B', D]]: A_g (covariates, per missing pattern) is
// factored once per pattern, each pair factors only its r x r Schur
// complement, and RSS is the downdate RSS0(trait) - |z_r|^2 of the
// covariate-only residual, plus the terms the reference's misplaced
// corrections add.  In exact arithmetic this is the reference's full-matrix
// Cholesky solve (same pivots, same failure rule; src/Algorithm.cpp:714-792,
// 1370-1399).
#pragma once
#include "oagpu_kernels.h"
#include "oagpu_geno.h"

namespace oa {

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
// register index is a compile-time constant.  vx points at the SNP's var_x
// row (p values); a caller whose SNP is wave-uniform passes it in the
// constant address space so the values come through scalar loads.
template <int R, int LB, typename VX>
__device__ void stats_emit(const DevScan& d, const TraitK& tk, int y, int s, const double* b, double rss, VX vx) {
  const int l = d.l, p = d.p;
  const int ymiss = tk.ymiss;
  const double vf = tk.vf;
  const double R2 = 1.0 - rss * tk.r2s;
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
  double sg2 = 0.0;   // sigma^2; its square root only for what is kept
  if (rss > 0.0 && (R2 > d.min_r2_store || d.store_all)) {
    sg2 = rss * tk.inv_nc;
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
      const double se = sqrt(sg2) * sqrt(vxh * vf);
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
    const double se = sqrt(sg2) * sqrt(vx[h] * vf);
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

template <int R, int LB>
__device__ __forceinline__ void stats_emit(const DevScan& d, const TraitK& tk, int y, int s, const double* b,
                                           double rss) {
  stats_emit<R, LB>(d, tk, y, s, b, rss, static_cast<const double*>(d.varx + (size_t)s * d.p));
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
// terms); yzero and guard (RSS_GUARD * yy) are the trait's.
template <int R, int LB, typename VX>
__device__ __forceinline__ void finish_pair(const DevScan& d, const TraitK& tk, int y, int s, bool ok, double* b,
                                            const double* rb, double rss, bool yzero, double guard, VX vx) {
  const bool fallback = !ok || yzero || !(rss >= guard);
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
      // coefficients in the reference's order h (covariates 0..l-1, then
      // the SNP terms): the fallback kernel may be compiled for another LB
#pragma unroll
      for (int h = 0; h < LB; ++h)
        if (h < d.l) f->b[h] = b[h];
#pragma unroll
      for (int c = 0; c < R; ++c) f->b[d.l + c] = b[LB + c];
      return;
    }
    rss = explicit_rss_thread<R, LB>(d, y, s, b);
  }
  stats_emit<R, LB>(d, tk, y, s, b, rss, vx);
}

template <int R, int LB>
__device__ __forceinline__ void finish_pair(const DevScan& d, const TraitK& tk, int y, int s, bool ok, double* b,
                                            const double* rb, double rss) {
  finish_pair<R, LB>(d, tk, y, s, ok, b, rb, rss, d.yzero[y] > 0, RSS_GUARD * d.yy[y],
                     static_cast<const double*>(d.varx + (size_t)s * d.p));
}

}  // namespace oa

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

