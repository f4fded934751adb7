// The pair solve when missing-value patterns are many (typically one per
// trait): each pair assembles its own S block from the per-SNP Gram pieces
// and its pattern's corrections.
#include "oagpu_kernels.h"
#include "oagpu_pair.h"

namespace oa {

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
  chunk_start(d);
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
// launcher
void launch_epilogue(const DevScan& d, int prec, int lb, hipStream_t st) {
  const int n_q = (d.t + 3) / 4, tiles = (d.mc + 63) / 64;
  dim3 grid((unsigned)((long long)((tiles + 7) / 8) * 8 * n_q)), block(256);
#define OA_EP(R_, LB_)                                                                                  \
  {                                                                                                       \
    if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue<R_, LB_, double>), grid, block, 0, st, d, n_q); \
    else hipLaunchKernelGGL((k_epilogue<R_, LB_, float>), grid, block, 0, st, d, n_q);                  \
  }
  // one or two columns per SNP and at most 8 covariates: compiled for the
  // exact covariate count (fewer registers and no padded multiply-adds)
  if (d.r <= 2 && d.l <= 8) {
    switch (d.r * 16 + d.l) {
      case 17: OA_EP(1, 1) break; case 18: OA_EP(1, 2) break;
      case 19: OA_EP(1, 3) break; case 20: OA_EP(1, 4) break;
      case 21: OA_EP(1, 5) break; case 22: OA_EP(1, 6) break;
      case 23: OA_EP(1, 7) break; case 24: OA_EP(1, 8) break;
      case 33: OA_EP(2, 1) break; case 34: OA_EP(2, 2) break;
      case 35: OA_EP(2, 3) break; case 36: OA_EP(2, 4) break;
      case 37: OA_EP(2, 5) break; case 38: OA_EP(2, 6) break;
      case 39: OA_EP(2, 7) break; default: OA_EP(2, 8) break;
    }
  } else {
    OA_DISPATCH_RLB(0, 0, OA_EP)
  }
#undef OA_EP
}

}  // namespace oa
