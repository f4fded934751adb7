// The pair solve with a shared (SNP, pattern) table: when patterns are few
// (4 G <= t), k_patsnp computes everything that depends only on (SNP,
// pattern) once and k_epilogue_pst finishes each (trait, SNP) pair from it.
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
  const int n_tb = (d.t + 63) / 64, n_sq = (d.mc + 3) / 4;
  dim3 grid((unsigned)((long long)n_tb * ((n_sq + 7) / 8) * 8)), block(256);
#define OA_EQ(R_, LB_)                                                                                      \
  {                                                                                                         \
    if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue_pst<R_, LB_, double>), grid, block, 0, st, d, n_tb); \
    else hipLaunchKernelGGL((k_epilogue_pst<R_, LB_, float>), grid, block, 0, st, d, n_tb);                  \
  }
  OA_PST_DISPATCH(OA_EQ)
#undef OA_EQ
}
#undef OA_PST_DISPATCH

}  // namespace oa
