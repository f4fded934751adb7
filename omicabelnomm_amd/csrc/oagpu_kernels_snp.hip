// SNP-side kernels of one chunk: the per-SNP Gram pieces Str, var_x
// (build_S + potrf/potri on the unmasked S_full, src/Algorithm.cpp:646-671),
// the explicit residual pass for the pairs the epilogues hand over (hpc_SSY,
// src/Algorithm.cpp:896-1024), the subset Sbr of the request server and the
// counter resets.  The other stages of the chunk live beside this file:
//   oagpu_kernels_prep.hip  replace_nans(_avgs) on XR, interaction columns,
//                           Sbr = XR_s'XR_s (src/Algorithm.cpp:618-622,
//                           src/Utility.cpp:103-127, src/AIOwrapper.cpp:767-840)
//   oagpu_kernels_corr.hip  generate_correction_missings_Str
//                           (src/Algorithm.cpp:1260-1315)
//   oagpu_kernels_pst.hip,  per-pair build_S/correct_missings_S/sposv, RSS and
//   oagpu_kernels_epi.hip   hpc_statistics (src/Algorithm.cpp:714-792,
//                           1034-1214, 1370-1399)
// Everything here is FP64.
#include <algorithm>

#include "oagpu_kernels.h"
#include "oagpu_geno.h"
#include "oagpu_pair.h"

namespace oa {

// ---------------------------------------------------------------------------
// One thread per SNP: Str from the contraction output, and the unmasked
// var_x = diag(S_full^-1) (src/Algorithm.cpp:659-671) through the block
// inverse with U0 = chol(Stl).
template <int R, int LB, typename CT>
__global__ void k_snp_gram(DevScan d) {
  chunk_start(d);
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

// Explicit residual pass for the pairs the epilogue handed over: one wave
// per pair.
template <int R, int LB>
__global__ __launch_bounds__(256) void k_fallback(DevScan d) {
  chunk_start(d);
  constexpr int P = LB + R;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long total = d.counters[CNT_FB_CHUNK];
  const int count = (int)(total < (unsigned long long)d.fb_cap ? total : (unsigned long long)d.fb_cap);
  for (int e = blockIdx.x * 4 + w; e < count; e += gridDim.x * 4) {
    const FallbackEntry* f = d.fb + e;
    const int y = f->y, s = f->s;
    double b[P];   // split layout from the entry's h order
#pragma unroll
    for (int h = 0; h < LB; ++h) b[h] = (h < d.l) ? f->b[h] : 0.0;
#pragma unroll
    for (int c = 0; c < R; ++c) b[LB + c] = f->b[d.l + c];
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
// words 32-bit zeros from p (4-byte aligned), grid-stride
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* p, size_t words) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256) p[i] = 0u;
}
__global__ void k_reset_counters(unsigned long long* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// Sbr = x'x over the subset's rows for r = 1 (a request that shares the
// chunk's operand but not its individuals); x is imputed with the chunk's
// column mean.  One block per SNP.
__global__ __launch_bounds__(256) void k_subset_sbr(DevScan d, double* out) {
  chunk_start(d);
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
void launch_snp_gram(const DevScan& d, int prec, int lb, hipStream_t st) {
  dim3 grid((d.mc + 63) / 64), block(64);
#define OA_SG(R_, LB_)                                                                     \
  if (prec == PREC_FP64) hipLaunchKernelGGL((k_snp_gram<R_, LB_, double>), grid, block, 0, st, d); \
  else hipLaunchKernelGGL((k_snp_gram<R_, LB_, float>), grid, block, 0, st, d);
  OA_DISPATCH_RLB(0, 0, OA_SG)
#undef OA_SG
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

void launch_zero(void* p, size_t bytes, hipStream_t st) {
  const size_t words = bytes / 4;
  if (words == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 1024);
  hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(256), 0, st, static_cast<uint32_t*>(p), words);
}

__global__ void k_set_int(int* p, int v) { *p = v; }

void launch_set_int(int* p, int v, hipStream_t st) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p, v); }

void launch_reset_counter(unsigned long long* c, int idx, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, c, idx);
}

}  // namespace oa
