// Trait-side kernels: missing-value scan, trait statistics, pattern
// factorisations.  They replace, per trait block:
//   replace_nans on Y + Ymiss ........ src/Algorithm.cpp:558-572, src/Utility.cpp:130-179
//   sumSquares (SST) ................. src/Algorithm.cpp:1222-1251
//   Ay_top = XL'Y .................... src/Algorithm.cpp:583-585
//   generate_correction_missings_Stl . src/Algorithm.cpp:1324-1362
// The reference rebuilds the per-trait XL'XL correction for every trait;
// here traits with the same missing rows share one pattern, and each
// pattern's corrected covariate block Stl - C_ll is factored once.
#include <algorithm>

#include "oagpu_kernels.h"

namespace oa {

// One block of 256 threads per trait.  Writes the trait's column of the
// contraction operand B (NaN -> 0, rows n..n_pad-1 zero), the sorted list
// of its missing rows (rows_out + y*n), and its statistics.
template <int OP>
__global__ __launch_bounds__(256) void k_ystats(DevScan d, void* Bhi, void* Blo, int* rows_out,
                                                int* ymiss, int* yzero, double* yy, double* sst,
                                                double* rt, unsigned long long* hash) {
  const int y = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const size_t ybase = (size_t)y * d.ldy;
  const size_t bbase = (size_t)y * d.ldop;
  __shared__ int wcnt[4];
  __shared__ double shd[4];
  __shared__ long long shl[4];

  double sum = 0.0, sq = 0.0;
  long long nzero = 0;
  unsigned long long h1 = 0, h2 = 0;
  int written = 0;
  for (int k0 = 0; k0 < d.n_pad; k0 += 256) {
    const int k = k0 + tid;
    double v = 0.0;
    bool miss = false;
    if (k < d.n && (!d.keep || d.keep[k])) {   // rows outside the subset stay 0
      v = load_elem(d.Y, d.ydt, ybase + k);
      miss = (v != v);
      if (miss) v = 0.0;
      else if (v == 0.0) ++nzero;
    }
    if (k < d.n_pad) {
      if (OP == PREC_FP64) {
        reinterpret_cast<double*>(Bhi)[bbase + k] = v;
      } else if (OP == PREC_FP32) {
        reinterpret_cast<float*>(Bhi)[bbase + k] = (float)v;
      } else {
        uint16_t hi, lo;
        split_bf16((float)v, hi, lo);
        reinterpret_cast<uint16_t*>(Bhi)[bbase + k] = hi;
        reinterpret_cast<uint16_t*>(Blo)[bbase + k] = lo;
      }
    }
    sum += v;
    sq += v * v;
    if (miss) {
      h1 += mix64a((uint64_t)k);
      h2 += mix64b((uint64_t)k);
    }
    // Ordered compaction of the missing rows (ascending, like the sorted
    // std::list of replace_nans).
    const unsigned long long ballot = __ballot(miss);
    const int pre = __popcll(ballot & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[w] = __popcll(ballot);
    __syncthreads();
    int off = written;
    for (int i = 0; i < w; ++i) off += wcnt[i];
    if (miss) rows_out[(size_t)y * d.n + off + pre] = k;
    written += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
  }
  sum = block_sum256(sum, shd);
  sq = block_sum256(sq, shd);
  nzero = block_sum256_ll(nzero, shl);
  const long long H1 = block_sum256_ll((long long)h1, shl);
  const long long H2 = block_sum256_ll((long long)h2, shl);
  if (tid == 0) {
    ymiss[y] = written;
    yzero[y] = (int)nzero;
    yy[y] = sq;
    // sumSquares: sxx - sx^2 / (rows - nans)
    sst[y] = sq - (sum * sum) / (double)(d.n_stat - written);
    hash[2 * y] = (unsigned long long)H1;
    hash[2 * y + 1] = (unsigned long long)H2;
  }
  // XL' y0 (Ay_top).
  for (int j = 0; j < d.l; ++j) {
    double acc = 0.0;
    for (int k = tid; k < d.n; k += 256) {
      if (d.keep && !d.keep[k]) continue;
      double v = load_elem(d.Y, d.ydt, ybase + k);
      if (v == v) acc += d.XL_cm[(size_t)j * d.n_pad + k] * v;
    }
    acc = block_sum256(acc, shd);
    if (tid == 0) rt[(size_t)j * d.ldt + y] = acc;
  }
}

// Checks that every trait's missing rows equal its pattern representative's
// (the grouping key is a 128-bit hash plus the count).
__global__ void k_verify_patterns(int t, int n, const int* rows, const int* rep_of, const int* ymiss,
                                  int* mismatch) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= t) return;
  const int rep = rep_of[y];
  int bad = 0;
  if (rep != y) {
    const int cnt = ymiss[y];
    if (ymiss[rep] != cnt) bad = 1;
    for (int i = 0; i < cnt && !bad; ++i)
      if (rows[(size_t)y * n + i] != rows[(size_t)rep * n + i]) bad = 1;
  }
  mismatch[y] = bad;
}

// Union of the patterns' missing rows: mark[k] = 1 for every row some
// pattern misses (every writer stores the same value).
__global__ void k_union_mark(int n, const int* rows, const int* pat_rep, const int* pat_cnt,
                             unsigned char* mark) {
  const int g = blockIdx.x;
  const int cnt = pat_cnt[g];
  const int* r = rows + (size_t)pat_rep[g] * n;
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) mark[r[i]] = 1;
}

// Pattern g's missing rows as positions in the union, at pu_list + off[g].
__global__ void k_union_list(int n, const int* rows, const int* pat_rep, const int* pat_cnt,
                             const int* off, const int* uidx, int* pu_list) {
  const int g = blockIdx.x;
  const int cnt = pat_cnt[g];
  const int* r = rows + (size_t)pat_rep[g] * n;
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) pu_list[off[g] + i] = uidx[r[i]];
}

// One block per pattern: C_ll = sum over the pattern's missing rows of
// xl xl', A = Stl - C_ll, then the upper Cholesky factor of A with LAPACK's
// failure rule (a pivot <= 0 or NaN).
__global__ __launch_bounds__(256) void k_pattern(DevScan d, const double* Stl_pk, int G,
                                                 const int* pat_rep, double* UA, double* UA_invd,
                                                 int* pat_pd) {
  const int g = blockIdx.x;
  const int l = d.l;
  const int ne = l * (l + 1) / 2;
  extern __shared__ double A[];  // ne doubles
  const int rep = pat_rep[g];
  const int cnt = d.pat_cnt[g];
  const int* rows = d.rows + (size_t)rep * d.n;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    // packed (i, j), i <= j
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= e) ++j;
    const int i = e - j * (j + 1) / 2;
    double acc = 0.0;
    for (int q = 0; q < cnt; ++q) {
      const int k = rows[q];
      acc += d.XL_rm[(size_t)k * l + i] * d.XL_rm[(size_t)k * l + j];
    }
    A[e] = Stl_pk[e] - acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int pd = 1;
    for (int j = 0; j < l && pd; ++j) {
      double ajj = A[upk(j, j)];
      for (int k = 0; k < j; ++k) ajj -= A[upk(k, j)] * A[upk(k, j)];
      if (!(ajj > 0.0)) { pd = 0; break; }
      ajj = sqrt(ajj);
      A[upk(j, j)] = ajj;
      for (int i = j + 1; i < l; ++i) {
        double aji = A[upk(j, i)];
        for (int k = 0; k < j; ++k) aji -= A[upk(k, j)] * A[upk(k, i)];
        A[upk(j, i)] = aji / ajj;
      }
    }
    pat_pd[g] = pd;
    for (int e = 0; e < ne; ++e) UA[(size_t)e * G + g] = pd ? A[e] : 0.0;
    for (int j = 0; j < l; ++j) UA_invd[(size_t)j * G + g] = pd ? 1.0 / A[upk(j, j)] : 0.0;
  }
}

// One thread per trait: zt = U_A^-T XL'y0 and RSS0 = y0'y0 - |zt|^2 (the
// covariate-only residual, from which every pair's RSS is a downdate).
template <int LB>
__global__ void k_trait_solve(DevScan d, double* zt, double* ay, double* rss0) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= d.t) return;
  const int g = d.ypat[y];
  const int l = d.l, G = d.G;
  double z[LB];
  double ss = 0.0;
#pragma unroll
  for (int j = 0; j < LB; ++j) {
    z[j] = 0.0;
    if (j < l) {
      double v = d.rt[(size_t)j * d.ldt + y];
#pragma unroll
      for (int i = 0; i < j; ++i) v -= d.UA[(size_t)upk(i, j) * G + g] * z[i];
      z[j] = v * d.UA_invd[(size_t)j * G + g];
      zt[(size_t)j * d.ldt + y] = z[j];
      ss += z[j] * z[j];
    }
  }
  rss0[y] = d.yy[y] - ss;
  // a = U_A^-1 z: the covariate-only coefficients
#pragma unroll
  for (int j = LB - 1; j >= 0; --j) {
    if (j < l) {
      double v = z[j];
#pragma unroll
      for (int i = j + 1; i < LB; ++i)
        if (i < l) v -= d.UA[(size_t)upk(j, i) * G + g] * z[i];
      z[j] = v * d.UA_invd[(size_t)j * G + g];
      ay[(size_t)j * d.ldt + y] = z[j];
    }
  }
}

// ---- host launchers ----
void launch_ystats(const DevScan& d, int prec, void* Bhi, void* Blo, int* rows_out, int* ymiss,
                   int* yzero, double* yy, double* sst, double* rt, unsigned long long* hash,
                   hipStream_t st) {
  dim3 grid(d.t), block(256);
  if (prec == PREC_FP64)
    hipLaunchKernelGGL(k_ystats<PREC_FP64>, grid, block, 0, st, d, Bhi, Blo, rows_out, ymiss, yzero, yy, sst, rt, hash);
  else if (prec == PREC_FP32)
    hipLaunchKernelGGL(k_ystats<PREC_FP32>, grid, block, 0, st, d, Bhi, Blo, rows_out, ymiss, yzero, yy, sst, rt, hash);
  else
    hipLaunchKernelGGL(k_ystats<PREC_BF16X3>, grid, block, 0, st, d, Bhi, Blo, rows_out, ymiss, yzero, yy, sst, rt, hash);
}

void launch_verify_patterns(int t, int n, const int* rows, const int* rep_of, const int* ymiss,
                            int* mismatch, hipStream_t st) {
  hipLaunchKernelGGL(k_verify_patterns, dim3((t + 255) / 256), dim3(256), 0, st, t, n, rows, rep_of,
                     ymiss, mismatch);
}

void launch_union_mark(int n, int G, const int* rows, const int* pat_rep, const int* pat_cnt,
                       unsigned char* mark, hipStream_t st) {
  hipLaunchKernelGGL(k_union_mark, dim3(G), dim3(256), 0, st, n, rows, pat_rep, pat_cnt, mark);
}

void launch_union_list(int n, int G, const int* rows, const int* pat_rep, const int* pat_cnt, const int* off,
                       const int* uidx, int* pu_list, hipStream_t st) {
  hipLaunchKernelGGL(k_union_list, dim3(G), dim3(256), 0, st, n, rows, pat_rep, pat_cnt, off, uidx, pu_list);
}

// XLp[i][j] = XLu[pu_list[i]][j]: the covariate rows of every pattern's
// missing rows in the patterns' own order, so the corrections read a
// pattern's rows as one contiguous run instead of gathering them by position.
__global__ __launch_bounds__(256) void k_pattern_rows(const int* pu_list, long long tot, int l, const double* XLu,
                                                      double* XLp) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < tot * l; q += (long long)gridDim.x * 256) {
    const long long i = q / l;
    const int j = (int)(q - i * l);
    XLp[q] = XLu[(size_t)pu_list[i] * l + j];
  }
}

void launch_pattern_rows(const int* pu_list, long long tot, int l, const double* XLu, double* XLp, hipStream_t st) {
  if (tot <= 0) return;
  const long long n = tot * l;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_pattern_rows, dim3(blocks), dim3(256), 0, st, pu_list, tot, l, XLu, XLp);
}

void launch_pattern(const DevScan& d, const double* Stl_pk, int G, const int* pat_rep, double* UA,
                    double* UA_invd, int* pat_pd, hipStream_t st) {
  const size_t sh = sizeof(double) * (size_t)(d.l * (d.l + 1) / 2);
  hipLaunchKernelGGL(k_pattern, dim3(G), dim3(256), sh, st, d, Stl_pk, G, pat_rep, UA, UA_invd, pat_pd);
}

void launch_trait_solve(const DevScan& d, int lb, double* zt, double* ay, double* rss0, hipStream_t st) {
  dim3 grid((d.t + 127) / 128), block(128);
  switch (lb) {
    case 8: hipLaunchKernelGGL(k_trait_solve<8>, grid, block, 0, st, d, zt, ay, rss0); break;
    case 16: hipLaunchKernelGGL(k_trait_solve<16>, grid, block, 0, st, d, zt, ay, rss0); break;
    case 24: hipLaunchKernelGGL(k_trait_solve<24>, grid, block, 0, st, d, zt, ay, rss0); break;
    default: hipLaunchKernelGGL(k_trait_solve<32>, grid, block, 0, st, d, zt, ay, rss0); break;
  }
}

}  // namespace oa
