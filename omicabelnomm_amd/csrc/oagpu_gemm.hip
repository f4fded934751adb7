// The bulk contraction of the scan, C = A' [Y0 | XL], on gfx950 MFMA.
// It replaces the reference's BLAS calls
//   Ay_bot = XR'Y  (cblas_sgemm, src/Algorithm.cpp:705-707)
//   Str    = XL'XR (cblas_sgemm, src/Algorithm.cpp:633-635)
// with one pass over the genotype operand: the covariate columns ride along
// as extra columns of B.
//
// Layout: A has M rows (genotype columns) of K contiguous values, B has N
// rows (traits, then covariates) of K contiguous values; both are padded to
// whole 128-row tiles and K is a multiple of 64 (zero rows).  C is row-major
// M x ldc.  One workgroup = 4 waves = a 128 x 128 tile, each wave 64 x 64
// (4 x 4 MFMA blocks of 16 x 16).  Each lane feeds the MFMAs straight from
// global memory: lane (i, g) = (lane & 15, lane >> 4) reads a contiguous
// run of K values of row i, so the 4 lane groups of a row read one full
// cache line.  The k-values a lane group supplies are chosen identically
// for A and B, so the contraction is exact whatever the hardware's internal
// k order.
//
// Workgroups are mapped to tiles XCD-major: the hardware hands consecutive
// workgroup ids to the 8 XCDs round-robin, and each XCD receives a
// contiguous range of tiles (row-tile major), so the tiles that share an A
// row block run on the same XCD and reuse it from that XCD's L2.
//
// launch_gemm times these kernels and the library GEMMs of
// oagpu_gemm_lib.hip on the first operands of each shape class and keeps the
// fastest; the choices persist in a cache file across processes.
#include <sys/stat.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "oagpu_kernels.h"
#include "oagpu_gemm_lib.h"

namespace oa {

__device__ __forceinline__ int xcd_tile(int id, int total) {
  const int q = total >> 3, rem = total & 7;
  const int xcd = id & 7, loc = id >> 3;
  return xcd * q + (xcd < rem ? xcd : rem) + loc;
}

// FP64: v_mfma_f64_16x16x4f64, fed from global memory.  4 waves in a 2 x 2
// grid, each wave AB x BB blocks of 16 x 16 (tile 32 AB x 32 BB); per k-step
// of KS, KS / 16 runs of 4 doubles per lane and block are loaded, lane
// group g supplying k = 4g + step of each 16.  The tuner times the tile
// shapes and k-steps of kVariants per shape class.
template <int AB, int BB, int KS>
__global__ __launch_bounds__(256, 2) void k_gemm_f64(const double* __restrict__ A,
                                                    const double* __restrict__ B,
                                                    double* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                    int ldc, int tiles_n, int total) {
  constexpr int TM = 32 * AB, TN = 32 * BB, KR = KS / 16;
  const int tile = xcd_tile(blockIdx.x, total);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int row0 = tm * TM + (w & 1) * (16 * AB);
  const int col0 = tn * TN + (w >> 1) * (16 * BB);
  d4 acc[AB][BB];
#pragma unroll
  for (int a = 0; a < AB; ++a)
#pragma unroll
    for (int b = 0; b < BB; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const double* Ap[AB];
  const double* Bp[BB];
#pragma unroll
  for (int a = 0; a < AB; ++a) Ap[a] = A + (size_t)(row0 + 16 * a + i) * lda + 4 * g;
#pragma unroll
  for (int b = 0; b < BB; ++b) Bp[b] = B + (size_t)(col0 + 16 * b + i) * ldb + 4 * g;
  for (int k0 = 0; k0 < K; k0 += KS) {
    d4 av[KR][AB], bv[KR][BB];
#pragma unroll
    for (int u = 0; u < KR; ++u) {
#pragma unroll
      for (int a = 0; a < AB; ++a) av[u][a] = *reinterpret_cast<const d4*>(Ap[a] + k0 + 16 * u);
#pragma unroll
      for (int b = 0; b < BB; ++b) bv[u][b] = *reinterpret_cast<const d4*>(Bp[b] + k0 + 16 * u);
    }
#pragma unroll
    for (int u = 0; u < KR; ++u)
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int a = 0; a < AB; ++a)
#pragma unroll
          for (int b = 0; b < BB; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][a][st], bv[u][b][st], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < AB; ++a)
#pragma unroll
    for (int b = 0; b < BB; ++b) {
      const int col = col0 + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // f64 16x16x4 output layout: lane (i, g), value v -> row 4v + g
        const int row = row0 + 16 * a + 4 * v + g;
        if (row < M && col < N) C[(size_t)row * ldc + col] = acc[a][b][v];
      }
    }
}

// FP32: v_mfma_f32_16x16x4f32, same structure with 4-float runs.
__global__ __launch_bounds__(256, 2) void k_gemm_f32(const float* __restrict__ A,
                                                    const float* __restrict__ B,
                                                    float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                    int ldc, int tiles_n, int total) {
  const int tile = xcd_tile(blockIdx.x, total);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int row0 = tm * 128 + (w & 1) * 64;
  const int col0 = tn * 128 + (w >> 1) * 64;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const float* Ap[4];
  const float* Bp[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    Ap[a] = A + (size_t)(row0 + 16 * a + i) * lda + 4 * g;
    Bp[a] = B + (size_t)(col0 + 16 * a + i) * ldb + 4 * g;
  }
  for (int k0 = 0; k0 < K; k0 += 16) {
    f4 av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      av[a] = *reinterpret_cast<const f4*>(Ap[a] + k0);
      bv[a] = *reinterpret_cast<const f4*>(Bp[a] + k0);
    }
#pragma unroll
    for (int st = 0; st < 4; ++st)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a][st], bv[b][st], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = col0 + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + 16 * a + 4 * g + v;
        if (row < M && col < N) C[(size_t)row * ldc + col] = acc[a][b][v];
      }
    }
}

// Split BF16, 3 terms: x*y ~= xh*yh + xh*yl + xl*yh with FP32 accumulation,
// v_mfma_f32_16x16x32_bf16.  K step 32; lane group g supplies k = 8g..8g+7.
typedef __attribute__((ext_vector_type(4))) unsigned int u4;

__device__ __forceinline__ bf16x8 as_bf16x8(u4 v) { return __builtin_bit_cast(bf16x8, v); }

__global__ __launch_bounds__(256, 2) void k_gemm_bf16x3(const uint16_t* __restrict__ Ah,
                                                       const uint16_t* __restrict__ Al,
                                                       const uint16_t* __restrict__ Bh,
                                                       const uint16_t* __restrict__ Bl,
                                                       float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                       int ldc, int tiles_n, int total) {
  const int tile = xcd_tile(blockIdx.x, total);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int row0 = tm * 128 + (w & 1) * 64;
  const int col0 = tn * 128 + (w >> 1) * 64;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  size_t ao[4], bo[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    ao[a] = (size_t)(row0 + 16 * a + i) * lda + 8 * g;
    bo[a] = (size_t)(col0 + 16 * a + i) * ldb + 8 * g;
  }
  for (int k0 = 0; k0 < K; k0 += 32) {
    u4 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      ah[a] = *reinterpret_cast<const u4*>(Ah + ao[a] + k0);
      al[a] = *reinterpret_cast<const u4*>(Al + ao[a] + k0);
      bh[a] = *reinterpret_cast<const u4*>(Bh + bo[a] + k0);
      bl[a] = *reinterpret_cast<const u4*>(Bl + bo[a] + k0);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(al[a]), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah[a]), as_bf16x8(bl[b]), acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah[a]), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
      }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = col0 + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + 16 * a + 4 * g + v;
        if (row < M && col < N) C[(size_t)row * ldc + col] = acc[a][b][v];
      }
    }
}


// Same contraction with LDS staging: the 128 x 128 workgroup tile's A and B
// slices (hi and lo) are loaded once per k-step into LDS (double
// buffered, rows padded to 80 bytes) and every wave reads its MFMA
// operands from there, so each global byte feeds two waves instead of one.
constexpr int LDS_ROW = 40;   // bf16 per LDS row: 32 k-values + 8 padding

__global__ __launch_bounds__(256, 1) void k_gemm_bf16x3_lds(const uint16_t* __restrict__ Ah,
                                                           const uint16_t* __restrict__ Al,
                                                           const uint16_t* __restrict__ Bh,
                                                           const uint16_t* __restrict__ Bl,
                                                           float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                           int ldc, int tiles_n, int total) {
  __shared__ __attribute__((aligned(16))) uint16_t sm[2][4][128 * LDS_ROW];   // [stage][Ah,Al,Bh,Bl]
  const int tile = xcd_tile(blockIdx.x, total);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;

  // global -> register staging: chunk c in {tid, tid + 256}: row c >> 2, 8 k-values at 8*(c & 3)
  size_t goff[2][2];   // [chunk][A/B]
  int soff[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = tid + q * 256, row = c >> 2, ch = c & 3;
    goff[q][0] = (size_t)(tm * 128 + row) * lda + 8 * ch;
    goff[q][1] = (size_t)(tn * 128 + row) * ldb + 8 * ch;
    soff[q] = row * LDS_ROW + 8 * ch;
  }
  u4 st[2][4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      st[q][0] = *reinterpret_cast<const u4*>(Ah + goff[q][0] + k0);
      st[q][1] = *reinterpret_cast<const u4*>(Al + goff[q][0] + k0);
      st[q][2] = *reinterpret_cast<const u4*>(Bh + goff[q][1] + k0);
      st[q][3] = *reinterpret_cast<const u4*>(Bl + goff[q][1] + k0);
    }
  };
  auto sstore = [&](int stage) {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int a = 0; a < 4; ++a) *reinterpret_cast<u4*>(&sm[stage][a][soff[q]]) = st[q][a];
  };

  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  gload(0);
  sstore(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bool more = k0 + 32 < K;
    if (more) gload(k0 + 32);
    u4 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int ra = (wr + 16 * a + i) * LDS_ROW + 8 * g;
      const int rb = (wc + 16 * a + i) * LDS_ROW + 8 * g;
      ah[a] = *reinterpret_cast<const u4*>(&sm[cur][0][ra]);
      al[a] = *reinterpret_cast<const u4*>(&sm[cur][1][ra]);
      bh[a] = *reinterpret_cast<const u4*>(&sm[cur][2][rb]);
      bl[a] = *reinterpret_cast<const u4*>(&sm[cur][3][rb]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(al[a]), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah[a]), as_bf16x8(bl[b]), acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah[a]), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
      }
    if (more) sstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = tn * 128 + wc + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tm * 128 + wr + 16 * a + 4 * g + v;
        if (row < M && col < N) C[(size_t)row * ldc + col] = acc[a][b][v];
      }
    }
}


// Larger tiles for the split-BF16 contraction: workgroup 256 x 128, wave
// 128 x 64 (8 x 4 MFMA blocks), so each LDS byte a wave reads feeds twice
// as many MFMAs as in the 128 x 128 kernel.  The three products are issued
// product-major, so consecutive MFMAs never share an accumulator.
template <int WM_BLK>
__global__ __launch_bounds__(256, 1) void k_gemm_bf16x3_big(const uint16_t* __restrict__ Ah,
                                                           const uint16_t* __restrict__ Al,
                                                           const uint16_t* __restrict__ Bh,
                                                           const uint16_t* __restrict__ Bl,
                                                           float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                           int ldc, int tiles_n, int total) {
  constexpr int BM = 32 * WM_BLK;          // rows of A per workgroup (2 waves x 16*WM_BLK)
  constexpr int BN = 128;
  constexpr int A_CH = BM * 4 / 256;       // 16-byte chunks per thread per array
  constexpr int B_CH = BN * 4 / 256;
  __shared__ __attribute__((aligned(16))) uint16_t sa[2][2][BM * LDS_ROW];
  __shared__ __attribute__((aligned(16))) uint16_t sb[2][2][BN * LDS_ROW];
  const int tile = xcd_tile(blockIdx.x, total);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int wr = (w & 1) * (16 * WM_BLK), wc = (w >> 1) * 64;

  size_t ga[A_CH], gb[B_CH];
  int sa_off[A_CH], sb_off[B_CH];
#pragma unroll
  for (int q = 0; q < A_CH; ++q) {
    const int c = tid + q * 256, row = c >> 2, ch = c & 3;
    ga[q] = (size_t)(tm * BM + row) * lda + 8 * ch;
    sa_off[q] = row * LDS_ROW + 8 * ch;
  }
#pragma unroll
  for (int q = 0; q < B_CH; ++q) {
    const int c = tid + q * 256, row = c >> 2, ch = c & 3;
    gb[q] = (size_t)(tn * BN + row) * ldb + 8 * ch;
    sb_off[q] = row * LDS_ROW + 8 * ch;
  }
  u4 sta[A_CH][2], stb[B_CH][2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < A_CH; ++q) {
      sta[q][0] = *reinterpret_cast<const u4*>(Ah + ga[q] + k0);
      sta[q][1] = *reinterpret_cast<const u4*>(Al + ga[q] + k0);
    }
#pragma unroll
    for (int q = 0; q < B_CH; ++q) {
      stb[q][0] = *reinterpret_cast<const u4*>(Bh + gb[q] + k0);
      stb[q][1] = *reinterpret_cast<const u4*>(Bl + gb[q] + k0);
    }
  };
  auto sstore = [&](int stage) {
#pragma unroll
    for (int q = 0; q < A_CH; ++q) {
      *reinterpret_cast<u4*>(&sa[stage][0][sa_off[q]]) = sta[q][0];
      *reinterpret_cast<u4*>(&sa[stage][1][sa_off[q]]) = sta[q][1];
    }
#pragma unroll
    for (int q = 0; q < B_CH; ++q) {
      *reinterpret_cast<u4*>(&sb[stage][0][sb_off[q]]) = stb[q][0];
      *reinterpret_cast<u4*>(&sb[stage][1][sb_off[q]]) = stb[q][1];
    }
  };

  f4 acc[WM_BLK][4];
#pragma unroll
  for (int a = 0; a < WM_BLK; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  gload(0);
  sstore(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bool more = k0 + 32 < K;
    if (more) gload(k0 + 32);
    u4 bh[4], bl[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int rb = (wc + 16 * b + i) * LDS_ROW + 8 * g;
      bh[b] = *reinterpret_cast<const u4*>(&sb[cur][0][rb]);
      bl[b] = *reinterpret_cast<const u4*>(&sb[cur][1][rb]);
    }
#pragma unroll
    for (int a = 0; a < WM_BLK; ++a) {
      const int ra = (wr + 16 * a + i) * LDS_ROW + 8 * g;
      const u4 ah = *reinterpret_cast<const u4*>(&sa[cur][0][ra]);
      const u4 al = *reinterpret_cast<const u4*>(&sa[cur][1][ra]);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(al), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah), as_bf16x8(bl[b]), acc[a][b], 0, 0, 0);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
    }
    if (more) sstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int a = 0; a < WM_BLK; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = tn * BN + wc + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tm * BM + wr + 16 * a + 4 * g + v;
        if (row < M && col < N) C[(size_t)row * ldc + col] = acc[a][b][v];
      }
    }
}


// 128 x 128 split-BF16 tiles with the global loads issued two k-steps
// ahead of their use (two register stages), so ~1500 cycles of MFMA work
// cover the memory latency instead of ~770.
__global__ __launch_bounds__(256, 2) void k_gemm_bf16x3_pf2(const uint16_t* __restrict__ Ah,
                                                           const uint16_t* __restrict__ Al,
                                                           const uint16_t* __restrict__ Bh,
                                                           const uint16_t* __restrict__ Bl,
                                                           float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                           int ldc, int tiles_n, int total) {
  __shared__ __attribute__((aligned(16))) uint16_t sm[2][4][128 * LDS_ROW];   // [stage][Ah,Al,Bh,Bl]
  const int tile = xcd_tile(blockIdx.x, total);
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
  size_t goff[2][2];
  int soff[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = tid + q * 256, row = c >> 2, ch = c & 3;
    goff[q][0] = (size_t)(tm * 128 + row) * lda + 8 * ch;
    goff[q][1] = (size_t)(tn * 128 + row) * ldb + 8 * ch;
    soff[q] = row * LDS_ROW + 8 * ch;
  }
  u4 ra[2][4], rb[2][4];   // two register stages
  auto gload = [&](u4 (&st)[2][4], int k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      st[q][0] = *reinterpret_cast<const u4*>(Ah + goff[q][0] + k0);
      st[q][1] = *reinterpret_cast<const u4*>(Al + goff[q][0] + k0);
      st[q][2] = *reinterpret_cast<const u4*>(Bh + goff[q][1] + k0);
      st[q][3] = *reinterpret_cast<const u4*>(Bl + goff[q][1] + k0);
    }
  };
  auto sstore = [&](const u4 (&st)[2][4], int stage) {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int a = 0; a < 4; ++a) *reinterpret_cast<u4*>(&sm[stage][a][soff[q]]) = st[q][a];
  };
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  gload(ra, 0);
  if (K > 32) gload(rb, 32);
  sstore(ra, 0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += 64) {
    // step k0 (stage cur, data of ra already in LDS); rb holds k0 + 32
    if (k0 + 64 < K) gload(ra, k0 + 64);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (half == 1 && k0 + 32 >= K) break;
      u4 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int rA = (wr + 16 * a + i) * LDS_ROW + 8 * g;
        const int rB = (wc + 16 * a + i) * LDS_ROW + 8 * g;
        ah[a] = *reinterpret_cast<const u4*>(&sm[cur][0][rA]);
        al[a] = *reinterpret_cast<const u4*>(&sm[cur][1][rA]);
        bh[a] = *reinterpret_cast<const u4*>(&sm[cur][2][rB]);
        bl[a] = *reinterpret_cast<const u4*>(&sm[cur][3][rB]);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(al[a]), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah[a]), as_bf16x8(bl[b]), acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ah[a]), as_bf16x8(bh[b]), acc[a][b], 0, 0, 0);
      // stage the next k-step: rb (k0+32) after the first half, ra (k0+64) after the second
      if (half == 0) {
        if (k0 + 32 < K) sstore(rb, cur ^ 1);
      } else {
        if (k0 + 64 < K) sstore(ra, cur ^ 1);
      }
      __syncthreads();
      cur ^= 1;
      if (half == 0 && k0 + 96 < K) gload(rb, k0 + 96);
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = tn * 128 + wc + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tm * 128 + wr + 16 * a + 4 * g + v;
        if (row < M && col < N) C[(size_t)row * ldc + col] = acc[a][b][v];
      }
    }
}

// ---------------------------------------------------------------------------
// Kernel variants and the autotuner.  Each precision has a list of
// variants; the first launch for a (precision, M, N, K) shape class times
// every variant on the real operands (HIP events on the caller's stream)
// and caches the fastest.  OAGPU_GEMM_VARIANT=<name> pins one.
struct GemmVariant {
  const char* name;
  int prec;
  int bm, bn;   // workgroup tile
  // timed by the autotuner (the others never won a measured shape and are
  // kept for OAGPU_GEMM_VARIANT and their parity tests)
  bool tuned;
  // false: the variant cannot run this shape and launched nothing
  bool (*launch)(const void*, const void*, const void*, const void*, void*, int, int, int, int, int, int,
                 hipStream_t);
};

template <int AB, int BB, int KS>
static bool v_f64(const void* Ah, const void*, const void* Bh, const void*, void* C, int M, int N, int K,
                  int lda, int ldb, int ldc, hipStream_t st) {
  constexpr int TM = 32 * AB, TN = 32 * BB;
  if (K % KS) return false;
  const int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN, total = tiles_m * tiles_n;
  hipLaunchKernelGGL((k_gemm_f64<AB, BB, KS>), dim3(total), dim3(256), 0, st, (const double*)Ah,
                     (const double*)Bh, (double*)C, M, N, K, lda, ldb, ldc, tiles_n, total);
  return true;
}
static bool v_f32(const void* Ah, const void*, const void* Bh, const void*, void* C, int M, int N, int K,
                  int lda, int ldb, int ldc, hipStream_t st) {
  const int tiles_m = (M + 127) / 128, tiles_n = (N + 127) / 128, total = tiles_m * tiles_n;
  hipLaunchKernelGGL(k_gemm_f32, dim3(total), dim3(256), 0, st, (const float*)Ah, (const float*)Bh,
                     (float*)C, M, N, K, lda, ldb, ldc, tiles_n, total);
  return true;
}
static bool v_bf_direct(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                        int K, int lda, int ldb, int ldc, hipStream_t st) {
  const int tiles_m = (M + 127) / 128, tiles_n = (N + 127) / 128, total = tiles_m * tiles_n;
  hipLaunchKernelGGL(k_gemm_bf16x3, dim3(total), dim3(256), 0, st, (const uint16_t*)Ah, (const uint16_t*)Al,
                     (const uint16_t*)Bh, (const uint16_t*)Bl, (float*)C, M, N, K, lda, ldb, ldc, tiles_n, total);
  return true;
}
static bool v_bf_lds(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                     int K, int lda, int ldb, int ldc, hipStream_t st) {
  const int tiles_m = (M + 127) / 128, tiles_n = (N + 127) / 128, total = tiles_m * tiles_n;
  hipLaunchKernelGGL(k_gemm_bf16x3_lds, dim3(total), dim3(256), 0, st, (const uint16_t*)Ah,
                     (const uint16_t*)Al, (const uint16_t*)Bh, (const uint16_t*)Bl, (float*)C, M, N, K, lda, ldb, ldc,
                     tiles_n, total);
  return true;
}
static bool v_bf_pf2(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                     int K, int lda, int ldb, int ldc, hipStream_t st) {
  const int tiles_m = (M + 127) / 128, tiles_n = (N + 127) / 128, total = tiles_m * tiles_n;
  hipLaunchKernelGGL(k_gemm_bf16x3_pf2, dim3(total), dim3(256), 0, st, (const uint16_t*)Ah,
                     (const uint16_t*)Al, (const uint16_t*)Bh, (const uint16_t*)Bl, (float*)C, M, N, K, lda, ldb, ldc,
                     tiles_n, total);
  return true;
}
static bool v_bf_big(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                     int K, int lda, int ldb, int ldc, hipStream_t st) {
  const int tiles_m = (M + 255) / 256, tiles_n = (N + 127) / 128, total = tiles_m * tiles_n;
  hipLaunchKernelGGL(k_gemm_bf16x3_big<8>, dim3(total), dim3(256), 0, st, (const uint16_t*)Ah,
                     (const uint16_t*)Al, (const uint16_t*)Bh, (const uint16_t*)Bl, (float*)C, M, N, K, lda, ldb, ldc,
                     tiles_n, total);
  return true;
}

static bool v_lib_bf16x3(const void* Ah, const void* Al, const void* Bh, const void* Bl, void* C, int M, int N,
                         int K, int lda, int ldb, int ldc, hipStream_t st) {
  return lib_gemm_bf16x3(Ah, Al, Bh, Bl, C, M, N, K, lda, ldb, ldc, st);
}
static bool v_lib_f32(const void* Ah, const void*, const void* Bh, const void*, void* C, int M, int N, int K,
                      int lda, int ldb, int ldc, hipStream_t st) {
  return lib_gemm_f32(Ah, Bh, C, M, N, K, lda, ldb, ldc, st);
}
static const GemmVariant kVariants[] = {
    {"f64_direct_128x128", PREC_FP64, 128, 128, true, v_f64<4, 4, 16>},
    {"f64_direct_128x128_k32", PREC_FP64, 128, 128, true, v_f64<4, 4, 32>},
    {"f64_direct_128x64", PREC_FP64, 128, 64, true, v_f64<4, 2, 16>},
    {"f64_direct_64x128", PREC_FP64, 64, 128, true, v_f64<2, 4, 16>},
    {"f64_direct_64x64_k32", PREC_FP64, 64, 64, true, v_f64<2, 2, 32>},
    {"f32_direct_128x128", PREC_FP32, 128, 128, true, v_f32},
    {"f32_hipblaslt", PREC_FP32, 0, 0, true, v_lib_f32},
    {"bf16x3_hipblaslt", PREC_BF16X3, 0, 0, true, v_lib_bf16x3},
    {"bf16x3_lds_pf2_128x128", PREC_BF16X3, 128, 128, false, v_bf_pf2},
    {"bf16x3_lds_256x128", PREC_BF16X3, 256, 128, false, v_bf_big},
    {"bf16x3_lds_128x128", PREC_BF16X3, 128, 128, true, v_bf_lds},
    {"bf16x3_direct_128x128", PREC_BF16X3, 128, 128, false, v_bf_direct},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

struct TuneKey {
  int prec, mclass, n, k;
  bool operator<(const TuneKey& o) const {
    return std::tie(prec, mclass, n, k) < std::tie(o.prec, o.mclass, o.n, o.k);
  }
};
static std::map<TuneKey, int> g_tuned;
static std::mutex g_tune_mu;
static int g_tunings = 0;   // shapes timed by this process

// The tuning cache persists across processes in a text file, one line per
// shape class: "prec mclass nclass K variant".  OAGPU_TUNE_CACHE names the
// file ("" switches persistence off); by default it is
// $HOME/.cache/omicabelnomm/gemm_tune_<arch>_<CUs>.txt, so a cache is only
// read on the kind of device that wrote it.
static std::string tune_cache_path() {
  const char* e = getenv("OAGPU_TUNE_CACHE");
  if (e) return e;
  const char* home = getenv("HOME");
  if (!home || !*home) return "";
  int dev = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "";
  std::string arch = prop.gcnArchName;
  for (char& ch : arch)
    if (!std::isalnum((unsigned char)ch)) ch = '_';
  return std::string(home) + "/.cache/omicabelnomm/gemm_tune_" + arch + "_" +
         std::to_string(prop.multiProcessorCount) + ".txt";
}

static bool g_cache_loaded = false;
static std::string g_cache_file;

// Reads the cache file once (under g_tune_mu); lines naming a variant this
// build does not have are skipped.
static void load_tune_cache() {
  if (g_cache_loaded) return;
  g_cache_loaded = true;
  g_cache_file = tune_cache_path();
  if (g_cache_file.empty()) return;
  FILE* f = std::fopen(g_cache_file.c_str(), "r");
  if (!f) return;
  int prec, mc, nc, k;
  char name[96];
  while (std::fscanf(f, "%d %d %d %d %95s", &prec, &mc, &nc, &k, name) == 5)
    for (int v = 0; v < kNumVariants; ++v)
      if (kVariants[v].prec == prec && kVariants[v].tuned && !std::strcmp(kVariants[v].name, name))
        g_tuned[TuneKey{prec, mc, nc, k}] = v;
  std::fclose(f);
}

static void save_tune_entry(const TuneKey& key, int v) {
  if (g_cache_file.empty() || v < 0) return;
  const size_t slash = g_cache_file.rfind('/');
  if (slash != std::string::npos && slash > 0) {
    // create the directory chain (errors surface as a failed fopen)
    for (size_t pos = 1; pos <= slash; ++pos)
      if (g_cache_file[pos] == '/' || pos == slash) (void)mkdir(g_cache_file.substr(0, pos).c_str(), 0755);
  }
  FILE* f = std::fopen(g_cache_file.c_str(), "a");
  if (!f) return;
  std::fprintf(f, "%d %d %d %d %s\n", key.prec, key.mclass, key.n, key.k, kVariants[v].name);
  std::fclose(f);
}

int gemm_tunings() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tunings;
}

static int pinned_variant(int prec) {
  const char* e = getenv("OAGPU_GEMM_VARIANT");
  if (!e) return -1;
  for (int v = 0; v < kNumVariants; ++v)
    if (kVariants[v].prec == prec && !strcmp(kVariants[v].name, e)) return v;
  return -1;
}

// Shape classes of the tuning cache: M in three bands, N rounded up to a
// power of two (the request server changes N whenever its batch changes).
static int m_class(int M) { return M >= 65536 ? 2 : (M >= 4096 ? 1 : 0); }
static int n_class(int N) {
  int c = 64;
  while (c < N) c *= 2;
  return c;
}

// The variant launch_gemm runs for this shape: the pinned one, the timed
// winner, or the only candidate of the precision.
const char* gemm_variant_name(int prec, int M, int N, int K) {
  const int pv = pinned_variant(prec);
  if (pv >= 0) return kVariants[pv].name;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tune_cache();
    auto it = g_tuned.find(TuneKey{prec, m_class(M), n_class(N), K});
    if (it != g_tuned.end()) return kVariants[it->second].name;
  }
  int nvar = 0, only = -1;
  for (int c = 0; c < kNumVariants; ++c)
    if (kVariants[c].prec == prec && kVariants[c].tuned) {
      ++nvar;
      only = c;
    }
  return nvar == 1 ? kVariants[only].name : "untuned";
}

int gemm_row_pad() { return 256; }
int gemm_fixed_row_pad() { return 128; }

void launch_gemm_fixed(int prec, const void* Ahi, const void* Alo, const void* Bhi, const void* Blo, void* C, int M,
                       int N, int K, int lda, int ldb, int ldc, hipStream_t st) {
  if (prec == PREC_FP64) v_f64<4, 4, 16>(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st);
  else if (prec == PREC_FP32) v_f32(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st);
  else v_bf_lds(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st);
}

void launch_gemm(int prec, const void* Ahi, const void* Alo, const void* Bhi, const void* Blo,
                 void* C, int M, int N, int K, int lda, int ldb, int ldc, hipStream_t st) {
  int v = pinned_variant(prec);
  if (v < 0) {
    const TuneKey key{prec, m_class(M), n_class(N), K};
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      load_tune_cache();
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) v = it->second;
    }
    int nvar = 0, only = -1;
    for (int c = 0; c < kNumVariants; ++c)
      if (kVariants[c].prec == prec && kVariants[c].tuned) {
        ++nvar;
        only = c;
      }
    if (v < 0 && nvar == 1) v = only;
    if (v < 0) {
      // time every variant of this precision on these operands; one that
      // cannot run the shape drops out
      float best = 1e30f;
      hipEvent_t e0, e1;
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      for (int c = 0; c < kNumVariants; ++c) {
        if (kVariants[c].prec != prec || !kVariants[c].tuned) continue;
        if (!kVariants[c].launch(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st)) continue;   // warm
        (void)hipEventRecord(e0, st);
        const bool ran = kVariants[c].launch(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ran && ms < best) {
          best = ms;
          v = c;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[key] = v;
      ++g_tunings;
      save_tune_entry(key, v);
    }
  }
  if (v >= 0 && kVariants[v].launch(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st)) return;
  // a pinned library variant that cannot run the shape: the first
  // hand-written kernel of this precision
  for (int c = 0; c < kNumVariants; ++c)
    if (kVariants[c].prec == prec && kVariants[c].bm > 0) {
      kVariants[c].launch(Ahi, Alo, Bhi, Blo, C, M, N, K, lda, ldb, ldc, st);
      return;
    }
}


}  // namespace oa
