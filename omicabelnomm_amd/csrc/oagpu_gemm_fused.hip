// Split-BF16 contraction fed straight from the genotype columns, for the
// reference's interaction layout with the SNP kept: each SNP gives two
// columns [E*G, G] from one genotype column G and one interaction column E
// (src/AIOwrapper.cpp:808-838).
//
// It computes C = A [Y0 | XL]' like the split-BF16 contraction of
// oagpu_gemm.hip -- the reference's
//   Ay_bot = XR'Y  (cblas_sgemm, src/Algorithm.cpp:705-707)
//   Str    = XL'XR (cblas_sgemm, src/Algorithm.cpp:633-635)
// -- but forms the operand itself: a genotype column is read once (float32
// or float16), a missing value replaced by its column's mean
// (replace_nans_avgs, src/Utility.cpp:103; the means come from the operand
// pass run before it, which then writes no operand), multiplied by E for the
// interaction row, and both rows are split in registers into hi = bf16(x),
// lo = bf16(x - hi).  No operand matrix is written to HBM.
//
// Tile: 256 A rows (128 SNPs) x 256 B rows, 8 waves in a 2 x 4 grid, each
// wave 128 x 64 (8 x 4 blocks of v_mfma_f32_16x16x32_bf16), k-steps of 32;
// per (a, b) block the three products accumulate small terms first (lo*hi,
// hi*lo, hi*hi) in FP32.  Lane (i, g) = (lane & 15, lane >> 4) supplies
// k = 8g..8g+7 of row i for both operands.
//
// LDS: two stages.  B arrives in its own copy (launch_fused_b_layout) where
// a row's hi and lo values of one k-step are one 128-byte run, so each
// k-step fetches whole cache lines (the interleaved [lo | hi] rows fetched a
// half-used line per half: 1.8 % slower, and 13 % more L2 misses); its
// stage rows are those 128 bytes with the 16-byte chunks XOR-swizzled by
// row & 7.  A's hi and lo halves are separate arrays of 64-byte rows, the
// tile's E*G rows in rows 0..127 and its G rows in 128..255 with their own
// swizzle (swa), so a wave's conversion stores fill consecutive rows; no
// LDS access of the kernel has a bank conflict (SQ_LDS_BANK_CONFLICT reads
// 0 per n = 10,000 launch; a padded 80-byte row measured 1.33x slower).
// Per step s:
//   1. convert the genotypes of step s+1 (loaded two steps earlier) into the
//      other stage;
//   2. DMA B's k-slice of step s+1 into the other stage (global_load_lds,
//      16 bytes per lane, no registers);
//   3. load the genotypes and E of step s+3;
//   4. the MFMAs of step s;
//   5. wait for this wave's DMA and LDS writes (the 3 newer loads may stay in
//      flight), barrier.
// Converting before the DMA issue matters: the compiler does not see the
// DMA (issued from asm, so it adds no conservative wait before LDS reads),
// and its wait for the genotype registers then covers only older DMAs.
// (Running 4 before 1-3 in half of the waves, so that the two waves of a
// SIMD convert and multiply at different times, measured 3.5 % slower.)
// A tile whose columns hold no missing genotype runs a loop without the
// mean substitution (4 of the ~11 VALU operations per genotype).
// Genotype runs past n re-read the column's last run: B is zero there.
//
// Workgroups map to tiles XCD-major with the B tiles of one A row block
// consecutive, so an XCD works on a few row blocks x all column blocks at a
// time and each genotype column and B k-slice is fetched into its L2 once.
#include <algorithm>
#include <type_traits>

#include "oagpu_kernels.h"

namespace oa {

namespace {

constexpr int FT = 256, FK = 32;
typedef __attribute__((ext_vector_type(4))) unsigned int u4;
typedef __attribute__((ext_vector_type(2))) float f2v;
typedef __attribute__((ext_vector_type(2))) __bf16 b2v;

__device__ __forceinline__ bf16x8 as_b8(u4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ int xcd_tile_f(int id, int total) {
  const int q = total >> 3, rem = total & 7;
  const int xcd = id & 7, loc = id >> 3;
  return xcd * q + (xcd < rem ? xcd : rem) + loc;
}

// The A stages hold the E*G rows of the tile's 128 SNPs in rows 0..127 and
// the G rows in 128..255 (MFMA row 2s + t of the tile is stage row 128 t +
// s), so the conversion's 8-byte stores of a wave fill consecutive rows.
// Chunk c of stage row R sits at c xor (2 ((R >> 2) & 1) + (R >> 7)).
// gfx950 serves a 16-byte LDS read in groups of 16 lanes, 256 bytes over
// the 64 banks per cycle: within each half wave, the lanes whose bits 2..4
// have even parity, then those with odd parity (measured: two active lanes
// reading one bank conflict exactly when they share a group).  For an
// operand read, lane (i, g) reads chunk g of row R(i), and a group holds
// i in {0-3, 12-15} with one g and i in {4-11} with the other: rows R and
// R ^ 4 meet in one group with the same R & 3, so bit 2 of R flips the
// chunk and the 16 lanes hit 16 different 16-byte bank groups.
__device__ __forceinline__ int swa(int R, int c) {
  return R * 32 + ((c ^ ((((R >> 2) & 1) << 1) | ((R >> 7) & 1))) << 3);
}

// 16 bytes per lane from global memory into LDS at lds + 16 * lane.  From
// asm: no output registers; the caller waits (vmcnt) before publishing.
__device__ __forceinline__ void dma16(const void* g, void* lds) {
  const uint32_t a =
      __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(__attribute__((address_space(3))) char*)lds);
  __asm__ volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(a) : "memory", "m0");
}

// Four genotypes as floats (float32: one 16-byte load; float16: 8 bytes).
template <typename GT> struct Run4;
template <> struct Run4<float> {
  typedef float4 T;
  static __device__ __forceinline__ T load(const float* p) { return *reinterpret_cast<const float4*>(p); }
  static __device__ __forceinline__ void get(const T& v, float* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
};
template <> struct Run4<_Float16> {
  typedef uint2 T;
  static __device__ __forceinline__ T load(const _Float16* p) { return *reinterpret_cast<const uint2*>(p); }
  static __device__ __forceinline__ void get(const T& v, float* o) {
    _Float16 h[4];
    __builtin_memcpy(h, &v, sizeof(h));
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = (float)h[e];
  }
};

}  // namespace

template <typename GT>
__global__ __launch_bounds__(512, 1) void k_gemm_bf16x3_fused(FusedGemmArgs a) {
  typedef typename Run4<GT>::T RV;
  __shared__ __attribute__((aligned(16))) uint16_t sA0[2][FT * 32], sA1[2][FT * 32];   // [hi, lo]
  // B stage: 128-byte rows [hi | lo] of the k-step (fused_b_layout)
  __shared__ __attribute__((aligned(16))) uint16_t sB0[FT * 64], sB1[FT * 64];
  auto SA = [&](auto ST) -> uint16_t(*)[FT * 32] {
    if constexpr (decltype(ST)::value == 0) return sA0; else return sA1;
  };
  auto SB = [&](auto ST) -> uint16_t* {
    if constexpr (decltype(ST)::value == 0) return sB0; else return sB1;
  };
  const int tile = xcd_tile_f(blockIdx.x, a.total);
  const int tm = tile / a.tiles_n, tn = tile - tm * a.tiles_n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, i = lane & 15, g = lane >> 4;
  const int wr = (w & 1) * 128, wc = (w >> 1) * 64;
  // producer map: SNPs col0 and col0 + 64 of the tile (rows 2 col, 2 col + 1),
  // genotype run 4 gc .. 4 gc + 3 of the step
  const int gc = tid & 7, col0 = tid >> 3;
  const int ms = a.M >> 1;
  const GT* gcol[2];
  float mE[2], mG[2];
  int tnan = 0;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int s = min(tm * 128 + col0 + 64 * j, ms - 1);
    gcol[j] = static_cast<const GT*>(a.G) + (size_t)a.col_src[(size_t)(a.s0 + s) * 2] * a.ldg;
    mE[j] = (float)a.col_mean[2 * s];
    mG[j] = (float)a.col_mean[2 * s + 1];
    tnan |= a.col_nan[2 * s + 1];
  }
  // missing genotypes in the tile's columns: the conversion substitutes
  // means (workgroup-uniform choice of the loop below)
  const bool anynan = __syncthreads_or(tnan) != 0;
  // B DMA: wave w fills rows 32w + 8q .. +7 of the stage; lane -> row
  // + lane / 8, physical 16-byte chunk lane & 7 (chunks 0..3 hi, 4..7 lo,
  // chunk c of row r stored at c xor (r & 7))
  int bsrc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = 32 * w + 8 * q + (lane >> 3);
    const int logical = (lane & 7) ^ (r & 7);
    bsrc[q] = (min(tn * FT + r, a.N - 1) - tn * FT) * (int)a.ldb + 8 * logical;
  }
  const uint16_t* bbase = a.B + (size_t)(tn * FT) * a.ldb;
  const int nsteps = a.K / FK;
  struct GE { RV g0, g1; float4 e; };
  auto geload = [&](int step, GE& r) {
    const int kk = min(step, nsteps - 1) * FK + 4 * gc;
    const int kc = min(kk, a.n - 4);   // past n: the column's last run (B is zero there)
    r.g0 = Run4<GT>::load(gcol[0] + kc);
    r.g1 = Run4<GT>::load(gcol[1] + kc);
    r.e = *reinterpret_cast<const float4*>(a.E + kk);
  };
  auto bdma = [&](auto ST, int step) {
    const int k0 = min(step, nsteps - 1) * 2 * FK;   // [hi | lo] of each k-step
#pragma unroll
    for (int q = 0; q < 4; ++q) dma16(bbase + bsrc[q] + k0, SB(ST) + (32 * w + 8 * q) * 64);
  };
  auto conv = [&](auto NANP, auto ST, const GE& P) {
    const float ee[4] = {P.e.x, P.e.y, P.e.z, P.e.w};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      float gg[4];
      Run4<GT>::get(j ? P.g1 : P.g0, gg);
      float xe[4], xg[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float ve = gg[e] * ee[e];
        if constexpr (decltype(NANP)::value) {
          xe[e] = (ve != ve) ? mE[j] : ve;
          xg[e] = (gg[e] != gg[e]) ? mG[j] : gg[e];
        } else {
          xe[e] = ve;
          xg[e] = gg[e];
        }
      }
      unsigned h[4], l[4];
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const f2v v = {xe[2 * p], xe[2 * p + 1]};
        const b2v hv = __builtin_convertvector(v, b2v);
        const b2v lv = __builtin_convertvector(v - __builtin_convertvector(hv, f2v), b2v);
        h[p] = __builtin_bit_cast(unsigned, hv);
        l[p] = __builtin_bit_cast(unsigned, lv);
        const f2v u = {xg[2 * p], xg[2 * p + 1]};
        const b2v hu = __builtin_convertvector(u, b2v);
        const b2v lu = __builtin_convertvector(u - __builtin_convertvector(hu, f2v), b2v);
        h[2 + p] = __builtin_bit_cast(unsigned, hu);
        l[2 + p] = __builtin_bit_cast(unsigned, lu);
      }
      const int snp = col0 + 64 * j;
      const int o0 = swa(snp, gc >> 1) + 4 * (gc & 1), o1 = swa(128 + snp, gc >> 1) + 4 * (gc & 1);
      *reinterpret_cast<uint2*>(&SA(ST)[0][o0]) = make_uint2(h[0], h[1]);
      *reinterpret_cast<uint2*>(&SA(ST)[1][o0]) = make_uint2(l[0], l[1]);
      *reinterpret_cast<uint2*>(&SA(ST)[0][o1]) = make_uint2(h[2], h[3]);
      *reinterpret_cast<uint2*>(&SA(ST)[1][o1]) = make_uint2(l[2], l[3]);
    }
  };
  f4 acc[8][4];
#pragma unroll
  for (int x = 0; x < 8; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};
  auto consume = [&](auto CUR) {
    u4 bh[4], bl[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int r = wc + 16 * b + i;
      bh[b] = *reinterpret_cast<const u4*>(SB(CUR) + r * 64 + ((g ^ (r & 7)) << 3));
      bl[b] = *reinterpret_cast<const u4*>(SB(CUR) + r * 64 + (((4 + g) ^ (r & 7)) << 3));
    }
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      const int L = wr + 16 * x + i;   // MFMA row of the tile: SNP L / 2, row type L & 1
      const int o = swa(((L & 1) << 7) + (L >> 1), g);
      const u4 ah = *reinterpret_cast<const u4*>(&SA(CUR)[0][o]);
      const u4 al = *reinterpret_cast<const u4*>(&SA(CUR)[1][o]);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[x][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_b8(al), as_b8(bh[b]), acc[x][b], 0, 0, 0);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[x][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_b8(ah), as_b8(bl[b]), acc[x][b], 0, 0, 0);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[x][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_b8(ah), as_b8(bh[b]), acc[x][b], 0, 0, 0);
    }
  };
  auto run = [&](auto NANP) {
    std::integral_constant<int, 0> S0;
    std::integral_constant<int, 1> S1;
    GE P, Q;
    {
      GE p0;
      geload(0, p0);
      conv(NANP, S0, p0);
      bdma(S0, 0);
    }
    geload(1, P);
    geload(2, Q);
    // the step-0 DMA and LDS writes done; the 6 loads of P and Q may fly
    __asm__ volatile("s_waitcnt vmcnt(6) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    auto step = [&](auto CUR, int s, GE& Pc) {
      constexpr int cur = decltype(CUR)::value;
      std::integral_constant<int, cur ^ 1> NXT;
      conv(NANP, NXT, Pc);    // genotypes of step s+1
      bdma(NXT, s + 1);       // B of step s+1 (the other stage is free since the last barrier)
      geload(s + 3, Pc);
      consume(CUR);
      // this wave's DMA and LDS writes done; its 3 newer loads may fly
      __asm__ volatile("s_waitcnt vmcnt(3) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };
    int s = 0;
    for (; s + 2 <= nsteps; s += 2) {
      step(S0, s, P);
      step(S1, s + 1, Q);
    }
    if (s < nsteps) step(S0, s, P);
  };
  // two copies of the loop, each branch-free inside
  if (anynan) run(std::true_type{});
  else run(std::false_type{});
  // lane (i, g), value v of block (x, b): A row 16x + 4g + v, B row 16b + i
#pragma unroll
  for (int x = 0; x < 8; ++x)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = tn * FT + wc + 16 * b + i;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tm * FT + wr + 16 * x + 4 * g + v;
        if (row < a.M && col < a.N) a.C[(size_t)row * a.ldc + col] = acc[x][b][v];
      }
    }
}

// The fused contraction's copy of B: per row, each k-step's 32 hi values then
// its 32 lo values (one 128-byte run), from rows interleaved [lo | hi].
__global__ __launch_bounds__(256) void k_fused_b_layout(const uint16_t* B, long long ldb, int K, int rows,
                                                        uint16_t* Bf) {
  const long long chunks = (long long)rows * (K / 4);   // 16-byte output chunks (2K values per row)
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < chunks; q += (long long)gridDim.x * 256) {
    const int j = (int)(q / (K / 4)), o = (int)(q - (long long)j * (K / 4));   // output chunk o of row j
    const int st = o >> 3, c = o & 7;                                          // k-step, chunk within it
    const uint16_t* src = B + (size_t)j * ldb + (c < 4 ? K : 0) + st * FK + (c & 3) * 8;
    *reinterpret_cast<u4*>(Bf + (size_t)j * ldb + (size_t)o * 8) = *reinterpret_cast<const u4*>(src);
  }
}

void launch_fused_b_layout(const uint16_t* B, long long ldb, int K, int rows, uint16_t* Bf, hipStream_t st) {
  const long long chunks = (long long)rows * (K / 4);
  const int blocks = (int)std::min<long long>((chunks + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fused_b_layout, dim3(blocks), dim3(256), 0, st, B, ldb, K, rows, Bf);
}

bool gemm_fused_supported(int gdt, const void* G, long long ldg, int n, int K) {
  const size_t esz = gdt == DT_F32 ? 4 : 2;
  // float32 or float16, every 4-element run of every column on a 4-element
  // boundary, and every run wholly inside or wholly past n
  return (gdt == DT_F32 || gdt == DT_F16) && (uintptr_t)G % (4 * esz) == 0 && ldg % 4 == 0 && n >= 4 &&
         n % 4 == 0 && K % FK == 0;
}

void launch_gemm_fused(const FusedGemmArgs& args, int gdt, hipStream_t st) {
  FusedGemmArgs a = args;
  a.tiles_n = (a.N + FT - 1) / FT;
  a.total = ((a.M + FT - 1) / FT) * a.tiles_n;
  if (gdt == DT_F32) hipLaunchKernelGGL(k_gemm_bf16x3_fused<float>, dim3(a.total), dim3(512), 0, st, a);
  else hipLaunchKernelGGL(k_gemm_bf16x3_fused<_Float16>, dim3(a.total), dim3(512), 0, st, a);
}

}  // namespace oa
