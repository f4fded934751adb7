// Operand pass of the scan, one launch per chunk of SNPs: mean imputation of
// missing genotypes (replace_nans_avgs, src/Utility.cpp:103, called at
// src/Algorithm.cpp:622), interaction columns (src/AIOwrapper.cpp:808-838),
// the contraction operand A (row s*R+c at stride ldop; entries n..n_pad-1
// zero), Sbr = XR_s'XR_s (src/Algorithm.cpp:490) and -- when the patterns'
// missing rows are few -- the imputed values at the union of those rows
// (u-major, so the corrections kernel reads them coalesced).
#include "oagpu_kernels.h"
#include "oagpu_geno.h"
#include <type_traits>

namespace oa {

// SNPs per block: about 4 output columns, so a thread's 4-row runs of every
// column stay in registers at 5 waves per SIMD.
template <int R>
constexpr int prep_spw() { return R == 1 ? 4 : (R == 2 ? 2 : 1); }

template <int OP>
__device__ __forceinline__ void store_op4(void* Ahi, void* Alo, size_t o, const double* v) {
  if constexpr (OP == PREC_FP64) {
    d4 x = {v[0], v[1], v[2], v[3]};
    *reinterpret_cast<d4*>(reinterpret_cast<double*>(Ahi) + o) = x;
  } else if (OP == PREC_FP32) {
    f4 x = {(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
    *reinterpret_cast<f4*>(reinterpret_cast<float*>(Ahi) + o) = x;
  } else {
    uint16_t h[4], lo[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) split_bf16((float)v[e], h[e], lo[e]);
    uint2 hv = {(uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
    uint2 lv = {(uint32_t)lo[0] | ((uint32_t)lo[1] << 16), (uint32_t)lo[2] | ((uint32_t)lo[3] << 16)};
    *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(Ahi) + o) = hv;
    *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(Alo) + o) = lv;
  }
}

typedef __attribute__((ext_vector_type(4))) _Float16 h4;

// XT: element type of the compact rows of U (void: no compact rows written).
__device__ __forceinline__ void store_xu(double* p, double v) { *p = v; }
__device__ __forceinline__ void store_xu(float* p, double v) { *p = (float)v; }
__device__ __forceinline__ void store_xu(_Float16* p, double v) { *p = (_Float16)(float)v; }
__device__ __forceinline__ void store_xu4(double* p, const double* v) {
  *reinterpret_cast<d4*>(p) = d4{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void store_xu4(float* p, const double* v) {
  *reinterpret_cast<f4*>(p) = f4{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
}
__device__ __forceinline__ void store_xu4(_Float16* p, const double* v) {
  *reinterpret_cast<h4*>(p) = h4{(_Float16)(float)v[0], (_Float16)(float)v[1], (_Float16)(float)v[2], (_Float16)(float)v[3]};
}

// Four raw genotypes of element type GT at rows k0..k0+3 of one source
// column; rows >= n read as 0.  vec (wave-uniform): every run of the wave is
// inside the column and aligned, so the wave takes one path and its loads
// of several columns stay in flight together.
template <typename GT> struct GVec;
template <> struct GVec<float> { typedef f4 T; };
template <> struct GVec<double> { typedef d4 T; };
template <> struct GVec<_Float16> { typedef h4 T; };

template <typename GT>
__device__ __forceinline__ typename GVec<GT>::T load_g4(const DevScan& d, int src, int k0, bool vec) {
  typedef typename GVec<GT>::T V;
  const GT* p = static_cast<const GT*>(d.Gm) + (size_t)src * d.ldg + k0;
  V v;
  if (vec) {
    v = *reinterpret_cast<const V*>(p);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (k0 + e < d.n) ? p[e] : (GT)0;
  }
  return v;
}

// Each thread handles runs of 4 consecutive rows, so loads and operand
// stores are vectors.  A run's loads (genotypes of every column, environment
// rows, subset flags, row-union positions) are all issued before any of them
// is used, so a run costs one memory round trip: the genotype element type
// is a template parameter, which keeps every load free of branches.
template <int R, int OP, typename XT, typename GT>
__global__ __launch_bounds__(256) void k_prep(DevScan d, void* Ahi, void* Alo) {
  chunk_start(d);
  constexpr bool CORR = !std::is_void<XT>::value;
  constexpr int RR = R * (R + 1) / 2;
  constexpr int PREP_SPW = prep_spw<R>();
  constexpr int NV = PREP_SPW * R;
  typedef typename GVec<GT>::T GV;
  const int tid = threadIdx.x;
  // Consecutive SNP blocks run on one XCD (blocks are dealt to the 8 XCDs in
  // turn), so the 32-byte pieces that neighbouring blocks write into one
  // compact-row cache line meet in the same L2.
  const int per = (gridDim.x + 7) >> 3;
  const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int sb = blk * PREP_SPW;
  if (sb >= d.mc) return;   // block-uniform, before any barrier
  __shared__ double shd[4 * (2 * NV + PREP_SPW * RR)];
  int src[NV], env[NV];
  bool live[PREP_SPW];
#pragma unroll
  for (int q = 0; q < PREP_SPW; ++q) {
    live[q] = sb + q < d.mc;
    const int gs = d.s0 + (live[q] ? sb + q : sb);
    col_desc<R>(d, gs, src + q * R, env + q * R);
  }
  // One pass that assumes no missing genotype (the common case): values go
  // into the sums unchecked, so a column with a missing value ends with a
  // NaN sum.  Only then does the block count its missing values (sums over
  // the others) and repeat the pass with the means imputed, overwriting
  // everything the first pass wrote.
  double sum[NV], cnt[NV], mean[NV], acc[PREP_SPW * RR];
#pragma unroll
  for (int v = 0; v < NV; ++v) { sum[v] = 0.0; cnt[v] = 0.0; mean[v] = 0.0; }
  // KEEP: a subset is set (a separate copy of the loop, so that without a
  // subset no flag load sits in front of the genotype loads)
  // CHECK: test every value for NaN (counting pass, imputing pass)
  auto run = [&](bool impute, auto KEEP, auto CHECK) {
#pragma unroll
    for (int e = 0; e < PREP_SPW * RR; ++e) acc[e] = 0.0;
    // A run's loads (subset flags, row-union positions, environment rows by
    // column position -- every SNP of an interaction set multiplies column c
    // by the same row --, the genotypes of every column) are issued one run
    // ahead, so the memory round trip of run i+1 overlaps the work of run i.
    // keep, uidx and the E rows are padded to n_pad, so a run is in bounds.
    uchar4 kvn = make_uchar4(1, 1, 1, 1);
    int4 uin = make_int4(-1, -1, -1, -1);
    d4 ecn[R];
    GV gn[NV];
    auto fetch = [&](int k) {
      if constexpr (decltype(KEEP)::value) kvn = *reinterpret_cast<const uchar4*>(d.keep + k);
      if constexpr (CORR) uin = *reinterpret_cast<const int4*>(d.uidx + k);
#pragma unroll
      for (int c = 0; c < R; ++c)
        ecn[c] = env[c] >= 0 ? *reinterpret_cast<const d4*>(d.E + (size_t)env[c] * d.lde + k) : d4{1.0, 1.0, 1.0, 1.0};
      const int kw = __builtin_amdgcn_readfirstlane(k - 4 * (tid & 63));   // the wave's first run
      const bool vec = d.g_vec && kw + 256 <= d.n;
#pragma unroll
      for (int v = 0; v < NV; ++v) gn[v] = load_g4<GT>(d, src[v], k, vec);
    };
    if (4 * tid < d.n_pad) fetch(4 * tid);
    for (int k0 = 4 * tid; k0 < d.n_pad; k0 += 1024) {
      const uchar4 kv = kvn;
      const int4 ui = uin;
      d4 ec[R];
#pragma unroll
      for (int c = 0; c < R; ++c) ec[c] = ecn[c];
      GV g[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) g[v] = gn[v];
      if (k0 + 1024 < d.n_pad) fetch(k0 + 1024);
      // rows in the analysis: inside the subset (rows >= n load as 0)
      const bool in[4] = {kv.x != 0, kv.y != 0, kv.z != 0, kv.w != 0};
      double x[NV][4];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        d4 ev = ec[v % R];
        if (v >= R && env[v] != env[v % R])   // block-uniform: a descriptor layout of another shape
          ev = env[v] >= 0 ? *reinterpret_cast<const d4*>(d.E + (size_t)env[v] * d.lde + k0) : d4{1.0, 1.0, 1.0, 1.0};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          double xv;
          if constexpr (std::is_same<GT, _Float16>::value) xv = (double)(float)g[v][e];
          else xv = (double)g[v][e];
          xv *= ev[e];   // 1 without an environment row
          if (!in[e]) {
            xv = 0.0;
          } else if constexpr (decltype(CHECK)::value) {
            if (xv == xv) {
              if (!impute) sum[v] += xv;
            } else {
              if (!impute) cnt[v] += 1.0;
              xv = mean[v];
            }
          } else {
            sum[v] += xv;
          }
          x[v][e] = xv;
        }
      }
      if constexpr (CORR) {
        // rows of U: their imputed values, u-major, for the corrections
        const int uu[4] = {ui.x, ui.y, ui.z, ui.w};
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (uu[e] >= 0) {
            XT* dst = static_cast<XT*>(d.Xu) + (size_t)uu[e] * d.ldxu + (size_t)sb * R;
            if (NV == 4 && live[PREP_SPW - 1]) {
              // one vector store: sb * R and ldxu are multiples of 4
              const double v4[4] = {x[0][e], x[NV > 1 ? 1 : 0][e], x[NV > 2 ? 2 : 0][e], x[NV > 3 ? 3 : 0][e]};
              store_xu4(dst, v4);
            } else {
#pragma unroll
              for (int v = 0; v < NV; ++v)
                if (live[v / R]) store_xu(dst + v, x[v][e]);
            }
          }
      }
#pragma unroll
      for (int q = 0; q < PREP_SPW; ++q) {
        if (!live[q]) continue;
#pragma unroll
        for (int c = 0; c < R; ++c)
          store_op4<OP>(Ahi, Alo, (size_t)((sb + q) * R + c) * d.ldop + k0, x[q * R + c]);
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int b = 0; b < R; ++b)
#pragma unroll
            for (int a = 0; a <= b; ++a) acc[q * RR + upk(a, b)] += x[q * R + a][e] * x[q * R + b][e];
      }
    }
  };
  std::false_type fast;
  std::true_type checked;
  if (d.keep) run(false, std::true_type{}, fast);
  else run(false, std::false_type{}, fast);
  // one block reduction for the column sums, NaN counts and Gram pieces; the
  // Gram pieces are final unless some column holds NaN
  constexpr int NA = PREP_SPW * RR;
  double red[2 * NV + NA];
  auto reduce = [&]() {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      red[v] = sum[v];
      red[NV + v] = cnt[v];
    }
#pragma unroll
    for (int e = 0; e < NA; ++e) red[2 * NV + e] = acc[e];
    block_sum256_n<2 * NV + NA>(red, shd);
  };
  reduce();
  bool nan_sum = false;
#pragma unroll
  for (int v = 0; v < NV; ++v) nan_sum = nan_sum || (live[v / R] && red[v] != red[v]);
  if (nan_sum) {   // block-uniform: count the missing values, sum the others
#pragma unroll
    for (int v = 0; v < NV; ++v) sum[v] = 0.0;
    if (d.keep) run(false, std::true_type{}, checked);
    else run(false, std::false_type{}, checked);
    reduce();
  }
  double missing = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    mean[v] = red[v] / ((double)d.n_stat - red[NV + v]);
    if (live[v / R]) missing += red[NV + v];
  }
  if (missing > 0.0) {   // block-uniform; each thread rewrites its own rows
    if (d.keep) run(true, std::true_type{}, checked);
    else run(true, std::false_type{}, checked);
    block_sum256_n<NA>(acc, shd);
  } else {
#pragma unroll
    for (int e = 0; e < NA; ++e) acc[e] = red[2 * NV + e];
  }
  if (tid == 0) {
#pragma unroll
    for (int e = 0; e < NA; ++e)
      if (live[e / RR]) d.Sbr[(size_t)(sb + e / RR) * RR + (e % RR)] = acc[e];
#pragma unroll
    for (int q = 0; q < PREP_SPW; ++q)
      if (live[q])
#pragma unroll
        for (int c = 0; c < R; ++c) {
          d.col_mean[(size_t)(sb + q) * R + c] = mean[q * R + c];
          if (d.col_nan) d.col_nan[(size_t)(sb + q) * R + c] = red[NV + q * R + c] > 0.0 ? 1 : 0;
        }
  }
}

// ---------------------------------------------------------------------------
// The operand pass below the fused contraction, for blocks of [E*G, G]
// pairs (every SNP: one genotype column G, output columns E*G and G, one
// interaction row E; no subset).  No operand is written, so the pass is a
// read of the genotypes: column means and missing flags, Sbr and the
// compact rows of U.  A block owns 16 SNPs, one wave 4 of them; a genotype
// is loaded once for both of its columns, and each value costs 7 FP64
// operations: the conversion, E*G, two sums, three Gram terms.  The four
// waves walk the rows in step, 256 at a time: the values at that range's
// rows of U go to LDS, and after a barrier the block writes each such row's
// 16 SNPs as one 128-byte line (one store instruction per 32 rows of U,
// where per-lane stores took one per row position with few lanes active).
// A block whose sums come out NaN repeats its walk: counting the missing
// values, then with the means imputed (replace_nans_avgs,
// src/Utility.cpp:103), overwriting what the first walk wrote.
constexpr int SPW_P = 4;   // SNPs per wave
static_assert(SPW_P == 4, "the compact-row staging below holds four SNPs per wave");

// CORR: write the compact rows (FP32, the type below the split-BF16
// contraction).
template <bool CORR, typename GT>
__global__ __launch_bounds__(256, 2) void k_prep_pair(DevScan d) {
  chunk_start(d);
  typedef typename GVec<GT>::T GV;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // blocks dealt to the XCDs in turn: consecutive blocks on one XCD, so the
  // lines of neighbouring blocks in one row of U meet in one L2
  const int per = (gridDim.x + 7) >> 3;
  const int blk = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int s_blk = blk * 16;                 // the block's first SNP
  if (s_blk >= d.mc) return;                  // block-uniform, before any barrier
  const int sb = s_blk + w * SPW_P;           // the wave's first SNP
  bool live[SPW_P];
  int src[SPW_P], env = -1;
#pragma unroll
  for (int q = 0; q < SPW_P; ++q) {
    live[q] = sb + q < d.mc;
    int s2[2], e2[2];
    col_desc<2>(d, d.s0 + min(sb + q, d.mc - 1), s2, e2);
    src[q] = s2[0];
    env = e2[0];   // the block's one interaction row
  }
  const double* Erow = d.E + (size_t)env * d.lde;   // padded to n_pad with zeros
  // the range's rows of U, 16 SNPs x (E*G, G) each, two buffers
  __shared__ __attribute__((aligned(16))) float xs[CORR ? 2 : 1][CORR ? 256 : 1][32];
  // per SNP: sums of E*G and G, missing counts, Gram terms (EG EG, EG G, G G)
  double sum[SPW_P][2], cnt[SPW_P][2], mean[SPW_P][2], gram[SPW_P][3];
#pragma unroll
  for (int q = 0; q < SPW_P; ++q)
#pragma unroll
    for (int c = 0; c < 2; ++c) { sum[q][c] = 0.0; cnt[q][c] = 0.0; mean[q][c] = 0.0; }
  // MODE 0: unchecked sums, 1: count the missing values, 2: impute
  auto walk = [&](auto MODE) {
    constexpr int mode = decltype(MODE)::value;
    constexpr bool STORE = CORR && mode != 1;
#pragma unroll
    for (int q = 0; q < SPW_P; ++q)
#pragma unroll
      for (int e = 0; e < 3; ++e) gram[q][e] = 0.0;
    // two ranges in flight per lane: a range's loads are issued two ahead
    struct Buf { GV g[SPW_P]; d4 e; int4 u; };
    auto fetch = [&](int k, Buf& b) {
      if (k >= d.n_pad) return;
      const bool vec = d.g_vec && __builtin_amdgcn_readfirstlane(k - 4 * lane) + 256 <= d.n;
#pragma unroll
      for (int q = 0; q < SPW_P; ++q) b.g[q] = load_g4<GT>(d, src[q], k, vec);
      b.e = *reinterpret_cast<const d4*>(Erow + k);
      if constexpr (CORR) b.u = *reinterpret_cast<const int4*>(d.uidx + k);
    };
    // the range of rows base .. base + 255 (every wave and lane takes the
    // same path: the loop and its barriers are block-uniform)
    auto body = [&](int base, Buf& b, int buf) {
      const int k0 = base + 4 * lane;
      const bool in = k0 < d.n_pad;
      GV gv[SPW_P];
#pragma unroll
      for (int q = 0; q < SPW_P; ++q) gv[q] = b.g[q];
      const d4 ev = b.e;
      const int4 ui = in ? b.u : make_int4(-1, -1, -1, -1);
      fetch(k0 + 512, b);
      // the range's rows of U: the first one's position and their count (U
      // is sorted, so the range holds positions umin .. umin + nr - 1)
      const int uu[4] = {ui.x, ui.y, ui.z, ui.w};
      int umin = 0x7fffffff, nr = 0;
      if constexpr (STORE) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (uu[e] >= 0) umin = min(umin, uu[e]);
          nr += __builtin_popcountll(__ballot(uu[e] >= 0));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) umin = min(umin, __shfl_xor(umin, off, 64));
      }
#pragma unroll
      for (int q = 0; q < SPW_P; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          double g;
          if constexpr (std::is_same<GT, _Float16>::value) g = (double)(float)gv[q][e];
          else g = (double)gv[q][e];
          if (!in) g = 0.0;
          double eg = g * ev[e];
          if (!in) eg = 0.0;
          if constexpr (mode == 0) {
            sum[q][0] += eg;
            sum[q][1] += g;
          } else if constexpr (mode == 1) {
            if (eg == eg) sum[q][0] += eg; else cnt[q][0] += 1.0;
            if (g == g) sum[q][1] += g; else cnt[q][1] += 1.0;
          } else {
            if (eg != eg) eg = mean[q][0];
            if (g != g) g = mean[q][1];
          }
          if constexpr (mode != 1) {
            gram[q][0] += eg * eg;
            gram[q][1] += eg * g;
            gram[q][2] += g * g;
          }
          // the compact rows keep FP32 values
          if constexpr (STORE)
            if (uu[e] >= 0)
              *reinterpret_cast<float2*>(&xs[buf][uu[e] - umin][8 * w + 2 * q]) = make_float2((float)eg, (float)g);
        }
      if constexpr (CORR) if constexpr (STORE) {
        // every range passes this barrier, with rows of U or without: the
        // stores below read this buffer, and the range two later refills it
        // only after the next range's barrier
        __syncthreads();
        if (nr > 0) {   // block-uniform: every wave sees the same rows
          // 8 threads per row of U: columns 2 s_blk + 4 c .. + 3, those of
          // SNPs past the chunk left alone
          const int c = tid & 7, col = 2 * s_blk + 4 * c;
          for (int r = tid >> 3; r < nr; r += 32) {
            const float4 v = *reinterpret_cast<const float4*>(xs[buf][r] + 4 * c);
            float* dst = static_cast<float*>(d.Xu) + (size_t)(umin + r) * d.ldxu + col;
            if (col + 4 <= 2 * d.mc) *reinterpret_cast<float4*>(dst) = v;
            else if (col < 2 * d.mc) *reinterpret_cast<float2*>(dst) = make_float2(v.x, v.y);   // 2 d.mc is even
          }
        }
      }
    };
    Buf b0, b1;
    b0.u = b1.u = make_int4(-1, -1, -1, -1);
    fetch(4 * lane, b0);
    fetch(4 * lane + 256, b1);
    // two buffers of the staged rows: a buffer is refilled two ranges later,
    // after the next range's barrier (taken by every range), by which time
    // every thread has stored it
    for (int base = 0; base < d.n_pad; base += 512) {
      body(base, b0, 0);
      if (base + 256 < d.n_pad) body(base + 256, b1, CORR ? 1 : 0);
    }
  };
  auto close = [&]() {
#pragma unroll
    for (int q = 0; q < SPW_P; ++q) {
#pragma unroll
      for (int c = 0; c < 2; ++c) { sum[q][c] = wave_sum(sum[q][c]); cnt[q][c] = wave_sum(cnt[q][c]); }
#pragma unroll
      for (int e = 0; e < 3; ++e) gram[q][e] = wave_sum(gram[q][e]);
    }
  };
  walk(std::integral_constant<int, 0>{});
  close();
  bool nan_sum = false;
#pragma unroll
  for (int q = 0; q < SPW_P; ++q) nan_sum = nan_sum || (live[q] && (sum[q][0] != sum[q][0] || sum[q][1] != sum[q][1]));
  if (__syncthreads_or(nan_sum)) {   // block-uniform: count the missing values, sum the others, then impute
#pragma unroll
    for (int q = 0; q < SPW_P; ++q)
#pragma unroll
      for (int c = 0; c < 2; ++c) sum[q][c] = 0.0;
    walk(std::integral_constant<int, 1>{});
    close();
#pragma unroll
    for (int q = 0; q < SPW_P; ++q)
#pragma unroll
      for (int c = 0; c < 2; ++c) mean[q][c] = sum[q][c] / ((double)d.n_stat - cnt[q][c]);
    walk(std::integral_constant<int, 2>{});
#pragma unroll
    for (int q = 0; q < SPW_P; ++q)
#pragma unroll
      for (int e = 0; e < 3; ++e) gram[q][e] = wave_sum(gram[q][e]);
  } else {
#pragma unroll
    for (int q = 0; q < SPW_P; ++q)
#pragma unroll
      for (int c = 0; c < 2; ++c) mean[q][c] = sum[q][c] / (double)d.n_stat;
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < SPW_P; ++q)
      if (live[q]) {
        const int s = sb + q;
#pragma unroll
        for (int e = 0; e < 3; ++e) d.Sbr[(size_t)s * 3 + e] = gram[q][e];   // upk: (0,0), (0,1), (1,1)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          d.col_mean[(size_t)s * 2 + c] = mean[q][c];
          if (d.col_nan) d.col_nan[(size_t)s * 2 + c] = cnt[q][c] > 0.0 ? 1 : 0;
        }
      }
  }
}

template <bool CORR>
static void prep_pair_x(const DevScan& d, hipStream_t st) {
  const int nb = (d.mc + 15) / 16;
  dim3 grid((nb + 7) / 8 * 8), block(256);   // a multiple of 8 for the XCD mapping
  if (d.gdt == DT_F32) hipLaunchKernelGGL((k_prep_pair<CORR, float>), grid, block, 0, st, d);
  else hipLaunchKernelGGL((k_prep_pair<CORR, _Float16>), grid, block, 0, st, d);
}

void launch_prep_pair(const DevScan& d, bool corr, hipStream_t st) {
  // the compact rows are FP32 below the split-BF16 contraction, the only
  // one this pass serves
  if (corr) prep_pair_x<true>(d, st);
  else prep_pair_x<false>(d, st);
}

// ---------------------------------------------------------------------------
// launchers
template <int R, typename XT, typename GT>
static void prep_rxg(const DevScan& d, int prec, void* Ahi, void* Alo, hipStream_t st) {
  const int nb = (d.mc + prep_spw<R>() - 1) / prep_spw<R>();
  dim3 grid((nb + 7) / 8 * 8), block(256);   // a multiple of 8 for the XCD mapping
  if (prec == PREC_FP64) hipLaunchKernelGGL((k_prep<R, PREC_FP64, XT, GT>), grid, block, 0, st, d, Ahi, Alo);
  else if (prec == PREC_FP32) hipLaunchKernelGGL((k_prep<R, PREC_FP32, XT, GT>), grid, block, 0, st, d, Ahi, Alo);
  else hipLaunchKernelGGL((k_prep<R, PREC_BF16X3, XT, GT>), grid, block, 0, st, d, Ahi, Alo);
}

// Narrow compact rows only below the FP32 and split-BF16 contractions.
template <int R, typename XT, typename GT>
static void prep_rxgo(const DevScan& d, int prec, void* Ahi, void* Alo, hipStream_t st) {
  const int nb = (d.mc + prep_spw<R>() - 1) / prep_spw<R>();
  dim3 grid((nb + 7) / 8 * 8), block(256);
  if (prec == PREC_FP32) hipLaunchKernelGGL((k_prep<R, PREC_FP32, XT, GT>), grid, block, 0, st, d, Ahi, Alo);
  else hipLaunchKernelGGL((k_prep<R, PREC_BF16X3, XT, GT>), grid, block, 0, st, d, Ahi, Alo);
}

template <int R, typename XT>
static void prep_rxo(const DevScan& d, int prec, void* Ahi, void* Alo, hipStream_t st) {
  if (d.gdt == DT_F32) prep_rxgo<R, XT, float>(d, prec, Ahi, Alo, st);
  else if (d.gdt == DT_F64) prep_rxgo<R, XT, double>(d, prec, Ahi, Alo, st);
  else prep_rxgo<R, XT, _Float16>(d, prec, Ahi, Alo, st);
}

template <int R, typename XT>
static void prep_rx(const DevScan& d, int prec, void* Ahi, void* Alo, hipStream_t st) {
  if (d.gdt == DT_F32) prep_rxg<R, XT, float>(d, prec, Ahi, Alo, st);
  else if (d.gdt == DT_F64) prep_rxg<R, XT, double>(d, prec, Ahi, Alo, st);
  else prep_rxg<R, XT, _Float16>(d, prec, Ahi, Alo, st);
}

// Compact rows: FP64, FP32 (the scan below the FP32 and split-BF16
// contractions) or the panel's own narrow type (the request server, one
// column per SNP, any contraction).  The element type is the one the
// buffer was sized for (d.xu_dt).
template <int R>
static void prep_r(const DevScan& d, int prec, void* Ahi, void* Alo, bool corr, hipStream_t st) {
  if (!corr) prep_rx<R, void>(d, prec, Ahi, Alo, st);
  else if (d.xu_dt == DT_F64) prep_rx<R, double>(d, prec, Ahi, Alo, st);
  else if constexpr (R == 1) {
    if (d.xu_dt == DT_F32) prep_rx<1, float>(d, prec, Ahi, Alo, st);
    else prep_rx<1, _Float16>(d, prec, Ahi, Alo, st);
  } else {
    prep_rxo<R, float>(d, prec, Ahi, Alo, st);   // FP32 rows with r > 1: never below an FP64 contraction
  }
}

// The request server's compact rows when the panel has no missing genotype:
// Xu[u][s] = G[s0 + s][u] for every individual u, the panel chunk
// transposed (exact in the panel's own type).  64 x 64 tiles through LDS, so
// the reads along individuals and the writes along SNPs are both 128-byte
// (float16) or 256-byte (float32) runs per wave, where the operand pass
// would write 8-byte pieces to 4 rows per lane.
template <typename T>
__global__ __launch_bounds__(256) void k_panel_rows(const T* __restrict__ G, long long ldg, int s0, int mc, int n,
                                                    T* __restrict__ Xu, long long ldxu, const int* s0_dev) {
  if (s0_dev) s0 = *s0_dev;   // replayed from a graph
  __shared__ T tile[64][65];
  const int u0 = blockIdx.x * 64, sb = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int i = ty; i < 64; i += 4) {   // SNP sb + i, individuals u0 + tx
    const int sn = sb + i, u = u0 + tx;
    tile[i][tx] = (sn < mc && u < n) ? G[(size_t)(s0 + sn) * ldg + u] : (T)0;
  }
  __syncthreads();
  for (int i = ty; i < 64; i += 4) {   // individual u0 + i, SNPs sb + tx
    const int u = u0 + i, sn = sb + tx;
    if (u < n && sn < mc) Xu[(size_t)u * ldxu + sn] = tile[tx][i];
  }
}

void launch_panel_rows(const void* G, int gdt, long long ldg, int s0, int mc, int n, void* Xu, long long ldxu,
                       hipStream_t st, const int* s0_dev) {
  dim3 grid((n + 63) / 64, (mc + 63) / 64), block(256);
  if (gdt == DT_F32)
    hipLaunchKernelGGL(k_panel_rows<float>, grid, block, 0, st, static_cast<const float*>(G), ldg, s0, mc, n,
                       static_cast<float*>(Xu), ldxu, s0_dev);
  else
    hipLaunchKernelGGL(k_panel_rows<_Float16>, grid, block, 0, st, static_cast<const _Float16*>(G), ldg, s0, mc, n,
                       static_cast<_Float16*>(Xu), ldxu, s0_dev);
}

// Dosage model columns, formed once per genotype block before the scan
// (src/AIOwrapper.cpp:1258-1290: an sgemm of the n_used probability
// columns with the factor vector, or each column times its factor).  Output
// columns on the grid's y axis (one division per column), individuals on x
// with four consecutive rows per thread: one vector load per source column
// and one 32-byte store when the block is aligned (vec), element by element
// otherwise.  HBM-bound, n_used reads and one FP64 write per element.
template <typename GT>
__global__ __launch_bounds__(256) void k_dosage(const GT* __restrict__ G, long long ldg, int n, long long ncol,
                                                DosageModel dm, double* __restrict__ out, long long ldo, int vec) {
  typedef typename GVec<GT>::T GV;
  const int per = dm.sum ? 1 : dm.n_used;
  const int k0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (k0 >= n) return;   // no barrier follows
  const bool full = vec && k0 + 4 <= n;
  for (long long col = blockIdx.y; col < ncol; col += gridDim.y) {
    const long long s = col / per;
    const int j = (int)(col - s * per);
    const GT* g = G + (size_t)(s * dm.file_r) * ldg + k0;
    double* o = out + (size_t)col * ldo + k0;
    const int i0 = dm.sum ? 0 : j, i1 = dm.sum ? dm.n_used : j + 1;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = i0; i < i1; ++i) {
      const GT* gi = g + (size_t)i * ldg;
      if (full) {
        const GV x = *reinterpret_cast<const GV*>(gi);
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] += dm.f[i] * (double)x[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (k0 + e < n) v[e] += dm.f[i] * (double)gi[e];
      }
    }
    if (full) {
      *reinterpret_cast<d4*>(o) = d4{v[0], v[1], v[2], v[3]};
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (k0 + e < n) o[e] = v[e];
    }
  }
}

void launch_dosage(const void* G, int gdt, long long ldg, int n, int n_snp, const DosageModel& dm, double* out,
                   long long ldo, hipStream_t st) {
  const long long ncol = (long long)n_snp * (dm.sum ? 1 : dm.n_used);
  if (ncol <= 0 || n <= 0) return;
  const size_t esz = gdt == DT_F64 ? 8 : (gdt == DT_F32 ? 4 : 2);
  // vector path: 4-row runs of every column start on a 4-element boundary
  // and the output rows (ldo, a multiple of 4) on a 32-byte one
  const int vec = ((uintptr_t)G % (4 * esz) == 0) && (ldg % 4 == 0) && ((uintptr_t)out % 32 == 0) && (ldo % 4 == 0);
  const dim3 grid((unsigned)((n + 1023) / 1024), (unsigned)std::min<long long>(ncol, 65535));
  if (gdt == DT_F64)
    hipLaunchKernelGGL(k_dosage<double>, grid, dim3(256), 0, st, static_cast<const double*>(G), ldg, n, ncol, dm, out,
                       ldo, vec);
  else if (gdt == DT_F32)
    hipLaunchKernelGGL(k_dosage<float>, grid, dim3(256), 0, st, static_cast<const float*>(G), ldg, n, ncol, dm, out,
                       ldo, vec);
  else
    hipLaunchKernelGGL(k_dosage<_Float16>, grid, dim3(256), 0, st, static_cast<const _Float16*>(G), ldg, n, ncol, dm,
                       out, ldo, vec);
}

void launch_prep(const DevScan& d, int prec, void* Ahi, void* Alo, bool corr, hipStream_t st) {
  switch (d.r) {
    case 1: prep_r<1>(d, prec, Ahi, Alo, corr, st); break;
    case 2: prep_r<2>(d, prec, Ahi, Alo, corr, st); break;
    case 3: prep_r<3>(d, prec, Ahi, Alo, corr, st); break;
    default: prep_r<4>(d, prec, Ahi, Alo, corr, st); break;
  }
}

}  // namespace oa
