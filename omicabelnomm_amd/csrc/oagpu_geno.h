// Genotype column access shared by the operand pass and the pair kernels:
// output column c of SNP s is source column src times environment row env
// (none when env < 0), NaN where the genotype is missing.
#pragma once
#include "oagpu_kernels.h"

namespace oa {

template <int R>
__device__ __forceinline__ void col_desc(const DevScan& d, int gs, int* src, int* env) {
#pragma unroll
  for (int c = 0; c < R; ++c) {
    const int idx = gs * R + c;
    src[c] = d.col_src ? d.col_src[idx] : idx;
    env[c] = d.col_env ? d.col_env[idx] : -1;
  }
}

// Genotype value of output column (src, env) at row k, NaN if missing.
__device__ __forceinline__ double geno_value(const DevScan& d, int src, int env, int k) {
  double v = load_elem(d.Gm, d.gdt, (size_t)src * d.ldg + k);
  if (env >= 0) v *= d.E[(size_t)env * d.lde + k];
  return v;
}

// Output column (src, env) at rows k0..k0+3 (k0 a multiple of 4), NaN if
// missing, 0 beyond n.  Interior runs are one vector load.
__device__ __forceinline__ void geno_value4(const DevScan& d, int src, int env, int k0, double* x) {
  const size_t base = (size_t)src * d.ldg + k0;
  if (k0 + 4 <= d.n) {
    load_elem4(d.Gm, d.gdt, d.g_vec != 0, base, x);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = (k0 + e < d.n) ? load_elem(d.Gm, d.gdt, base + e) : 0.0;
  }
  if (env >= 0) {   // E rows are padded to n_pad, so the run is always in bounds
    const double* ep = d.E + (size_t)env * d.lde + k0;
    const double2 a = *reinterpret_cast<const double2*>(ep);
    const double2 b = *reinterpret_cast<const double2*>(ep + 2);
    x[0] *= a.x; x[1] *= a.y; x[2] *= b.x; x[3] *= b.y;
  }
}

}  // namespace oa
