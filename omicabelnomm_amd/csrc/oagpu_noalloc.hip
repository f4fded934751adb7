// oagpu_scan_noalloc: the scan of oagpu_scan for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The chunk buffers the scan's
// kernels work in -- operand rows, contraction output, Gram pieces,
// corrections, per-(SNP, pattern) solve pieces, compact missing-row values,
// counters and the fallback list -- are carved from one file-scope device
// array, `scratch`, instead of being allocated; chunks are sized to fit it.
// The context (set up beforehand with oagpu_create and the loads) holds the
// inputs and the hit arena.
//
// Every call on a device shares that device's array and leaves its contents
// behind: the counters are cleared by a kernel at the start of each call and
// the compact rows' zero row is rewritten each call; everything else is
// written before it is read.  Calls that overlap on one device (from two
// host threads, or on two contexts' streams) race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "oagpu_ctx.h"

using namespace oa;

static __device__ __align__(256) unsigned char scratch[OAGPU_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDev = 64;
// Fallback entries a chunk can hand to the explicit residual pass; pairs
// beyond it compute their residual inside the pair kernel.
constexpr int kFallbackCap = 4096;

struct ScratchDev {
  bool looked_up = false;
  hipError_t status = hipSuccess;
  unsigned char* ptr = nullptr;   // device address of this device's `scratch`
};
std::mutex g_mu;
ScratchDev g_dev[kMaxDev];

// 256-byte aligned regions of the array, laid out in order; with base ==
// nullptr it only measures.
struct Carver {
  unsigned char* base;
  size_t off = 0;
  void* take(size_t bytes) {
    off = (off + 255) & ~size_t(255);
    void* p = base ? base + off : nullptr;
    off += bytes;
    return p;
  }
};

struct Layout {
  unsigned long long* counters;
  FallbackEntry* fb;
  void* A;
  void* C;
  double *col_mean, *Sbr, *Str, *varx, *Ccorr, *pst;
  void* Xu;
  long long ldxu;
};

// The chunk buffers of mc SNPs (the sizes ensure_chunk and
// ensure_pair_buffers allocate), in order; returns the bytes used.
size_t layout(const oagpu_ctx* c, long long mc, unsigned char* base, Layout* L) {
  const int r = c->r, l = c->l, p = c->p, G = c->G, Q = c->Q;
  Carver cv{base};
  Layout o{};
  o.counters = static_cast<unsigned long long*>(cv.take(sizeof(unsigned long long) * CNT_N));
  o.fb = static_cast<FallbackEntry*>(cv.take(sizeof(FallbackEntry) * kFallbackCap));
  const size_t pad = (size_t)gemm_fixed_row_pad();
  const size_t arows = ((size_t)mc * r + pad - 1) / pad * pad;
  o.A = cv.take(arows * c->ldop * c->op_size);
  o.C = cv.take((size_t)mc * r * out_cols(c) * c->c_size);
  o.col_mean = static_cast<double*>(cv.take(sizeof(double) * (size_t)mc * r));
  o.Sbr = static_cast<double*>(cv.take(sizeof(double) * (size_t)mc * (r * (r + 1) / 2)));
  o.Str = static_cast<double*>(cv.take(sizeof(double) * (size_t)mc * l * r));
  o.varx = static_cast<double*>(cv.take(sizeof(double) * (size_t)mc * p));
  o.Ccorr = static_cast<double*>(cv.take(sizeof(double) * (size_t)mc * Q * std::max(G, 1)));
  o.pst = c->use_pst ? static_cast<double*>(cv.take(sizeof(double) * (size_t)mc * G * patsnp_width(r, l, c->lb)))
                     : nullptr;
  o.ldxu = ((long long)mc * r + 7) / 8 * 8;
  // rows 0..nU-1 of U, then a zero row (nU) for the corrections' padding lanes
  o.Xu = c->compact_corr ? cv.take(xu_size(c) * ((size_t)c->nU + 1) * o.ldxu) : nullptr;
  if (L) *L = o;
  return cv.off;
}

const char* fixed_variant_name(int prec) {
  return prec == OAGPU_FP64 ? "f64_direct_128x128"
                            : (prec == OAGPU_FP32 ? "f32_direct_128x128" : "bf16x3_lds_128x128");
}

}  // namespace

// Device address of this device's scratch array, looked up once per device
// with that device current (called from oagpu_create).  A failed lookup is
// remembered and reported by oagpu_scan_noalloc; the other entry points do
// not use the array.
void oa::noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDev) return;
  std::lock_guard<std::mutex> lk(g_mu);
  ScratchDev& s = g_dev[device];
  if (s.looked_up) return;
  s.looked_up = true;
  unsigned char* ptr = nullptr;
  s.status = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  s.ptr = s.status == hipSuccess ? ptr : nullptr;
}

extern "C" int oagpu_scan_noalloc(oagpu_ctx* c, oagpu_hit* hits, long long hit_cap, oagpu_stats* stats) {
  if (!c) return OAGPU_ERR_ARG;
  if (c->pending) return fail(c, OAGPU_ERR_STATE, "a scan is already in flight (oagpu_scan_wait first)");
  if (!c->have_traits || !c->have_geno) return fail(c, OAGPU_ERR_STATE, "load traits and genotypes before scanning");
  const int dev = c->prm.device;
  if (dev < 0 || dev >= kMaxDev) return fail(c, OAGPU_ERR_ARG, "device ordinal beyond the scratch table");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    const ScratchDev& s = g_dev[dev];
    if (!s.looked_up || s.status != hipSuccess || !s.ptr)
      return fail(c, OAGPU_ERR_HIP,
                  std::string("no-allocation scratch array unavailable on this device: ") +
                      (s.looked_up ? hipGetErrorString(s.status) : "not set up"));
    base = s.ptr;
  }
  HIPCK(c, hipSetDevice(dev));
  const int r = c->r, m = c->m;
  c->use_pst = (long long)c->G * 4 <= c->t;
  // the largest chunk (at most max_snps when set) whose buffers fit the array
  long long hi = c->prm.max_snps > 0 ? std::min<long long>(c->prm.max_snps, m) : m;
  hi = std::min(hi, (long long)(0x7fffffff / std::max(1, r * out_cols(c))));
  const size_t cap = OAGPU_NOALLOC_SCRATCH_BYTES;
  if (layout(c, 1, nullptr, nullptr) > cap)
    return fail(c, OAGPU_ERR_ARG,
                "workload too large for oagpu_scan_noalloc: one SNP's chunk buffers need " +
                    std::to_string(layout(c, 1, nullptr, nullptr)) + " bytes, the scratch array holds " +
                    std::to_string(cap) + " (n, traits or covariate patterns too large; use oagpu_scan)");
  long long lo = 1;
  while (lo < hi) {
    const long long mid = lo + (hi - lo + 1) / 2;
    if (layout(c, mid, nullptr, nullptr) <= cap) lo = mid;
    else hi = mid - 1;
  }
  const long long mc = lo;
  Layout L{};
  layout(c, mc, base, &L);

  if (c->compact_corr) launch_zero(static_cast<char*>(L.Xu) + xu_size(c) * (size_t)c->nU * L.ldxu,
                                   xu_size(c) * (size_t)L.ldxu, c->st);
  if (c->upload_pending) {
    HIPCK(c, hipStreamWaitEvent(c->st, c->up_ev[c->gslot], 0));
    c->upload_pending = false;
  }
  DevScan d = make_dev(c);
  d.counters = L.counters;
  d.fb = L.fb;
  d.fb_cap = kFallbackCap;
  d.Cbuf = L.C;
  d.col_mean = L.col_mean;
  d.col_nan = nullptr;   // only the fused contraction reads the flags
  d.Sbr = L.Sbr;
  d.Str = L.Str;
  d.varx = L.varx;
  d.Ccorr = L.Ccorr;
  d.pst = L.pst;
  d.Xu = L.Xu;
  d.ldxu = L.ldxu;
  void* Ahi = L.A;
  void* Alo = c->prm.precision == OAGPU_BF16X3 ? static_cast<void*>(static_cast<uint16_t*>(L.A) + c->n_pad) : nullptr;
  int rc = enqueue_chunks(c, d, mc, Ahi, Alo, nullptr, true);
  if (rc) return rc;
  c->gemm_variant = fixed_variant_name(c->prm.precision);
  HIPCK(c, hipEventRecord(c->free_ev[c->gslot], c->st));
  c->free_valid[c->gslot] = true;
  HIPCK(c, hipMemcpyAsync(c->cnt_h, L.counters, sizeof(unsigned long long) * CNT_N, hipMemcpyDeviceToHost, c->st));
  HIPCK(c, hipEventRecord(c->done_ev, c->st));
  c->pending = true;
  c->p_timed = false;
  c->p_dense = false;
  c->p_nchunks = (int)((m + mc - 1) / mc);
  c->p_t = c->t;
  c->p_m = m;
  return scan_collect(c, nullptr, hits, hit_cap, stats);
}
