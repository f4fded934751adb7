// Resident genotype panel and request server (include/oagpu_serve.h).
//
// A server thread walks the panel chunk by chunk.  Per chunk: one operand
// pass over the panel (shared by every active request), one contraction of
// that operand against the concatenated [Y | XL] rows of all active
// requests, then each request's own Gram pieces, corrections, solves and
// statistics on its slice of the contraction output.  Requests are admitted
// at chunk boundaries within an HBM budget and a slot limit, and leave after
// one full cycle of chunks, so the set of active requests changes while the
// panel streams (continuous batching).  Each request keeps its trait-side
// state in an ordinary oagpu_ctx.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/oagpu_serve.h"
#include "oagpu_ctx.h"

using namespace oa;

struct oagpu_panel {
  int device = 0, n = 0, n_pad = 0;
  long long m = 0;
  int dtype = OAGPU_F32;
  void* G = nullptr;
  long long ldg = 0;
  bool own = false;
  bool has_nan = true;   // any missing genotype
};

namespace {
// flag = 1 if any of the panel's n x m values is NaN
__global__ void k_any_nan(const void* G, int dt, long long ld, int n, long long m, int* flag) {
  const long long total = (long long)n * m;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long s = i / n, k = i - s * n;
    const double v = load_elem(G, dt, (size_t)(s * ld + k));
    if (v != v) *flag = 1;
  }
}
}  // namespace

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t dt_bytes(int dt) { return dt == OAGPU_F64 ? 8 : (dt == OAGPU_F32 ? 4 : 2); }

enum { REQ_QUEUED = 0, REQ_ACTIVE = 1, REQ_DONE = 2 };

struct Request {
  long long id = 0;
  int t = 0, l = 0;
  std::vector<double> Y, XL;
  std::vector<uint8_t> keep;
  float min_p_store = 0.f, min_p_disp = 0.f, min_r2_store = 0.f;
  int disp_cov = 0, store_all = 0;
  long long max_hits = 0;

  oagpu_ctx* ctx = nullptr;
  int state = REQ_QUEUED;
  int status = OAGPU_OK;
  std::string err;
  int chunks_done = 0;
  int overtaken = 0;   // requests admitted ahead of it while it did not fit
  int col_off = 0;
  long long bytes = 0;
  std::vector<oagpu_hit> hits;
  oagpu_stats stats{};
  double t_submit = 0.0, t_done = 0.0;
};

}  // namespace

// Times a request that does not fit the HBM budget may be passed by later
// requests that do.
constexpr int kMaxOvertaken = 8;

// Streams the per-request kernels of a chunk are spread over (the server's
// own stream and kLanes - 1 more): a request with few traits launches grids
// far smaller than the chip, so several requests run side by side.
constexpr int kLanes = 4;

struct oagpu_server {
  oagpu_panel* panel = nullptr;
  oagpu_server_params prm{};
  std::string err;
  hipStream_t st = nullptr;
  hipStream_t lane_st[kLanes - 1] = {};
  hipEvent_t ev_operand = nullptr, ev_lane[kLanes - 1] = {};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;        // work arrived / stop
  std::condition_variable done_cv;   // a request finished
  bool stop = false;
  std::deque<Request*> queue;
  std::vector<Request*> active;      // touched by the server thread only
  std::map<long long, Request*> all;
  long long next_id = 1;
  long long used_bytes = 0;

  // chunk state
  int mc = 0, nchunks = 0, cur = 0;
  size_t op_size = 8, c_size = 8;
  long long ldop = 0;
  int n_pad = 0;
  void *Abuf = nullptr, *Ahi = nullptr, *Alo = nullptr;
  double *col_mean = nullptr, *Sbr = nullptr;
  // the chunk's genotype values, row-major, for the missing-row corrections:
  // in the panel's own type when that holds them exactly (no missing value)
  void* Xu = nullptr;
  int xu_dt = DT_F64;
  long long ldxu = 0;
  int* uidx = nullptr;               // identity row numbering for the operand pass
  void* Barena = nullptr;
  size_t B_cap = 0;
  void* Cbuf = nullptr;
  size_t C_cap = 0;
  int ldc = 0;
  bool arena_dirty = true;

  // graphs of a chunk's launches: one per chunk size (slot 0 the full
  // chunks, slot 1 the last, shorter one), captured for the active set of
  // epoch `epoch` and replayed with the chunk start written to s0_dev
  long long epoch = 0;
  int* s0_dev = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  long long g_epoch[2] = {-1, -1};   // active set a slot's graph was captured for
  long long g_seen[2] = {-1, -1};    // active set a slot last ran eagerly for
  bool graphs_off = false;           // replay not asked for, or a capture failed
  long long graph_chunks = 0;        // chunks replayed from a graph

  // chunk size chosen by timing (chunk_snps < 0): buffers are sized for the
  // largest candidate, the choice is made when the first requests arrive
  bool tune_chunk = false;
  int chunk_now = 0;                 // S->mc, readable under mu
};

#define SRVCK(S, expr)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (S)->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr;    \
      return OAGPU_ERR_HIP;                                                            \
    }                                                                                  \
  } while (0)

extern "C" int oagpu_panel_create(int device, int n, long long m, const void* G, int dtype, int where,
                                  long long ld, oagpu_panel** out) {
  if (!out || !G || n <= 0 || m <= 0 || ld < n || dtype < 0 || dtype > 2) return OAGPU_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return OAGPU_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return OAGPU_ERR_HIP;
  auto* P = new oagpu_panel();
  P->device = device;
  P->n = n;
  P->n_pad = (n + 63) / 64 * 64;
  P->m = m;
  P->dtype = dtype;
  P->ldg = ld;
  if (where == OAGPU_HOST) {
    const size_t bytes = dt_bytes(dtype) * ((size_t)ld * (m - 1) + n);
    if (hipMalloc(&P->G, bytes) != hipSuccess || hipMemcpy(P->G, G, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      if (P->G) (void)hipFree(P->G);
      delete P;
      return OAGPU_ERR_HIP;
    }
    P->own = true;
  } else {
    P->G = const_cast<void*>(G);
  }
  int* flag = nullptr;
  int h = 1;
  if (hipMalloc(&flag, sizeof(int)) == hipSuccess && hipMemset(flag, 0, sizeof(int)) == hipSuccess) {
    hipLaunchKernelGGL(k_any_nan, dim3(4096), dim3(256), 0, 0, P->G, dtype, ld, n, m, flag);
    if (hipMemcpy(&h, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) h = 1;
  }
  if (flag) (void)hipFree(flag);
  P->has_nan = h != 0;
  *out = P;
  return OAGPU_OK;
}

extern "C" void oagpu_panel_destroy(oagpu_panel* P) {
  if (!P) return;
  if (P->own && P->G) {
    (void)hipSetDevice(P->device);
    (void)hipFree(P->G);
  }
  delete P;
}

namespace {

// HBM a request holds while active: its context (traits, patterns, hit
// arena) plus its per-chunk pair buffers and its share of the output.
long long request_bytes(const oagpu_server* S, const Request* R) {
  const long long n = S->panel->n, t = R->t, l = R->l, mc = S->mc;
  const long long Q = l + 1;
  long long b = 0;
  b += (t + l + 128) * S->ldop * (long long)S->op_size;   // B rows
  b += t * n * 4 + n * t * 8;                               // missing-row lists, Y copy
  b += (R->max_hits > 0 ? R->max_hits : (1ll << 20)) * (long long)sizeof(oagpu_hit);
  b += (1ll << 16) * (long long)sizeof(FallbackEntry);
  const long long lb = l <= 8 ? 8 : (l <= 16 ? 16 : 32);
  b += mc * Q * t * 8;                                      // corrections, at most one pattern per trait
  b += mc * (t / 4) * (3 * lb + 3) * 8;                     // (SNP, pattern) table: 4 patterns <= t
  b += mc * (t + l) * (long long)S->c_size;                 // share of the contraction output
  b += mc * (2 * l + 4) * 8;                                // Str, var_x, Sbr, means
  return b;
}

void finish(oagpu_server* S, Request* R, int status, const std::string& err) {
  R->status = status;
  R->err = err;
  R->t_done = now_ms();
  if (R->ctx) {
    oagpu_destroy(R->ctx);
    R->ctx = nullptr;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  R->state = REQ_DONE;
  S->used_bytes -= R->bytes;
  S->done_cv.notify_all();
}

// Trait-side setup of an admitted request: an oagpu_ctx over the panel's
// individuals, its subset, covariates and traits.
int setup_request(oagpu_server* S, Request* R) {
  oagpu_params p{};
  p.n = S->panel->n;
  p.l = R->l;
  p.r = 1;
  p.real_n = S->panel->n;
  p.precision = S->prm.precision;
  p.disp_cov = R->disp_cov;
  p.store_all = R->store_all;
  p.min_r2_store = R->min_r2_store;
  p.min_p_store = R->min_p_store;
  p.min_p_disp = R->min_p_disp;
  p.max_traits = R->t;
  p.max_snps = S->mc;
  p.max_hits = R->max_hits > 0 ? R->max_hits : (1ll << 20);
  p.device = S->panel->device;
  int rc = oagpu_create(&p, &R->ctx);
  if (rc) {
    R->err = R->ctx ? oagpu_last_error(R->ctx) : "oagpu_create failed";
    return rc;
  }
  oagpu_ctx* c = R->ctx;
  c->union_identity = true;   // corrections read the chunk's shared transposed values
  if (!R->keep.empty() && (rc = oagpu_set_subset(c, R->keep.data()))) return rc;
  if ((rc = oagpu_set_covariates(c, R->XL.data(), S->panel->n))) return rc;
  if ((rc = oagpu_load_traits(c, R->Y.data(), OAGPU_F64, OAGPU_HOST, S->panel->n, R->t, 0))) return rc;
  c->use_pst = (long long)c->G * 4 <= R->t;
  if ((rc = ensure_pair_buffers(c, S->mc))) return rc;
  launch_reset_counters(c->counters, CNT_N, c->st);
  HIPCK(c, hipStreamSynchronize(c->st));
  // the inputs are on the device now
  R->Y.clear();
  R->Y.shrink_to_fit();
  R->XL.clear();
  R->XL.shrink_to_fit();
  return OAGPU_OK;
}

int rebuild_arena(oagpu_server* S) {
  int total = 0;
  for (Request* R : S->active) {
    R->col_off = total;
    total += R->t + R->l;
  }
  S->ldc = (total + 63) / 64 * 64;
  const size_t rows = (size_t)(S->ldc + 127) / 128 * 128 + 128;
  const size_t rowb = (size_t)S->ldop * S->op_size;
  if (!S->Barena || S->B_cap < rows * rowb) {
    SRVCK(S, hipStreamSynchronize(S->st));
    SRVCK(S, grow(&S->Barena, &S->B_cap, rows * rowb));
    SRVCK(S, hipMemsetAsync(S->Barena, 0, S->B_cap, S->st));
  }
  for (Request* R : S->active)
    SRVCK(S, hipMemcpyAsync((char*)S->Barena + (size_t)R->col_off * rowb, R->ctx->Bbuf,
                            (size_t)(R->t + R->l) * rowb, hipMemcpyDeviceToDevice, S->st));
  const size_t cbytes = (size_t)S->mc * S->ldc * S->c_size;
  if (S->C_cap < cbytes) {
    SRVCK(S, hipStreamSynchronize(S->st));
    SRVCK(S, grow(&S->Cbuf, &S->C_cap, cbytes));
  }
  S->arena_dirty = false;
  ++S->epoch;   // a new active set: captured graphs no longer apply
  return OAGPU_OK;
}

// The launches of one chunk (operand pass, contraction, every active
// request's pair side on the lane streams, joined back into S->st).  With
// s0_dev the kernels read the chunk start from there (graph capture).
int enqueue_chunk(oagpu_server* S, int s0, int mcc, const int* s0_dev) {
  const oagpu_panel* P = S->panel;
  const int prec = S->prm.precision;
  bool any_missing = false;
  for (Request* R : S->active) any_missing = any_missing || R->ctx->compact_corr;

  // operand pass over the chunk (shared)
  DevScan dp{};
  dp.n = P->n;
  dp.n_pad = S->n_pad;
  dp.ldop = S->ldop;
  dp.n_stat = P->n;
  dp.r = 1;
  dp.l = 1;
  dp.Gm = P->G;
  dp.gdt = P->dtype;
  dp.ldg = P->ldg;
  dp.g_vec = ((uintptr_t)P->G % 16 == 0) && (P->ldg % 4 == 0);
  dp.s0 = s0;
  dp.s0_dev = s0_dev;
  dp.mc = mcc;
  dp.col_mean = S->col_mean;
  dp.Sbr = S->Sbr;
  dp.uidx = S->uidx;
  dp.Xu = S->Xu;
  dp.xu_dt = S->xu_dt;
  dp.ldxu = S->ldxu;
  if (any_missing && S->xu_dt != DT_F64) {
    // no missing genotype in the panel: the compact rows are the chunk itself,
    // transposed, written by a tiled kernel instead of the operand pass
    launch_prep(dp, prec, S->Ahi, S->Alo, false, S->st);
    launch_panel_rows(P->G, S->xu_dt, P->ldg, s0, mcc, P->n, S->Xu, S->ldxu, S->st, s0_dev);
  } else {
    launch_prep(dp, prec, S->Ahi, S->Alo, any_missing, S->st);
  }

  // one contraction for every active request's columns
  const bool split = prec == OAGPU_BF16X3;
  const void* Blo = split ? S->Barena : nullptr;
  const void* Bhi = split ? (const void*)(static_cast<uint16_t*>(S->Barena) + S->n_pad) : S->Barena;
  launch_gemm(prec, S->Ahi, S->Alo, Bhi, Blo, S->Cbuf, mcc, S->ldc, S->n_pad, (int)S->ldop, (int)S->ldop, S->ldc,
              S->st);

  // each request on its slice; the requests read the shared operand, Xu,
  // Sbr, column means and contraction output and write only their own
  // buffers, so they run on several streams, each request on the one with
  // the fewest traits so far; all of them join the server's stream before
  // the next chunk overwrites the shared buffers
  SRVCK(S, hipEventRecord(S->ev_operand, S->st));
  for (hipStream_t w : S->lane_st) SRVCK(S, hipStreamWaitEvent(w, S->ev_operand, 0));
  long long lane_load[kLanes] = {0};
  for (Request* R : S->active) {
    oagpu_ctx* c = R->ctx;
    const int lane = (int)(std::min_element(lane_load, lane_load + kLanes) - lane_load);
    lane_load[lane] += c->t;
    hipStream_t st = lane == 0 ? S->st : S->lane_st[lane - 1];
    DevScan d = make_dev(c);
    d.s0 = s0;
    d.s0_dev = s0_dev;
    d.mc = mcc;
    d.Gm = P->G;
    d.gdt = P->dtype;
    d.ldg = P->ldg;
    d.g_vec = dp.g_vec;
    d.col_src = d.col_env = nullptr;
    d.col_mean = S->col_mean;
    d.Sbr = c->keep ? c->Sbr : S->Sbr;
    d.Cbuf = (const char*)S->Cbuf + (size_t)R->col_off * S->c_size;
    d.ldc = S->ldc;
    d.Xu = S->Xu;
    d.xu_dt = S->xu_dt;
    d.ldxu = S->ldxu;
    d.m_total = (int)P->m;
    d.snp_offset = 0;
    d.trait_offset = 0;
    d.need_bl = c->prm.disp_cov;
    if (c->keep) launch_subset_sbr(d, c->Sbr, st);
    launch_snp_gram(d, prec, c->lb, st);
    if (c->compact_corr) launch_corr(d, c->lb, st);
    if (c->use_pst) launch_patsnp(d, c->lb, st);
    launch_reset_counter(c->counters, CNT_FB_CHUNK, st);
    if (c->use_pst) launch_epilogue_pst(d, prec, c->lb, st);
    else launch_epilogue(d, prec, c->lb, st);
    launch_fallback(d, c->lb, st);
  }
  for (int w = 0; w < kLanes - 1; ++w) {
    SRVCK(S, hipEventRecord(S->ev_lane[w], S->lane_st[w]));
    SRVCK(S, hipStreamWaitEvent(S->st, S->ev_lane[w], 0));
  }
  SRVCK(S, hipGetLastError());
  return OAGPU_OK;
}

bool graph_has_memset(hipGraph_t g) {
  size_t n = 0;
  if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess) return true;
  std::vector<hipGraphNode_t> nodes(n);
  if (n && hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return true;
  for (hipGraphNode_t x : nodes) {
    hipGraphNodeType ty;
    if (hipGraphNodeGetType(x, &ty) != hipSuccess || ty == hipGraphNodeTypeMemset) return true;
  }
  return false;
}

// Captures a chunk's launches into a graph of slot `slot`; false (and
// graphs switched off for this server) when the capture or its
// instantiation fails or the graph holds a memset node.
bool capture_chunk(oagpu_server* S, int slot, int s0, int mcc) {
  if (S->gexec[slot]) {
    (void)hipGraphExecDestroy(S->gexec[slot]);
    S->gexec[slot] = nullptr;
  }
  if (hipStreamBeginCapture(S->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    S->graphs_off = true;
    return false;
  }
  const int rc = enqueue_chunk(S, s0, mcc, S->s0_dev);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(S->st, &g);
  bool ok = rc == OAGPU_OK && e == hipSuccess && g && !graph_has_memset(g);
  if (ok) ok = hipGraphInstantiate(&S->gexec[slot], g, nullptr, nullptr, 0) == hipSuccess;
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  if (!ok) {
    S->gexec[slot] = nullptr;
    S->graphs_off = true;
    return false;
  }
  S->g_epoch[slot] = S->epoch;
  return true;
}

// One chunk: replayed from the slot's graph when one was captured for the
// current active set, else launched one by one.  A slot's first chunk of an
// active set runs eagerly (it also settles the contraction autotuner, whose
// timing runs cannot be captured); its second is captured and replayed.
int run_chunk(oagpu_server* S) {
  const oagpu_panel* P = S->panel;
  const int s0 = S->cur * S->mc;
  const int mcc = (int)std::min<long long>(S->mc, P->m - s0);
  const int slot = mcc == S->mc ? 0 : 1;
  bool replay = false;
  if (!S->graphs_off) {
    if (S->g_epoch[slot] == S->epoch && S->gexec[slot]) {
      replay = true;
    } else if (S->g_seen[slot] == S->epoch) {
      replay = capture_chunk(S, slot, s0, mcc);
    } else {
      S->g_seen[slot] = S->epoch;
    }
  }
  if (replay) {
    launch_set_int(S->s0_dev, s0, S->st);
    SRVCK(S, hipGraphLaunch(S->gexec[slot], S->st));
    std::lock_guard<std::mutex> lk(S->mu);
    ++S->graph_chunks;
  } else {
    const int rc = enqueue_chunk(S, s0, mcc, nullptr);
    if (rc) return rc;
  }
  SRVCK(S, hipGetLastError());
  SRVCK(S, hipStreamSynchronize(S->st));
  return OAGPU_OK;
}

// Chunk size by timing: one chunk of each candidate size (4,096 to 32,768
// SNPs, at most the panel) with the first active set, after a run of it
// that settles the contraction autotuner; the size with the least time per
// SNP wins and is remembered per panel shape and precision.  The timing
// runs write nothing the requests keep: their counters are cleared after,
// and they start at chunk 0.
std::mutex g_chunk_mu;
std::map<std::tuple<int, int, long long, int, int>, int> g_chunk_choice;   // device, n, m, dtype, precision

int tune_chunk_size(oagpu_server* S) {
  const oagpu_panel* P = S->panel;
  const auto key = std::make_tuple(P->device, P->n, P->m, P->dtype, S->prm.precision);
  int best = 0;
  {
    std::lock_guard<std::mutex> lk(g_chunk_mu);
    auto it = g_chunk_choice.find(key);
    if (it != g_chunk_choice.end()) best = it->second;
  }
  if (best == 0) {
    const int cap = S->mc;   // buffers hold this many SNPs
    std::vector<int> cands;
    for (int c : {4096, 8192, 16384, 32768})
      if (c <= cap) cands.push_back(c);
    if (cands.empty()) cands.push_back(cap);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SRVCK(S, hipEventCreate(&e0));
    SRVCK(S, hipEventCreate(&e1));
    double best_per = 1e300;
    int rc = OAGPU_OK;
    for (int c : cands) {
      const int mcc = (int)std::min<long long>(c, P->m);
      if ((rc = enqueue_chunk(S, 0, mcc, nullptr))) break;   // settles the autotuner for this shape
      if (hipEventRecord(e0, S->st) != hipSuccess) { rc = OAGPU_ERR_HIP; break; }
      if ((rc = enqueue_chunk(S, 0, mcc, nullptr))) break;
      if (hipEventRecord(e1, S->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        rc = OAGPU_ERR_HIP;
        break;
      }
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = OAGPU_ERR_HIP; break; }
      const double per = ms / mcc;
      if (per < best_per) {
        best_per = per;
        best = c;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
      S->err = "chunk size tuning failed: " + S->err;
      return rc;
    }
    std::lock_guard<std::mutex> lk(g_chunk_mu);
    g_chunk_choice[key] = best;
  }
  // the timing runs' hits and counters go; every active request starts over
  for (Request* R : S->active) {
    launch_reset_counters(R->ctx->counters, CNT_N, S->st);
    R->chunks_done = 0;
  }
  SRVCK(S, hipGetLastError());
  SRVCK(S, hipStreamSynchronize(S->st));
  std::lock_guard<std::mutex> lk(S->mu);
  S->mc = best;
  S->chunk_now = best;
  S->nchunks = (int)((P->m + best - 1) / best);
  S->cur = 0;
  ++S->epoch;
  return OAGPU_OK;
}

// Counters and hits of a request that has seen every chunk.
void collect(oagpu_server* S, Request* R) {
  oagpu_ctx* c = R->ctx;
  unsigned long long cnt[CNT_N];
  if (hipMemcpy(cnt, c->counters, sizeof(cnt), hipMemcpyDeviceToHost) != hipSuccess) {
    finish(S, R, OAGPU_ERR_HIP, "HIP error reading the counters");
    return;
  }
  const long long nh = (long long)cnt[CNT_HITS];
  R->stats.n_hits = nh;
  R->stats.n_emitted = (long long)cnt[CNT_EMITTED];
  R->stats.n_sig = (long long)cnt[CNT_SIG];
  R->stats.n_fallback = (long long)cnt[CNT_FALLBACK];
  R->stats.n_patterns = c->G;
  R->stats.n_chunks = S->nchunks;
  R->stats.t_star = oagpu_tvalue(c->prm.min_p_store);
  if (cnt[CNT_ERR] & 1ull) {
    finish(S, R, OAGPU_ERR_NUMERIC, "negative RSS or variance (the reference exits here)");
    return;
  }
  if (nh > c->hit_cap) {
    finish(S, R, OAGPU_ERR_HIT_OVERFLOW, "more hits than the request's max_hits");
    return;
  }
  R->hits.resize((size_t)nh);
  if (nh > 0 &&
      hipMemcpy(R->hits.data(), c->hits, sizeof(oagpu_hit) * (size_t)nh, hipMemcpyDeviceToHost) != hipSuccess) {
    finish(S, R, OAGPU_ERR_HIP, "HIP error reading the hits");
    return;
  }
  std::sort(R->hits.begin(), R->hits.end(), [](const oagpu_hit& a, const oagpu_hit& b) {
    if (a.trait != b.trait) return a.trait < b.trait;
    if (a.snp != b.snp) return a.snp < b.snp;
    return a.h > b.h;
  });
  finish(S, R, OAGPU_OK, "");
}

void server_loop(oagpu_server* S) {
  (void)hipSetDevice(S->panel->device);
  for (;;) {
    std::vector<Request*> admitted;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->active.empty(); });
      if (S->stop && S->active.empty()) break;
      // first fit in arrival order: a request that does not fit the
      // budget beside the active ones lets later, smaller ones pass, at most
      // kMaxOvertaken times; then nothing behind it is admitted until it fits
      // (a request larger than the budget runs alone)
      std::vector<Request*> skipped;
      for (auto it = S->queue.begin();
           !S->stop && it != S->queue.end() && (int)(S->active.size() + admitted.size()) < S->prm.max_active;) {
        Request* R = *it;
        R->bytes = request_bytes(S, R);
        const bool empty = S->active.empty() && admitted.empty();
        if (empty || S->used_bytes + R->bytes <= S->prm.hbm_budget) {
          for (Request* W : skipped) ++W->overtaken;
          it = S->queue.erase(it);
          S->used_bytes += R->bytes;
          admitted.push_back(R);
          continue;
        }
        if (R->overtaken >= kMaxOvertaken) break;
        skipped.push_back(R);
        ++it;
      }
    }
    for (Request* R : admitted) {
      const int rc = setup_request(S, R);
      if (rc) {
        finish(S, R, rc, R->err.empty() && R->ctx ? oagpu_last_error(R->ctx) : R->err);
        continue;
      }
      R->state = REQ_ACTIVE;
      R->chunks_done = 0;
      S->active.push_back(R);
      S->arena_dirty = true;
    }
    if (S->active.empty()) continue;
    int rc = S->arena_dirty ? rebuild_arena(S) : OAGPU_OK;
    if (!rc && S->tune_chunk) {
      S->tune_chunk = false;
      rc = tune_chunk_size(S);
    }
    if (!rc) rc = run_chunk(S);
    if (rc) {   // a device error ends every active request
      for (Request* R : S->active) finish(S, R, rc, S->err);
      S->active.clear();
      S->arena_dirty = true;
      continue;
    }
    S->cur = (S->cur + 1) % S->nchunks;
    std::vector<Request*> still;
    for (Request* R : S->active) {
      if (++R->chunks_done == S->nchunks) {
        collect(S, R);
        S->arena_dirty = true;
      } else {
        still.push_back(R);
      }
    }
    S->active.swap(still);
  }
  // requests still queued at shutdown
  std::lock_guard<std::mutex> lk(S->mu);
  for (Request* R : S->queue) {
    R->status = OAGPU_ERR_STATE;
    R->err = "server stopped before the request ran";
    R->state = REQ_DONE;
    R->t_done = now_ms();
  }
  S->queue.clear();
  S->done_cv.notify_all();
}

}  // namespace

extern "C" int oagpu_server_create(oagpu_panel* P, const oagpu_server_params* prm, oagpu_server** out) {
  if (!P || !prm || !out) return OAGPU_ERR_ARG;
  if (prm->precision < 0 || prm->precision > 2) return OAGPU_ERR_ARG;
  *out = nullptr;
  auto* S = new oagpu_server();
  *out = S;
  S->panel = P;
  S->prm = *prm;
  if (S->prm.hbm_budget <= 0) S->prm.hbm_budget = 64ll << 30;
  if (S->prm.max_active <= 0) S->prm.max_active = 64;
  S->tune_chunk = S->prm.chunk_snps < 0;
  if (S->prm.chunk_snps == 0) S->prm.chunk_snps = 16384;
  if (S->tune_chunk) S->prm.chunk_snps = 32768;   // buffers for the largest candidate
  SRVCK(S, hipSetDevice(P->device));
  S->mc = (int)std::min<long long>(S->prm.chunk_snps, P->m);
  S->chunk_now = S->mc;
  S->nchunks = (int)((P->m + S->mc - 1) / S->mc);
  S->op_size = prm->precision == OAGPU_FP64 ? 8 : (prm->precision == OAGPU_FP32 ? 4 : 2);
  S->c_size = prm->precision == OAGPU_FP64 ? 8 : 4;
  S->n_pad = P->n_pad;
  S->ldop = (long long)P->n_pad * (prm->precision == OAGPU_BF16X3 ? 2 : 1);
  const size_t pad = (size_t)gemm_row_pad();
  const size_t arows = ((size_t)S->mc + pad - 1) / pad * pad;
  SRVCK(S, hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
  for (hipStream_t& w : S->lane_st) SRVCK(S, hipStreamCreateWithFlags(&w, hipStreamNonBlocking));
  SRVCK(S, hipEventCreateWithFlags(&S->ev_operand, hipEventDisableTiming));
  for (hipEvent_t& e : S->ev_lane) SRVCK(S, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  SRVCK(S, hipMalloc(&S->Abuf, arows * S->ldop * S->op_size));
  S->Ahi = S->Abuf;
  S->Alo = prm->precision == OAGPU_BF16X3 ? static_cast<uint16_t*>(S->Abuf) + P->n_pad : nullptr;
  SRVCK(S, hipMalloc(&S->col_mean, sizeof(double) * S->mc));
  SRVCK(S, hipMalloc(&S->Sbr, sizeof(double) * S->mc));
  S->ldxu = ((long long)S->mc + 7) / 8 * 8;
  S->xu_dt = (!P->has_nan && P->dtype != OAGPU_F64) ? (P->dtype == OAGPU_F32 ? DT_F32 : DT_F16) : DT_F64;
  // rows of all n individuals, then a zero row (n) for the corrections' padding lanes
  const size_t xu_el = dt_bytes(S->xu_dt == DT_F64 ? OAGPU_F64 : P->dtype);
  SRVCK(S, hipMalloc(&S->Xu, xu_el * ((size_t)P->n + 1) * S->ldxu));
  SRVCK(S, hipMemset(static_cast<char*>(S->Xu) + xu_el * (size_t)P->n * S->ldxu, 0, xu_el * S->ldxu));
  std::vector<int> id(P->n_pad, -1);
  for (int k = 0; k < P->n; ++k) id[k] = k;
  SRVCK(S, hipMalloc(&S->s0_dev, sizeof(int)));
  S->graphs_off = S->prm.graphs == 0;
  SRVCK(S, hipMalloc(&S->uidx, sizeof(int) * id.size()));
  SRVCK(S, hipMemcpy(S->uidx, id.data(), sizeof(int) * id.size(), hipMemcpyHostToDevice));
  S->th = std::thread(server_loop, S);
  return OAGPU_OK;
}

extern "C" void oagpu_server_destroy(oagpu_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->th.joinable()) S->th.join();
  for (auto& kv : S->all) {
    if (kv.second->ctx) oagpu_destroy(kv.second->ctx);
    delete kv.second;
  }
  (void)hipSetDevice(S->panel->device);
  for (hipGraphExec_t& x : S->gexec)
    if (x) (void)hipGraphExecDestroy(x);
  void* bufs[] = {S->Abuf, S->col_mean, S->Sbr, S->Xu, S->uidx, S->Barena, S->Cbuf, S->s0_dev};
  for (void* b : bufs) dfree(b);
  for (hipStream_t w : S->lane_st)
    if (w) (void)hipStreamDestroy(w);
  for (hipEvent_t e : S->ev_lane)
    if (e) (void)hipEventDestroy(e);
  if (S->ev_operand) (void)hipEventDestroy(S->ev_operand);
  if (S->st) (void)hipStreamDestroy(S->st);
  delete S;
}

extern "C" const char* oagpu_server_last_error(const oagpu_server* S) { return S ? S->err.c_str() : "null server"; }

extern "C" int oagpu_server_chunk_snps(const oagpu_server* S) {
  if (!S) return 0;
  std::lock_guard<std::mutex> lk(const_cast<oagpu_server*>(S)->mu);
  return S->chunk_now;
}

extern "C" long long oagpu_server_graph_chunks(const oagpu_server* S) {
  if (!S) return 0;
  std::lock_guard<std::mutex> lk(const_cast<oagpu_server*>(S)->mu);
  return S->graph_chunks;
}

extern "C" int oagpu_server_submit(oagpu_server* S, const oagpu_request* q, long long* id) {
  if (!S || !q || !id) return OAGPU_ERR_ARG;
  const int n = S->panel->n;
  if (q->t <= 0 || q->l <= 0 || q->l > 32 || !q->Y || !q->XL || q->ldy < n || q->ldxl < n) {
    S->err = "bad request (t >= 1, 1 <= l <= 32, ldy and ldxl >= n)";
    return OAGPU_ERR_ARG;
  }
  if (q->keep && S->panel->has_nan && std::find(q->keep, q->keep + n, (uint8_t)0) != q->keep + n) {
    // the shared operand pass imputes with the panel-wide column means; the
    // subset's own means would be needed
    S->err = "an individual subset needs a panel without missing genotypes";
    return OAGPU_ERR_UNSUPPORTED;
  }
  auto* R = new Request();
  R->t = q->t;
  R->l = q->l;
  R->Y.resize((size_t)n * q->t);
  for (int y = 0; y < q->t; ++y) std::memcpy(&R->Y[(size_t)y * n], q->Y + (size_t)y * q->ldy, sizeof(double) * n);
  R->XL.resize((size_t)n * q->l);
  for (int j = 0; j < q->l; ++j)
    std::memcpy(&R->XL[(size_t)j * n], q->XL + (size_t)j * q->ldxl, sizeof(double) * n);
  if (q->keep) R->keep.assign(q->keep, q->keep + n);
  R->min_p_store = q->min_p_store;
  R->min_p_disp = q->min_p_disp;
  R->min_r2_store = q->min_r2_store;
  R->disp_cov = q->disp_cov;
  R->store_all = q->store_all;
  R->max_hits = q->max_hits;
  R->t_submit = now_ms();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      delete R;
      return OAGPU_ERR_STATE;
    }
    R->id = S->next_id++;
    S->all[R->id] = R;
    S->queue.push_back(R);
    *id = R->id;
  }
  S->cv.notify_all();
  return OAGPU_OK;
}

extern "C" int oagpu_server_poll(oagpu_server* S, long long id) {
  if (!S) return OAGPU_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->all.find(id);
  if (it == S->all.end()) return OAGPU_ERR_ARG;
  return it->second->state == REQ_DONE ? 1 : 0;
}

extern "C" int oagpu_server_wait(oagpu_server* S, long long id, oagpu_hit* hits, long long hit_cap,
                                 oagpu_stats* stats, double* latency_ms, char* errbuf, int errlen) {
  if (!S) return OAGPU_ERR_ARG;
  Request* R = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return OAGPU_ERR_ARG;
    R = it->second;
    S->done_cv.wait(lk, [&] { return R->state == REQ_DONE; });
    S->all.erase(it);
  }
  int rc = R->status;
  if (stats) *stats = R->stats;
  if (latency_ms) *latency_ms = R->t_done - R->t_submit;
  if (rc == OAGPU_OK) {
    const long long nh = (long long)R->hits.size();
    if (hits && nh > 0) {
      if (nh > hit_cap) rc = OAGPU_ERR_HIT_OVERFLOW;
      else std::memcpy(hits, R->hits.data(), sizeof(oagpu_hit) * (size_t)nh);
    }
  }
  if (errbuf && errlen > 0) {
    const std::string& e = rc == OAGPU_ERR_HIT_OVERFLOW && R->err.empty() ? std::string("hit_cap too small") : R->err;
    std::snprintf(errbuf, (size_t)errlen, "%s", e.c_str());
  }
  delete R;
  return rc;
}
