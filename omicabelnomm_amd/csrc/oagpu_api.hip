// C ABI of liboagpu.so (include/oagpu.h): context, uploads, pattern
// grouping and the chunked scan driver.  The loop structure follows the
// reference's partialNEQ_Blocked_STL_MD (src/Algorithm.cpp:322-886): a
// trait block is prepared once (src/Algorithm.cpp:550-592), then SNP chunks
// stream through prep -> contraction -> Gram -> corrections -> pair solve
// -> statistics (src/Algorithm.cpp:598-827).  The reference's SNP blocks of
// 500 become chunks sized to fill the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "oagpu_ctx.h"

using namespace oa;

static double erfinv_host(double x) {
  if (x <= -1.0) return -INFINITY;
  if (x >= 1.0) return INFINITY;
  // Giles (2010) single-precision start, refined by Newton steps on erf.
  double w = -std::log((1.0 - x) * (1.0 + x)), y;
  if (w < 5.0) {
    w -= 2.5;
    double q = 2.81022636e-08;
    q = 3.43273939e-07 + q * w; q = -3.5233877e-06 + q * w;
    q = -4.39150654e-06 + q * w; q = 0.00021858087 + q * w;
    q = -0.00125372503 + q * w; q = -0.00417768164 + q * w;
    q = 0.246640727 + q * w; q = 1.50140941 + q * w;
    y = q * x;
  } else {
    w = std::sqrt(w) - 3.0;
    double q = -0.000200214257;
    q = 0.000100950558 + q * w; q = 0.00134934322 + q * w;
    q = -0.00367342844 + q * w; q = 0.00573950773 + q * w;
    q = -0.0076224613 + q * w; q = 0.00943887047 + q * w;
    q = 1.00167406 + q * w; q = 2.83297682 + q * w;
    y = q * x;
  }
  for (int it = 0; it < 4; ++it) {
    const double e = std::erf(y) - x;
    y -= e / (1.1283791670955126 * std::exp(-y * y));
  }
  return y;
}

extern "C" double oagpu_tvalue(float pval) {
  return (double)(float)(erfinv_host(1.0 - (double)pval) * std::sqrt(2.0));
}

// Upper Cholesky of a packed symmetric matrix (LAPACK potrf('U') rule).
static bool chol_packed(std::vector<double>& A, int l) {
  for (int j = 0; j < l; ++j) {
    double ajj = A[upk(j, j)];
    for (int k = 0; k < j; ++k) ajj -= A[upk(k, j)] * A[upk(k, j)];
    if (!(ajj > 0.0)) return false;
    ajj = std::sqrt(ajj);
    A[upk(j, j)] = ajj;
    for (int i = j + 1; i < l; ++i) {
      double aji = A[upk(j, i)];
      for (int k = 0; k < j; ++k) aji -= A[upk(k, j)] * A[upk(k, i)];
      A[upk(j, i)] = aji / ajj;
    }
  }
  return true;
}

// Smallest singular value of XL (n x l, column stride ld), the reference's
// linear-dependence test (LAPACKE_sgesvd, src/Algorithm.cpp:508-522).  A
// Householder QR of XL in FP64 gives R with the same singular values; a
// one-sided (Hestenes) Jacobi on R's columns then converges to them with
// relative accuracy.  Working on XL'XL instead would square the condition
// number: at n = 10,000 its eigenvalues near (1e-6)^2 are below the
// rounding of the largest.
static double min_singular(const double* XL, int n, int l, long long ld) {
  std::vector<double> A((size_t)n * l);
  for (int j = 0; j < l; ++j)
    for (int k = 0; k < n; ++k) A[(size_t)j * n + k] = XL[(size_t)j * ld + k];
  std::vector<double> R((size_t)l * l, 0.0);   // column-major, upper
  for (int k = 0; k < l && k < n; ++k) {
    double* a = &A[(size_t)k * n];
    double nrm = 0.0;
    for (int i = k; i < n; ++i) nrm = std::hypot(nrm, a[i]);
    const double alpha = a[k] > 0 ? -nrm : nrm;
    if (nrm > 0.0) {            // (a zero remainder needs no reflection)
      a[k] -= alpha;            // v = x - alpha e_k, stored in place
      const double vtv = -2.0 * alpha * a[k];   // |v|^2 = 2 nrm (nrm + |x_k|)
      for (int j = k + 1; j < l; ++j) {
        double* b = &A[(size_t)j * n];
        double d = 0.0;
        for (int i = k; i < n; ++i) d += a[i] * b[i];
        const double f = 2.0 * d / vtv;
        for (int i = k; i < n; ++i) b[i] -= f * a[i];
      }
    }
    R[(size_t)k * l + k] = alpha;
    for (int j = k + 1; j < l; ++j) R[(size_t)j * l + k] = A[(size_t)j * n + k];
  }
  // one-sided Jacobi: rotate column pairs of R until all are orthogonal
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < l; ++p)
      for (int q = p + 1; q < l; ++q) {
        double* cp = &R[(size_t)p * l];
        double* cq = &R[(size_t)q * l];
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = 0; i < l; ++i) {
          a += cp[i] * cp[i];
          b += cq[i] * cq[i];
          g += cp[i] * cq[i];
        }
        if (g == 0.0 || std::fabs(g) <= 1e-15 * std::sqrt(a * b)) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < l; ++i) {
          const double x = cp[i], y = cq[i];
          cp[i] = cs * x - sn * y;
          cq[i] = sn * x + cs * y;
        }
      }
    if (!rotated) break;
  }
  double mn = INFINITY;
  for (int j = 0; j < l; ++j) {
    double nrm = 0.0;
    for (int i = 0; i < l; ++i) nrm = std::hypot(nrm, R[(size_t)j * l + i]);
    mn = std::min(mn, nrm);
  }
  return mn;
}

// Row stride of the contraction output: traits then covariates, padded to
// 64 columns (the padding columns are computed from zero or stale B rows and
// never read), so library GEMMs see aligned shapes.
int oa::out_cols(const oagpu_ctx* c) { return (c->t + c->l + 63) / 64 * 64; }

DevScan oa::make_dev(const oagpu_ctx* c) {
  DevScan d{};
  d.n = c->n; d.n_pad = c->n_pad; d.ldop = c->ldop;
  d.keep = c->keep; d.n_stat = c->keep ? c->n_stat : c->n; d.l = c->l; d.r = c->r; d.p = c->p; d.Q = c->Q;
  d.t = c->t; d.ldt = c->ldt; d.ldc = out_cols(c); d.G = c->G;
  d.start_h = c->prm.disp_cov ? 1 : c->l;
  d.disp_cov = c->prm.disp_cov; d.store_all = c->prm.store_all; d.need_bl = c->prm.disp_cov;
  d.min_r2_store = (double)c->prm.min_r2_store;
  d.t_star = oagpu_tvalue(c->prm.min_p_store);
  d.min_p_disp = (double)c->prm.min_p_disp;
  d.XL_cm = c->XL_cm; d.XL_rm = c->XL_rm; d.U0 = c->U0; d.Stl_inv_diag = c->Stl_inv_d;
  d.E = c->E; d.lde = c->lde;
  d.Y = c->Y; d.ydt = c->ydt; d.ldy = c->ldy;
  d.ymiss = c->ymiss; d.yzero = c->yzero; d.yy = c->yy; d.sst = c->sst; d.rss0 = c->rss0;
  d.ypat = c->ypat; d.rt = c->rt; d.zt = c->zt; d.ay = c->ay;
  d.pat_cnt = c->pat_cnt; d.pat_rep = c->pat_rep; d.rows = c->rows;
  d.UA = c->UA; d.UA_invd = c->UA_invd; d.pat_pd = c->pat_pd;
  d.wtrait = c->wtrait; d.wpat = c->wpat; d.n_waves = c->n_waves;
  d.nU = c->compact_corr ? c->nU : 0;
  d.uidx = c->uidx; d.XLu = c->XLu; d.pu_off = c->pu_off; d.pu_list = c->pu_list; d.XLp = c->XLp;
  d.Xu = c->Xu; d.xu_dt = xu_dtype(c); d.ldxu = c->ldxu; d.pst = c->pst;
  d.Gm = c->Gm; d.gdt = c->gdt; d.ldg = c->ldg; d.col_src = c->col_src; d.col_env = c->col_env;
  d.g_vec = c->Gm && ((uintptr_t)c->Gm % 16 == 0) && (c->ldg % 4 == 0);
  d.col_mean = c->col_mean; d.col_nan = c->col_nan; d.Sbr = c->Sbr; d.Str = c->Str; d.varx = c->varx; d.Ccorr = c->Ccorr;
  d.Cbuf = c->Cbuf;
  d.hits = c->hits; d.hit_cap = c->hit_cap; d.counters = c->counters;
  d.trait_offset = c->trait_offset; d.snp_offset = c->snp_offset; d.m_total = c->m;
  d.fb = c->fb; d.fb_cap = c->fb_cap;
  return d;
}

extern "C" int oagpu_create(const oagpu_params* prm, oagpu_ctx** out) {
  if (!prm || !out) return OAGPU_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return OAGPU_ERR_NO_DEVICE;
  auto* c = new oagpu_ctx();
  c->prm = *prm;
  *out = c;
  const int n = prm->n, l = prm->l, r = prm->r;
  if (n <= 0 || l <= 0 || r <= 0 || prm->max_traits <= 0)
    return fail(c, OAGPU_ERR_ARG, "n, l, r and max_traits must be positive");
  if (l + r >= n)
    return fail(c, OAGPU_ERR_ARG,
                "Error! The amount of unknows (individuals|elements) must be greater than total of "
                "variables (#covariates+#ind.variables)!");
  if (r > 4 || (r == 1 && l > 32) || (r == 2 && l > 16) || (r >= 3 && l > 8))
    return fail(c, OAGPU_ERR_UNSUPPORTED, "unsupported (l, r): compiled for r<=4 with l<=32/16/8");
  if (prm->precision < 0 || prm->precision > 2) return fail(c, OAGPU_ERR_ARG, "bad precision");
  if (prm->device < 0 || prm->device >= ndev) return fail(c, OAGPU_ERR_ARG, "bad device");
  HIPCK(c, hipSetDevice(prm->device));
  noalloc_setup(prm->device);
  HIPCK(c, hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  HIPCK(c, hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) {
    HIPCK(c, hipEventCreateWithFlags(&c->up_ev[i], hipEventDisableTiming));
    HIPCK(c, hipEventCreateWithFlags(&c->free_ev[i], hipEventDisableTiming));
  }
  HIPCK(c, hipEventCreate(&c->done_ev));
  HIPCK(c, hipHostMalloc((void**)&c->cnt_h, sizeof(unsigned long long) * CNT_N, hipHostMallocDefault));
  c->n = n; c->l = l; c->r = r; c->p = l + r;
  c->Q = l * r + r * (r + 1) / 2;
  c->n_pad = (n + 63) / 64 * 64;
  c->lb = l <= 8 ? 8 : (l <= 16 ? 16 : (l <= 24 ? 24 : 32));   // 24 only occurs with r = 1
  c->op_size = prm->precision == OAGPU_FP64 ? 8 : (prm->precision == OAGPU_FP32 ? 4 : 2);
  c->c_size = prm->precision == OAGPU_FP64 ? 8 : 4;
  c->ldop = (long long)c->n_pad * (prm->precision == OAGPU_BF16X3 ? 2 : 1);
  c->hit_cap = prm->max_hits > 0 ? prm->max_hits : (1ll << 22);
  // fallback entries per chunk; pairs beyond them compute the explicit
  // residual inside the pair kernel (OAGPU_FB_CAP sets the capacity)
  c->fb_cap = 1 << 16;
  if (const char* e = getenv("OAGPU_FB_CAP")) c->fb_cap = std::max(1, atoi(e));
  {
    std::vector<float> ones((size_t)c->n_pad, 1.0f);
    HIPCK(c, hipMalloc(&c->E32, sizeof(float) * ones.size()));
    HIPCK(c, hipMemcpy(c->E32, ones.data(), sizeof(float) * ones.size(), hipMemcpyHostToDevice));
  }
  HIPCK(c, hipMalloc(&c->hits, sizeof(oagpu_hit) * (size_t)c->hit_cap));
  HIPCK(c, hipMalloc(&c->counters, sizeof(unsigned long long) * CNT_N));
  HIPCK(c, hipMalloc(&c->fb, sizeof(FallbackEntry) * (size_t)c->fb_cap));
  // Trait-side buffers sized by max_traits.
  const int T = prm->max_traits;
  c->ldt = (T + 63) / 64 * 64;
  const size_t brows = (size_t)((T + l + 127) / 128 * 128);
  HIPCK(c, hipMalloc(&c->Bbuf, brows * c->ldop * c->op_size));
  HIPCK(c, hipMemset(c->Bbuf, 0, brows * c->ldop * c->op_size));
  if (prm->precision == OAGPU_BF16X3) {
    c->Blo = c->Bbuf;
    c->Bhi = static_cast<uint16_t*>(c->Bbuf) + c->n_pad;
  } else {
    c->Bhi = c->Bbuf;
  }
  HIPCK(c, hipMalloc(&c->rows, sizeof(int) * (size_t)T * n));
  HIPCK(c, hipMalloc(&c->ymiss, sizeof(int) * T));
  HIPCK(c, hipMalloc(&c->yzero, sizeof(int) * T));
  HIPCK(c, hipMalloc(&c->ypat, sizeof(int) * T));
  HIPCK(c, hipMalloc(&c->scratch_i, sizeof(int) * T));
  HIPCK(c, hipMalloc(&c->yy, sizeof(double) * T));
  HIPCK(c, hipMalloc(&c->sst, sizeof(double) * T));
  HIPCK(c, hipMalloc(&c->rss0, sizeof(double) * T));
  HIPCK(c, hipMalloc(&c->rt, sizeof(double) * (size_t)l * c->ldt));
  HIPCK(c, hipMalloc(&c->zt, sizeof(double) * (size_t)l * c->ldt));
  HIPCK(c, hipMalloc(&c->ay, sizeof(double) * (size_t)l * c->ldt));
  HIPCK(c, hipMalloc(&c->hash, sizeof(unsigned long long) * 2 * T));
  return OAGPU_OK;
}

extern "C" void oagpu_destroy(oagpu_ctx* c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->cst) (void)hipStreamSynchronize(c->cst);
  for (int i = 0; i < 2; ++i) {
    dfree(c->G_slot[i]);
    dfree(c->G_dos[i]);
    dfree(c->d_src[i]);
    dfree(c->d_env[i]);
    if (c->up_ev[i]) (void)hipEventDestroy(c->up_ev[i]);
    if (c->free_ev[i]) (void)hipEventDestroy(c->free_ev[i]);
  }
  for (auto& e : c->pev) (void)hipEventDestroy(e);
  if (c->done_ev) (void)hipEventDestroy(c->done_ev);
  if (c->cnt_h) (void)hipHostFree(c->cnt_h);
  void* dense_bufs[] = {c->pd_beta, c->pd_se, c->pd_t, c->pd_p, c->pd_kept, c->pd_r2, c->pd_rss, c->pd_em,
                        c->pd_varx};
  for (void* b : dense_bufs) dfree(b);
  void* bufs[] = {c->XL_cm, c->XL_rm, c->U0, c->Stl_inv_d, c->Stl_pk, c->XLop_hi, c->XLop_lo, c->E, c->E32,
                  c->Y_own, c->Bbuf, c->rows, c->ymiss, c->yzero, c->ypat, c->yy, c->sst,
                  c->rss0, c->rt, c->zt, c->ay, c->hash, c->scratch_i, c->pat_cnt, c->pat_rep, c->pat_pd, c->keep,
                  c->UA, c->UA_invd, c->wtrait, c->wpat, c->Abuf, c->Cbuf,
                  c->col_mean, c->col_nan, c->Sbr, c->Str, c->varx, c->Ccorr, c->hits, c->counters, c->fb,
                  c->uidx, c->pu_off, c->pu_list, c->XLu, c->XLp, c->Xu, c->pst, c->Bf};
  for (void* b : bufs) dfree(b);
  if (c->st) (void)hipStreamDestroy(c->st);
  if (c->cst) (void)hipStreamDestroy(c->cst);
  delete c;
}

extern "C" const char* oagpu_last_error(const oagpu_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int oagpu_set_subset(oagpu_ctx* c, const uint8_t* keep) {
  if (!c) return OAGPU_ERR_ARG;
  if (c->pending) return fail(c, OAGPU_ERR_STATE, "a scan is in flight (oagpu_scan_wait first)");
  HIPCK(c, hipSetDevice(c->prm.device));
  c->have_cov = c->have_traits = false;
  if (!keep) {
    c->keep_h.clear();
    dfree(c->keep);
    c->keep = nullptr;
    c->n_stat = c->n;
    return OAGPU_OK;
  }
  c->keep_h.assign(keep, keep + c->n);
  int nk = 0;
  for (auto& v : c->keep_h) {
    v = v ? 1 : 0;
    nk += v;
  }
  if (nk <= c->p + 1) return fail(c, OAGPU_ERR_ARG, "the subset holds too few individuals");
  if (!c->keep) HIPCK(c, hipMalloc(&c->keep, (size_t)c->n_pad));
  HIPCK(c, hipMemset(c->keep, 0, (size_t)c->n_pad));
  HIPCK(c, hipMemcpy(c->keep, c->keep_h.data(), (size_t)c->n, hipMemcpyHostToDevice));
  c->n_stat = nk;
  return OAGPU_OK;
}

extern "C" int oagpu_set_covariates(oagpu_ctx* c, const double* XL, long long ld) {
  if (!c || !XL || ld < c->n) return c ? fail(c, OAGPU_ERR_ARG, "bad covariates") : OAGPU_ERR_ARG;
  if (c->pending) return fail(c, OAGPU_ERR_STATE, "a scan is in flight (oagpu_scan_wait first)");
  HIPCK(c, hipSetDevice(c->prm.device));
  const int n = c->n, l = c->l, np = c->n_pad;
  // with a subset, the rows outside it are zero (they then drop out of
  // every covariate product)
  std::vector<double> XLm;
  if (!c->keep_h.empty()) {
    XLm.assign((size_t)n * l, 0.0);
    for (int j = 0; j < l; ++j)
      for (int k = 0; k < n; ++k)
        if (c->keep_h[k]) XLm[(size_t)j * n + k] = XL[(size_t)j * ld + k];
    XL = XLm.data();
    ld = n;
  }
  std::vector<double> S(l * l);
  for (int i = 0; i < l; ++i)
    for (int j = 0; j < l; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += XL[(size_t)i * ld + k] * XL[(size_t)j * ld + k];
      S[i * l + j] = s;
    }
  for (int j = 0; j < l; ++j)
    for (int k = 0; k < n; ++k)
      if (std::isnan(XL[(size_t)j * ld + k]))
        return fail(c, OAGPU_ERR_ARG, "covariates contain NaN (rows must be excluded first)");
  if (min_singular(XL, n, l, ld) < 0.000001)
    return fail(c, OAGPU_ERR_COVARIATES,
                "Error, the covariate matrix contains linearly dependent columns. Please review and "
                "change it accordingly.");
  std::vector<double> pk(l * (l + 1) / 2);
  for (int j = 0; j < l; ++j)
    for (int i = 0; i <= j; ++i) pk[upk(i, j)] = S[i * l + j];
  std::vector<double> U = pk;
  if (!chol_packed(U, l)) return fail(c, OAGPU_ERR_COVARIATES, "covariate Gram matrix not positive definite");
  // diag(Stl^-1) from the inverse of U.
  std::vector<double> Ui(l * l, 0.0);
  for (int j = 0; j < l; ++j) {
    Ui[j * l + j] = 1.0 / U[upk(j, j)];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += U[upk(i, k)] * Ui[j * l + k];
      Ui[j * l + i] = -s / U[upk(i, i)];
    }
  }
  c->Stl_inv_diag.assign(l, 0.0);
  for (int i = 0; i < l; ++i) {
    double s = 0.0;
    for (int j = i; j < l; ++j) s += Ui[j * l + i] * Ui[j * l + i];
    c->Stl_inv_diag[i] = s;
  }
  std::vector<double> cm((size_t)l * np, 0.0), rm((size_t)n * l);
  for (int j = 0; j < l; ++j)
    for (int k = 0; k < n; ++k) {
      cm[(size_t)j * np + k] = XL[(size_t)j * ld + k];
      rm[(size_t)k * l + j] = XL[(size_t)j * ld + k];
    }
  if (!c->XL_cm) {
    HIPCK(c, hipMalloc(&c->XL_cm, sizeof(double) * cm.size()));
    HIPCK(c, hipMalloc(&c->XL_rm, sizeof(double) * rm.size()));
    HIPCK(c, hipMalloc(&c->U0, sizeof(double) * pk.size()));
    HIPCK(c, hipMalloc(&c->Stl_pk, sizeof(double) * pk.size()));
    HIPCK(c, hipMalloc(&c->Stl_inv_d, sizeof(double) * l));
    HIPCK(c, hipMalloc(&c->XLop_hi, c->op_size * (size_t)l * np));
    if (c->prm.precision == OAGPU_BF16X3) HIPCK(c, hipMalloc(&c->XLop_lo, c->op_size * (size_t)l * np));
  }
  HIPCK(c, hipMemcpy(c->XL_cm, cm.data(), sizeof(double) * cm.size(), hipMemcpyHostToDevice));
  HIPCK(c, hipMemcpy(c->XL_rm, rm.data(), sizeof(double) * rm.size(), hipMemcpyHostToDevice));
  c->XL_rm_h = rm;
  HIPCK(c, hipMemcpy(c->U0, U.data(), sizeof(double) * U.size(), hipMemcpyHostToDevice));
  HIPCK(c, hipMemcpy(c->Stl_pk, pk.data(), sizeof(double) * pk.size(), hipMemcpyHostToDevice));
  HIPCK(c, hipMemcpy(c->Stl_inv_d, c->Stl_inv_diag.data(), sizeof(double) * l, hipMemcpyHostToDevice));
  // Covariate rows of the contraction operand B.
  if (c->prm.precision == OAGPU_FP64) {
    HIPCK(c, hipMemcpy(c->XLop_hi, cm.data(), sizeof(double) * cm.size(), hipMemcpyHostToDevice));
  } else if (c->prm.precision == OAGPU_FP32) {
    std::vector<float> f(cm.size());
    for (size_t i = 0; i < cm.size(); ++i) f[i] = (float)cm[i];
    HIPCK(c, hipMemcpy(c->XLop_hi, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
  } else {
    std::vector<uint16_t> hi(cm.size()), lo(cm.size());
    for (size_t i = 0; i < cm.size(); ++i) {
      const float x = (float)cm[i];
      uint32_t u;
      std::memcpy(&u, &x, 4);
      u += 0x7FFFu + ((u >> 16) & 1u);
      hi[i] = (uint16_t)(u >> 16);
      uint32_t hb = ((uint32_t)hi[i]) << 16;
      float hf;
      std::memcpy(&hf, &hb, 4);
      const float rest = x - hf;
      std::memcpy(&u, &rest, 4);
      u += 0x7FFFu + ((u >> 16) & 1u);
      lo[i] = (uint16_t)(u >> 16);
    }
    HIPCK(c, hipMemcpy(c->XLop_hi, hi.data(), 2 * hi.size(), hipMemcpyHostToDevice));
    HIPCK(c, hipMemcpy(c->XLop_lo, lo.data(), 2 * lo.size(), hipMemcpyHostToDevice));
  }
  c->have_cov = true;
  c->have_traits = false;   // trait statistics depend on XL
  return OAGPU_OK;
}

extern "C" int oagpu_set_interactions(oagpu_ctx* c, const double* E, long long ld, int n_env) {
  if (!c || !E || ld < c->n || n_env <= 0) return c ? fail(c, OAGPU_ERR_ARG, "bad interactions") : OAGPU_ERR_ARG;
  if (c->pending) return fail(c, OAGPU_ERR_STATE, "a scan is in flight (oagpu_scan_wait first)");
  HIPCK(c, hipSetDevice(c->prm.device));
  dfree(c->E);
  c->E = nullptr;
  // rows padded to n_pad with zeros, so kernels load 4-row runs as vectors
  const size_t np = (size_t)c->n_pad;
  std::vector<double> e((size_t)n_env * np, 0.0);
  for (int j = 0; j < n_env; ++j)
    for (int k = 0; k < c->n; ++k) {
      e[(size_t)j * np + k] = E[(size_t)j * ld + k];
      if (std::isnan(e[(size_t)j * np + k]))
        return fail(c, OAGPU_ERR_ARG, "interaction covariates contain NaN (rows must be excluded first)");
    }
  HIPCK(c, hipMalloc(&c->E, sizeof(double) * e.size()));
  HIPCK(c, hipMemcpy(c->E, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice));
  std::vector<float> e32((size_t)(n_env + 1) * np, 0.0f);
  for (size_t q = 0; q < e.size(); ++q) e32[q] = (float)e[q];
  for (size_t k = 0; k < np; ++k) e32[(size_t)n_env * np + k] = 1.0f;
  dfree(c->E32);
  c->E32 = nullptr;
  HIPCK(c, hipMalloc(&c->E32, sizeof(float) * e32.size()));
  HIPCK(c, hipMemcpy(c->E32, e32.data(), sizeof(float) * e32.size(), hipMemcpyHostToDevice));
  c->lde = (long long)np;
  c->n_env = n_env;
  return OAGPU_OK;
}

static size_t dt_size(int dt) { return dt == OAGPU_F64 ? 8 : (dt == OAGPU_F32 ? 4 : 2); }

static int build_union(oagpu_ctx* c, const std::vector<int>& pat_rep, const std::vector<int>& pat_cnt);

extern "C" int oagpu_load_traits(oagpu_ctx* c, const void* Y, int dtype, int where, long long ld,
                                 int t, int trait_offset) {
  if (!c) return OAGPU_ERR_ARG;
  if (!c->have_cov) return fail(c, OAGPU_ERR_STATE, "oagpu_set_covariates must come first");
  if (c->pending) return fail(c, OAGPU_ERR_STATE, "a scan is in flight (oagpu_scan_wait first)");
  if (!Y || t <= 0 || t > c->prm.max_traits || ld < c->n || dtype < 0 || dtype > 2 || dtype == OAGPU_F16)
    return fail(c, OAGPU_ERR_ARG, "bad trait block (t <= max_traits, ld >= n, dtype F64/F32)");
  HIPCK(c, hipSetDevice(c->prm.device));
  const int n = c->n, l = c->l;
  if ((long long)t * n > 0x7fffffffll) return fail(c, OAGPU_ERR_ARG, "t*n too large");
  c->t = t;
  c->trait_offset = trait_offset;
  c->ydt = dtype;
  c->ldy = ld;
  if (where == OAGPU_HOST) {
    const size_t bytes = dt_size(dtype) * ((size_t)ld * (t - 1) + n);
    HIPCK(c, grow(&c->Y_own, &c->Y_own_bytes, bytes));
    HIPCK(c, hipMemcpyAsync(c->Y_own, Y, bytes, hipMemcpyHostToDevice, c->st));
    c->Y = c->Y_own;
  } else {
    c->Y = Y;
  }
  c->G = 0;
  DevScan d = make_dev(c);
  launch_ystats(d, c->prm.precision, c->Bhi, c->Blo, c->rows, c->ymiss, c->yzero, c->yy, c->sst, c->rt,
                c->hash, c->st);
  HIPCK(c, hipGetLastError());
  // covariate rows of B after the traits
  const size_t rowb = (size_t)c->n_pad * c->op_size, pitch = (size_t)c->ldop * c->op_size;
  HIPCK(c, hipMemcpy2DAsync((char*)c->Bhi + pitch * t, pitch, c->XLop_hi, rowb, rowb, l, hipMemcpyDeviceToDevice,
                            c->st));
  if (c->Blo)
    HIPCK(c, hipMemcpy2DAsync((char*)c->Blo + pitch * t, pitch, c->XLop_lo, rowb, rowb, l,
                              hipMemcpyDeviceToDevice, c->st));

  // Group traits by missing-row pattern: key = (count, 128-bit hash),
  // verified row by row on the device.
  std::vector<int> ymiss(t);
  std::vector<unsigned long long> hash(2 * (size_t)t);
  HIPCK(c, hipMemcpyAsync(ymiss.data(), c->ymiss, sizeof(int) * t, hipMemcpyDeviceToHost, c->st));
  HIPCK(c, hipMemcpyAsync(hash.data(), c->hash, sizeof(unsigned long long) * 2 * t, hipMemcpyDeviceToHost, c->st));
  HIPCK(c, hipStreamSynchronize(c->st));
  std::map<std::tuple<int, unsigned long long, unsigned long long>, int> key2rep;
  std::vector<int> rep_of(t);
  for (int y = 0; y < t; ++y) {
    auto key = std::make_tuple(ymiss[y], hash[2 * y], hash[2 * y + 1]);
    auto it = key2rep.find(key);
    if (it == key2rep.end()) {
      key2rep.emplace(key, y);
      rep_of[y] = y;
    } else {
      rep_of[y] = it->second;
    }
  }
  // Only a trait grouped with another and missing some rows can mismatch
  // its group's rows; without one (nothing missing, or every pattern
  // distinct) the device check and its round trip are skipped.
  bool verify = false;
  for (int y = 0; y < t && !verify; ++y) verify = rep_of[y] != y && ymiss[y] > 0;
  std::vector<int> mism(t, 0);
  if (verify) {
    HIPCK(c, hipMemcpyAsync(c->scratch_i, rep_of.data(), sizeof(int) * t, hipMemcpyHostToDevice, c->st));
    int* mism_d = c->ypat;   // ypat is rewritten below
    launch_verify_patterns(t, n, c->rows, c->scratch_i, c->ymiss, mism_d, c->st);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipMemcpyAsync(mism.data(), mism_d, sizeof(int) * t, hipMemcpyDeviceToHost, c->st));
    HIPCK(c, hipStreamSynchronize(c->st));
  }
  std::vector<int> pat_of_rep(t, -1), ypat(t), pat_rep, pat_cnt;
  for (int y = 0; y < t; ++y) {
    const int rep = mism[y] ? y : rep_of[y];
    if (pat_of_rep[rep] < 0) {
      pat_of_rep[rep] = (int)pat_rep.size();
      pat_rep.push_back(rep);
      pat_cnt.push_back(ymiss[rep]);
    }
    ypat[y] = pat_of_rep[rep];
  }
  const int G = (int)pat_rep.size();
  c->G = G;
  c->max_pat_cnt = 0;
  for (int g = 0; g < G; ++g) c->max_pat_cnt = std::max(c->max_pat_cnt, pat_cnt[g]);
  if ((size_t)G > c->pat_cap) {
    dfree(c->pat_cnt); dfree(c->pat_rep); dfree(c->pat_pd); dfree(c->UA); dfree(c->UA_invd);
    c->pat_cap = (size_t)G;
    HIPCK(c, hipMalloc(&c->pat_cnt, sizeof(int) * G));
    HIPCK(c, hipMalloc(&c->pat_rep, sizeof(int) * G));
    HIPCK(c, hipMalloc(&c->pat_pd, sizeof(int) * G));
    HIPCK(c, hipMalloc(&c->UA, sizeof(double) * (size_t)G * (l * (l + 1) / 2)));
    HIPCK(c, hipMalloc(&c->UA_invd, sizeof(double) * (size_t)G * l));
  }
  HIPCK(c, hipMemcpyAsync(c->pat_cnt, pat_cnt.data(), sizeof(int) * G, hipMemcpyHostToDevice, c->st));
  HIPCK(c, hipMemcpyAsync(c->pat_rep, pat_rep.data(), sizeof(int) * G, hipMemcpyHostToDevice, c->st));
  HIPCK(c, hipMemcpyAsync(c->ypat, ypat.data(), sizeof(int) * t, hipMemcpyHostToDevice, c->st));
  {
    // Trait waves of the pattern-table pair kernel.  A pattern's traits fill
    // waves of 64 of their own, and its remainder too when it holds 32 or
    // more traits, so such a wave reads its (SNP, pattern) rows with scalar
    // loads; the smaller remainders share mixed waves, which read per lane.
    std::vector<std::vector<int>> by(G);
    for (int y = 0; y < t; ++y) by[ypat[y]].push_back(y);
    std::vector<int> wtrait, wpat, pool;
    auto emit = [&](const int* ys, size_t k, int g) {
      wpat.push_back(g);
      for (size_t i = 0; i < 64; ++i) wtrait.push_back(i < k ? ys[i] : -1);
    };
    for (int g = 0; g < G; ++g) {
      const std::vector<int>& v = by[g];
      size_t i = 0;
      for (; i + 64 <= v.size(); i += 64) emit(v.data() + i, 64, g);
      if (v.size() - i >= 32) emit(v.data() + i, v.size() - i, g);
      else pool.insert(pool.end(), v.begin() + i, v.end());
    }
    for (size_t i = 0; i < pool.size(); i += 64) emit(pool.data() + i, std::min<size_t>(64, pool.size() - i), -1);
    c->n_waves = (int)wpat.size();
    if (wpat.size() > c->wave_cap) {
      HIPCK(c, hipStreamSynchronize(c->st));
      dfree(c->wtrait); dfree(c->wpat);
      c->wave_cap = wpat.size();
      HIPCK(c, hipMalloc(&c->wtrait, sizeof(int) * 64 * c->wave_cap));
      HIPCK(c, hipMalloc(&c->wpat, sizeof(int) * c->wave_cap));
    }
    HIPCK(c, hipMemcpyAsync(c->wtrait, wtrait.data(), sizeof(int) * wtrait.size(), hipMemcpyHostToDevice, c->st));
    HIPCK(c, hipMemcpyAsync(c->wpat, wpat.data(), sizeof(int) * wpat.size(), hipMemcpyHostToDevice, c->st));
    // the host vectors go out of scope: finish the copies first
    HIPCK(c, hipStreamSynchronize(c->st));
  }
  d = make_dev(c);
  launch_pattern(d, c->Stl_pk, G, c->pat_rep, c->UA, c->UA_invd, c->pat_pd, c->st);
  HIPCK(c, hipGetLastError());
  {
    int rc = build_union(c, pat_rep, pat_cnt);
    if (rc) return rc;
  }
  d = make_dev(c);
  launch_trait_solve(d, c->lb, c->zt, c->ay, c->rss0, c->st);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipStreamSynchronize(c->st));
  c->have_traits = true;
  c->bf_ok = false;
  return OAGPU_OK;
}

// U = union of the patterns' missing rows, found on the device; uidx maps a
// row to its position in U, the covariate rows of U are copied, and each
// pattern's rows are listed as positions in U.  The corrections then read
// the genotype values at U from the compact u-major buffer the operand pass
// writes.
static int build_union(oagpu_ctx* c, const std::vector<int>& pat_rep, const std::vector<int>& pat_cnt) {
  const int n = c->n, l = c->l, G = (int)pat_rep.size();
  c->compact_corr = false;
  c->nU = 0;
  if (c->max_pat_cnt == 0) return OAGPU_OK;
  std::vector<int> off(G + 1, 0);
  for (int g = 0; g < G; ++g) off[g + 1] = off[g] + pat_cnt[g];
  const size_t tot = (size_t)off[G];
  HIPCK(c, grow((void**)&c->uidx, &c->uidx_cap, sizeof(int) * (size_t)c->n_pad));
  HIPCK(c, grow((void**)&c->pu_off, &c->pu_off_cap, sizeof(int) * off.size()));
  HIPCK(c, grow((void**)&c->pu_list, &c->pu_list_cap, sizeof(int) * std::max<size_t>(tot, 1)));
  HIPCK(c, hipMemcpyAsync(c->pu_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, c->st));
  std::vector<unsigned char> mk(n, 1);
  if (!c->union_identity) {
    // mark the union in the uidx buffer's bytes, then number it on the host
    unsigned char* mark = reinterpret_cast<unsigned char*>(c->uidx);
    HIPCK(c, hipMemsetAsync(mark, 0, (size_t)n, c->st));
    launch_union_mark(n, G, c->rows, c->pat_rep, c->pat_cnt, mark, c->st);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipMemcpyAsync(mk.data(), mark, (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPCK(c, hipStreamSynchronize(c->st));
  }
  std::vector<int> uidx(c->n_pad, -1);
  int nU = 0;
  for (int k = 0; k < n; ++k)
    if (mk[k]) uidx[k] = nU++;
  std::vector<double> XLu((size_t)nU * l);
  for (int k = 0; k < n; ++k)
    if (uidx[k] >= 0)
      for (int j = 0; j < l; ++j) XLu[(size_t)uidx[k] * l + j] = c->XL_rm_h[(size_t)k * l + j];
  HIPCK(c, grow((void**)&c->XLu, &c->XLu_cap, sizeof(double) * std::max<size_t>(XLu.size(), 1)));
  HIPCK(c, hipMemcpyAsync(c->XLu, XLu.data(), sizeof(double) * XLu.size(), hipMemcpyHostToDevice, c->st));
  HIPCK(c, hipMemcpyAsync(c->uidx, uidx.data(), sizeof(int) * uidx.size(), hipMemcpyHostToDevice, c->st));
  launch_union_list(n, G, c->rows, c->pat_rep, c->pat_cnt, c->pu_off, c->uidx, c->pu_list, c->st);
  HIPCK(c, grow((void**)&c->XLp, &c->XLp_cap, sizeof(double) * std::max<size_t>(tot * l, 1)));
  launch_pattern_rows(c->pu_list, (long long)tot, l, c->XLu, c->XLp, c->st);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipStreamSynchronize(c->st));   // the host vectors above are released on return
  c->nU = nU;
  c->compact_corr = true;
  return OAGPU_OK;
}

extern "C" int oagpu_load_genotypes(oagpu_ctx* c, const void* G, int dtype, int where, long long ld,
                                    int n_src, int m, const int32_t* col_src, const int32_t* col_env,
                                    int snp_offset) {
  if (!c) return OAGPU_ERR_ARG;
  if (!G || m <= 0 || n_src <= 0 || ld < c->n || dtype < 0 || dtype > 2)
    return fail(c, OAGPU_ERR_ARG, "bad genotype block");
  HIPCK(c, hipSetDevice(c->prm.device));
  const int r = c->r;
  // Under a dosage model the descriptors index the combined columns.
  int n_col = n_src;
  if (c->dos_on) {
    if (n_src % c->dos.file_r != 0)
      return fail(c, OAGPU_ERR_ARG, "dosage model: the block is not a whole number of file SNPs");
    n_col = (n_src / c->dos.file_r) * (c->dos.sum ? 1 : c->dos.n_used);
  }
  if (!col_src && (long long)m * r > n_col) return fail(c, OAGPU_ERR_ARG, "m*r exceeds the source columns");
  if (col_src)
    for (long long i = 0; i < (long long)m * r; ++i)
      if (col_src[i] < 0 || col_src[i] >= n_col) return fail(c, OAGPU_ERR_ARG, "column descriptor out of range");
  if (col_env) {
    for (long long i = 0; i < (long long)m * r; ++i)
      if (col_env[i] >= c->n_env) return fail(c, OAGPU_ERR_ARG, "interaction column out of range");
  }
  // The other slot; wait until the scan that last read it has finished.
  const int slot = c->gslot ^ 1;
  if (c->free_valid[slot]) HIPCK(c, hipStreamWaitEvent(c->cst, c->free_ev[slot], 0));
  c->gdt = dtype;
  c->ldg = ld;
  c->n_src = n_src;
  c->m = m;
  c->snp_offset = snp_offset;
  if (where == OAGPU_HOST) {
    const size_t bytes = dt_size(dtype) * ((size_t)ld * (n_src - 1) + c->n);
    if (c->G_slot_cap[slot] < bytes) {
      HIPCK(c, hipStreamSynchronize(c->cst));
      HIPCK(c, grow(&c->G_slot[slot], &c->G_slot_cap[slot], bytes));
    }
    HIPCK(c, hipMemcpyAsync(c->G_slot[slot], G, bytes, hipMemcpyHostToDevice, c->cst));
    c->Gm = c->G_slot[slot];
  } else {
    c->Gm = G;
  }
  if (c->dos_on) {   // combine on the copy stream; the scan waits for up_ev
    const long long ldo = ((long long)c->n + 3) & ~3ll;
    const size_t bytes = sizeof(double) * (size_t)ldo * n_col;
    if (c->G_dos_cap[slot] < bytes) {
      HIPCK(c, hipStreamSynchronize(c->cst));
      HIPCK(c, grow(&c->G_dos[slot], &c->G_dos_cap[slot], bytes));
    }
    launch_dosage(c->Gm, dtype, ld, c->n, n_src / c->dos.file_r, c->dos, static_cast<double*>(c->G_dos[slot]), ldo,
                  c->cst);
    HIPCK(c, hipGetLastError());
    c->Gm = c->G_dos[slot];
    c->gdt = OAGPU_F64;
    c->ldg = ldo;
    c->n_src = n_col;
  }
  const size_t nd = (size_t)m * r;
  if (col_src || col_env) {
    if (nd > c->d_desc_cap[slot]) {
      HIPCK(c, hipStreamSynchronize(c->cst));
      dfree(c->d_src[slot]);
      dfree(c->d_env[slot]);
      c->d_src[slot] = c->d_env[slot] = nullptr;
      HIPCK(c, hipMalloc(&c->d_src[slot], sizeof(int) * nd));
      HIPCK(c, hipMalloc(&c->d_env[slot], sizeof(int) * nd));
      c->d_desc_cap[slot] = nd;
    }
    std::vector<int>& h = c->h_desc[slot];
    h.resize(2 * nd);
    for (size_t i = 0; i < nd; ++i) {
      h[i] = col_src ? col_src[i] : (int)i;
      h[nd + i] = col_env ? col_env[i] : -1;
    }
    HIPCK(c, hipMemcpyAsync(c->d_src[slot], h.data(), sizeof(int) * nd, hipMemcpyHostToDevice, c->cst));
    HIPCK(c, hipMemcpyAsync(c->d_env[slot], h.data() + nd, sizeof(int) * nd, hipMemcpyHostToDevice, c->cst));
    c->col_src = c->d_src[slot];
    c->col_env = c->d_env[slot];
  } else {
    // identity descriptors: kernels read col_src == nullptr as s*r + c
    c->col_src = c->col_env = nullptr;
  }
  c->pair_env = -1;
  if (r == 2 && col_src && col_env && col_env[0] >= 0) {
    bool pair = true;
    for (long long s = 0; s < m && pair; ++s)
      pair = col_src[2 * s] == col_src[2 * s + 1] && col_env[2 * s] == col_env[0] && col_env[2 * s + 1] < 0;
    if (pair) c->pair_env = col_env[0];
  }
  HIPCK(c, hipEventRecord(c->up_ev[slot], c->cst));
  c->gslot = slot;
  c->upload_pending = true;
  c->have_geno = true;
  return OAGPU_OK;
}

extern "C" int oagpu_set_dosage_model(oagpu_ctx* c, int mode, int file_r, int n_used, const double* factors) {
  if (!c) return OAGPU_ERR_ARG;
  if (mode == OAGPU_DOSAGE_NONE) {
    c->dos_on = false;
    return OAGPU_OK;
  }
  if (mode != OAGPU_DOSAGE_SUM && mode != OAGPU_DOSAGE_SCALE)
    return fail(c, OAGPU_ERR_ARG, "unknown dosage model mode");
  if (!factors || file_r < 1 || n_used < 1 || n_used > file_r || n_used > OAGPU_MAX_DOSAGE_FACTORS)
    return fail(c, OAGPU_ERR_ARG, "dosage model: need 1 <= n_used <= file_r and n_used <= 16 factors");
  static_assert(sizeof(c->dos.f) / sizeof(double) == OAGPU_MAX_DOSAGE_FACTORS, "factor capacity");
  c->dos = DosageModel{};
  c->dos.file_r = file_r;
  c->dos.n_used = n_used;
  c->dos.sum = mode == OAGPU_DOSAGE_SUM;
  for (int i = 0; i < n_used; ++i) c->dos.f[i] = factors[i];
  c->dos_on = true;
  return OAGPU_OK;
}

extern "C" int oagpu_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return OAGPU_ERR_ARG;
  *ptr = nullptr;
  return hipHostMalloc(ptr, bytes > 0 ? bytes : 64, hipHostMallocDefault) == hipSuccess ? OAGPU_OK : OAGPU_ERR_HIP;
}

extern "C" void oagpu_host_free(void* ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int oa::ensure_pair_buffers(oagpu_ctx* c, int mc) {
  const int r = c->r, l = c->l, p = c->p, G = c->G, Q = c->Q;
  const size_t qbytes = sizeof(double) * (size_t)mc * Q * std::max(G, 1);
  if (mc > c->mc_cap) {
    dfree(c->col_mean); dfree(c->col_nan); dfree(c->Sbr); dfree(c->Str); dfree(c->varx);
    c->col_nan = nullptr;
    HIPCK(c, hipMalloc(&c->col_mean, sizeof(double) * (size_t)mc * r));
    HIPCK(c, hipMalloc(&c->col_nan, sizeof(int) * (size_t)mc * r));
    HIPCK(c, hipMalloc(&c->Sbr, sizeof(double) * (size_t)mc * (r * (r + 1) / 2)));
    HIPCK(c, hipMalloc(&c->Str, sizeof(double) * (size_t)mc * l * r));
    HIPCK(c, hipMalloc(&c->varx, sizeof(double) * (size_t)mc * p));
    c->mc_cap = mc;
  }
  HIPCK(c, grow((void**)&c->Ccorr, &c->Ccorr_cap, qbytes));
  if (c->use_pst)
    HIPCK(c, grow((void**)&c->pst, &c->pst_cap,
                  sizeof(double) * (size_t)mc * G * patsnp_width(r, c->l, c->lb)));
  return OAGPU_OK;
}

static int ensure_chunk(oagpu_ctx* c, int mc) {
  const int r = c->r;
  const size_t pad = (size_t)gemm_row_pad();
  const size_t arows = ((size_t)mc * r + pad - 1) / pad * pad;
  const size_t abytes = arows * c->ldop * c->op_size;
  const size_t cbytes = (size_t)mc * r * out_cols(c) * c->c_size;
  // 32 elements of padding: the corrections read 512-byte pieces of many
  // rows at once, and a row stride off the chunk size's round multiple
  // measured 1.5 % faster at n = 5,000 (100,000 SNPs per chunk)
  c->ldxu = ((long long)mc * r + 7) / 8 * 8 + 32;
  // rows 0..nU-1 of U, then a zero row (nU) for the corrections' padding lanes
  const size_t xubytes = c->compact_corr ? xu_size(c) * ((size_t)c->nU + 1) * c->ldxu : 0;
  int rc = ensure_pair_buffers(c, mc);
  if (rc) return rc;
  HIPCK(c, grow(&c->Abuf, &c->A_cap, abytes));
  c->Ahi = c->Abuf;
  c->Alo = c->prm.precision == OAGPU_BF16X3 ? static_cast<uint16_t*>(c->Abuf) + c->n_pad : nullptr;
  HIPCK(c, grow(&c->Cbuf, &c->C_cap, cbytes));
  if (xubytes) HIPCK(c, grow((void**)&c->Xu, &c->Xu_cap, xubytes));
  return OAGPU_OK;
}

// Contraction path per shape class: the fused split-BF16 kernel (the
// operand pass then writes no operand) or the operand pass plus the
// autotuned contraction.  Chosen by timing both once on the real operands;
// OAGPU_GEMM_VARIANT pins one (the fused kernel's name, or any other).
namespace {
struct PathKey {
  int mclass, nclass, k, gdt, r;
  bool operator<(const PathKey& o) const {
    return std::tie(mclass, nclass, k, gdt, r) < std::tie(o.mclass, o.nclass, o.k, o.gdt, o.r);
  }
};
std::map<PathKey, bool> g_path;
std::mutex g_path_mu;
PathKey path_key(const oagpu_ctx* c, int M, int N) {
  int nc = 64;
  while (nc < N) nc *= 2;
  return PathKey{M >= 65536 ? 2 : (M >= 4096 ? 1 : 0), nc, c->n_pad, c->gdt, c->r};
}
}  // namespace

static bool fused_capable(const oagpu_ctx* c) {
  return c->prm.precision == OAGPU_BF16X3 && !c->keep && c->pair_env >= 0 && c->col_src &&
         gemm_fused_supported(c->gdt, c->Gm, c->ldg, c->n, c->n_pad);
}

static FusedGemmArgs fused_args(const oagpu_ctx* c, const DevScan& d, int M) {
  FusedGemmArgs a{};
  a.G = d.Gm; a.ldg = d.ldg; a.n = c->n; a.K = c->n_pad;
  a.col_src = d.col_src; a.s0 = d.s0;
  a.E = c->E32 + (size_t)c->pair_env * c->n_pad;
  a.col_mean = d.col_mean;
  a.col_nan = d.col_nan;
  a.B = c->Bf; a.ldb = c->ldop;
  a.C = static_cast<float*>(const_cast<void*>(d.Cbuf)); a.ldc = out_cols(c);
  a.M = M; a.N = out_cols(c);
  return a;
}

int oa::enqueue_chunks(oagpu_ctx* c, DevScan& d, long long mc, void* Ahi, void* Alo, hipEvent_t* E,
                       bool fixed_gemm) {
  const int r = c->r, m = c->m;
  launch_reset_counters(d.counters, CNT_N, c->st);
  const int nchunks = (int)((m + mc - 1) / mc);
  for (int ci = 0; ci < nchunks; ++ci) {
    const int s0 = (int)(ci * mc);
    const int mcc = (int)std::min<long long>(mc, m - s0);
    d.s0 = s0;
    d.mc = mcc;
    hipEvent_t* Ec = E ? E + (size_t)ci * 7 : nullptr;
    void* C = const_cast<void*>(d.Cbuf);
    const int M = mcc * r, N = out_cols(c);
    auto run_lib = [&]() {
      launch_prep(d, c->prm.precision, Ahi, Alo, c->compact_corr, c->st);
      launch_gemm(c->prm.precision, Ahi, Alo, c->Bhi, c->Blo, C, M, N, c->n_pad, (int)c->ldop, (int)c->ldop, N,
                  c->st);
    };
    auto run_fused = [&]() {
      launch_prep_pair(d, c->compact_corr, c->st);
      launch_gemm_fused(fused_args(c, d, M), c->gdt, c->st);
    };
    bool fused = false;
    if (!fixed_gemm && fused_capable(c)) {
      if (!c->bf_ok) {
        const int rows = out_cols(c);
        HIPCK(c, grow((void**)&c->Bf, &c->Bf_cap, (size_t)rows * c->ldop * sizeof(uint16_t)));
        launch_fused_b_layout(static_cast<const uint16_t*>(c->Blo), c->ldop, c->n_pad, rows, c->Bf, c->st);
        HIPCK(c, hipGetLastError());
        c->bf_ok = true;
      }
      const char* pin = getenv("OAGPU_GEMM_VARIANT");
      if (pin) {
        fused = !strcmp(pin, kFusedVariant);
      } else {
        const PathKey key = path_key(c, M, N);
        int known = -1;
        {
          std::lock_guard<std::mutex> lk(g_path_mu);
          auto it = g_path.find(key);
          if (it != g_path.end()) known = it->second ? 1 : 0;
        }
        if (known < 0) {
          // time both paths once (each after a warm-up run, which also lets
          // the contraction autotuner settle), keep the faster
          hipEvent_t t0, t1, t2, t3;
          HIPCK(c, hipEventCreate(&t0));
          HIPCK(c, hipEventCreate(&t1));
          HIPCK(c, hipEventCreate(&t2));
          HIPCK(c, hipEventCreate(&t3));
          run_lib();
          HIPCK(c, hipEventRecord(t0, c->st));
          run_lib();
          HIPCK(c, hipEventRecord(t1, c->st));
          run_fused();
          HIPCK(c, hipEventRecord(t2, c->st));
          run_fused();
          HIPCK(c, hipEventRecord(t3, c->st));
          HIPCK(c, hipEventSynchronize(t3));
          float ms_lib = 0.f, ms_fused = 0.f;
          HIPCK(c, hipEventElapsedTime(&ms_lib, t0, t1));
          HIPCK(c, hipEventElapsedTime(&ms_fused, t2, t3));
          for (hipEvent_t e : {t0, t1, t2, t3}) (void)hipEventDestroy(e);
          known = ms_fused <= ms_lib ? 1 : 0;
          std::lock_guard<std::mutex> lk(g_path_mu);
          g_path[key] = known == 1;
        }
        fused = known == 1;
      }
    }
    c->last_fused = fused;
    if (Ec) HIPCK(c, hipEventRecord(Ec[0], c->st));
    if (fused) {
      launch_prep_pair(d, c->compact_corr, c->st);
      if (Ec) HIPCK(c, hipEventRecord(Ec[1], c->st));
      launch_gemm_fused(fused_args(c, d, M), c->gdt, c->st);
    } else {
      launch_prep(d, c->prm.precision, Ahi, Alo, c->compact_corr, c->st);
      if (Ec) HIPCK(c, hipEventRecord(Ec[1], c->st));
      if (fixed_gemm)
        launch_gemm_fixed(c->prm.precision, Ahi, Alo, c->Bhi, c->Blo, C, M, N, c->n_pad, (int)c->ldop,
                          (int)c->ldop, N, c->st);
      else
        launch_gemm(c->prm.precision, Ahi, Alo, c->Bhi, c->Blo, C, M, N, c->n_pad, (int)c->ldop, (int)c->ldop, N,
                    c->st);
    }
    if (Ec) HIPCK(c, hipEventRecord(Ec[2], c->st));
    launch_snp_gram(d, c->prm.precision, c->lb, c->st);
    if (Ec) HIPCK(c, hipEventRecord(Ec[3], c->st));
    if (c->compact_corr) launch_corr(d, c->lb, c->st);
    if (c->use_pst) launch_patsnp(d, c->lb, c->st);
    if (Ec) HIPCK(c, hipEventRecord(Ec[4], c->st));
    launch_reset_counter(d.counters, CNT_FB_CHUNK, c->st);
    if (c->use_pst) launch_epilogue_pst(d, c->prm.precision, c->lb, c->st);
    else launch_epilogue(d, c->prm.precision, c->lb, c->st);
    if (Ec) HIPCK(c, hipEventRecord(Ec[5], c->st));
    launch_fallback(d, c->lb, c->st);
    if (Ec) HIPCK(c, hipEventRecord(Ec[6], c->st));
    HIPCK(c, hipGetLastError());
  }
  return OAGPU_OK;
}

static int scan_enqueue(oagpu_ctx* c, bool want_dense) {
  if (c->pending) return fail(c, OAGPU_ERR_STATE, "a scan is already in flight (oagpu_scan_wait first)");
  if (!c->have_traits || !c->have_geno) return fail(c, OAGPU_ERR_STATE, "load traits and genotypes before scanning");
  HIPCK(c, hipSetDevice(c->prm.device));
  const int r = c->r, p = c->p, t = c->t, m = c->m, l = c->l;
  // The per-(SNP, pattern) pieces pay off when patterns are shared by
  // several traits.
  c->use_pst = (long long)c->G * 4 <= t;
  // Chunk size: bounded by max_snps and by a buffer budget of a quarter of
  // the free device memory, at most 32 GiB (the 200,000-SNP G x E scan at
  // n = 10,000 fits one ~22 GB chunk on a 288 GB MI355X; one chunk instead
  // of two measured 1.5 % faster there) and at least 1 GiB.  Fixed at the
  // context's first scan, so later scans (after its buffers are allocated)
  // chunk the same way.
  if (c->chunk_budget == 0) {
    size_t free_b = 0, total_b = 0;
    HIPCK(c, hipMemGetInfo(&free_b, &total_b));
    c->chunk_budget = std::max<size_t>(1ull << 30, std::min<size_t>(32ull << 30, free_b / 4));
  }
  const size_t budget = c->chunk_budget;
  const size_t per_snp = (size_t)r * c->n_pad * c->op_size * (c->prm.precision == OAGPU_BF16X3 ? 2 : 1) +
                         (size_t)r * out_cols(c) * c->c_size +
                         sizeof(double) * (size_t)c->Q * std::max(c->G, 1) +
                         sizeof(double) * (size_t)(r + r * (r + 1) / 2 + l * r + p) +
                         (c->compact_corr ? xu_size(c) * (size_t)c->nU * r : 0) +
                         (c->use_pst ? sizeof(double) * (size_t)c->G * patsnp_width(r, c->l, c->lb) : 0);
  long long mc = c->prm.max_snps > 0 ? c->prm.max_snps : (long long)(budget / per_snp);
  mc = std::max(1ll, std::min(mc, (long long)m));
  mc = std::min(mc, (long long)(0x7fffffff / std::max(1, r * out_cols(c))));
  int rc = ensure_chunk(c, (int)mc);
  if (rc) return rc;
  if (c->compact_corr)
    HIPCK(c, hipMemsetAsync(static_cast<char*>(c->Xu) + xu_size(c) * (size_t)c->nU * c->ldxu, 0,
                            xu_size(c) * (size_t)c->ldxu, c->st));

  const size_t npair = (size_t)t * m;
  if (want_dense) {
    HIPCK(c, hipMalloc(&c->pd_beta, sizeof(double) * npair * p));
    HIPCK(c, hipMalloc(&c->pd_se, sizeof(double) * npair * p));
    HIPCK(c, hipMalloc(&c->pd_t, sizeof(double) * npair * p));
    HIPCK(c, hipMalloc(&c->pd_p, sizeof(double) * npair * p));
    HIPCK(c, hipMalloc(&c->pd_kept, npair * p));
    HIPCK(c, hipMalloc(&c->pd_r2, sizeof(double) * npair));
    HIPCK(c, hipMalloc(&c->pd_rss, sizeof(double) * npair));
    HIPCK(c, hipMalloc(&c->pd_em, npair));
    HIPCK(c, hipMalloc(&c->pd_varx, sizeof(double) * (size_t)m * p));
  }
  if (c->upload_pending) {
    HIPCK(c, hipStreamWaitEvent(c->st, c->up_ev[c->gslot], 0));
    c->upload_pending = false;
  }
  const int nchunks = (int)((m + mc - 1) / mc);
  // timing events, kept across scans
  while (c->pev.size() < (size_t)nchunks * 7) {
    hipEvent_t e = nullptr;
    HIPCK(c, hipEventCreate(&e));
    c->pev.push_back(e);
  }
  DevScan d = make_dev(c);
  d.need_bl = c->prm.disp_cov || want_dense;
  d.d_beta = c->pd_beta; d.d_se = c->pd_se; d.d_t = c->pd_t; d.d_p = c->pd_p; d.d_kept = c->pd_kept;
  d.d_r2 = c->pd_r2; d.d_rss = c->pd_rss; d.d_emitted = c->pd_em; d.d_varx = c->pd_varx;
  rc = enqueue_chunks(c, d, mc, c->Ahi, c->Alo, c->pev.data(), false);
  if (rc) return rc;
  c->gemm_variant = c->last_fused ? kFusedVariant
                                  : gemm_variant_name(c->prm.precision, (int)std::min<long long>(mc, m) * r,
                                                      out_cols(c), c->n_pad);
  // the genotype slot may be overwritten once everything above has run
  HIPCK(c, hipEventRecord(c->free_ev[c->gslot], c->st));
  c->free_valid[c->gslot] = true;
  HIPCK(c, hipMemcpyAsync(c->cnt_h, c->counters, sizeof(unsigned long long) * CNT_N, hipMemcpyDeviceToHost, c->st));
  HIPCK(c, hipEventRecord(c->done_ev, c->st));
  c->pending = true;
  c->p_timed = true;
  c->p_dense = want_dense;
  c->p_nchunks = nchunks;
  c->p_t = t;
  c->p_m = m;
  return OAGPU_OK;
}

int oa::scan_collect(oagpu_ctx* c, oagpu_dense* dense, oagpu_hit* hits, long long hit_cap,
                     oagpu_stats* stats) {
  if (!c->pending) return fail(c, OAGPU_ERR_STATE, "no scan in flight");
  c->pending = false;
  HIPCK(c, hipEventSynchronize(c->done_ev));
  const int p = c->p, t = c->p_t, m = c->p_m, nchunks = c->p_nchunks;
  for (int i = 0; i < 8; ++i) c->timings[i] = 0.0;
  for (int ci = 0; ci < (c->p_timed ? nchunks : 0); ++ci) {
    hipEvent_t* E = &c->pev[(size_t)ci * 7];
    for (int k = 0; k < 6; ++k) {
      float ms = 0.f;
      HIPCK(c, hipEventElapsedTime(&ms, E[k], E[k + 1]));
      c->timings[k] += ms;
    }
  }
  if (c->p_timed) {
    float ms = 0.f;
    HIPCK(c, hipEventElapsedTime(&ms, c->pev[0], c->pev[(size_t)(nchunks - 1) * 7 + 6]));
    c->timings[6] = ms;
  }
  unsigned long long cnt[CNT_N];
  std::memcpy(cnt, c->cnt_h, sizeof(cnt));
  const long long nh = (long long)cnt[CNT_HITS];
  if (stats) {
    stats->n_hits = nh;
    stats->n_emitted = (long long)cnt[CNT_EMITTED];
    stats->n_sig = (long long)cnt[CNT_SIG];
    stats->n_fallback = (long long)cnt[CNT_FALLBACK];
    stats->n_patterns = c->G;
    stats->n_chunks = nchunks;
    stats->t_star = oagpu_tvalue(c->prm.min_p_store);
  }
  int status = OAGPU_OK;
  if (cnt[CNT_ERR] & 1ull)
    status = fail(c, OAGPU_ERR_NUMERIC, "Error with Syy < 0 or var_x < 0 (the reference exits here)");
  if (c->p_dense) {
    const size_t npair = (size_t)t * m;
    auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
      if (!dst || !dense) return hipSuccess;
      return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    };
    if (dense) {
      HIPCK(c, cp(dense->beta, c->pd_beta, sizeof(double) * npair * p));
      HIPCK(c, cp(dense->se, c->pd_se, sizeof(double) * npair * p));
      HIPCK(c, cp(dense->tstat, c->pd_t, sizeof(double) * npair * p));
      HIPCK(c, cp(dense->pval, c->pd_p, sizeof(double) * npair * p));
      HIPCK(c, cp(dense->kept, c->pd_kept, npair * p));
      HIPCK(c, cp(dense->r2, c->pd_r2, sizeof(double) * npair));
      HIPCK(c, cp(dense->rss, c->pd_rss, sizeof(double) * npair));
      HIPCK(c, cp(dense->emitted, c->pd_em, npair));
      HIPCK(c, cp(dense->var_x, c->pd_varx, sizeof(double) * (size_t)m * p));
      HIPCK(c, cp(dense->n_used, c->ymiss, sizeof(int) * t));
      if (dense->n_used)
        for (int y = 0; y < t; ++y) dense->n_used[y] = (c->keep ? c->n_stat : c->n) - dense->n_used[y];
    }
    void* bufs[] = {c->pd_beta, c->pd_se, c->pd_t, c->pd_p, c->pd_kept, c->pd_r2, c->pd_rss, c->pd_em, c->pd_varx};
    for (void* b : bufs) dfree(b);
    c->pd_beta = c->pd_se = c->pd_t = c->pd_p = c->pd_r2 = c->pd_rss = c->pd_varx = nullptr;
    c->pd_kept = c->pd_em = nullptr;
  }
  if (nh > c->hit_cap || (hits && nh > hit_cap)) {
    return fail(c, OAGPU_ERR_HIT_OVERFLOW,
                "hit arena too small: " + std::to_string(nh) + " hits (capacity " +
                    std::to_string(std::min(c->hit_cap, hits ? hit_cap : c->hit_cap)) + ")");
  }
  if (hits && nh > 0) {
    HIPCK(c, hipMemcpy(hits, c->hits, sizeof(oagpu_hit) * (size_t)nh, hipMemcpyDeviceToHost));
    std::sort(hits, hits + nh, [](const oagpu_hit& a, const oagpu_hit& b) {
      if (a.trait != b.trait) return a.trait < b.trait;
      if (a.snp != b.snp) return a.snp < b.snp;
      return a.h > b.h;
    });
  }
  return status;
}

static bool dense_requested(const oagpu_dense* dense) {
  return dense && (dense->beta || dense->se || dense->tstat || dense->pval || dense->kept || dense->r2 ||
                   dense->rss || dense->emitted || dense->var_x || dense->n_used);
}

extern "C" int oagpu_scan(oagpu_ctx* c, oagpu_dense* dense, oagpu_hit* hits, long long hit_cap,
                          oagpu_stats* stats) {
  if (!c) return OAGPU_ERR_ARG;
  const int rc = scan_enqueue(c, dense_requested(dense));
  if (rc != OAGPU_OK) return rc;
  return scan_collect(c, dense, hits, hit_cap, stats);
}

extern "C" int oagpu_scan_async(oagpu_ctx* c, int want_dense) {
  if (!c) return OAGPU_ERR_ARG;
  return scan_enqueue(c, want_dense != 0);
}

extern "C" int oagpu_scan_wait(oagpu_ctx* c, oagpu_dense* dense, oagpu_hit* hits, long long hit_cap,
                               oagpu_stats* stats) {
  if (!c) return OAGPU_ERR_ARG;
  return scan_collect(c, dense, hits, hit_cap, stats);
}

extern "C" int oagpu_last_timings(const oagpu_ctx* c, double* out, int n) {
  if (!c || !out) return 0;
  const int k = std::min(n, 7);
  for (int i = 0; i < k; ++i) out[i] = c->timings[i];
  return k;
}

extern "C" const char* oagpu_gemm_variant(const oagpu_ctx* c) { return c ? c->gemm_variant.c_str() : ""; }

extern "C" int oagpu_gemm_tunings(void) { return gemm_tunings(); }
