/*
 * FP32 restatement of the reference's blocked CPU algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the
 * GPU scan (its cpu_baseline leg), held to the FP64 oracle by
 * tests/test_ref32.py.  Nothing in the product path links or loads it.
 *
 * It follows the structure and the precision of partialNEQ_Blocked_STL_MD
 * (src/Algorithm.cpp:322-886), where the FP64 oracle (oracle.cpp) follows
 * only its arithmetic:
 *   Stl = XL'XL (cblas_ssyrk) ............. src/Algorithm.cpp:490-491
 *   per trait block of tb traits:
 *     NaN -> 0, missing-row lists ........ src/Utility.cpp:130-179
 *     Ay_top = XL'Y (cblas_sgemm) ........ src/Algorithm.cpp:583-585
 *     SST, Stl corrections ............... src/Algorithm.cpp:1222-1251, 1324-1362
 *     per SNP block of mb SNPs:
 *       XR NaN -> column mean ............ src/Utility.cpp:103-127
 *       Str = XL'XR, Sbr (sgemm/ssyrk) ... src/Algorithm.cpp:633-656
 *       var_x (spotrf/spotri) ............ src/Algorithm.cpp:659-671
 *       Str/Sbr corrections .............. src/Algorithm.cpp:1260-1315
 *       Ay_bot = XR'Y (cblas_sgemm) ...... src/Algorithm.cpp:705-707
 *       per (trait, SNP), OpenMP over the block's SNPs: build_S,
 *       correct_missings_S, sposv('U'), info ignored
 *                                          src/Algorithm.cpp:714-769, 60-95, 1370-1399
 *       explicit RSS over rows with y != 0, Xb accumulated in float, the
 *       sum in long double (singleSSY) .. src/Algorithm.cpp:976-1024
 *       hpc_statistics ................... src/Algorithm.cpp:1034-1214
 * Every value is a float, as in the reference, except the RSS sum.  The
 * BLAS products run through the image's MKL (cblas_sgemm of
 * /opt/conda/lib/libmkl_rt.so.1, loaded at run time, GNU OpenMP threading),
 * or, where it cannot be loaded or OA_REF32_BLAS=builtin is set, through a
 * cache-blocked FP32 kernel of this file (GCC vector extensions, OpenMP);
 * oa_ref32_blas() names the one in use.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <dlfcn.h>
#include <omp.h>
#include <cstdlib>

namespace {

typedef float f8 __attribute__((vector_size(32)));

inline float hsum(f8 v) {
    float s = 0.0f;
    for (int i = 0; i < 8; ++i) s += v[i];
    return s;
}

/* A run-time CBLAS: MKL's cblas_sgemm, loaded once. */
typedef void (*sgemm_fn)(int, int, int, int, int, int, float, const float *, int, const float *, int, float,
                         float *, int);
struct Blas {
    sgemm_fn sgemm = nullptr;
    void (*set_threads)(int) = nullptr;
    const char *name = "FP32 cache-blocked kernel of oracle/ref32.cpp (AVX2/FMA vectors, OpenMP)";
};
const Blas &blas() {
    static const Blas b = [] {
        Blas r;
        const char *force = std::getenv("OA_REF32_BLAS");
        if (force && !std::strcmp(force, "builtin")) return r;
        for (const char *path : {"/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so.1", "libmkl_rt.so"}) {
            void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
            if (!h) continue;
            auto layer = reinterpret_cast<int (*)(int)>(dlsym(h, "MKL_Set_Threading_Layer"));
            auto gemm = reinterpret_cast<sgemm_fn>(dlsym(h, "cblas_sgemm"));
            if (!gemm) continue;
            if (layer) layer(3);   // MKL_THREADING_GNU: the same OpenMP runtime as this file
            r.sgemm = gemm;
            r.set_threads = reinterpret_cast<void (*)(int)>(dlsym(h, "MKL_Set_Num_Threads"));
            r.name = "MKL cblas_sgemm (/opt/conda/lib/libmkl_rt.so.1, GNU OpenMP threads)";
            break;
        }
        return r;
    }();
    return b;
}

/* C[i + j*ldc] = sum_k A[k + i*lda] B[k + j*ldb], i < M, j < N, k < K:
 * A'B for column-major A (K x M) and B (K x N), the shape of every product
 * of the scan (columns of length n).  With a CBLAS: one sgemm (A transposed,
 * B not).  Else 4 x 4 blocks of dot products, eight k-values per vector;
 * OpenMP over the blocks. */
void gemm_tn(int M, int N, int K, const float *A, long lda, const float *B, long ldb, float *C, long ldc,
             int threads) {
    const Blas &b = blas();
    if (b.sgemm) {
        if (b.set_threads) b.set_threads(threads);
        b.sgemm(102 /* column major */, 112 /* A transposed */, 111 /* B not */, M, N, K, 1.0f, A, (int)lda, B,
                (int)ldb, 0.0f, C, (int)ldc);
        return;
    }
    const int K8 = K / 8 * 8;
    const int mb = (M + 3) / 4, nb = (N + 3) / 4;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads) collapse(2)
    for (int bi = 0; bi < mb; ++bi)
        for (int bj = 0; bj < nb; ++bj) {
            const int i0 = bi * 4, j0 = bj * 4;
            const float *a[4], *b[4];
            for (int q = 0; q < 4; ++q) {
                a[q] = A + (long)std::min(i0 + q, M - 1) * lda;
                b[q] = B + (long)std::min(j0 + q, N - 1) * ldb;
            }
            f8 acc[4][4];
            for (int x = 0; x < 4; ++x)
                for (int y = 0; y < 4; ++y) acc[x][y] = f8{};
            for (int k = 0; k < K8; k += 8) {
                f8 va[4], vb[4];
                for (int q = 0; q < 4; ++q) {
                    std::memcpy(&va[q], a[q] + k, 32);
                    std::memcpy(&vb[q], b[q] + k, 32);
                }
                for (int x = 0; x < 4; ++x)
                    for (int y = 0; y < 4; ++y) acc[x][y] += va[x] * vb[y];
            }
            for (int x = 0; x < 4; ++x)
                for (int y = 0; y < 4; ++y) {
                    if (i0 + x >= M || j0 + y >= N) continue;
                    float s = hsum(acc[x][y]);
                    for (int k = K8; k < K; ++k) s += a[x][k] * b[y][k];
                    C[(long)(i0 + x) + (long)(j0 + y) * ldc] = s;
                }
        }
}

/* spotrf('U') on a column-major p x p matrix (upper triangle: A[i + j*p],
 * i <= j); 0 or the failing column + 1. */
int potrf_u(float *A, int p) {
    for (int j = 0; j < p; ++j) {
        float ajj = A[j + j * p];
        for (int k = 0; k < j; ++k) ajj -= A[k + j * p] * A[k + j * p];
        if (!(ajj > 0.0f)) return j + 1;
        ajj = std::sqrt(ajj);
        A[j + j * p] = ajj;
        for (int i = j + 1; i < p; ++i) {
            float aji = A[j + i * p];
            for (int k = 0; k < j; ++k) aji -= A[k + j * p] * A[k + i * p];
            A[j + i * p] = aji / ajj;
        }
    }
    return 0;
}

/* sposv('U', nrhs = 1): b is overwritten by the solution, or left as it is
 * when the factorisation fails (info ignored by the reference). */
void sposv_u(float *A, int p, float *b) {
    if (potrf_u(A, p)) return;
    for (int i = 0; i < p; ++i) {
        float s = b[i];
        for (int k = 0; k < i; ++k) s -= A[k + i * p] * b[k];
        b[i] = s / A[i + i * p];
    }
    for (int i = p - 1; i >= 0; --i) {
        float s = b[i];
        for (int k = i + 1; k < p; ++k) s -= A[i + k * p] * b[k];
        b[i] = s / A[i + i * p];
    }
}

/* diag(inv(S)) from spotrf + spotri ('U'); NaN when S is not PD. */
void inv_diag(const float *S, int p, float *d) {
    std::vector<float> U(S, S + p * p), Ui(p * p, 0.0f);
    if (potrf_u(U.data(), p)) {
        for (int h = 0; h < p; ++h) d[h] = NAN;
        return;
    }
    for (int j = 0; j < p; ++j) {
        Ui[j + j * p] = 1.0f / U[j + j * p];
        for (int i = j - 1; i >= 0; --i) {
            float s = 0.0f;
            for (int k = i + 1; k <= j; ++k) s += U[i + k * p] * Ui[k + j * p];
            Ui[i + j * p] = -s / U[i + i * p];
        }
    }
    for (int i = 0; i < p; ++i) {
        float s = 0.0f;
        for (int j = i; j < p; ++j) s += Ui[i + j * p] * Ui[i + j * p];
        d[i] = s;
    }
}

double pvalue(double t) {
    /* src/Algorithm.cpp:1159-1167 */
    if (t < 1.6) return 0.5 - 0.1 * (4.4 - t) * t;
    return std::erfc(t * 0.70710678118654752440084436210484903928);
}

}  // namespace

extern "C" const char *oa_ref32_blas() { return blas().name; }

extern "C" int oa_ref32_scan(const oa_oracle_params *prm, const float *XL, const float *XR, const float *Y,
                             int tb, int mb, oa_oracle_out *out) {
    if (!prm || !XL || !XR || !Y || !out || tb <= 0 || mb <= 0) return -1;
    const int n = prm->n, l = prm->l, r = prm->r, t = prm->t, m = prm->m, p = l + r;
    const long N = n;
    const int threads = prm->threads > 0 ? prm->threads : 1;
    const int start_h = prm->disp_cov ? 1 : l;
    const float t_star = (float)prm->t_star, min_r2 = (float)prm->min_r2_store;

    std::vector<float> Stl(l * l);   /* Stl[i + j*l] */
    gemm_tn(l, l, n, XL, N, XL, N, Stl.data(), l, threads);

    int64_t n_sig = 0, n_kept = 0, n_emitted = 0;
    int status = 0;
    std::vector<float> Yb, Ay_top, Stl_corr, sst_b, XRb, Str, Sbr, Ay_bot, varx, Str_c, Sbr_c, B;
    std::vector<std::vector<int>> miss;
    for (int y0 = 0; y0 < t; y0 += tb) {
        const int yb = std::min(tb, t - y0);
        /* NaN -> 0 and the missing-row lists (replace_nans) */
        Yb.assign((size_t)yb * N, 0.0f);
        miss.assign(yb, {});
        for (int jj = 0; jj < yb; ++jj)
            for (long k = 0; k < N; ++k) {
                const float v = Y[(size_t)(y0 + jj) * N + k];
                if (std::isnan(v)) miss[jj].push_back((int)k);
                else Yb[(size_t)jj * N + k] = v;
            }
        Ay_top.assign((size_t)l * yb, 0.0f);
        gemm_tn(l, yb, n, XL, N, Yb.data(), N, Ay_top.data(), l, threads);
        /* SST over the observed rows (sumSquares, long double) */
        sst_b.assign(yb, 0.0f);
        Stl_corr.assign((size_t)yb * l * l, 0.0f);
#pragma omp parallel for num_threads(threads)
        for (int jj = 0; jj < yb; ++jj) {
            long double s = 0, s2 = 0;
            const int nn = n - (int)miss[jj].size();
            for (long k = 0; k < N; ++k) {
                s += Yb[(size_t)jj * N + k];
                s2 += (long double)Yb[(size_t)jj * N + k] * Yb[(size_t)jj * N + k];
            }
            sst_b[jj] = (float)(s2 - s * s / nn);
            float *c = &Stl_corr[(size_t)jj * l * l];
            for (int i = 0; i < l; ++i)
                for (int j = 0; j < l; ++j) {
                    float a = 0.0f;
                    for (int k : miss[jj]) a += XL[i * N + k] * XL[j * N + k];
                    c[i + j * l] = a;
                }
            if (out->n_used) out->n_used[y0 + jj] = nn;
        }

        for (int s0 = 0; s0 < m; s0 += mb) {
            const int sb = std::min(mb, m - s0);
            const int cols = sb * r;
            /* XR block, NaN -> column mean */
            XRb.assign((size_t)cols * N, 0.0f);
#pragma omp parallel for num_threads(threads)
            for (int c = 0; c < cols; ++c) {
                const float *src = XR + ((size_t)s0 * r + c) * N;
                float *dst = &XRb[(size_t)c * N];
                float sum = 0.0f;
                int nn = 0;
                for (long k = 0; k < N; ++k) {
                    if (std::isnan(src[k])) ++nn;
                    else sum += src[k];
                }
                const float avg = sum / (float)(n - nn);
                for (long k = 0; k < N; ++k) dst[k] = std::isnan(src[k]) ? avg : src[k];
            }
            Str.assign((size_t)l * cols, 0.0f);   /* Str[j + c*l] */
            gemm_tn(l, cols, n, XL, N, XRb.data(), N, Str.data(), l, threads);
            Sbr.assign((size_t)sb * r * r, 0.0f);
            Ay_bot.assign((size_t)cols * yb, 0.0f);   /* Ay_bot[c + jj*cols] */
            gemm_tn(cols, yb, n, XRb.data(), N, Yb.data(), N, Ay_bot.data(), cols, threads);
            varx.assign((size_t)sb * p, 0.0f);
            B.assign((size_t)yb * sb * p, 0.0f);
#pragma omp parallel num_threads(threads)
            {
                std::vector<float> S(p * p), Sfull(p * p), b(p), sc(l * r), sbc(r * r);
#pragma omp for schedule(dynamic, 1)
                for (int ii = 0; ii < sb; ++ii) {
                    const float *xr = &XRb[(size_t)ii * r * N];
                    float *sbr = &Sbr[(size_t)ii * r * r];
                    for (int i = 0; i < r; ++i)
                        for (int j = 0; j < r; ++j) {
                            float a = 0.0f;
                            for (long k = 0; k < N; ++k) a += xr[j * N + k] * xr[i * N + k];
                            sbr[i * r + j] = a;
                        }
                    /* build_S: upper triangle of [[Stl, Str], [., Sbr]] */
                    std::fill(Sfull.begin(), Sfull.end(), 0.0f);
                    for (int j = 0; j < l; ++j)
                        for (int i = 0; i <= j; ++i) Sfull[i + j * p] = Stl[i + j * l];
                    for (int c = 0; c < r; ++c) {
                        for (int j = 0; j < l; ++j) Sfull[j + (l + c) * p] = Str[j + (size_t)(ii * r + c) * l];
                        for (int d = 0; d <= c; ++d) Sfull[(l + d) + (l + c) * p] = sbr[c * r + d];
                    }
                    inv_diag(Sfull.data(), p, &varx[(size_t)ii * p]);
                    for (int jj = 0; jj < yb; ++jj) {
                        /* generate_correction_missings_Str: XR-XR terms only
                         * for j < min(l, r) */
                        std::fill(sc.begin(), sc.end(), 0.0f);
                        std::fill(sbc.begin(), sbc.end(), 0.0f);
                        for (int i = 0; i < r; ++i)
                            for (int j = 0; j < l; ++j)
                                for (int k : miss[jj]) {
                                    sc[i * l + j] += XL[j * N + k] * xr[i * N + k];
                                    if (j < r) sbc[i * r + j] += xr[j * N + k] * xr[i * N + k];
                                }
                        /* correct_missings_S, the reference's layout: every
                         * column's XL-XR correction lands in the first SNP
                         * column */
                        S = Sfull;
                        const float *c_l = &Stl_corr[(size_t)jj * l * l];
                        for (int j = 0; j < l; ++j)
                            for (int i = 0; i <= j; ++i) S[i + j * p] -= c_l[i + j * l];
                        for (int c = 0; c < r; ++c)
                            for (int d = 0; d <= c; ++d) S[(l + d) + (l + c) * p] -= sbc[c * r + d];
                        for (int c = 0; c < r; ++c)
                            for (int j = 0; j < l; ++j) S[j + l * p] -= sc[c * l + j];
                        for (int j = 0; j < l; ++j) b[j] = Ay_top[j + (size_t)jj * l];
                        for (int c = 0; c < r; ++c) b[l + c] = Ay_bot[(size_t)(ii * r + c) + (size_t)jj * cols];
                        sposv_u(S.data(), p, b.data());
                        std::copy(b.begin(), b.end(), &B[((size_t)jj * sb + ii) * p]);
                    }
                }
            }
            /* hpc_SSY + hpc_statistics */
#pragma omp parallel num_threads(threads) reduction(+ : n_sig, n_kept, n_emitted)
            {
                std::vector<float> sy(N);
#pragma omp for schedule(dynamic, 4) collapse(2)
                for (int jj = 0; jj < yb; ++jj)
                    for (int ii = 0; ii < sb; ++ii) {
                        const float *b = &B[((size_t)jj * sb + ii) * p];
                        const float *yv = &Yb[(size_t)jj * N];
                        const float *xr = &XRb[(size_t)ii * r * N];
                        std::fill(sy.begin(), sy.end(), 0.0f);
                        for (int h = 0; h < l; ++h) {
                            const float bv = b[h];
                            const float *x = XL + h * N;
                            for (long k = 0; k < N; ++k) sy[k] += x[k] * bv;
                        }
                        for (int c = 0; c < r; ++c) {
                            const float bv = b[l + c];
                            const float *x = xr + c * N;
                            for (long k = 0; k < N; ++k) sy[k] += x[k] * bv;
                        }
                        long double sum = 0;
                        for (long k = 0; k < N; ++k)
                            if (yv[k] != 0.0f) {
                                const float e = sy[k] - yv[k];
                                sum += (long double)e * e;
                            }
                        const float rss = (float)sum;
                        const int nn = n - (int)miss[jj].size();
                        const int y = y0 + jj, s = s0 + ii;
                        const size_t pair = (size_t)y * m + s;
                        /* hpc_statistics: R2 with the reference's degrees
                         * of freedom, SE from the unmasked var_x */
                        const float n_c = (float)(nn - p);
                        const float R2 = 1.0f - (rss / (n_c - 1.0f)) / (sst_b[jj] / (float)(n - 1));
                        const float sig2 = rss / n_c;
                        const float vf = (float)(nn) / (float)n;
                        if (out->r2) out->r2[pair] = R2;
                        if (out->rss) out->rss[pair] = rss;
                        bool has_snp = false;
                        int kept_here = 0, sig_here = 0;
                        for (int h = 0; h < p; ++h) {
                            const size_t o = pair * p + h;
                            if (out->beta) out->beta[o] = b[h];
                            if (out->se) out->se[o] = NAN;
                            if (out->tstat) out->tstat[o] = NAN;
                            if (out->pval) out->pval[o] = NAN;
                            if (out->kept) out->kept[o] = 0;
                        }
                        if (rss > 0.0f && (R2 > min_r2 || prm->store_all)) {
                            for (int h = p - 1; h >= start_h; --h) {
                                const float vx = varx[(size_t)ii * p + h];
                                if (vx < 0.0f) status = -2;
                                const float se = std::sqrt(sig2 * vx * vf);
                                const float tv = std::fabs(b[h] / se);
                                const size_t o = pair * p + h;
                                if (out->se) out->se[o] = se;
                                if (out->tstat) out->tstat[o] = tv;
                                if (tv > t_star || prm->store_all) {
                                    if (h >= l) has_snp = true;
                                    const double pv = pvalue(tv);
                                    if (out->pval) out->pval[o] = pv;
                                    if (out->kept) out->kept[o] = 1;
                                    ++kept_here;
                                    if (prm->min_p_disp > pv) ++sig_here;
                                }
                            }
                        }
                        if (out->emitted) out->emitted[pair] = has_snp ? 1 : 0;
                        n_sig += sig_here;
                        if (has_snp) {
                            ++n_emitted;
                            n_kept += kept_here;
                        }
                    }
            }
        }
    }
    out->n_sig = n_sig;
    out->n_kept = n_kept;
    out->n_emitted = n_emitted;
    return status;
}
