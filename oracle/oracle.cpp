/*
 * FP64 CPU oracle for the blocked normal-equation OLS scan.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP path's results are compared with,
 * and the CPU restatement timed by bench.py's cpu_baseline leg.  Nothing in the
 * product path links or loads it.
 *
 * What it restates (all in double precision; the reference computes in
 * float with a few long-double accumulators):
 *   Stl = XL'XL ............................ src/Algorithm.cpp:490-491
 *   Y NaN -> 0 + sorted NaN row lists ...... src/Utility.cpp:130-179
 *   SST (sumSquares) ....................... src/Algorithm.cpp:1222-1251
 *   Ay_top = XL'Y .......................... src/Algorithm.cpp:583-585
 *   Stl_corr (per trait) ................... src/Algorithm.cpp:1324-1362
 *   XR NaN -> column mean .................. src/Utility.cpp:103-127
 *   Str = XL'XR, Sbr = XR_s'XR_s ........... src/Algorithm.cpp:633-656
 *   build_S ................................ src/Algorithm.cpp:60-95
 *   var_x = diag(inv(S_full)) (potrf/potri)  src/Algorithm.cpp:659-671
 *   Str_corr / Sbr_corr .................... src/Algorithm.cpp:1260-1315
 *   correct_missings_S (with its quirks) ... src/Algorithm.cpp:1370-1399
 *   b = S\rhs (sposv, info ignored) ........ src/Algorithm.cpp:734-768
 *   RSS over rows with y != 0 .............. src/Algorithm.cpp:976-1024
 *   hpc_statistics ......................... src/Algorithm.cpp:1034-1214
 *   getTvalue .............................. src/Utility.cpp:244-250
 *   exclusion list + splitpair ............. src/AIOwrapper.cpp:1994-2233
 *
 * Pinning (tests/test_oracle.py): b, RSS, R2 against numpy lstsq on the
 * observed rows (r = 1, and r > 1 without missing values); var_x against
 * numpy inv; the column misplacement against a case where the misplaced
 * terms vanish and a case where they do not; SST against numpy; P
 * against scipy erfc; T* against scipy erfinv; the sposv failure rule
 * (b = rhs) on an indefinite system; the exclusion parser against the
 * shipped example file; the reference's own known-answer lower bounds on
 * its interaction fixtures (tests/test.cpp:212-308).
 * Parity unpinned: var_x when S_full is not positive definite (the
 * reference's potri result is undefined; here NaN).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Upper Cholesky of a column-major p*p matrix, LAPACK potrf('U') rule:
 * fails on the first pivot that is <= 0 or NaN.  Only the upper triangle
 * (row <= col) is read and overwritten.  Returns 0 or the failing column
 * + 1, like LAPACK's info. */
int chol_upper(double *A, int p) {
    for (int j = 0; j < p; ++j) {
        double ajj = A[j * p + j];
        for (int k = 0; k < j; ++k) ajj -= A[j * p + k] * A[j * p + k];
        if (!(ajj > 0.0)) return j + 1;
        ajj = std::sqrt(ajj);
        A[j * p + j] = ajj;
        for (int i = j + 1; i < p; ++i) {
            double aji = A[i * p + j];
            for (int k = 0; k < j; ++k) aji -= A[j * p + k] * A[i * p + k];
            A[i * p + j] = aji / ajj;
        }
    }
    return 0;
}

/* Solve U'U x = b in place (potrs('U')). */
void chol_solve_upper(const double *U, int p, double *b) {
    for (int i = 0; i < p; ++i) {          /* U' z = b */
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= U[i * p + k] * b[k];
        b[i] = s / U[i * p + i];
    }
    for (int i = p - 1; i >= 0; --i) {     /* U x = z */
        double s = b[i];
        for (int k = i + 1; k < p; ++k) s -= U[k * p + i] * b[k];
        b[i] = s / U[i * p + i];
    }
}

/* diag(inv(U'U)) from the Cholesky factor (potri('U') then the diagonal,
 * src/Algorithm.cpp:667-671). */
void chol_inv_diag(const double *U, int p, double *d) {
    std::vector<double> Ui(p * p, 0.0);    /* inverse of U, upper */
    for (int j = 0; j < p; ++j) {
        Ui[j * p + j] = 1.0 / U[j * p + j];
        for (int i = j - 1; i >= 0; --i) {
            double s = 0.0;
            for (int k = i + 1; k <= j; ++k) s += U[k * p + i] * Ui[j * p + k];
            Ui[j * p + i] = -s / U[i * p + i];
        }
    }
    for (int i = 0; i < p; ++i) {          /* (Ui Ui')_ii = sum_j Ui(i,j)^2 */
        double s = 0.0;
        for (int j = i; j < p; ++j) s += Ui[j * p + i] * Ui[j * p + i];
        d[i] = s;
    }
}

struct TraitData {
    std::vector<double> y0;        /* NaN -> 0                         */
    std::vector<int> miss;         /* sorted NaN rows                  */
    double sst = 0.0;
    std::vector<double> ay_top;    /* XL' y0 (l)                       */
    std::vector<double> stl_corr;  /* l*l, index i*l+j                 */
};

}  // namespace

extern "C" double oa_oracle_erfinv(double x) {
    /* Inverse of erf by bisection: erf(y) = x for |x| < 0.5, erfc(y) = 1 - x
     * above (1 - x is exact there and erfc keeps the relative accuracy that
     * erf loses near 1).  The bracket is halved until it no longer shrinks,
     * i.e. to adjacent doubles. */
    if (x < 0.0) return -oa_oracle_erfinv(-x);
    if (x >= 1.0) return INFINITY;
    if (x == 0.0) return 0.0;
    const bool small = x < 0.5;
    const double target = small ? x : 1.0 - x;
    double lo = 0.0, hi = small ? 0.5 : 27.3;   /* erf(0.5) > 0.5, erfc(27.3) < 1e-300 */
    for (int it = 0; it < 2000; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        const bool below = small ? (std::erf(mid) < target) : (std::erfc(mid) > target);
        if (below) lo = mid; else hi = mid;
    }
    /* the end of the bracket whose function value is closer */
    const double flo = small ? std::erf(lo) : std::erfc(lo), fhi = small ? std::erf(hi) : std::erfc(hi);
    return std::fabs(flo - target) <= std::fabs(fhi - target) ? lo : hi;
}

extern "C" double oa_oracle_tvalue(float pval) {
    /* src/Utility.cpp:247: erf_inv((long double)(1.0 - pval)) * sqrt(2.0),
     * stored into a float. */
    double v = oa_oracle_erfinv(1.0 - (double)pval) * std::sqrt(2.0);
    return (double)(float)v;
}

extern "C" double oa_oracle_pvalue(double t) {
    /* src/Algorithm.cpp:1159-1167 */
    if (t < 1.6) return 0.5 - 0.1 * (4.4 - t) * t;
    return std::erfc(t * 0.70710678118654752440084436210484903928);
}

extern "C" double oa_oracle_min_singular(const double *XL, int n, int l) {
    /* Smallest singular value of XL (n x l), the reference's test
     * (LAPACKE_sgesvd, src/Algorithm.cpp:508-522): one-sided (Hestenes)
     * Jacobi on the columns of XL itself, rotating pairs until every pair
     * is orthogonal; the singular values are then the column norms (no
     * squaring of the condition number). */
    std::vector<double> A(XL, XL + (size_t)n * l);
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < l; ++p)
            for (int q = p + 1; q < l; ++q) {
                double *cp = &A[(size_t)p * n], *cq = &A[(size_t)q * n];
                double a = 0.0, b = 0.0, g = 0.0;
                for (int k = 0; k < n; ++k) {
                    a += cp[k] * cp[k];
                    b += cq[k] * cq[k];
                    g += cp[k] * cq[k];
                }
                if (g == 0.0 || std::fabs(g) <= 1e-15 * std::sqrt(a * b)) continue;
                rotated = true;
                const double zeta = (b - a) / (2.0 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int k = 0; k < n; ++k) {
                    const double x = cp[k], y = cq[k];
                    cp[k] = c * x - s * y;
                    cq[k] = s * x + c * y;
                }
            }
        if (!rotated) break;
    }
    double mn = INFINITY;
    for (int j = 0; j < l; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += A[(size_t)j * n + k] * A[(size_t)j * n + k];
        mn = std::min(mn, std::sqrt(s));
    }
    return mn;
}

extern "C" int oa_oracle_exclusions(const char *text, int n, uint8_t *keep) {
    /* read_excludeList (src/AIOwrapper.cpp:2146-2233).  The reference keeps
     * a list of kept (start,len) ranges and removes one row at a time with
     * splitpair; the set it ends with is "all rows minus the removed ones",
     * which is what keep[] holds.  Line parsing follows the reference: a
     * line with one number leaves `second` at its previous value, which is
     * how single rows are recognised. */
    for (int i = 0; i < n; ++i) keep[i] = 1;
    std::istringstream in(text ? text : "");
    std::string line;
    int first = 0, second = 0, prev_second = -1, excluded = 0;
    while (std::getline(in, line) && second < n) {
        std::istringstream iss(line);
        iss >> first;
        iss >> second;
        if (prev_second > first) return -2;
        if (second == prev_second) {
            second = first;
        } else {
            if (first > second) return -1;
            if (first < 1 || second < 1) return -3;
        }
        prev_second = second;
        for (int i = first - 1; i <= second - 1; ++i) {
            if (i >= 0 && i < n && keep[i]) {
                keep[i] = 0;
                ++excluded;
            }
        }
    }
    if (excluded >= n) return -4;
    return excluded;
}

extern "C" int oa_oracle_scan(const oa_oracle_params *prm, const double *XL,
                              const double *XR, const double *Y,
                              oa_oracle_out *out) {
    if (!prm || !XL || !XR || !Y || !out) return -1;
    const int n = prm->n, l = prm->l, r = prm->r, t = prm->t, m = prm->m;
    const int p = l + r;
    if (n <= 0 || l <= 0 || r <= 0 || t <= 0 || m <= 0 || p >= n) return -1;
    const size_t N = (size_t)n;

    /* Stl = XL'XL (src/Algorithm.cpp:490), full l*l. */
    std::vector<double> Stl(l * l);
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < l; ++j) {
            double s = 0.0;
            for (size_t k = 0; k < N; ++k) s += XL[i * N + k] * XL[j * N + k];
            Stl[i * l + j] = s;
        }

    /* Per trait: NaN -> 0 with sorted NaN rows (src/Utility.cpp:130-179),
     * SST (src/Algorithm.cpp:1222-1251), Ay_top (src/Algorithm.cpp:583),
     * Stl_corr (src/Algorithm.cpp:1324-1362). */
    std::vector<TraitData> td(t);
    for (int y = 0; y < t; ++y) {
        TraitData &d = td[y];
        d.y0.resize(N);
        for (size_t k = 0; k < N; ++k) {
            double v = Y[y * N + k];
            if (std::isnan(v)) { d.y0[k] = 0.0; d.miss.push_back((int)k); }
            else d.y0[k] = v;
        }
        double sx = 0.0, sxx = 0.0;
        for (size_t k = 0; k < N; ++k) { sx += d.y0[k]; sxx += d.y0[k] * d.y0[k]; }
        d.sst = sxx - (sx * sx) / (double)(n - (int)d.miss.size());
        d.ay_top.assign(l, 0.0);
        for (int j = 0; j < l; ++j) {
            double s = 0.0;
            for (size_t k = 0; k < N; ++k) s += XL[j * N + k] * d.y0[k];
            d.ay_top[j] = s;
        }
        d.stl_corr.assign(l * l, 0.0);
        for (int i = 0; i < l; ++i)
            for (int j = 0; j < l; ++j) {
                double s = 0.0;
                for (int k : d.miss) s += XL[i * N + k] * XL[j * N + k];
                d.stl_corr[i * l + j] = s;
            }
        if (out->n_used) out->n_used[y] = n - (int)d.miss.size();
    }

    const double one_over_sqrt2 = 0.70710678118654752440084436210484903928;
    (void)one_over_sqrt2;
    const int start_h = prm->disp_cov ? 1 : l;
    int64_t n_sig = 0, n_kept = 0, n_emitted = 0;
    int status = 0;

#ifdef _OPENMP
    int threads = prm->threads > 0 ? prm->threads : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) \
    reduction(+ : n_sig, n_kept, n_emitted)
#endif
    for (int s = 0; s < m; ++s) {
        /* XR columns of this SNP with NaN -> column mean
         * (src/Utility.cpp:103-127). */
        std::vector<double> xr(N * r);
        for (int c = 0; c < r; ++c) {
            const double *src = XR + ((size_t)s * r + c) * N;
            double sum = 0.0;
            int nn = 0;
            for (size_t k = 0; k < N; ++k) {
                if (std::isnan(src[k])) { ++nn; xr[c * N + k] = 0.0; }
                else { xr[c * N + k] = src[k]; sum += src[k]; }
            }
            double avg = sum / (double)(n - nn);
            for (size_t k = 0; k < N; ++k)
                if (std::isnan(src[k])) xr[c * N + k] = avg;
        }
        /* Str (l x r, index i*l+j) and Sbr (r x r, index i*r+j)
         * (src/Algorithm.cpp:633-656). */
        std::vector<double> Str(l * r), Sbr(r * r);
        for (int i = 0; i < r; ++i) {
            for (int j = 0; j < l; ++j) {
                double acc = 0.0;
                for (size_t k = 0; k < N; ++k) acc += XL[j * N + k] * xr[i * N + k];
                Str[i * l + j] = acc;
            }
            for (int j = 0; j < r; ++j) {
                double acc = 0.0;
                for (size_t k = 0; k < N; ++k) acc += xr[j * N + k] * xr[i * N + k];
                Sbr[i * r + j] = acc;
            }
        }
        /* build_S (src/Algorithm.cpp:60-95): upper triangle, column-major. */
        std::vector<double> Sfull(p * p, 0.0);
        for (int i = 0; i < l; ++i)
            for (int j = 0; j <= i; ++j) Sfull[i * p + j] = Stl[j + i * l];
        for (int i = 0; i < r; ++i)
            for (int j = 0; j < l; ++j) Sfull[l * p + p * i + j] = Str[j + i * l];
        for (int i = 0; i < r; ++i)
            for (int j = 0; j <= i; ++j) Sfull[l * p + l + p * i + j] = Sbr[j + i * r];

        /* var_x = diag(inv(S_full)) (src/Algorithm.cpp:659-671). */
        std::vector<double> varx(p);
        {
            std::vector<double> U = Sfull;
            if (chol_upper(U.data(), p) == 0) chol_inv_diag(U.data(), p, varx.data());
            else for (int h = 0; h < p; ++h) varx[h] = NAN;
        }
        if (out->var_x)
            for (int h = 0; h < p; ++h) out->var_x[(size_t)s * p + h] = varx[h];

        std::vector<double> S(p * p), b(p), Str_c(l * r), Sbr_c(r * r);
        for (int y = 0; y < t; ++y) {
            const TraitData &d = td[y];
            /* generate_correction_missings_Str (src/Algorithm.cpp:1260-1315):
             * Sbr_corr is accumulated only for j < min(l, r). */
            std::fill(Str_c.begin(), Str_c.end(), 0.0);
            std::fill(Sbr_c.begin(), Sbr_c.end(), 0.0);
            for (int i = 0; i < r; ++i)
                for (int j = 0; j < l; ++j)
                    for (int k : d.miss) {
                        Str_c[i * l + j] += XL[j * N + k] * xr[i * N + k];
                        if (j < r) Sbr_c[i * r + j] += xr[j * N + k] * xr[i * N + k];
                    }
            /* build_S + correct_missings_S (src/Algorithm.cpp:742,
             * 1370-1399), literally, including the subtraction of every
             * Str correction from column l. */
            S = Sfull;
            for (int i = 0; i < l; ++i)
                for (int j = 0; j < l; ++j) S[i * p + j] -= d.stl_corr[i * l + j];
            for (int i = 0; i < r; ++i)
                for (int j = 0; j < r; ++j) S[(l * p + l) + i * p + j] -= Sbr_c[i * r + j];
            for (int i = 0; i < r; ++i)
                for (int j = 0; j < l; ++j) S[l * p + j] -= Str_c[i * l + j];
            /* rhs = [Ay_top; Ay_bot] (src/Algorithm.cpp:734-736). */
            for (int j = 0; j < l; ++j) b[j] = d.ay_top[j];
            for (int c = 0; c < r; ++c) {
                double acc = 0.0;
                for (size_t k = 0; k < N; ++k) acc += xr[c * N + k] * d.y0[k];
                b[l + c] = acc;
            }
            /* sposv('U'): on a failed factorisation LAPACK leaves the
             * right-hand side untouched and the reference ignores info
             * (src/Algorithm.cpp:767-773). */
            if (chol_upper(S.data(), p) == 0) chol_solve_upper(S.data(), p, b.data());

            /* RSS over rows with y != 0 (src/Algorithm.cpp:976-1024). */
            double rss = 0.0;
            for (size_t k = 0; k < N; ++k) {
                double sy = 0.0;
                for (int h = 0; h < l; ++h) sy += XL[h * N + k] * b[h];
                for (int h = l; h < p; ++h) sy += xr[(h - l) * N + k] * b[h];
                if (d.y0[k] != 0.0) {
                    sy -= d.y0[k];
                    rss += sy * sy;
                }
            }

            /* hpc_statistics (src/Algorithm.cpp:1034-1214). */
            const int ymiss = (int)d.miss.size();
            const int n_corr = n - ymiss - p;
            const double var_factor = (double)(n_corr + p) / (double)n;
            const double R2 = 1.0 - ((rss / (n_corr - 1)) / (d.sst / (n - 1)));
            const size_t pair = (size_t)y * m + s;
            if (out->r2) out->r2[pair] = R2;
            if (out->rss) out->rss[pair] = rss;
            if (rss < 0.0) status = -2;
            const double sig = std::sqrt(rss / n_corr);
            bool has_snp = false;
            int64_t kept_here = 0, sig_here = 0;
            for (int h = 0; h < p; ++h) {
                size_t o = pair * p + h;
                if (out->beta) out->beta[o] = b[h];
                if (out->se) out->se[o] = NAN;
                if (out->tstat) out->tstat[o] = NAN;
                if (out->pval) out->pval[o] = NAN;
                if (out->kept) out->kept[o] = 0;
            }
            if (rss > 0.0 && (R2 > prm->min_r2_store || prm->store_all)) {
                for (int h = p - 1; h >= start_h; --h) {
                    if (varx[h] < 0.0) status = -2;
                    double se = sig * std::sqrt(varx[h] * var_factor);
                    double tv = std::fabs(b[h] / se);
                    size_t o = pair * p + h;
                    if (out->se) out->se[o] = se;
                    if (out->tstat) out->tstat[o] = tv;
                    if (tv > prm->t_star || prm->store_all) {
                        if (h >= l) has_snp = true;
                        double pv = oa_oracle_pvalue(tv);
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
    out->n_sig = n_sig;
    out->n_kept = n_kept;
    out->n_emitted = n_emitted;
    return status;
}
