/*
 * FP64 CPU oracle for the blocked normal-equation OLS scan.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * reference the HIP results are compared with (or as the timed CPU
 * restatement).  The product path never calls it.
 *
 * It restates, in double precision, the arithmetic of OmicABELnoMM's
 * partialNEQ_Blocked_STL_MD (src/Algorithm.cpp:322-886) including the
 * reference's quirks (missing-row corrections with the column misplacement
 * at src/Algorithm.cpp:1390-1396, the partial Sbr correction at
 * src/Algorithm.cpp:1300, SE from the unmasked inverse scaled by
 * (n-Ymiss)/n, RSS over rows with y!=0, the normal-approximation p-value).
 * Every function cites the reference lines it follows.
 *
 * Parity status per function is listed in oracle.cpp's header comment.
 */
#ifndef OA_ORACLE_H
#define OA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oa_oracle_params {
    int n;              /* rows after exclusion                          */
    int real_n;         /* rows before exclusion (nUsedPct denominator)  */
    int l;              /* covariate columns, intercept included         */
    int r;              /* columns per SNP                               */
    int t;              /* traits                                        */
    int m;              /* SNPs                                          */
    int disp_cov;       /* -i : report covariate coefficients too        */
    int store_all;      /* -f : storePInd, ignore every threshold        */
    double min_r2_store;/* minR2store (float value widened)              */
    double t_star;      /* minTstore = getTvalue(minPstore), float value */
    double min_p_disp;  /* minPdisp (float value widened)                */
    int threads;        /* OpenMP threads (<=0: 1)                       */
} oa_oracle_params;

/* Dense per-(trait, SNP) outputs; any pointer may be NULL.
 * beta/se/tstat/pval/kept : t*m*p, index ((y*m)+s)*p + h
 * r2/rss/emitted          : t*m,   index y*m + s
 * n_used                  : t
 * var_x                   : m*p    (diag of the unmasked inverse)      */
typedef struct oa_oracle_out {
    double *beta, *se, *tstat, *pval;
    uint8_t *kept;
    double *r2, *rss;
    uint8_t *emitted;
    int32_t *n_used;
    double *var_x;
    int64_t n_sig;      /* count of kept results with P < minPdisp       */
    int64_t n_kept;     /* kept results inside emitted (trait,SNP) pairs */
    int64_t n_emitted;  /* emitted (trait,SNP) records                   */
} oa_oracle_out;

/* XL: n*l column-major, no NaN.  XR: n*(r*m) column-major, NaN = missing
 * genotype (imputed with the column mean).  Y: n*t column-major, NaN =
 * missing phenotype.  Returns 0, or -1 for bad arguments, -2 when the
 * reference would exit (negative RSS or negative var_x). */
int oa_oracle_scan(const oa_oracle_params *prm, const double *XL,
                   const double *XR, const double *Y, oa_oracle_out *out);

/* getTvalue (src/Utility.cpp:244-250): sqrt(2)*erfinv(1-p), rounded to
 * float as the reference stores it. */
double oa_oracle_tvalue(float pval);

/* erfinv in double precision (used by oa_oracle_tvalue). */
double oa_oracle_erfinv(double x);

/* p-value rule of hpc_statistics (src/Algorithm.cpp:1159-1167). */
double oa_oracle_pvalue(double t);

/* Rank check of the covariates (src/Algorithm.cpp:497-527): returns the
 * smallest singular value of XL (n*l column-major). */
double oa_oracle_min_singular(const double *XL, int n, int l);

/* Exclusion list semantics (src/AIOwrapper.cpp:2146-2233 with splitpair
 * 1994-2085): parses the text of an exclusion file against n rows and
 * writes keep[n] (1 = row kept).  Returns the number of excluded rows,
 * or -1 "Please provide ordered pairs!", -2 "valid positive ordered
 * indixes", -3 "valid positive indixes", -4 everything excluded. */
int oa_oracle_exclusions(const char *text, int n, uint8_t *keep);

/* FP32 restatement of the reference's blocked algorithm (ref32.cpp): the
 * same inputs in float, tb traits and mb SNPs per block, the same outputs
 * (widened to double).  The CPU baseline of the benchmark. */
int oa_ref32_scan(const oa_oracle_params *prm, const float *XL, const float *XR, const float *Y,
                  int tb, int mb, oa_oracle_out *out);
/* The BLAS the FP32 restatement's products run through (MKL loaded at run
 * time, or the file's own kernel). */
const char *oa_ref32_blas(void);

#ifdef __cplusplus
}
#endif
#endif
