/*
 * oagpu.h -- C ABI of liboagpu.so, the MI355X (gfx950) implementation of
 * OmicABELnoMM's blocked normal-equation OLS scan.
 *
 * The reference's compute boundary is
 *     Algorithm::solve(Settings params, Outputs &out, int type)
 *         (src/Algorithm.h:41-47, src/Algorithm.cpp:37-48)
 * which runs partialNEQ_Blocked_STL_MD (src/Algorithm.cpp:322-886) over
 * blocks that AIOwrapper hands it:
 *     load_AL      (src/AIOwrapper.cpp:2088-2130)  -> oagpu_set_covariates
 *     load_Yblock  (src/AIOwrapper.cpp:1723-1769)  -> oagpu_load_traits
 *     load_ARblock (src/AIOwrapper.cpp:1674-1720)  -> oagpu_load_genotypes
 *     interaction synthesis (src/AIOwrapper.cpp:767-840)
 *                                                  -> oagpu_set_interactions
 *                                                     + column descriptors
 *     hpc_statistics + getCurrentWriteBuffers/write_OutFiles
 *         (src/Algorithm.cpp:797-819,1034-1214)    -> oagpu_scan (hits)
 *     applyDefaultParams / threshold normalisation
 *         (src/Algorithm.cpp:277-311,338-360)      -> oagpu_params
 *
 * Conventions: plain pointers and sizes, int status (OAGPU_OK = 0, < 0 on
 * error, message in oagpu_last_error).  Matrices are column-major with a
 * leading dimension `ld` >= n (DatABEL .fvd order).  NaN marks a missing
 * value.  One context per host thread; every call is ordered on the
 * context's HIP streams.  Device pointers passed with OAGPU_DEVICE, and
 * host genotype blocks passed with OAGPU_HOST, must stay valid and
 * unchanged until the scan that reads them has returned.
 */
#ifndef OAGPU_H
#define OAGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OAGPU_OK 0
#define OAGPU_ERR_ARG (-1)          /* bad argument or size                    */
#define OAGPU_ERR_HIP (-2)          /* HIP runtime error                       */
#define OAGPU_ERR_STATE (-3)        /* call order (e.g. scan before loads)     */
#define OAGPU_ERR_HIT_OVERFLOW (-4) /* more hits than capacity; stats hold the
                                       needed count                           */
#define OAGPU_ERR_NUMERIC (-5)      /* the reference would exit: RSS < 0 or
                                       var_x < 0 (src/Algorithm.cpp:1096,1130) */
#define OAGPU_ERR_UNSUPPORTED (-6)  /* shape outside the compiled kernels     */
#define OAGPU_ERR_COVARIATES (-7)   /* linearly dependent covariates
                                       (src/Algorithm.cpp:510-522)             */
#define OAGPU_ERR_NO_DEVICE (-8)

/* Arithmetic of the bulk contraction XR'[Y | XL].  Everything else (Gram
 * pieces XR_s'XR_s, missing-row corrections, Cholesky, solve, RSS, t, p) is
 * FP64 in every mode. */
enum { OAGPU_FP64 = 0, OAGPU_FP32 = 1, OAGPU_BF16X3 = 2 };

/* Element types of trait and genotype inputs. */
enum { OAGPU_F64 = 0, OAGPU_F32 = 1, OAGPU_F16 = 2 };

/* Where an input pointer lives. */
enum { OAGPU_HOST = 0, OAGPU_DEVICE = 1 };

typedef struct oagpu_params {
    int n;              /* individuals after exclusion (Settings::n)          */
    int l;              /* covariate columns incl. intercept (Settings::l)    */
    int r;              /* columns per SNP (Settings::r)                       */
    int real_n;         /* individuals before exclusion (AIOwrapper::realN)   */
    int precision;      /* OAGPU_FP64 / OAGPU_FP32 / OAGPU_BF16X3              */
    int disp_cov;       /* -i (Settings::disp_cov)                             */
    int store_all;      /* -f (Settings::storePInd)                            */
    float min_r2_store; /* Settings::minR2store after normalisation            */
    float min_p_store;  /* Settings::minPstore after normalisation             */
    float min_p_disp;   /* Settings::minPdisp                                  */
    int max_traits;     /* capacity of one trait block                         */
    int max_snps;       /* SNPs per device chunk (0 = choose from memory)      */
    long long max_hits; /* device hit arena capacity (0 = 1<<22)               */
    int device;         /* HIP device ordinal                                  */
} oagpu_params;

/* One stored coefficient (reference `result` inside a `resultH`,
 * src/Definitions.h:260-280).  h >= l: SNP term, AR index = snp*r + h - l;
 * h < l: covariate term (AL index h). */
typedef struct oagpu_hit {
    int32_t trait;      /* trait index: trait_offset + local index            */
    int32_t snp;        /* SNP index: snp_offset + local index                */
    int32_t h;
    int32_t n_used;     /* n - Ymiss                                           */
    double b, se, t, p, r2;
} oagpu_hit;

/* Optional dense outputs (host pointers, any may be NULL), for parity
 * checks on small blocks.  (t, m, p) arrays are indexed ((y*m)+s)*p + h. */
typedef struct oagpu_dense {
    double *beta, *se, *tstat, *pval;
    uint8_t *kept;
    double *r2, *rss;   /* (t, m) */
    uint8_t *emitted;   /* (t, m): a resultH would be emitted */
    int32_t *n_used;    /* (t) */
    double *var_x;      /* (m, p) */
} oagpu_dense;

typedef struct oagpu_stats {
    long long n_hits;       /* stored coefficients in emitted records         */
    long long n_emitted;    /* emitted (trait, SNP) records                    */
    long long n_sig;        /* kept coefficients with P < minPdisp
                               (Outputs::total_sig_results)                     */
    long long n_fallback;   /* pairs whose RSS came from the explicit residual */
    int n_patterns;         /* distinct missing-value patterns in the traits   */
    int n_chunks;
    double t_star;          /* minTstore = getTvalue(minPstore)                */
} oagpu_stats;

typedef struct oagpu_ctx oagpu_ctx;

/* Create a context on params->device; allocates the trait-side buffers. */
int oagpu_create(const oagpu_params *params, oagpu_ctx **ctx);
void oagpu_destroy(oagpu_ctx *ctx);
const char *oagpu_last_error(const oagpu_ctx *ctx);

/* Individual subset: keep[k] != 0 keeps individual k (n entries; NULL
 * keeps all).  Rows outside the subset take no part in any sum, count or
 * statistic, as if the reference had excluded them (src/AIOwrapper.cpp:
 * 1947-2233), so one loaded panel serves analyses of different subsets.
 * Call before oagpu_set_covariates; it clears the loaded covariates and
 * traits.  No counterpart call in the reference, which removes excluded
 * rows while reading the files. */
int oagpu_set_subset(oagpu_ctx *ctx, const uint8_t *keep);

/* Covariates XL (n x l, FP64 host, no NaN).  Runs the reference's rank
 * check (src/Algorithm.cpp:497-527) and returns OAGPU_ERR_COVARIATES when
 * a singular value is below 1e-6. */
int oagpu_set_covariates(oagpu_ctx *ctx, const double *XL, long long ld);

/* Interaction covariates E (n x n_env, FP64 host, no NaN): genotype
 * columns whose descriptor names an env column are multiplied by it. */
int oagpu_set_interactions(oagpu_ctx *ctx, const double *E, long long ld,
                           int n_env);

/* One block of t traits (n x t, NaN = missing).  Computes the missing-row
 * lists, their deduplicated patterns, SST and XL'y. */
int oagpu_load_traits(oagpu_ctx *ctx, const void *Y, int dtype, int where,
                      long long ld, int t, int trait_offset);

/* Dosage model (src/AIOwrapper.cpp:256-340 sets it up, 1208-1297 applies it
 * to every genotype block): the genotype file holds file_r columns per SNP
 * (e.g. P(AA), P(AB), P(BB)) and only the first n_used of them are read.
 * OAGPU_DOSAGE_SUM makes one column per SNP, sum_i f[i] * G[:, s*file_r+i]
 * (the reference's additive, dominant, recessive and -y models, r = 1);
 * OAGPU_DOSAGE_SCALE makes n_used columns, f[i] * G[:, s*file_r+i] (the -z
 * linear model).  The columns are formed on the device, in FP64, before
 * NaN imputation, as the reference combines before replace_nans_avgs.
 * Afterwards oagpu_load_genotypes takes blocks of whole file SNPs (n_src a
 * multiple of file_r) and its column descriptors index the combined
 * columns.  mode OAGPU_DOSAGE_NONE switches the model off.  At most
 * OAGPU_MAX_DOSAGE_FACTORS factors. */
#define OAGPU_MAX_DOSAGE_FACTORS 16
enum { OAGPU_DOSAGE_NONE = 0, OAGPU_DOSAGE_SUM = 1, OAGPU_DOSAGE_SCALE = 2 };
int oagpu_set_dosage_model(oagpu_ctx *ctx, int mode, int file_r, int n_used,
                           const double *factors);

/* One block of m SNPs.  G holds the source genotype columns (n rows,
 * leading dimension ld, NaN = missing, imputed with the column mean).
 * Output column c of SNP s (c < r) is G[:, col_src[s*r+c]], times
 * E[:, col_env[s*r+c]] when that is >= 0.  NULL descriptors mean
 * col_src[s*r+c] = s*r+c and no interaction.  n_src = number of columns
 * in G. */
int oagpu_load_genotypes(oagpu_ctx *ctx, const void *G, int dtype, int where,
                         long long ld, int n_src, int m,
                         const int32_t *col_src, const int32_t *col_env,
                         int snp_offset);

/* Scan the loaded traits x SNPs.  Hits are written to hits[0..hit_cap)
 * sorted by (trait, snp, h descending), the reference's order inside one
 * block.  dense and hits may be NULL.  Equivalent to oagpu_scan_async
 * followed by oagpu_scan_wait. */
int oagpu_scan(oagpu_ctx *ctx, oagpu_dense *dense, oagpu_hit *hits,
               long long hit_cap, oagpu_stats *stats);

/* The same scan in two halves: oagpu_scan_async enqueues the kernels and
 * returns; the next genotype block can then be uploaded (its copy overlaps
 * this scan) before oagpu_scan_wait collects the results.  This pair
 * replaces the reference's pthread hand-off between the compute loop and
 * async_io (src/AIOwrapper.cpp:1674-1868). */
int oagpu_scan_async(oagpu_ctx *ctx, int want_dense);
int oagpu_scan_wait(oagpu_ctx *ctx, oagpu_dense *dense, oagpu_hit *hits,
                    long long hit_cap, oagpu_stats *stats);

/* oagpu_scan without device allocation.  No counterpart in the reference
 * (which scans on the CPU): it is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  It runs the same kernels as oagpu_scan on the
 * context's loaded covariates, traits and genotypes (set up beforehand) and
 * returns the same hits and statistics, but its chunk buffers (operand rows,
 * contraction output, Gram pieces, corrections, counters, fallback list) live
 * in one static device array of OAGPU_NOALLOC_SCRATCH_BYTES per GPU, held by
 * the library until the process ends; SNP chunks are sized to fit it.  The
 * contraction is the precision's default hand-written kernel (no autotuner,
 * no library GEMM).  A workload whose single-SNP chunk does not fit returns
 * OAGPU_ERR_ARG with the sizes in oagpu_last_error.  No dense outputs.  All
 * calls on one GPU share its array: calls that overlap on one GPU (other
 * host threads, other contexts) race. */
#define OAGPU_NOALLOC_SCRATCH_BYTES (128ull << 20)
int oagpu_scan_noalloc(oagpu_ctx *ctx, oagpu_hit *hits, long long hit_cap,
                       oagpu_stats *stats);

/* Page-locked host buffers for genotype and trait blocks (hipHostMalloc):
 * uploads from them run at full PCIe rate and asynchronously.  A host block
 * passed to oagpu_load_genotypes may be refilled once the scan that reads
 * it has returned from oagpu_scan / oagpu_scan_wait. */
int oagpu_host_alloc(size_t bytes, void **ptr);
void oagpu_host_free(void *ptr);

/* Device-side timing of the last oagpu_scan, per kernel family (ms):
 * out[0] prep, [1] gemm, [2] snp gram, [3] corrections and per-(SNP,
 * pattern) solve pieces, [4] epilogue,
 * [5] fallback, [6] whole scan.  Returns the number of values written. */
int oagpu_last_timings(const oagpu_ctx *ctx, double *out, int n);

/* Name of the contraction kernel variant the autotuner chose for the last
 * oagpu_scan (no counterpart in the reference, which calls BLAS). */
const char *oagpu_gemm_variant(const oagpu_ctx *ctx);

/* Shape classes the contraction autotuner has timed in this process (no
 * counterpart in the reference).  Its choices persist in a cache file,
 * OAGPU_TUNE_CACHE or $HOME/.cache/omicabelnomm/gemm_tune_<arch>_<CUs>.txt,
 * read before a shape is timed. */
int oagpu_gemm_tunings(void);

/* getTvalue (src/Utility.cpp:244-250): sqrt(2) erfinv(1 - p) as a float. */
double oagpu_tvalue(float pval);

#ifdef __cplusplus
}
#endif
#endif
