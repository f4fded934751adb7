/*
 * oagpu_serve.h -- a genotype panel resident in HBM and a request server
 * that scans many trait sets against it (liboagpu.so).
 *
 * No counterpart in the reference, which scans one trait file per process
 * (src/main.cpp:446-487).  Each request is a full OLS scan with the
 * reference's semantics -- its own traits, covariates and individual
 * subset, the same statistics as oagpu_scan -- over every SNP of the panel
 * (r = 1 column per SNP, no interaction terms).
 *
 * Scheduling: a server thread walks the panel in chunks of SNPs.  At every
 * chunk boundary it admits queued requests while the active set stays
 * within max_active requests and the HBM budget; each active request joins
 * at the current chunk and leaves after one full cycle.  One contraction per
 * chunk covers the trait columns of every active request, and each request
 * then runs its own corrections, solves and statistics on its columns, so
 * requests that arrive while others run share the device.  A request's
 * results do not depend on what else is active (up to the summation order of
 * the chosen contraction kernel).
 *
 * Requests whose individual subset is not everyone need a panel without
 * missing genotypes: the shared operand pass imputes a missing genotype with
 * the panel-wide mean, where a scan of the subset uses the subset's mean, so
 * oagpu_server_submit refuses such a request (OAGPU_ERR_UNSUPPORTED).
 */
#ifndef OAGPU_SERVE_H
#define OAGPU_SERVE_H

#include "oagpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oagpu_panel oagpu_panel;

/* n individuals x m SNPs, column-major with leading dimension ld >= n
 * (dtype OAGPU_F16/F32/F64, NaN = missing).  OAGPU_HOST copies the panel
 * to the device; OAGPU_DEVICE uses it in place (it must outlive the
 * panel). */
int oagpu_panel_create(int device, int n, long long m, const void *G, int dtype, int where, long long ld,
                       oagpu_panel **panel);
void oagpu_panel_destroy(oagpu_panel *panel);

typedef struct oagpu_server_params {
    int precision;          /* OAGPU_FP64 / OAGPU_FP32 / OAGPU_BF16X3              */
    long long hbm_budget;   /* bytes for active requests (0 = 64 GiB)               */
    int max_active;         /* requests in one batch (0 = 64)                       */
    int chunk_snps;         /* SNPs per chunk (0 = 16384; < 0: chosen by timing one
                               chunk of 4,096, 8,192, 16,384 and 32,768 SNPs with
                               the first requests, remembered per panel shape)     */
    int graphs;             /* 1: replay each chunk's launches from a HIP graph
                               captured once per active set and chunk size;
                               0: launch them one by one (the default)              */
} oagpu_server_params;

typedef struct oagpu_request {
    int t;                  /* traits                                               */
    int l;                  /* covariate columns incl. intercept (<= 32)            */
    const double *Y;        /* n x t, column-major, leading dimension ldy, NaN = missing */
    long long ldy;
    const double *XL;       /* n x l, leading dimension ldxl (no NaN inside the subset) */
    long long ldxl;
    const uint8_t *keep;    /* n flags, individual k used when keep[k] != 0 (NULL = all) */
    float min_p_store, min_p_disp, min_r2_store;   /* after threshold normalisation */
    int disp_cov, store_all;
    long long max_hits;     /* hit capacity on the device (0 = 1<<20)              */
} oagpu_request;

typedef struct oagpu_server oagpu_server;

int oagpu_server_create(oagpu_panel *panel, const oagpu_server_params *params, oagpu_server **server);
/* Stops the server after the active requests finish; unfinished queued
 * requests fail with OAGPU_ERR_STATE. */
void oagpu_server_destroy(oagpu_server *server);
const char *oagpu_server_last_error(const oagpu_server *server);

/* Chunks this server has replayed from a captured graph so far. */
long long oagpu_server_graph_chunks(const oagpu_server *server);

/* SNPs per chunk: the configured size, or the timed choice once the first
 * requests have been admitted (the largest candidate before that). */
int oagpu_server_chunk_snps(const oagpu_server *server);

/* Copies the request's inputs and queues it; returns at once.
 * OAGPU_ERR_UNSUPPORTED: a subset (keep with a zero flag) on a panel that
 * holds missing genotypes. */
int oagpu_server_submit(oagpu_server *server, const oagpu_request *req, long long *id);
/* 1 when the request has finished, 0 while it is queued or running, < 0 on
 * an unknown id. */
int oagpu_server_poll(oagpu_server *server, long long id);
/* Blocks until the request has finished, then returns its hits (sorted by
 * trait, SNP, h descending), statistics and latency from submission to
 * completion, and forgets the request.  Returns the request's status
 * (OAGPU_ERR_HIT_OVERFLOW when hit_cap is too small: stats->n_hits holds the
 * count). */
int oagpu_server_wait(oagpu_server *server, long long id, oagpu_hit *hits, long long hit_cap,
                      oagpu_stats *stats, double *latency_ms, char *errbuf, int errlen);

#ifdef __cplusplus
}
#endif
#endif
