/*
 * omicabelnomm.h -- C ABI of the drop-in scan driver (libomicabelnomm.so):
 * the file-level equivalent of OmicABELnoMM's
 *     Algorithm::applyDefaultParams(Settings&)   src/Algorithm.cpp:277-311
 *     Algorithm::solve(Settings, Outputs&, int)  src/Algorithm.cpp:37-48
 * reading DatABEL .fvi/.fvd inputs (src/AIOwrapper.cpp:17-764,2294-2317)
 * and writing <out>_dis.txt, <out>_sto.ibin and <out>_sto.dbin
 * (src/AIOwrapper.cpp:679-733,936-945,1440-1558), with the scan itself on
 * the GPU through liboagpu (include/oagpu.h).
 */
#ifndef OMICABELNOMM_H
#define OMICABELNOMM_H

#ifdef __cplusplus
extern "C" {
#endif

/* Settings (src/Definitions.h:168-219); strings are paths without the
 * .fvi/.fvd extension, NULL or "" when unused. */
typedef struct omicabelnomm_settings {
    const char *fnameAL, *fnameAR, *fnameY, *fnameOutFiles;
    const char *fname_excludelist;
    const char *fname_interactions;
    int r;                       /* -n: columns per SNP in the genotype file   */
    int threads;                 /* -t: accepted, the scan runs on the GPU     */
    int storeBin;                /* -b */
    int storePInd;               /* -f */
    int disp_cov;                /* -i */
    float minPdisp, minPstore, minR2disp, minR2store;   /* -d -s -r -e */
    int use_interactions, keep_depVar, use_multiple_interaction_sets;
    int ini_IDX_interactions, end_IDX_interactions;
    int limit_m, limit_t, limit_n;
    int precision;               /* OAGPU_FP64 / OAGPU_FP32 / OAGPU_BF16X3     */
    int device;
    int verbose;                 /* 1: the reference's console messages       */
    int dosages;                 /* -j/-k/-l/-z/-y given (Settings::dosages)   */
    int model;                   /* Settings::model: -1 none, 0 additive,
                                    1 dominant, 2 recessive, 3 linear from
                                    fname_dosages, 4 additive from it         */
    const char *fname_dosages;   /* factor list of models 3 and 4            */
} omicabelnomm_settings;

/* Outputs (src/Definitions.h:222-257): the fields this driver fills. */
typedef struct omicabelnomm_outputs {
    double duration;             /* scan loop wall time (s), reference scope   */
    double acc_loady, acc_loadxr, acc_gpu, acc_storeb;
    long long total_sig_results; /* "Number of Significant Results"           */
    long long n_txt_lines;
    long long n_bin_results;
    int n, l, r, t, m, realN;    /* problem sizes after exclusion              */
} omicabelnomm_outputs;

void omicabelnomm_apply_defaults(omicabelnomm_settings *s);

/* Returns 0, or a nonzero code where the reference prints and exits;
 * the message is copied to errbuf. */
int omicabelnomm_solve(const omicabelnomm_settings *s, omicabelnomm_outputs *out,
                       char *errbuf, int errlen);

#ifdef __cplusplus
}
#endif
#endif
