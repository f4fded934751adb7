// Drop-in scan driver: the file-level path of partialNEQ_Blocked_STL_MD
// (src/Algorithm.cpp:322-886) and AIOwrapper (src/AIOwrapper.cpp) on top of
// liboagpu.  Inputs: DatABEL files; outputs: <out>_dis.txt and, with -b,
// <out>_sto.ibin / <out>_sto.dbin.  The scan runs on the GPU; a reader
// thread prefetches the next genotype block while the current one is
// scanned (the reference's async_io pthread, src/AIOwrapper.cpp:895-1437).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "databel.h"
#include "oagpu.h"
#include "omicabelnomm.h"
#include "writers.h"

using namespace oahost;

namespace {

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

std::string str_or(const char* p) { return p ? std::string(p) : std::string(); }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// read_excludeList + splitpair (src/AIOwrapper.cpp:1994-2233): removes rows
// from keep[]; returns the number of rows removed.
int apply_exclusion_file(const std::string& fname, int n, std::vector<uint8_t>& keep, bool verbose) {
  std::ifstream in(fname);
  if (!in) throw Fail(1, "Error reading exclude list file.");
  std::string line;
  int first = 0, second = 0, prev_second = -1, excluded = 0;
  if (verbose) std::cout << "Excluding Ids: \n";
  while (std::getline(in, line) && second < n) {
    std::istringstream iss(line);
    iss >> first;
    iss >> second;
    if (prev_second > first) {
      std::ostringstream m;
      m << "\nPlease provide valid positive ordered indixes!\n" << prev_second << " > " << first << "\n";
      throw Fail(1, m.str());
    }
    if (second == prev_second) {
      if (verbose) std::cout << first << ", ";
      second = first;
    } else {
      if (verbose) std::cout << first << "-" << second << ", ";
      if (first > second) throw Fail(1, "\nPlease provide ordered pairs!\n");
      if (first < 1 || second < 1) throw Fail(1, "\nPlease provide valid positive indixes!\n");
    }
    prev_second = second;
    for (int i = first - 1; i <= second - 1; ++i)
      if (i >= 0 && i < n && keep[i]) {
        keep[i] = 0;
        ++excluded;
      }
  }
  if (excluded >= n) {
    std::ostringstream m;
    m << "\nExclusion List excluded all data!\nTotal Ids: " << n << "\nExcluded Ids: " << excluded;
    throw Fail(1, m.str());
  }
  if (verbose) std::cout << "Excluded: " << excluded << " Using: " << n - excluded << std::endl;
  return excluded;
}

// read_dosages (src/AIOwrapper.cpp:2236-2256): `expected` factors read with
// operator>>.  The reference also counts a read that fails (at the end of
// the file after a trailing blank, or on a token that is not a number) and
// leaves that factor uninitialised; here the list must hold `expected`
// numbers.
std::vector<double> read_dosage_factors(const std::string& fname, int expected) {
  std::ifstream in(fname);
  if (!in) throw Fail(1, "Error reading dosages file.");
  std::vector<double> f;
  int i = 0;
  for (; i < expected; ++i) {
    float v = 0.0f;
    if (!(in >> v)) break;
    f.push_back(v);
  }
  if (i != expected) {
    std::ostringstream m;
    m << "The number of factors for the dosage model! does not coincide with the source data #facctors:" << i
      << ", expected:" << expected;
    throw Fail(1, m.str());
  }
  return f;
}

// Dosage model set-up (src/AIOwrapper.cpp:256-340).  Fills the factors and
// whether they are summed into one column per SNP.  The recessive model is
// the one its message states, 1*AA (the reference reads an uninitialised
// second factor for three columns per SNP and no column for two).
void dosage_setup(const omicabelnomm_settings& s, int file_r, std::vector<double>& f, bool& summed) {
  static const char* names[] = {"Additive", "Dominant", "Recessive"};
  switch (s.model) {
    case 0: case 1: case 2:
      if (file_r != 2 && file_r != 3)
        throw Fail(1, std::string("The amount of columns per Independent Variable (--ngpred) is not 2 or 3 for a "
                                  "valid ") + names[s.model] + " Model!");
      f = s.model == 0 ? std::vector<double>{2.0, 1.0}
                       : (s.model == 1 ? std::vector<double>{1.0, 1.0} : std::vector<double>{1.0});
      summed = true;
      return;
    case 3: case 4:
      if (file_r > OAGPU_MAX_DOSAGE_FACTORS)
        throw Fail(1, "dosage model: at most " + std::to_string(OAGPU_MAX_DOSAGE_FACTORS) + " columns per SNP");
      f = read_dosage_factors(str_or(s.fname_dosages), file_r);
      summed = s.model == 4;
      return;
    default:
      throw Fail(1, "Requested dosages model wihtout a valid model!");
  }
}

struct Plan {
  int n_file = 0, n = 0, realN = 0, l = 0, r_file = 1, r = 1, t = 0, m_file = 0, m = 0;
  int r_model = 1;                     // genotype columns per file SNP after the dosage model
  bool dos = false, dos_sum = false;
  std::vector<double> dos_f;
  int tb = 0, mb = 0;                  // reference block sizes (output grouping)
  bool inter = false, multi = false, keep = false;
  int num_inter = 0, ini_inter = 0, file_block = 0;   // multi-set: file SNPs per block
  std::vector<uint8_t> keep_rows;
  std::vector<std::string> y_names, al_names, ar_names;
};

// Virtual SNP layout: for SNP v, output column c reads genotype file column
// src(v, c) times interaction column env(v, c).  Columns are relative to
// the genotype block [f0, f0 + nf) of file SNPs.
void descriptors(const Plan& P, int f0, int nf, std::vector<int32_t>& src, std::vector<int32_t>& env,
                 int& v0, int& nv) {
  src.clear();
  env.clear();
  if (!P.inter) {
    v0 = f0;
    nv = nf;
    for (int s = 0; s < nf; ++s)
      for (int c = 0; c < P.r; ++c) {
        src.push_back(s * P.r_model + c);
        env.push_back(-1);
      }
    return;
  }
  if (!P.multi) {   // [I_1*X, ..., I_c*X, (X)] per SNP (src/AIOwrapper.cpp:796-840)
    v0 = f0;
    nv = nf;
    for (int s = 0; s < nf; ++s) {
      for (int h = 0; h < P.num_inter; ++h) {
        src.push_back(s);
        env.push_back(h);
      }
      if (P.keep) {
        src.push_back(s);
        env.push_back(-1);
      }
    }
    return;
  }
  // One analysis per interaction: file blocks of file_block SNPs; inside a
  // block, interaction h outer and SNP inner (src/AIOwrapper.cpp:767-794).
  v0 = f0 * P.num_inter;
  nv = nf * P.num_inter;
  for (int b0 = 0; b0 < nf; b0 += P.file_block) {
    const int bs = std::min(P.file_block, nf - b0);
    for (int h = 0; h < P.num_inter; ++h)
      for (int i = 0; i < bs; ++i) {
        src.push_back(b0 + i);
        env.push_back(h);
        if (P.keep) {
          src.push_back(b0 + i);
          env.push_back(-1);
        }
      }
  }
}

}  // namespace

extern "C" void omicabelnomm_apply_defaults(omicabelnomm_settings* s) {
  // src/Algorithm.cpp:277-311
  std::memset(s, 0, sizeof(*s));
  s->fname_excludelist = "";
  s->fname_interactions = "";
  s->threads = 1;
  s->r = 1;
  s->ini_IDX_interactions = -1;
  s->end_IDX_interactions = -1;
  s->minR2store = 0.00001f;
  s->minR2disp = 0.000001f;
  s->minPstore = 0.1f;
  s->minPdisp = 0.05f;
  s->limit_m = 2147483647;
  s->limit_t = 2147483647;
  s->limit_n = 2147483647;
  s->precision = OAGPU_FP64;
  s->device = 0;
  s->verbose = 1;
  s->dosages = 0;
  s->model = -1;
  s->fname_dosages = "";
}

static int solve_impl(const omicabelnomm_settings* sin, omicabelnomm_outputs* out) {
  omicabelnomm_settings s = *sin;
  const bool verbose = s.verbose != 0;
  // Threshold normalisation (src/Algorithm.cpp:338-342).
  if (s.minR2disp < s.minR2store || s.storeBin) s.minR2store = s.minR2disp;
  if (s.minPdisp > s.minPstore || s.storeBin) s.minPstore = s.minPdisp;

  const std::string fY = str_or(s.fnameY), fAL = str_or(s.fnameAL), fAR = str_or(s.fnameAR),
                    fOut = str_or(s.fnameOutFiles);
  Fvi yfvi = read_fvi(fY + ".fvi");
  Fvi alfvi = read_fvi(fAL + ".fvi");
  Fvi arfvi = read_fvi(fAR + ".fvi");
  if (alfvi.nobs != arfvi.nobs || arfvi.nobs != yfvi.nobs || alfvi.nobs != yfvi.nobs) {
    std::ostringstream m;
    m << "The number of elements/individuals from the input files do not coincide with each other! See:\n"
      << fY << ":" << yfvi.nobs << "\n" << fAL << ":" << alfvi.nobs << "\n" << fAR << ":" << arfvi.nobs;
    throw Fail(1, m.str());
  }
  Plan P;
  P.r_file = std::max(1, s.r);
  P.n_file = std::min((int)alfvi.nobs, s.limit_n);
  P.realN = P.n_file;
  if ((int)arfvi.nvars % P.r_file != 0) {
    std::ostringstream m;
    m << "Number of elements in " << fAR << ".fvi:" << arfvi.nvars
      << " is not consistent with the value of --ngpred (-n)" << P.r_file << " provided";
    throw Fail(1, m.str());
  }
  P.m_file = std::min((int)arfvi.nvars / P.r_file, s.limit_m / P.r_file);
  P.t = std::min((int)yfvi.nvars, s.limit_t);
  P.l = (int)alfvi.nvars;
  for (int i = 0; i < P.n_file; ++i) {
    const std::string yn = yfvi.obs_name(i), an = alfvi.obs_name(i), rn = arfvi.obs_name(i);
    if (yn != an || yn != rn) {
      if (verbose)
        std::cout << "Warning, ID names between the -p, -c and -g files do not coincide! The files must "
                     "have the same meaningful order"
                  << std::endl;
      break;
    }
  }
  for (int i = 0; i < P.t; ++i) P.y_names.push_back(yfvi.var_name(i));
  for (int i = 0; i < P.l; ++i) P.al_names.push_back(alfvi.var_name(i));

  // Row exclusion: exclusion list, then covariate NaN, then interaction NaN
  // (src/AIOwrapper.cpp:229-253, 434-470).
  P.keep_rows.assign(P.n_file, 1);
  const std::string fexcl = str_or(s.fname_excludelist);
  if (!fexcl.empty()) apply_exclusion_file(fexcl, P.n_file, P.keep_rows, verbose);
  std::vector<double> xl_all;
  {
    FvdReader rd(fAL + ".fvd", alfvi);
    rd.read_columns(0, P.l, xl_all);
    int almiss = 0;
    for (int h = 0; h < P.l; ++h)
      for (int i = 0; i < P.n_file; ++i)
        if (std::isnan(xl_all[(size_t)h * alfvi.nobs + i]) && P.keep_rows[i]) {
          P.keep_rows[i] = 0;
          ++almiss;
        }
    if (almiss && verbose) std::cout << "Excluding " << almiss << " from covariate missings" << std::endl;
  }
  // Dosage model (src/AIOwrapper.cpp:33-38,256-340).
  P.r_model = P.r_file;
  if (s.dosages || s.model != -1) {
    if (s.model == -1) throw Fail(1, "Requested dosages model wihtout a valid model!");
    dosage_setup(s, P.r_file, P.dos_f, P.dos_sum);
    P.dos = true;
    P.r_model = P.dos_sum ? 1 : (int)P.dos_f.size();
  }
  std::vector<double> e_all;
  Fvi ifvi;
  if (s.use_interactions) {
    P.inter = true;
    P.multi = s.use_multiple_interaction_sets != 0;
    P.keep = s.keep_depVar != 0;
    if (P.r_model != 1)   // src/AIOwrapper.cpp:352-366
      throw Fail(1, std::string(P.dos ? "Cannot use Interactions with the current dosage model! #Of factors can be "
                                        "at most 1!\n"
                                      : "") +
                        "Cannot use Interactions with --ngpred set to: " + std::to_string(P.r_file) +
                        ". Required is 1!");
    const std::string fI = str_or(s.fname_interactions);
    ifvi = read_fvi(fI + ".fvi");
    if (ifvi.nobs != arfvi.nobs)
      throw Fail(1, "The number of elements/individuals from the input files do not coincide with each other!");
    if (s.end_IDX_interactions != -1 || s.ini_IDX_interactions != -1) {
      P.num_inter = s.end_IDX_interactions - s.ini_IDX_interactions + 1;
      P.ini_inter = s.ini_IDX_interactions;
    } else {
      P.num_inter = (int)ifvi.nvars;
      P.ini_inter = 0;
    }
    if (P.num_inter <= 0 || P.ini_inter < 0 || P.ini_inter + P.num_inter > (int)ifvi.nvars)
      throw Fail(1, "The specified range for data to be read from the Interactions file exceeds its content dimensions!");
    FvdReader rd(fI + ".fvd", ifvi);
    rd.read_columns(P.ini_inter, P.num_inter, e_all);
    int imiss = 0;
    for (int h = 0; h < P.num_inter; ++h)
      for (int i = 0; i < P.n_file; ++i)
        if (std::isnan(e_all[(size_t)h * ifvi.nobs + i]) && P.keep_rows[i]) {
          P.keep_rows[i] = 0;
          ++imiss;
        }
    if (imiss && verbose) std::cout << "Excluding " << imiss << " from Interaction missings" << std::endl;
  }
  P.n = 0;
  for (int i = 0; i < P.n_file; ++i) P.n += P.keep_rows[i];

  // Model shape and AR names (src/AIOwrapper.cpp:421-446, 563-657).
  const int mb_file = std::min(P.m_file, 500);
  // Under a summed dosage model a SNP is named by its first column.
  std::vector<std::string> file_names;
  for (int i = 0; i < P.m_file * P.r_file; i += (P.dos && P.dos_sum) ? P.r_file : 1)
    file_names.push_back(arfvi.var_name(i));
  if (!P.inter) {
    P.r = P.r_model;
    P.m = P.m_file;
    P.mb = mb_file;
    P.ar_names = file_names;
  } else if (!P.multi) {
    P.r = P.num_inter + (P.keep ? 1 : 0);
    P.m = P.m_file;
    P.mb = mb_file;
    for (int i = 0; i < P.m_file; ++i) {
      for (int h = 0; h < P.num_inter; ++h)
        P.ar_names.push_back("I" + std::to_string(h + 1) + "x" + file_names[i]);
      if (P.keep) P.ar_names.push_back(file_names[i]);
    }
  } else {
    P.r = 1 + (P.keep ? 1 : 0);
    P.file_block = std::max(1, mb_file / P.num_inter);
    P.m = P.m_file * P.num_inter;
    P.mb = P.file_block * P.num_inter;
    for (int b0 = 0; b0 < P.m_file; b0 += P.file_block) {
      const int bs = std::min(P.file_block, P.m_file - b0);
      for (int h = 0; h < P.num_inter; ++h) {
        const std::string pre = "I" + std::to_string(h + 1) + "_";
        for (int i = 0; i < bs; ++i) {
          P.ar_names.push_back(pre + file_names[b0 + i]);
          if (P.keep) P.ar_names.push_back(pre + "o_" + file_names[b0 + i]);
        }
      }
    }
  }
  P.tb = std::min(P.t, 1000);
  const int p = P.l + P.r;
  if (p >= P.n)
    throw Fail(1, "Error! The amount of unknows (individuals|elements) must be greater than total of "
                  "variables (#covariates+#ind.variables)!");
  if (P.m <= 0 || P.t <= 0) throw Fail(1, "Nothing to scan: no SNPs or no traits.");
  if (verbose && (P.l > 255 || p > 255 || P.n > 65535))
    std::cout << "Warning, output binary format does not yet support current problem sizes for the "
                 "provided p, l, r and n.\nOmitting outputfile."
              << std::endl;

  if (s.storeBin) {
    std::vector<const char*> aln, arn, yn;
    for (int i = 1; i < P.l; ++i) aln.push_back(alfvi.var_name_raw(i));
    for (int i = 0; i < P.m_file * P.r_file; i += (P.dos && P.dos_sum) ? P.r_file : 1)
      arn.push_back(arfvi.var_name_raw(i));
    for (int i = 0; i < P.t; ++i) yn.push_back(yfvi.var_name_raw(i));
    write_ibin(fOut, P.n, P.l, P.r, P.t, P.m, arfvi.namelength, s.disp_cov, s.storePInd, aln, arn, yn);
  }

  WriterConfig wc;
  wc.out = fOut;
  wc.store_bin = s.storeBin;
  wc.store_all = s.storePInd;
  wc.disp_cov = s.disp_cov;
  wc.min_p_disp = s.minPdisp;
  wc.min_r2_disp = s.minR2disp;
  wc.l = P.l;
  wc.r = P.r;
  wc.real_n = P.realN;
  wc.y_names = P.y_names;
  wc.al_names = P.al_names;
  wc.ar_names = P.ar_names;
  ResultWriter writer(wc);

  // GPU context.
  oagpu_params gp{};
  gp.n = P.n;
  gp.l = P.l;
  gp.r = P.r;
  gp.real_n = P.realN;
  gp.precision = s.precision;
  gp.disp_cov = s.disp_cov;
  gp.store_all = s.storePInd;
  gp.min_r2_store = s.minR2store;
  gp.min_p_store = s.minPstore;
  gp.min_p_disp = s.minPdisp;
  gp.max_traits = P.tb;
  gp.max_snps = 0;
  gp.max_hits = 1ll << 22;
  gp.device = s.device;
  oagpu_ctx* ctx = nullptr;
  int rc = oagpu_create(&gp, &ctx);
  std::unique_ptr<oagpu_ctx, void (*)(oagpu_ctx*)> guard(ctx, oagpu_destroy);
  if (rc != OAGPU_OK) throw Fail(rc, ctx ? oagpu_last_error(ctx) : "no GPU device");

  std::vector<double> XL((size_t)P.n * P.l);
  for (int h = 0; h < P.l; ++h) {
    size_t o = 0;
    for (int i = 0; i < P.n_file; ++i)
      if (P.keep_rows[i]) XL[(size_t)h * P.n + o++] = xl_all[(size_t)h * alfvi.nobs + i];
  }
  rc = oagpu_set_covariates(ctx, XL.data(), P.n);
  if (rc != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
  if (P.dos) {
    rc = oagpu_set_dosage_model(ctx, P.dos_sum ? OAGPU_DOSAGE_SUM : OAGPU_DOSAGE_SCALE, P.r_file,
                                (int)P.dos_f.size(), P.dos_f.data());
    if (rc != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
  }
  if (P.inter) {
    std::vector<double> E((size_t)P.n * P.num_inter);
    for (int h = 0; h < P.num_inter; ++h) {
      size_t o = 0;
      for (int i = 0; i < P.n_file; ++i)
        if (P.keep_rows[i]) E[(size_t)h * P.n + o++] = e_all[(size_t)h * ifvi.nobs + i];
    }
    rc = oagpu_set_interactions(ctx, E.data(), P.n, P.num_inter);
    if (rc != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
  }

  // Genotype block: a whole number of reference SNP blocks (and of
  // interaction file blocks), about 512 MiB of host data.
  const bool g64 = arfvi.type == 6;
  const size_t gelem = g64 ? 8 : 4;
  const int unit = P.multi ? P.file_block : mb_file;
  int gb_file = (int)std::max<size_t>(1, ((512ull << 20) / ((size_t)P.n * gelem * P.r_file)) / unit) * unit;
  if (const char* e = std::getenv("OMICABELNOMM_BLOCK_SNPS"))   // smaller blocks (tests)
    gb_file = std::max(unit, (std::atoi(e) / unit) * unit);
  gb_file = std::min(gb_file, ((P.m_file + unit - 1) / unit) * unit);
  const bool y64 = yfvi.type == 6;

  FvdReader yrd(fY + ".fvd", yfvi);
  FvdReader grd(fAR + ".fvd", arfvi);
  std::vector<oagpu_hit> hits((size_t)gp.max_hits);
  long long total_sig = 0;
  double acc_loady = 0, acc_loadxr = 0, acc_gpu = 0, acc_store = 0;
  const double t_start = now_s();

  // Two page-locked genotype blocks: while the GPU scans block f, the
  // host reads block f+1 from disk into the other buffer and its upload is
  // queued on the library's copy stream.
  const size_t gblock_bytes = (size_t)gb_file * P.r_file * P.n * gelem;
  void* hbuf[2] = {nullptr, nullptr};
  for (int i = 0; i < 2; ++i)
    if (oagpu_host_alloc(gblock_bytes, &hbuf[i]) != OAGPU_OK) throw Fail(1, "cannot allocate pinned host memory");
  struct HostBufs {
    void* p[2];
    ~HostBufs() { oagpu_host_free(p[0]); oagpu_host_free(p[1]); }
  } hold{{hbuf[0], hbuf[1]}};

  auto read_geno = [&](int f0, void* buf) {
    const int nf = std::min(gb_file, P.m_file - f0);
    if (g64)
      grd.read_kept<double>((size_t)f0 * P.r_file, (size_t)nf * P.r_file, P.n_file, P.keep_rows, P.n,
                            reinterpret_cast<double*>(buf));
    else
      grd.read_kept<float>((size_t)f0 * P.r_file, (size_t)nf * P.r_file, P.n_file, P.keep_rows, P.n,
                           reinterpret_cast<float*>(buf));
  };
  auto upload = [&](int f0, void* buf) {
    const int nf = std::min(gb_file, P.m_file - f0);
    std::vector<int32_t> src, env;
    int v0 = 0, nv = 0;
    descriptors(P, 0, nf, src, env, v0, nv);
    v0 = P.multi ? f0 * P.num_inter : f0;
    const int r2 = oagpu_load_genotypes(ctx, buf, g64 ? OAGPU_F64 : OAGPU_F32, OAGPU_HOST, P.n, nf * P.r_file,
                                        nv, src.data(), P.inter ? env.data() : nullptr, v0);
    if (r2 != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
  };
  // Collect a finished scan's hits (the host buffer matches the device
  // arena, so an overflow is an arena overflow).
  auto collect = [&](int f0, std::vector<oagpu_hit>& out_hits) -> size_t {
    oagpu_stats st{};
    const int r2 = oagpu_scan_wait(ctx, nullptr, hits.data(), (long long)hits.size(), &st);
    if (r2 == OAGPU_ERR_HIT_OVERFLOW)
      throw Fail(1, std::string(oagpu_last_error(ctx)) + " -- raise the thresholds (-s/-e) or split the input");
    if (r2 != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
    total_sig += st.n_sig;
    out_hits.assign(hits.begin(), hits.begin() + (size_t)st.n_hits);
    (void)f0;
    return (size_t)st.n_hits;
  };
  // Regroup a block's hits into the reference's (trait block, SNP block) lists.
  auto write_hits = [&](std::vector<oagpu_hit>& h) {
    std::stable_sort(h.begin(), h.end(), [&](const oagpu_hit& a, const oagpu_hit& b) {
      return a.snp / P.mb < b.snp / P.mb;
    });
    for (size_t i = 0; i < h.size();) {
      const int blk = h[i].snp / P.mb;
      size_t j = i;
      while (j < h.size() && h[j].snp / P.mb == blk) ++j;
      writer.write_block(h.data() + i, j - i, blk * P.mb);
      i = j;
    }
  };

  for (int y0 = 0; y0 < P.t; y0 += P.tb) {
    const int tcur = std::min(P.tb, P.t - y0);
    double t0 = now_s();
    std::vector<unsigned char> ybuf((size_t)tcur * P.n * (y64 ? 8 : 4));
    if (y64)
      yrd.read_kept<double>(y0, tcur, P.n_file, P.keep_rows, P.n, reinterpret_cast<double*>(ybuf.data()));
    else
      yrd.read_kept<float>(y0, tcur, P.n_file, P.keep_rows, P.n, reinterpret_cast<float*>(ybuf.data()));
    rc = oagpu_load_traits(ctx, ybuf.data(), y64 ? OAGPU_F64 : OAGPU_F32, OAGPU_HOST, P.n, tcur, y0);
    if (rc != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
    acc_loady += now_s() - t0;

    t0 = now_s();
    read_geno(0, hbuf[0]);
    acc_loadxr += now_s() - t0;
    upload(0, hbuf[0]);
    rc = oagpu_scan_async(ctx, 0);
    if (rc != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
    std::vector<oagpu_hit> done;
    int bi = 0;
    for (int f0 = 0; f0 < P.m_file; f0 += gb_file, ++bi) {
      const int f1 = f0 + gb_file;
      if (f1 < P.m_file) {
        t0 = now_s();
        read_geno(f1, hbuf[(bi + 1) & 1]);   // overlaps the scan of block f0
        acc_loadxr += now_s() - t0;
        upload(f1, hbuf[(bi + 1) & 1]);      // its copy overlaps that scan too
      }
      t0 = now_s();
      collect(f0, done);
      if (f1 < P.m_file) {
        rc = oagpu_scan_async(ctx, 0);
        if (rc != OAGPU_OK) throw Fail(1, oagpu_last_error(ctx));
      }
      acc_gpu += now_s() - t0;
      t0 = now_s();
      write_hits(done);                      // overlaps the next scan
      acc_store += now_s() - t0;
    }
  }
  if (out) {
    out->duration = now_s() - t_start;
    out->acc_loady = acc_loady;
    out->acc_loadxr = acc_loadxr;
    out->acc_gpu = acc_gpu;
    out->acc_storeb = acc_store;
    out->total_sig_results = total_sig;
    out->n_txt_lines = writer.txt_lines();
    out->n_bin_results = writer.bin_results();
    out->n = P.n;
    out->l = P.l;
    out->r = P.r;
    out->t = P.t;
    out->m = P.m;
    out->realN = P.realN;
  }
  return 0;
}

extern "C" int omicabelnomm_solve(const omicabelnomm_settings* s, omicabelnomm_outputs* out, char* errbuf,
                                  int errlen) {
  if (errbuf && errlen > 0) errbuf[0] = 0;
  try {
    return solve_impl(s, out);
  } catch (const Fail& f) {
    if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", f.what());
    return f.code ? f.code : 1;
  } catch (const std::exception& e) {
    if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", e.what());
    return 1;
  }
}
