// omicabelnomm: command line of the scan (src/main.cpp:102-500), running the
// scan on an MI355X through libomicabelnomm/liboagpu.
//
// Flags and messages follow the reference.  Differences:
//   -v/--simpleinter and -w/--multinter take the interaction file path (the
//     reference assigns atof(optarg) to the path, src/main.cpp:370,381);
//   -u/--keepinter is a plain flag (the reference requires an argument);
//   -z/--mylinear and -y/--myaddit take the factor file path (the reference
//     never stores it, src/main.cpp:341-358, so its read fails);
//   -l/--recessive is the model its message states, 1*AA;
//   -m/--mem is accepted and ignored (the reference aborts);
//   --precision fp64|fp32|bf16x3 and --device N select the GPU path.
#include <getopt.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "oagpu.h"
#include "omicabelnomm.h"

static const char* kUsage =
    "usage: omicabelnomm -c <path/fname> --geno <path/fname> -p <path/fname> -o <path/fname> "
    "-x <path/fname> -n <#SNPcols> -t <#CPUs> -d <0.0~1.0> -r <-10.0~1.0> -b -s <0.0~1.0> "
    "-e <-10.0~1.0> -i -f [-j|-k|-l|-z <factors>|-y <factors>] [--precision fp64|fp32|bf16x3] "
    "[--device N]";

static std::string strip_ext(const char* arg) {
  std::string s(arg);
  const size_t pos = s.find_last_of('.');
  return pos == std::string::npos ? s : s.substr(0, pos);
}

int main(int argc, char* argv[]) {
  omicabelnomm_settings s;
  omicabelnomm_apply_defaults(&s);
  std::string fY, fAR, fAL, fOut, fExcl, fInter, fDos;
  bool phe = false, snp = false, cov = false, bout = false;
  static struct option opts[] = {
      {"phe", required_argument, 0, 'p'},       {"geno", required_argument, 0, 'g'},
      {"cov", required_argument, 0, 'c'},       {"out", required_argument, 0, 'o'},
      {"ngpred", required_argument, 0, 'n'},    {"thr", required_argument, 0, 't'},
      {"mem", required_argument, 0, 'm'},       {"excl", required_argument, 0, 'x'},
      {"stobin", no_argument, 0, 'b'},          {"pdisp", required_argument, 0, 'd'},
      {"psto", required_argument, 0, 's'},      {"rdisp", required_argument, 0, 'r'},
      {"rsto", required_argument, 0, 'e'},      {"fdcov", no_argument, 0, 'i'},
      {"fdgen", no_argument, 0, 'f'},           {"additive", no_argument, 0, 'j'},
      {"dominant", no_argument, 0, 'k'},        {"recessive", no_argument, 0, 'l'},
      {"mylinear", required_argument, 0, 'z'},  {"myaddit", required_argument, 0, 'y'},
      {"simpleinter", required_argument, 0, 'v'}, {"multinter", required_argument, 0, 'w'},
      {"keepinter", no_argument, 0, 'u'},       {"help", no_argument, 0, 'h'},
      {"precision", required_argument, 0, 1000}, {"device", required_argument, 0, 1001},
      {0, 0, 0, 0}};
  for (;;) {
    int idx = 0;
    const int c = getopt_long(argc, argv, "c:p:g:o:n:t:m:x:d:s:r:e:z:y:v:w:ufibhjkl", opts, &idx);
    if (c == -1) break;
    switch (c) {
      case 'p':
        phe = !phe;
        fY = strip_ext(optarg);
        std::cout << "-p Reading phenotypes from file " << optarg << std::endl;
        break;
      case 'g':
        snp = !snp;
        fAR = strip_ext(optarg);
        std::cout << "-g Reading with genotype data from file " << optarg << std::endl;
        break;
      case 'c':
        cov = !cov;
        fAL = strip_ext(optarg);
        std::cout << "-c Reading covariates from file " << optarg << std::endl;
        break;
      case 'o':
        bout = !bout;
        fOut = strip_ext(optarg);
        std::cout << "-o Writing output files to " << optarg << std::endl;
        break;
      case 'n':
        s.r = std::max(std::atoi(optarg), 1);
        std::cout << "-n Using columns per snp as " << s.r << std::endl;
        break;
      case 't':
        s.threads = std::min(std::max(std::atoi(optarg), 1), 32);
        std::cout << "-t Using " << s.threads << " CPU threads for parallelism" << std::endl;
        break;
      case 'm':
        std::cout << "-m ignored: device memory is managed by the GPU scan" << std::endl;
        break;
      case 'x':
        fExcl = optarg;
        std::cout << "-x Excluding ids on " << fExcl << std::endl;
        break;
      case 'd':
        s.minPdisp = (float)std::fabs(std::atof(optarg));
        std::cout << "-d Significance to display in .txt will be P-val's below " << s.minPdisp << std::endl;
        break;
      case 's':
        s.minPstore = (float)std::fabs(std::atof(optarg));
        std::cout << "-s Significance to store in .bin will be P-val's below " << s.minPstore << std::endl;
        break;
      case 'r':
        s.minR2disp = (float)std::atof(optarg);
        std::cout << "-r Minimum R2 to display in .txt will be above " << s.minR2disp << std::endl;
        break;
      case 'e':
        s.minR2store = (float)std::atof(optarg);
        std::cout << "-e Minimum R2 to store in .bin will be above " << s.minR2store << std::endl;
        break;
      case 'i':
        s.disp_cov = 1;
        std::cout << "-i Covariate results (significant) will be included in results whenever their "
                     "respective snps are also significant"
                  << std::endl;
        break;
      case 'f':
        s.storePInd = 1;
        std::cout << "-f Forcing all included results to be considered independently of max P-val or "
                     "min R2. (SLOW!)"
                  << std::endl;
        break;
      case 'b':
        s.storeBin = 1;
        std::cout << "-b Results will be stored in binary format too" << std::endl;
        break;
      case 'v':
      case 'w':
        fInter = strip_ext(optarg);
        s.use_interactions = 1;
        s.keep_depVar = 1;
        s.use_multiple_interaction_sets = (c == 'w');
        std::cout << (c == 'v' ? "-v File containing single interactions " : "-w File containing multiple interactions ")
                  << fInter << std::endl;
        break;
      case 'u':
        s.keep_depVar = 1;
        std::cout << "-u Keeping independent variable for interaction analysis " << std::endl;
        break;
      case 'j':   // src/main.cpp:314-358
        s.model = 0;
        s.dosages = 1;
        std::cout << "-j Using Additive Model with (2*AA, 1*AB, 0*BB) effects" << std::endl;
        break;
      case 'k':
        s.model = 1;
        s.dosages = 1;
        std::cout << "-k Using Dominant Model with (1*AA, 1*AB, 0*BB) effects" << std::endl;
        break;
      case 'l':
        s.model = 2;
        s.dosages = 1;
        std::cout << "-l Using Recessive Model with (1*AA, 0*AB, 0*BB) effects" << std::endl;
        break;
      case 'z':
      case 'y':
        s.model = c == 'z' ? 3 : 4;
        s.dosages = 1;
        fDos = optarg;
        std::cout << (c == 'z' ? "-z Using Custom Linear Model with parameters read from the file "
                               : "-y Using Custom Additive Model with parameters read from the file ")
                  << fDos << std::endl;
        break;
      case 1000:
        if (!std::strcmp(optarg, "fp64")) s.precision = OAGPU_FP64;
        else if (!std::strcmp(optarg, "fp32")) s.precision = OAGPU_FP32;
        else if (!std::strcmp(optarg, "bf16x3")) s.precision = OAGPU_BF16X3;
        else {
          std::cerr << "--precision must be fp64, fp32 or bf16x3" << std::endl;
          return 1;
        }
        break;
      case 1001:
        s.device = std::atoi(optarg);
        break;
      case 'h':
        std::cout << kUsage << std::endl;
        return 1;
      case '?':
      default:
        break;
    }
  }
  if (optind < argc) {
    std::printf("non-option ARGV-elements: ");
    while (optind < argc) std::printf("%s ", argv[optind++]);
    std::putchar('\n');
  }
  if (!bout || !phe || !cov || !snp) {
    std::cout << "Missing arguments! p:" << phe << " g:" << snp << " c:" << cov << " o:" << bout << std::endl;
    return 1;
  }
  s.fnameY = fY.c_str();
  s.fnameAR = fAR.c_str();
  s.fnameAL = fAL.c_str();
  s.fnameOutFiles = fOut.c_str();
  s.fname_excludelist = fExcl.c_str();
  s.fname_interactions = fInter.c_str();
  s.fname_dosages = fDos.c_str();
  std::cout << "omicabelnomm (MI355X)" << std::endl;
  omicabelnomm_outputs out;
  std::memset(&out, 0, sizeof(out));
  char err[4096];
  const int rc = omicabelnomm_solve(&s, &out, err, sizeof(err));
  if (rc != 0) {
    std::cout << err << std::endl;
    return 1;
  }
  std::cout << std::endl << "Number of Significant Results: " << out.total_sig_results << std::endl;
  std::cout << std::endl;
  return 0;
}
