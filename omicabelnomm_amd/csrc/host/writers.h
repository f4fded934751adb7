// Result sinks of the scan: <out>_dis.txt, <out>_sto.ibin, <out>_sto.dbin
// in the reference's formats (src/AIOwrapper.cpp:679-733, 936-945,
// 1440-1558), readable by OmicOUT (OmicOUT/main.cpp:280-461).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "oagpu.h"

namespace oahost {

struct WriterConfig {
  std::string out;          // output prefix
  bool store_bin = false;   // -b
  bool store_all = false;   // -f (storePInd)
  bool disp_cov = false;    // -i
  float min_p_disp = 0.05f;
  float min_r2_disp = 0.000001f;
  int l = 0, r = 1;
  int real_n = 1;           // rows before exclusion
  std::vector<std::string> y_names, al_names, ar_names;
};

class ResultWriter {
 public:
  explicit ResultWriter(const WriterConfig& cfg);
  ~ResultWriter();
  ResultWriter(const ResultWriter&) = delete;
  ResultWriter& operator=(const ResultWriter&) = delete;

  // One (trait block, SNP block) list, like one write_OutFiles call:
  // hits sorted by (trait, snp, h descending), all with snp in the block
  // starting at snp_block_start; ar_offset = snp_block_start * r.
  void write_block(const oagpu_hit* hits, size_t count, int snp_block_start);

  long long txt_lines() const { return txt_lines_; }
  long long bin_results() const { return bin_results_; }

 private:
  WriterConfig cfg_;
  FILE* txt_ = nullptr;
  FILE* bin_ = nullptr;
  std::vector<char> line_;
  long long txt_lines_ = 0, bin_results_ = 0;
};

// The _sto.ibin header and name tables (src/AIOwrapper.cpp:679-733).
void write_ibin(const std::string& out, int n, int l, int r, int t, int m, uint32_t namelength,
                bool disp_cov, bool store_all, const std::vector<const char*>& al_names_raw,
                const std::vector<const char*>& ar_names_raw, const std::vector<const char*>& y_names_raw);

// bf16-by-truncation of a float (src/Utility.cpp:262-268).
uint16_t float16_trunc(float x);

}  // namespace oahost
