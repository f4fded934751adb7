#include "writers.h"

#include <cstring>
#include <stdexcept>

namespace oahost {

uint16_t float16_trunc(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return (uint16_t)(u >> 16);
}

ResultWriter::ResultWriter(const WriterConfig& cfg) : cfg_(cfg) {
  const std::string txt = cfg.out + "_dis.txt";
  txt_ = std::fopen(txt.c_str(), "w");
  if (!txt_) throw std::runtime_error("Error Creating File " + txt);
  std::fputs("Phe\tSNP\tn\tnPct\tB\tR2\tSE\tT\tP\n", txt_);
  if (cfg.store_bin) {
    const std::string bin = cfg.out + "_sto.dbin";
    bin_ = std::fopen(bin.c_str(), "wb");
    if (!bin_) throw std::runtime_error("Error Creating File " + cfg.out + "_sto.bin");
  }
  line_.resize(4096);
}

ResultWriter::~ResultWriter() {
  if (txt_) std::fclose(txt_);
  if (bin_) std::fclose(bin_);
}

void ResultWriter::write_block(const oagpu_hit* hits, size_t count, int snp_block_start) {
  if (count == 0) return;
  const int l = cfg_.l, r = cfg_.r;
  const int ar_offset = snp_block_start * r;
  // Group consecutive hits of one (trait, SNP) into a record (resultH).
  std::vector<std::pair<size_t, size_t>> recs;
  for (size_t i = 0; i < count;) {
    size_t j = i + 1;
    while (j < count && hits[j].trait == hits[i].trait && hits[j].snp == hits[i].snp) ++j;
    recs.emplace_back(i, j);
    i = j;
  }
  if (bin_ && !cfg_.store_all) {
    const int32_t hdr[3] = {hits[0].trait, (int32_t)recs.size(), ar_offset};
    std::fwrite(hdr, sizeof(int32_t), 3, bin_);
  }
  for (const auto& rc : recs) {
    const oagpu_hit& h0 = hits[rc.first];
    const uint16_t n_used = (uint16_t)h0.n_used;
    const float r2f = (float)h0.r2;
    if (bin_) {
      const uint16_t r2b = float16_trunc(r2f);
      std::fwrite(&n_used, 2, 1, bin_);
      std::fwrite(&r2b, 2, 1, bin_);
      if (!cfg_.store_all) {
        const uint8_t size_p = (uint8_t)(rc.second - rc.first);
        std::fwrite(&size_p, 1, 1, bin_);
      }
    }
    for (size_t k = rc.first; k < rc.second; ++k) {
      const oagpu_hit& h = hits[k];
      const int ii = h.snp - snp_block_start;
      const bool snp_term = h.h >= l;
      const int ar_idx = ii * r + (snp_term ? h.h - l : 0);
      const uint8_t al_idx = snp_term ? (uint8_t)255 : (uint8_t)h.h;
      const float B = (float)h.b, SE = (float)h.se, T = (float)h.t;
      const long double P = (long double)h.p;
      if ((P <= cfg_.min_p_disp && h.r2 >= (double)cfg_.min_r2_disp) || cfg_.store_all) {
        std::string aname = " ";
        if (al_idx < l) aname = cfg_.al_names[al_idx] + ":";
        aname += cfg_.ar_names[(size_t)ar_idx + ar_offset];
        // nUsedPct = (n - Ymiss) / realN (src/Algorithm.cpp:1058)
        const float pct = (float)h.n_used / (float)cfg_.real_n;
        std::fprintf(txt_, "%s\t%s\t%u\t%.2f\t%.7f\t%.4f\t%.7f\t%.7f\t%.7Le\n",
                     cfg_.y_names[h.trait].c_str(), aname.c_str(), (unsigned)n_used, (double)pct,
                     (double)B, (double)r2f, (double)SE, (double)T, P);
        ++txt_lines_;
      }
      if (bin_) {
        if (cfg_.disp_cov && !cfg_.store_all) std::fwrite(&al_idx, 1, 1, bin_);
        if (!cfg_.store_all) {
          const uint16_t a16 = (uint16_t)ar_idx;
          std::fwrite(&a16, 2, 1, bin_);
        }
        std::fwrite(&B, 4, 1, bin_);
        std::fwrite(&SE, 4, 1, bin_);
        ++bin_results_;
      }
    }
  }
}

void write_ibin(const std::string& out, int n, int l, int r, int t, int m, uint32_t namelength,
                bool disp_cov, bool store_all, const std::vector<const char*>& al_names_raw,
                const std::vector<const char*>& ar_names_raw, const std::vector<const char*>& y_names_raw) {
  const std::string path = out + "_sto.ibin";
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Error Creating File: " + path);
  const int32_t hdr[6] = {n, l, r, t, m, (int32_t)namelength};
  std::fwrite(hdr, 4, 6, f);
  const uint8_t b0 = disp_cov ? 1 : 0, b1 = store_all ? 1 : 0;
  std::fwrite(&b0, 1, 1, f);
  std::fwrite(&b1, 1, 1, f);
  for (const auto* list : {&al_names_raw, &ar_names_raw, &y_names_raw})
    for (const char* p : *list) std::fwrite(p, 1, namelength, f);
  std::fclose(f);
}

}  // namespace oahost
