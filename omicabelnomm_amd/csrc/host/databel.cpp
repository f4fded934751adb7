#include "databel.h"

#include <cstring>
#include <stdexcept>

namespace oahost {

std::string Fvi::obs_name(size_t i) const {
  const char* p = names.data() + i * (size_t)namelength;
  return std::string(p, strnlen(p, namelength));
}

std::string Fvi::var_name(size_t j) const {
  const char* p = var_name_raw(j);
  return std::string(p, strnlen(p, namelength));
}

size_t Fvi::elem_size() const {
  switch (type) {
    case 1: case 2: return 2;
    case 3: case 4: return 4;
    case 5: return 4;
    case 6: return 8;
    case 7: case 8: return 1;
    default: return 4;
  }
}

Fvi read_fvi(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("\nError reading file: " + path);
  unsigned char h[48];
  if (std::fread(h, 1, 48, f) != 48) {
    std::fclose(f);
    throw std::runtime_error("\nError reading file: " + path + " (short header)");
  }
  Fvi v;
  auto u32 = [&](int off) {
    uint32_t x;
    std::memcpy(&x, h + off, 4);
    return x;
  };
  std::memcpy(&v.type, h, 2);
  v.nelements = u32(4);
  v.nobs = u32(8);
  v.nvars = u32(12);
  v.bytes_per_record = u32(16);
  v.bits_per_record = u32(20);
  v.namelength = u32(24);
  const size_t nn = ((size_t)v.nobs + v.nvars) * v.namelength;
  v.names.assign(nn + 1, 0);
  const size_t got = std::fread(v.names.data(), 1, nn, f);
  std::fclose(f);
  if (got != nn) throw std::runtime_error("\nError reading file: " + path + " (short name table)");
  return v;
}

FvdReader::FvdReader(const std::string& path, const Fvi& fvi) : fvi_(fvi), path_(path) {
  f_ = std::fopen(path.c_str(), "rb");
  if (!f_) throw std::runtime_error("Error opening File " + path);
}

FvdReader::~FvdReader() {
  if (f_) std::fclose(f_);
}

static inline double widen(const unsigned char* p, uint16_t type) {
  switch (type) {
    case 1: { uint16_t x; std::memcpy(&x, p, 2); return x; }
    case 2: { int16_t x; std::memcpy(&x, p, 2); return x; }
    case 3: { uint32_t x; std::memcpy(&x, p, 4); return x; }
    case 4: { int32_t x; std::memcpy(&x, p, 4); return x; }
    case 6: { double x; std::memcpy(&x, p, 8); return x; }
    case 7: return (double)(int8_t)p[0];
    case 8: return (double)p[0];
    default: { float x; std::memcpy(&x, p, 4); return x; }
  }
}

void FvdReader::read_columns(size_t col0, size_t ncols, std::vector<double>& out) {
  const size_t es = fvi_.elem_size(), nobs = fvi_.nobs;
  out.resize(ncols * nobs);
  buf_.resize(ncols * nobs * es);
  if (std::fseek(f_, (long)(col0 * nobs * es), SEEK_SET) != 0 ||
      std::fread(buf_.data(), es, ncols * nobs, f_) != ncols * nobs)
    throw std::runtime_error("Error reading File " + path_);
  for (size_t i = 0; i < ncols * nobs; ++i) out[i] = widen(buf_.data() + i * es, fvi_.type);
}

template <typename T>
void FvdReader::read_kept(size_t col0, size_t ncols, size_t nrows, const std::vector<uint8_t>& keep,
                          size_t nkeep, T* dst) {
  const size_t es = fvi_.elem_size(), nobs = fvi_.nobs;
  buf_.resize(ncols * nobs * es);
  if (std::fseek(f_, (long)(col0 * nobs * es), SEEK_SET) != 0 ||
      std::fread(buf_.data(), es, ncols * nobs, f_) != ncols * nobs)
    throw std::runtime_error("Error reading File " + path_);
  for (size_t c = 0; c < ncols; ++c) {
    const unsigned char* src = buf_.data() + c * nobs * es;
    T* d = dst + c * nkeep;
    size_t o = 0;
    for (size_t k = 0; k < nrows; ++k)
      if (keep[k]) d[o++] = (T)widen(src + k * es, fvi_.type);
  }
}

template void FvdReader::read_kept<float>(size_t, size_t, size_t, const std::vector<uint8_t>&, size_t, float*);
template void FvdReader::read_kept<double>(size_t, size_t, size_t, const std::vector<uint8_t>&, size_t, double*);

void write_fvi(const std::string& path, uint16_t type, uint32_t nobs, uint32_t nvars,
               const std::vector<std::string>& obs_names, const std::vector<std::string>& var_names,
               uint32_t namelength) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Error Creating File: " + path);
  unsigned char h[48] = {0};
  std::memcpy(h, &type, 2);
  const uint32_t es = (type == 6) ? 8 : 4;
  const uint32_t vals[6] = {nobs * nvars, nobs, nvars, es, es * 8, namelength};
  std::memcpy(h + 4, vals, sizeof(vals));
  std::fwrite(h, 1, 48, f);
  std::vector<char> nm(namelength);
  for (const auto* list : {&obs_names, &var_names})
    for (const auto& s : *list) {
      std::fill(nm.begin(), nm.end(), 0);
      std::memcpy(nm.data(), s.data(), std::min<size_t>(s.size(), namelength - 1));
      std::fwrite(nm.data(), 1, namelength, f);
    }
  std::fclose(f);
}

}  // namespace oahost
