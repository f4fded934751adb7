// DatABEL filevector inputs (.fvi header + names, .fvd column-major data).
// Header layout as read by the reference (src/AIOwrapper.h:145-175,
// src/AIOwrapper.cpp:2294-2317): u16 type, 2 pad bytes, u32 nelements,
// numObservations, numVariables, bytesPerRecord, bitsPerRecord,
// namelength, reserved[5]; then (numObservations + numVariables) names of
// namelength bytes each, observations first.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace oahost {

struct Fvi {
  uint16_t type = 5;
  uint32_t nelements = 0, nobs = 0, nvars = 0, bytes_per_record = 0, bits_per_record = 0;
  uint32_t namelength = 32;
  std::vector<char> names;   // (nobs + nvars) * namelength

  std::string obs_name(size_t i) const;
  std::string var_name(size_t j) const;
  const char* var_name_raw(size_t j) const { return names.data() + (nobs + j) * (size_t)namelength; }
  size_t elem_size() const;
};

// Throws std::runtime_error with the reference's message on failure.
Fvi read_fvi(const std::string& path);

// Column reader of a .fvd file.  Type codes 1..8 = u16, i16, u32, i32,
// float, double, i8, u8 (src/AIOwrapper.h:145-152); values are widened to
// double.  The reference always reads float; this reader honours the type.
class FvdReader {
 public:
  FvdReader(const std::string& path, const Fvi& fvi);
  ~FvdReader();
  FvdReader(const FvdReader&) = delete;
  FvdReader& operator=(const FvdReader&) = delete;
  // Reads `ncols` full columns starting at `col0` (each nobs values).
  void read_columns(size_t col0, size_t ncols, std::vector<double>& out);
  // Same, keeping only rows with keep[row] (keep.size() >= rows used) and
  // the first `nrows` rows of each column, into dst (stride nkeep).
  template <typename T>
  void read_kept(size_t col0, size_t ncols, size_t nrows, const std::vector<uint8_t>& keep,
                 size_t nkeep, T* dst);

 private:
  FILE* f_ = nullptr;
  const Fvi& fvi_;
  std::string path_;
  std::vector<unsigned char> buf_;
};

void write_fvi(const std::string& path, uint16_t type, uint32_t nobs, uint32_t nvars,
               const std::vector<std::string>& obs_names, const std::vector<std::string>& var_names,
               uint32_t namelength = 32);

}  // namespace oahost
