// Minimal .npy (version 1.0, little-endian f8, C order) writer: the format
// xt::dump_npy produces for the reference's monitor outputs
// (src/monitor/monitor.cpp:42-54).
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace npy {

inline void write(const std::string& path, const std::vector<std::size_t>& shape, const double* data,
                  const char* descr = "<f8", std::size_t item_doubles = 1) {
  // the shape as a Python tuple: "(n,)" for one dimension, "(a, b, c)" else
  std::string sh = "(";
  for (std::size_t i = 0; i < shape.size(); ++i) sh += (i ? ", " : "") + std::to_string(shape[i]);
  sh += shape.size() == 1 ? ",)" : ")";
  std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
  const std::size_t unpadded = 10 + hdr.size() + 1;
  hdr += std::string((64 - unpadded % 64) % 64, ' ');
  hdr += '\n';
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + path);
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  std::fwrite(magic, 1, 8, f);
  const uint16_t len = static_cast<uint16_t>(hdr.size());
  std::fwrite(&len, sizeof(len), 1, f);
  std::fwrite(hdr.data(), 1, hdr.size(), f);
  std::size_t n = 1;
  for (auto s : shape) n *= s;
  if (n) std::fwrite(data, sizeof(double), n * item_doubles, f);
  std::fclose(f);
}

// Complex128 ('<c16') array: interleaved (re, im) doubles, as xt::dump_npy
// writes std::complex<double>.
inline void writeComplex(const std::string& path, const std::vector<std::size_t>& shape, const double* data) {
  write(path, shape, data, "<c16", 2);
}

}  // namespace npy
