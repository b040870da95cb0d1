// Shared helpers of the command-line tools: load a matrix from a CSR1 file
// or from a generator spec "gen:<name>:<args...>".
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/appt_spgemm.h"

inline std::vector<std::string> split(const std::string& s, char d) {
  std::vector<std::string> out;
  size_t p = 0, q;
  while ((q = s.find(d, p)) != std::string::npos) { out.push_back(s.substr(p, q - p)); p = q + 1; }
  out.push_back(s.substr(p));
  return out;
}

// gen:laplace2d:G | gen:stencil27:G:DROP:SEED | gen:rmat:SCALE:EF:SEED |
// gen:banded:N:BW:FILL:SEED | gen:er:N:MEAN:SEED
inline appt_status load_matrix(const char* spec, appt_host_csr* h) {
  if (std::strncmp(spec, "gen:", 4) != 0) return appt_csr1_read(spec, h);
  auto f = split(spec + 4, ':');
  auto I = [&](size_t k, long d) { return k < f.size() ? std::atol(f[k].c_str()) : d; };
  auto D = [&](size_t k, double d) { return k < f.size() ? std::atof(f[k].c_str()) : d; };
  if (f[0] == "laplace2d") return appt_gen_laplace2d((int32_t)I(1, 64), h);
  if (f[0] == "stencil27") return appt_gen_stencil27((int32_t)I(1, 96), D(2, 0.0), (uint64_t)I(3, 1), h);
  if (f[0] == "rmat") return appt_gen_rmat((int32_t)I(1, 16), (int32_t)I(2, 16), 0.57, 0.19, 0.19, (uint64_t)I(3, 1), h);
  if (f[0] == "banded") return appt_gen_banded((int32_t)I(1, 4096), (int32_t)I(2, 16), D(3, 0.7), (uint64_t)I(4, 0), h);
  if (f[0] == "er") return appt_gen_erdos_renyi((int32_t)I(1, 4096), D(2, 8.0), (uint64_t)I(3, 0), h);
  return APPT_ERR_INVALID;
}

#define CLI_CHECK(x)                                                                  \
  do {                                                                                \
    appt_status st_ = (x);                                                            \
    if (st_ != APPT_OK) {                                                             \
      std::fprintf(stderr, "%s failed: status %d: %s\n", #x, (int)st_, appt_last_error()); \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)
