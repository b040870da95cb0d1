// gen.cpp -- host CSR: CSR1 files and the seeded synthetic generators.
//
// CSR1 layout (src/export_bin.py:6-8, writer :20-29; reader
// src/spgemm_engine.cuh:22-32): "CSR1", int32 n, int32 pad, int64 nnz,
// int32 indptr[n+1], int32 indices[nnz], f64 data[nnz], little-endian.
//
// Generators are deterministic functions of their arguments (xoshiro256**
// streams seeded through splitmix64), so the same matrix is produced on any
// host.  Rows are emitted with strictly ascending columns.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/appt_spgemm.h"

namespace appt {
void set_last_error(const std::string& m);
}

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double sym1() { return 2.0 * u01() - 1.0; }            // U(-1, 1)
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

appt_status alloc_csr(int32_t n, int64_t nnz, appt_host_csr* out) {
  out->n = n;
  out->nnz = nnz;
  out->rowptr = (int32_t*)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  out->val = (double*)std::malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  if (!out->rowptr || !out->col || !out->val) {
    std::free(out->rowptr); std::free(out->col); std::free(out->val);
    std::memset(out, 0, sizeof(*out));
    appt::set_last_error("host allocation failed");
    return APPT_ERR_OOM;
  }
  return APPT_OK;
}

// Build CSR from (row, col) pairs: group by row, sort, merge duplicates.
// Values are drawn afterwards in row-major order from `vals`.
template <typename ValFn>
appt_status from_pairs(int32_t n, std::vector<uint64_t>& keys, ValFn&& val, appt_host_csr* out) {
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  appt_status st = alloc_csr(n, (int64_t)keys.size(), out);
  if (st != APPT_OK) return st;
  int64_t p = 0;
  for (int32_t i = 0; i <= n; i++) {
    while (p < (int64_t)keys.size() && (int64_t)(keys[p] >> 32) < i) p++;
    out->rowptr[i] = (int32_t)p;
  }
  for (size_t t = 0; t < keys.size(); t++) {
    out->col[t] = (int32_t)(keys[t] & 0xffffffffu);
    out->val[t] = val();
  }
  return APPT_OK;
}

}  // namespace

namespace appt {
// Square CSR structure: rowptr[0] = 0, rowptr[n] = nnz, non-decreasing row
// pointers, columns in [0, n).
bool host_csr_valid(const appt_host_csr& m) {
  if (m.n < 0 || m.nnz < 0 || m.nnz > INT32_MAX || !m.rowptr) return false;
  if (m.nnz > 0 && !m.col) return false;
  if (m.rowptr[0] != 0 || m.rowptr[m.n] != m.nnz) return false;
  for (int32_t i = 0; i < m.n; i++)
    if (m.rowptr[i + 1] < m.rowptr[i]) return false;
  for (int64_t p = 0; p < m.nnz; p++)
    if (m.col[p] < 0 || m.col[p] >= m.n) return false;
  return true;
}

}  // namespace appt

extern "C" {

void appt_host_csr_free(appt_host_csr* m) {
  if (!m) return;
  std::free(m->rowptr); std::free(m->col); std::free(m->val);
  std::memset(m, 0, sizeof(*m));
}

appt_status appt_csr1_read(const char* path, appt_host_csr* out) {
  if (!path || !out) return APPT_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "rb");
  if (!f) { appt::set_last_error(std::string("cannot open ") + path); return APPT_ERR_IO; }
  char magic[4];
  int32_t n = 0, pad = 0;
  int64_t nnz = 0;
  if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "CSR1", 4) != 0 || std::fread(&n, 4, 1, f) != 1 ||
      std::fread(&pad, 4, 1, f) != 1 || std::fread(&nnz, 8, 1, f) != 1 || n < 0 || nnz < 0 || nnz > INT32_MAX) {
    std::fclose(f);
    appt::set_last_error("bad CSR1 header");
    return APPT_ERR_IO;
  }
  std::fseek(f, 0, SEEK_END);
  const long long size = std::ftell(f);
  const long long want = 20LL + 4LL * ((long long)n + 1) + 12LL * nnz;
  if (size != want) {
    std::fclose(f);
    appt::set_last_error("CSR1 size mismatch: " + std::to_string(size) + " bytes, expected " + std::to_string(want));
    return APPT_ERR_IO;
  }
  std::fseek(f, 20, SEEK_SET);
  appt_status st = alloc_csr(n, nnz, out);
  if (st != APPT_OK) { std::fclose(f); return st; }
  bool ok = std::fread(out->rowptr, 4, (size_t)n + 1, f) == (size_t)n + 1 &&
            (int64_t)std::fread(out->col, 4, (size_t)nnz, f) == nnz &&
            (int64_t)std::fread(out->val, 8, (size_t)nnz, f) == nnz;
  std::fclose(f);
  if (!ok) { appt_host_csr_free(out); appt::set_last_error("CSR1 truncated"); return APPT_ERR_IO; }
  if (!appt::host_csr_valid(*out)) {
    appt_host_csr_free(out);
    appt::set_last_error("CSR1 structure invalid");
    return APPT_ERR_INVALID;
  }
  return APPT_OK;
}

appt_status appt_csr1_write(const char* path, const appt_host_csr* m) {
  if (!path || !m || m->n < 0 || m->nnz < 0) return APPT_ERR_INVALID;
  FILE* f = std::fopen(path, "wb");
  if (!f) { appt::set_last_error(std::string("cannot create ") + path); return APPT_ERR_IO; }
  const int32_t pad = 0;
  bool ok = std::fwrite("CSR1", 1, 4, f) == 4 && std::fwrite(&m->n, 4, 1, f) == 1 && std::fwrite(&pad, 4, 1, f) == 1 &&
            std::fwrite(&m->nnz, 8, 1, f) == 1 &&
            std::fwrite(m->rowptr, 4, (size_t)m->n + 1, f) == (size_t)m->n + 1 &&
            (int64_t)std::fwrite(m->col, 4, (size_t)m->nnz, f) == m->nnz &&
            (int64_t)std::fwrite(m->val, 8, (size_t)m->nnz, f) == m->nnz;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) { appt::set_last_error("CSR1 write failed"); return APPT_ERR_IO; }
  return APPT_OK;
}

appt_status appt_gen_laplace2d(int32_t g, appt_host_csr* out) {
  if (!out || g < 0 || (int64_t)g * g > INT32_MAX / 5) return APPT_ERR_INVALID;
  const int32_t n = g * g;
  int64_t nnz = (int64_t)n * 5 - (g > 0 ? 4LL * g : 0);
  appt_status st = alloc_csr(n, nnz, out);
  if (st != APPT_OK) return st;
  int64_t p = 0;
  for (int32_t y = 0; y < g; y++)
    for (int32_t x = 0; x < g; x++) {
      const int32_t i = x + g * y;
      out->rowptr[i] = (int32_t)p;
      if (y > 0) { out->col[p] = i - g; out->val[p++] = -1.0; }
      if (x > 0) { out->col[p] = i - 1; out->val[p++] = -1.0; }
      out->col[p] = i; out->val[p++] = 4.0;
      if (x < g - 1) { out->col[p] = i + 1; out->val[p++] = -1.0; }
      if (y < g - 1) { out->col[p] = i + g; out->val[p++] = -1.0; }
    }
  out->rowptr[n] = (int32_t)p;
  out->nnz = p;
  return APPT_OK;
}

appt_status appt_gen_stencil27(int32_t g, double drop_p, uint64_t seed, appt_host_csr* out) {
  if (!out || g < 0 || (int64_t)g * g * g * 27 > INT32_MAX || drop_p < 0 || drop_p >= 1) return APPT_ERR_INVALID;
  const int32_t n = g * g * g;
  const int64_t cap = (int64_t)n * 27;
  appt_status st = alloc_csr(n, cap, out);
  if (st != APPT_OK) return st;
  Rng vr(seed), dr(seed ^ 0x5DEECE66Dull);
  int64_t p = 0;
  for (int32_t z = 0; z < g; z++)
    for (int32_t y = 0; y < g; y++)
      for (int32_t x = 0; x < g; x++) {
        const int32_t i = x + g * y + g * g * z;
        out->rowptr[i] = (int32_t)p;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              const int32_t xx = x + dx, yy = y + dy, zz = z + dz;
              if (xx < 0 || yy < 0 || zz < 0 || xx >= g || yy >= g || zz >= g) continue;
              const double v = vr.sym1();
              const bool diag = dx == 0 && dy == 0 && dz == 0;
              if (!diag && drop_p > 0 && dr.u01() < drop_p) continue;
              out->col[p] = xx + g * yy + g * g * zz;
              out->val[p++] = v;
            }
      }
  out->rowptr[n] = (int32_t)p;
  out->nnz = p;
  return APPT_OK;
}

appt_status appt_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                          appt_host_csr* out) {
  if (!out || scale < 0 || scale > 30 || edge_factor < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return APPT_ERR_INVALID;
  const int32_t n = (int32_t)1 << scale;
  const int64_t m = (int64_t)edge_factor << scale;
  if (m > INT32_MAX) return APPT_ERR_INVALID;
  std::vector<uint64_t> keys((size_t)m);
  Rng er(seed);
  const double ab = a + b, abc = a + b + c;
  for (int64_t e = 0; e < m; e++) {
    uint32_t r = 0, col = 0;
    for (int lvl = 0; lvl < scale; lvl++) {
      const double u = er.u01();
      uint32_t rb = 0, cb = 0;
      if (u < a) {} else if (u < ab) { cb = 1; } else if (u < abc) { rb = 1; } else { rb = 1; cb = 1; }
      r = (r << 1) | rb;
      col = (col << 1) | cb;
    }
    keys[(size_t)e] = ((uint64_t)r << 32) | col;
  }
  Rng vr(seed ^ 0xA5A5A5A5DEADBEEFull);
  return from_pairs(n, keys, [&]() { return vr.u01(); }, out);
}

appt_status appt_gen_banded(int32_t n, int32_t bw, double fill, uint64_t seed, appt_host_csr* out) {
  if (!out || n < 0 || bw < 0 || fill < 0 || fill > 1 || (int64_t)n * (2LL * bw + 1) > INT32_MAX) return APPT_ERR_INVALID;
  const int64_t cap = (int64_t)n * (2LL * bw + 1);
  appt_status st = alloc_csr(n, cap, out);
  if (st != APPT_OK) return st;
  Rng kr(seed), vr(seed ^ 0x1234567887654321ull);
  int64_t p = 0;
  for (int32_t i = 0; i < n; i++) {
    out->rowptr[i] = (int32_t)p;
    const int32_t j0 = i - bw < 0 ? 0 : i - bw;
    const int32_t j1 = i + bw > n - 1 ? n - 1 : i + bw;
    for (int32_t j = j0; j <= j1; j++) {
      if (j != i && kr.u01() >= fill) continue;
      out->col[p] = j;
      out->val[p++] = vr.sym1();
    }
  }
  out->rowptr[n] = (int32_t)p;
  out->nnz = p;
  return APPT_OK;
}

appt_status appt_gen_erdos_renyi(int32_t n, double mean_deg, uint64_t seed, appt_host_csr* out) {
  if (!out || n < 0 || mean_deg < 0 || mean_deg * (double)n > (double)INT32_MAX) return APPT_ERR_INVALID;
  const int64_t m = n > 0 ? (int64_t)std::llround(mean_deg * (double)n) : 0;
  std::vector<uint64_t> keys((size_t)m);
  Rng er(seed);
  for (int64_t e = 0; e < m; e++) {
    const uint32_t i = er.below((uint32_t)n), j = er.below((uint32_t)n);
    keys[(size_t)e] = ((uint64_t)i << 32) | j;
  }
  Rng vr(seed ^ 0x0F0F0F0FF0F0F0F0ull);
  return from_pairs(n, keys, [&]() { return vr.sym1(); }, out);
}

}  // extern "C"
