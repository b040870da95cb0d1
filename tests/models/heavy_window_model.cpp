// CPU model of the numeric heavy-row window sequence (test infrastructure).
//
// Restates the window loop of k_heavy_num (appt_spgemm_amd/csrc/kernels.hip)
// over a row's output columns, using the product's own shape rules
// (heavy_shape.hpp: the fixed shape under a tuning cap, else each sparse
// window sized from the row's remaining outputs and columns), and checks that
// every window advances, that a sparse window never holds more than its value
// slots, that a window's bits, prefix counts and slots fit the launch's LDS,
// and that the windows together emit every output exactly once.  Random rows: span up to 2^22 columns,
// densities from one column in 10^4 to fully dense, clustered or uniform.
// Prints one line per configuration and "OK <rows>" at the end; exits 1 on
// the first violated invariant.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "heavy_shape.hpp"

using namespace appt;

static int fail(const char* what, int64_t a, int64_t b) {
  std::printf("FAIL %s (%lld, %lld)\n", what, (long long)a, (long long)b);
  return 1;
}

struct Row {
  int64_t lo, hi;
  std::vector<int32_t> pre;  // pre[c]: outputs in [lo, lo + c)
};

int main() {
  // dynamic LDS of a heavy block: the device's LDS per block (64 KB .. 160 KB)
  // less the kernel's static staging (24 KB)
  const int64_t lds_blocks[] = {65536, 163840};
  const int32_t caps[] = {0, 1, 31, 32, 33, 63, 64, 65, 96, 100, 128, 1000, 4096, 65536, 131072, 1 << 20};
  std::mt19937_64 rng(12345);
  // rows: spans up to 5000 columns (dense windows) and up to 2^21 (sparse),
  // densities from one column in 10^4 to fully dense, clustered or uniform
  std::vector<Row> rows;
  const double dens[] = {1e-4, 1e-2, 0.3, 1.0};
  for (int t = 0; t < 32; t++) {
    Row r;
    const int64_t span = 1 + (int64_t)(rng() % (t < 20 ? 5000 : (1u << 21)));
    r.lo = (int64_t)(rng() % 1000000) & ~(int64_t)31;  // the pre-pass aligns lo to 32 columns
    r.hi = r.lo + span - 1;
    const double d = dens[t % 4];
    const bool clustered = (t / 4) % 2;
    std::uniform_real_distribution<double> u(0, 1);
    r.pre.assign((size_t)span + 1, 0);
    for (int64_t c = 0; c < span; c++) {
      const double p = clustered ? (((c / 4096) % 3 == 0) ? 1.0 : d * 0.01) : d;
      const bool on = c == 0 || c == span - 1 || u(rng) < p;  // lo and hi are output columns
      r.pre[(size_t)c + 1] = r.pre[(size_t)c] + (on ? 1 : 0);
    }
    rows.push_back(std::move(r));
  }
  int64_t checked = 0;
  for (int64_t lds : lds_blocks) {
    const int64_t dyn = lds - 24 * 1024;
    for (int32_t cap : caps) {
      int32_t wcols, vcap, dcols;
      heavy_num_shape(dyn, cap, &wcols, &vcap, &dcols);
      if (vcap < kHeavyMinSlots || dcols < kHeavyMinSlots || wcols < kHeavyMinCols || (wcols & 31) || (dcols & 31))
        return fail("shape floor", vcap, dcols);
      if (cap == 0 && heavy_lds_bytes(true, wcols, vcap, dcols) > dyn) return fail("shape exceeds LDS", wcols, dyn);
      // the launch's dynamic LDS, which the per-row shapes divide
      const int64_t lbytes = heavy_lds_bytes(true, wcols, vcap, dcols);
      int64_t nwin = 0;
      for (const Row& r : rows) {
        const int64_t lo = r.lo, hi = r.hi, span = hi - lo + 1;
        auto outputs = [&](int64_t a, int64_t b) { return (int64_t)r.pre[(size_t)(b - lo)] - r.pre[(size_t)(a - lo)]; };
        const int64_t total = r.pre[(size_t)span];
        const bool sparse = hi - lo + 1 > kHeavySparseSpan * dcols;
        int64_t w0 = lo, emitted = 0, windows = 0, wlim = 0;
        while (w0 <= hi) {
          int32_t wcr = wcols, vcr = vcap;
          if (sparse && cap == 0) {
            heavy_row_shape(lbytes, total - emitted, hi + 1 - w0, wlim, &wcr, &vcr);
            if (wcr < kHeavyMinCols || (wcr & 31) || vcr < kHeavyMinSlots) return fail("row shape floor", wcr, vcr);
            if ((int64_t)(wcr / 32) * 8 + (int64_t)vcr * 8 > lbytes) return fail("row shape exceeds LDS", wcr, vcr);
          }
          // dense windows without a cap: the width the LDS left by a row of
          // 1024 resident entries (12 KB) gives, as the kernel computes it
          const int32_t dcr = cap == 0 ? heavy_dense_cols(lbytes - 12 * 1024) : dcols;
          if (!sparse && cap == 0 && ((int64_t)(dcr / 32) * 4 + (int64_t)dcr * 8 > lbytes - 12 * 1024 || dcr < 64))
            return fail("dense width exceeds LDS", dcr, lbytes);
          const int64_t wspan = sparse ? wcr : dcr;
          const int64_t w1 = (w0 + wspan) < (hi + 1) ? (w0 + wspan) : (hi + 1);
          const int64_t words = (w1 - w0 + 31) >> 5;
          int64_t wcut = w1;
          if (sparse) {
            // the cut: the first word whose outputs would not all fit in vcap
            int64_t cut = words, run = 0;
            for (int64_t w = 0; w < words; w++) {
              const int64_t a = w0 + w * 32, b = (a + 32) < w1 ? (a + 32) : w1;
              const int64_t pc = outputs(a, b);
              if (run <= vcr && run + pc > vcr) { cut = w; break; }
              run += pc;
            }
            if (cut == 0) return fail("sparse window cannot advance", w0, vcap);
            if (cut < words) wcut = w0 + cut * 32;
          }
          if (wcut <= w0) return fail("window does not advance", w0, wcut);
          const int64_t in = outputs(w0, wcut);
          if (sparse && in > vcr) return fail("sparse window over capacity", in, vcr);
          if (!sparse && wcut - w0 > dcr) return fail("dense window over width", wcut - w0, dcr);
          emitted += in;
          if (sparse) wlim = heavy_next_wlim(wcut < w1, wcut - w0, wlim);
          w0 = wcut;
          if (++windows > span / 32 + 2) return fail("too many windows", windows, span);
        }
        if (emitted != total) return fail("outputs lost or repeated", emitted, total);
        nwin += windows;
        checked++;
      }
      std::printf("lds %lld cap %d: wcols %d vcap %d dcols %d, %lld windows over %zu rows\n", (long long)lds, cap,
                  wcols, vcap, dcols, (long long)nwin, rows.size());
    }
  }
  std::printf("OK %lld\n", (long long)checked);
  return 0;
}
