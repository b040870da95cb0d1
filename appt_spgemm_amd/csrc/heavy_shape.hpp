// heavy_shape.hpp -- window shapes of the heavy-row bin (host code, no HIP
// dependency: the kernels' launcher and a CPU model of the window sequence
// share these rules).
//
// Numeric heavy rows walk their output span [lo, hi] in windows.  A dense
// window covers dcols columns (one f64 slot per column).  A sparse window
// covers at most wcols columns (1 bit and a 32-bit prefix count per 32-column
// word) and holds at most vcap outputs: it is cut before the first word whose
// outputs would not all fit, and the next window starts at the cut.  A word
// holds at most 32 outputs, so vcap >= 32 guarantees the first word always
// fits and every window advances by at least 32 columns.
#pragma once
#include <stdint.h>

namespace appt {

constexpr int32_t kHeavyMinSlots = 32;   // vcap and dcols floor: one word's outputs
constexpr int32_t kHeavyMinCols = 64;    // wcols floor (two words)

// Symbolic heavy window for a dynamic LDS allowance of `bytes`: 1 bit per
// column, a multiple of 32 (capped by `cap` when non-zero).
inline int32_t heavy_window_cols(int64_t bytes, int32_t cap) {
  int64_t wc = bytes * 8;
  if (cap > 0 && wc > cap) wc = cap;
  wc &= ~(int64_t)31;
  if (wc < 32) wc = 32;
  if (wc > ((int64_t)1 << 30)) wc = (int64_t)1 << 30;
  return (int32_t)wc;
}

// Numeric heavy windows in `bytes` of LDS: sparse windows of `wcols` columns
// with `vcap` f64 slots, dense windows of `dcols` columns.  A cap (testing)
// sets wcols = cap, vcap = max(32, cap / 3) and dcols = max(32, cap / 2).
inline void heavy_num_shape(int64_t bytes, int32_t cap, int32_t* wcols, int32_t* vcap, int32_t* dcols) {
  int64_t wc = cap > 0 ? cap : (bytes < 131072 ? bytes : 131072);
  wc &= ~(int64_t)31;
  if (wc < kHeavyMinCols) wc = kHeavyMinCols;
  int64_t vc = (bytes - wc / 4) / 8;
  int64_t dc = ((bytes - wc / 8) / 8) & ~(int64_t)31;
  if (cap > 0) {
    vc = (cap / 3) < vc ? (cap / 3) : vc;
    dc = ((cap / 2) < dc ? (cap / 2) : dc) & ~(int64_t)31;
  }
  if (vc < kHeavyMinSlots) vc = kHeavyMinSlots;
  if (dc < kHeavyMinSlots) dc = kHeavyMinSlots;
  if (dc > wc) dc = wc;
  *wcols = (int32_t)wc;
  *vcap = (int32_t)vc;
  *dcols = (int32_t)dc;
}

inline int64_t heavy_lds_bytes(bool vals, int32_t wc, int32_t vcap, int32_t dcols) {
  if (!vals) return (int64_t)(wc / 32) * 4;
  const int64_t sparse = (int64_t)(wc / 32) * 8 + (int64_t)vcap * 8;
  const int64_t dense = (int64_t)(wc / 32) * 4 + (int64_t)dcols * 8;
  return sparse > dense ? sparse : dense;
}

// Rows spanning more than this many dense windows take sparse windows.
constexpr int64_t kHeavySparseSpan = 8;

}  // namespace appt
