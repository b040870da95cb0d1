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

// the row-shape rule below runs in the kernel and in the CPU model
#if defined(__HIPCC__)
#define APPT_HD __host__ __device__
#else
#define APPT_HD
#endif

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

// Sparse window of a row that still has c outputs over `span` columns, in
// `bytes` of LDS (no tuning cap).  Bits and prefix counts cost 2 bits per
// column and a value slot 8 bytes, so a window of W columns at the row's mean
// density c / span fills 60 % of its slots when W = 4 bytes span / (32 c 5/3
// + span): scattered rows (a few outputs per thousand columns) get windows of
// up to a few 10^5 columns instead of a fixed width, dense rows narrow ones
// whose structure pass reaches little past the cut.  `wlim` (0: none) caps W
// after a cut showed the local density above the mean.  (A CPU model over
// R-MAT 2^21 rows: 60 % fill gives fewer windows and less structure work
// than 75 % or 90 %, where local density above the mean cuts more windows.)
// W is a multiple of 32 within [kHeavyMinCols, (bytes - 8 kHeavyMinSlots) *
// 4] and no wider than the span needs, and vcap takes the rest of the LDS, so
// vcap >= 32 and (W / 32) * 8 + vcap * 8 <= bytes.
APPT_HD inline void heavy_row_shape(int64_t bytes, int64_t c, int64_t span, int64_t wlim, int32_t* wcols,
                                    int32_t* vcap) {
  if (c < 1) c = 1;
  if (span < 1) span = 1;
  int64_t w = (12 * bytes * span) / (160 * c + 3 * span);
  const int64_t span32 = (span + 31) & ~(int64_t)31;
  int64_t wmax = (bytes - 8 * kHeavyMinSlots) * 4;
  if (wmax > ((int64_t)1 << 30)) wmax = (int64_t)1 << 30;
  if (w > span32) w = span32;
  if (wlim > 0 && w > wlim) w = wlim;
  if (w > wmax) w = wmax;
  w &= ~(int64_t)31;
  if (w < kHeavyMinCols) w = kHeavyMinCols;
  *wcols = (int32_t)w;
  *vcap = (int32_t)((bytes - (w / 32) * 8) / 8);
}

// Dense window width in `bytes` of LDS (no tuning cap): a bitmap of the
// window's own words and one f64 slot per column, 65 / 8 bytes per column,
// a multiple of 64 columns (so the slots after the bitmap stay 8-byte
// aligned), at least 64.
APPT_HD inline int32_t heavy_dense_cols(int64_t bytes) {
  const int64_t d = (bytes * 8 / 65) & ~(int64_t)63;
  return (int32_t)(d < 64 ? 64 : d);
}

// Cap on the next sparse window after one that took `taken` columns (cut:
// the window held more outputs than its slots), from the cap `wlim` the
// window had: a cut window caps the next at 3/2 of the columns it took, an
// uncut one doubles the cap.
APPT_HD inline int64_t heavy_next_wlim(bool cut, int64_t taken, int64_t wlim) {
  if (cut) {
    const int64_t w = ((taken * 3) / 2 + 31) & ~(int64_t)31;
    return w < kHeavyMinCols ? kHeavyMinCols : w;
  }
  return wlim > 0 && wlim < ((int64_t)1 << 30) ? 2 * wlim : 0;
}

}  // namespace appt
