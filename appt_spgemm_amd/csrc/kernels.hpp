// kernels.hpp -- host-side launchers of the gfx950 SpGEMM kernels.
//
// Row bins (one kernel configuration each):
//   0                 empty output row
//   1 .. 24           LDS group bins: 1 + (cls*2 + mode)*4 + gidx
//                       cls  : products (symbolic) / row nnz (numeric) class
//                       mode : 0 dense bitmap over the row's output span,
//                              1 word hash (32-column words as keys)
//                       gidx : lanes per row G = 8, 16, 32, 64
//   25 / 26           one workgroup per row, dense / word hash in LDS
//   27                heavy rows: one workgroup per row, the output span
//                     walked in LDS column windows
//   28 .. 39          numeric only: wave-level groups for rows whose symbolic
//                     pass left a sorted word list, 28 + cls*4 + gidx; the
//                     LDS holds a byte index per window word and the listed
//                     words instead of a bitmap and prefix counts
//   40 .. 47          register sort: rows of at most N = 16 << (bin - 40)
//                     products (and A entries) whose output is too scattered
//                     for a window; the products are sorted by column in
//                     registers (bitonic network over G lanes x E slots)
//   48                row templates: rows whose structure is a shifted copy of
//                     a shared pattern (templates.hip); no symbolic pass,
//                     products go straight to their precomputed slots
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "heavy_shape.hpp"

namespace appt {

constexpr int kNumCls = 3;
constexpr int kNumBins = 49;
constexpr int kBinWgBitmap = 25, kBinWgHash = 26, kBinHeavy = 27, kBinList0 = 28, kBinSort0 = 40, kBinTemplate = 48;
constexpr int kSortMax = 2048;  // products (and A entries) of the largest sort bin
constexpr int kHeavyT = 1024;   // threads of a heavy-row workgroup
constexpr int kStatFields = 3;  // count, max metric, max words
constexpr int kStatsMaxBlocks = 4096;
// int32 elements of the per-block statistics scratch (plus u64 product counts)
int64_t stats_part_elems();

struct BinParams {
  int32_t ncls;
  int32_t cls_max[kNumCls];       // metric bound of class c (products, or row nnz)
  int32_t cls_bytes[kNumCls];     // LDS bytes one row of class c may use
  int32_t wg_bytes;               // LDS bytes of a workgroup-per-row bin
  int32_t glen_max[3];            // mean B-row length bounds for G = 8, 16, 32
  int32_t dense_ratio;            // numeric: dense if words(span) <= max(ratio * distinct words, 64)
  int32_t lds_budget;             // LDS bytes a block of the wave-level bins may take
  int32_t sort_ratio;             // register sort first when span words > sort_ratio * products (0: never);
                                  // numeric pass: scattered rows only
  int32_t sort_max;               // slots of the largest register-sort bin (16 .. kSortMax; 0: no sort bins)
};

inline bool bin_is_sort(int bin) { return bin >= kBinSort0 && bin < kBinSort0 + 8; }
inline int sort_bin_slots(int bin) { return 16 << (bin - kBinSort0); }
inline bool bin_is_list(int bin) { return bin >= kBinList0 && bin < kBinSort0; }
inline int bin_group(int bin) { return bin_is_list(bin) ? 8 << ((bin - kBinList0) % 4) : 8 << ((bin - 1) % 4); }
inline int bin_cls(int bin) { return bin_is_list(bin) ? (bin - kBinList0) / 4 : (bin - 1) / 8; }
inline bool bin_is_bitmap(int bin) { return bin < kBinWgBitmap && ((bin - 1) / 4) % 2 == 0; }

// Symbolic hash table slots for a row with f products (np2l((4f)/3+1) of the
// reference sizing rule); numeric hash slots for a row with c outputs.
inline int64_t np2_host(int64_t x) { if (x < 2) return 2; int64_t p = 2; while (p < x) p <<= 1; return p; }
inline int64_t sym_slots(int64_t f) { return np2_host((f * 4) / 3 + 1); }
// symbolic bin "words" statistic: span words (dense) or key bound (word hash)
inline int64_t num_slots(int64_t c) { return np2_host((c * 3) / 2 + 1); }

// ---- launchers (all asynchronous on `s`) ----
// gate (optional) on a launcher: its kernels run only when *gate != 0
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s, const int32_t* gate = nullptr);
// Raise a kernel's dynamic-LDS cap to at least `bytes` on the current device.
// The cap only has to cover each launch, so it is set once per thread,
// device and kernel for the largest size asked so far: between short
// kernels, the host call otherwise costs more than the kernels do.
void set_dyn_lds(const void* fn, size_t bytes);
// B-row metadata {start, end, min column, max column} (min = INT32_MAX,
// max = -1 for an empty row).
// Row k is [ptr[k], ptr[k+1]) of col, or [off2 + ptr2[k], off2 + ptr2[k+1])
// where sel (optional) is non-zero.  flag[0] is set to 1 when some row's
// columns are not in nondecreasing order (flag must be zero beforehand).
// (rows of at most min_len entries are skipped: another kernel describes them;
// gate (optional): the kernel does nothing while *gate is zero)
void launch_rowstats(int32_t n, const int32_t* ptr, const int32_t* ptr2, const uint8_t* sel, int64_t off2,
                     const int32_t* col, int4* meta, int32_t* flag, hipStream_t s, int32_t min_len = -1,
                     const int32_t* gate = nullptr);
void launch_colcount(int64_t nnz, const int32_t* col, int32_t* cnt, hipStream_t s, const int32_t* gate = nullptr);
// Structure check of a square CSR (flags zero beforehand): flag[0] = 1 when
// ptr[0] != 0 or ptr[n] != nnz, flag[1] = 1 when row pointers decrease,
// flag[2] = 1 when a column lies outside [0, n).  Columns are read only
// within [0, nnz).
void launch_csr_check(int32_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, int32_t* flag, hipStream_t s);
// flag[0] = 1 (zero beforehand) when some row of (ptr, col) is not strictly
// increasing
// gate (optional): nothing runs while *gate is zero
void launch_rows_strict(int32_t n, const int32_t* ptr, const int32_t* col, int32_t* flag, hipStream_t s,
                        const int32_t* gate = nullptr);
// Radix transpose: digits of at most kRadixMaxBits bits (the fewest passes
// that cover the column range, bits split evenly), one 256-thread block per
// tile of kRadixTile entries (kRadixChunks chunks of 64 per wave).
constexpr int kRadixMaxBits = 8, kRadixChunks = 16;
constexpr int64_t kRadixTile = 4 * 64 * kRadixChunks;
int64_t radix_tiles(int64_t nnz);
int radix_passes(int32_t n);
int radix_bits(int32_t n);  // digit bits per pass
// Scratch of the radix transpose: keys / rows / vals (nnz each) for the
// ping-pong passes, keys2 (nnz) for the final keys, hist (2^bits x tiles
// int32), hoff (2^bits x tiles + 1 int64) and htmp (scan_tmp_elems of that).
struct RadixScratch {
  int32_t* keys;
  int32_t* rows;
  double* vals;
  int32_t* keys2;
  int32_t* hist;
  int64_t* hoff;
  int64_t* htmp;
};
// B = A^T with sorted rows into (tp, bc, bv), decided on the device: a
// symmetric-pattern gather in A's layout (tp = ap); when A is not strictly
// sorted with a symmetric pattern it sets *asym (zero beforehand) and B is
// built by a stable radix sort of A's entries by column (rx given), or by the
// column histogram (cnt, n + 1 int32), scan (off64, tmp), atomic-cursor
// scatter and row sort (slist) otherwise.
void launch_transpose(int32_t n, int64_t nnz, const int32_t* ap, const int32_t* ac, const double* av,
                      int32_t* tp, int32_t* bc, double* bv, int32_t* cnt, int64_t* off64, int64_t* tmp,
                      int32_t* slist, const RadixScratch* rx, int32_t* asym, hipStream_t s);
// The same in two steps: the symmetric-pattern gather (sets *asym when it
// fails), then the general transpose, which runs on the device only when
// *asym is set -- a caller that has read *asym back skips it when zero.
void launch_transpose_gather(int32_t n, const int32_t* ap, const int32_t* ac, const double* av, int32_t* tp,
                             int32_t* bc, double* bv, int32_t* asym, hipStream_t s);
void launch_transpose_general(int32_t n, int64_t nnz, const int32_t* ap, const int32_t* ac, const double* av,
                              int32_t* tp, int32_t* bc, double* bv, int32_t* cnt, int64_t* off64, int64_t* tmp,
                              int32_t* slist, const RadixScratch* rx, int32_t* asym, hipStream_t s);
// AAtS mirror by radix sort: per-row counts of entries left / right of the
// diagonal, then (with their exclusive scans lo_off / up_off and the upper
// total `count`) the strictly-upper entries sorted by column into the
// lower halves.  rx sized for `count` entries (rows unused); vals2: count f64.
void launch_mirror_counts(int32_t n, const int64_t* off, const int32_t* ccol, int32_t* lower, int32_t* upper,
                          hipStream_t s);
void launch_mirror_radix(int32_t n, int64_t count, const int64_t* off, const int32_t* ccol, double* cval,
                         const int32_t* upper, const int64_t* up_off, const int64_t* lo_off, const RadixScratch& rx,
                         double* vals2, hipStream_t s);
// Atomic-cursor scatter; gate (optional): runs only when *gate != 0.
void launch_transpose_scatter(int32_t n, const int32_t* ap, const int32_t* ac, const double* av,
                              int32_t* cursor, int32_t* bc, double* bv, const int32_t* gate, hipStream_t s);
// Sort every row of (bp, bc, bv) by column, values moving with their keys;
// scratch: sort_rows_scratch_elems(n) int32; gate as above.
int64_t sort_rows_scratch_elems(int32_t n);
void launch_sort_rows(int32_t n, const int32_t* bp, int32_t* bc, double* bv, int32_t* scratch, const int32_t* gate,
                      hipStream_t s);
// Exclusive scan of n int32 counts into out64[0..n] (out64[n] = total); out32
// (optional) receives the same values truncated to int32.  `tmp` needs
// scan_tmp_elems(n) int64.
int64_t scan_tmp_elems(int64_t n);
void launch_scan(const int32_t* in, int64_t n, int64_t* out64, int32_t* out32, int64_t* tmp,
                 hipStream_t s, const int32_t* gate = nullptr);
struct TplArgs;
// T (optional): template candidates are checked against their template's
// representative and take the template bin, with rowcnt / rowwords written.
void launch_prepass(int32_t n, const int32_t* ap, const int32_t* ac, const int4* bmeta, int upper,
                    const BinParams& P,
                    int64_t* rowf, int32_t* rowlo, int32_t* rowhi, uint8_t* bin, int32_t* stats,
                    unsigned long long* total_products, int32_t* part, const TplArgs* T, int32_t* rowcnt,
                    int32_t* rowwords, hipStream_t s);
void launch_numbin(int32_t n, const int32_t* ap, const int64_t* rowf, const int32_t* rowcnt,
                   const int32_t* rowwords, const int32_t* rowlo, const int32_t* rowhi, const BinParams& P,
                   uint8_t* bin, int32_t* stats, int32_t* part, const TplArgs* T, hipStream_t s);

// Zero three int32 arrays in one launch.
void launch_fill3_i32(int32_t* a, int64_t na, int32_t* b, int64_t nb, int32_t* c, int64_t nc, hipStream_t s);
// Row list by bin: bin b's rows start at the exclusive scan of the counts
// stats[b * kStatFields] (bins 1..; bin 0 is not listed), in no particular
// order within a bin; taken: kNumBins int32, zero beforehand (blocks claim
// their slots there).  total_src (optional): *total_dst = *total_src on the
// way.
void launch_bin_scatter(int32_t n, const uint8_t* bin, const int32_t* stats, int32_t* taken, int32_t* rows,
                        const int64_t* total_src, int64_t* total_dst, hipStream_t s);

struct RowArgs {
  const int32_t* rows;   // row ids of this bin
  int32_t nrows;
  const int32_t* ap; const int32_t* ac; const double* av;
  const int32_t* bc; const double* bv;   // B rows are addressed through bmeta
  const int64_t* rowf; const int32_t* rowlo; const int32_t* rowhi;
  int32_t* rowwords;     // distinct 32-column words per output row (symbolic writes, numeric reads)
  int upper;
  int64_t bnnz;          // nnz of B (selects 32- or 64-bit gather offsets)
  const int4* bmeta;     // per B row: {start, end, min column, max column}
  uint2* wlist;          // per output row: up to kWordListCap sorted (word, mask) pairs from the symbolic pass
  int32_t lds_budget;    // LDS bytes a block of the wave-level bins may take
  int32_t ncols;         // bound on every output column index
  // bitmaps of the symbolic workgroup-bitmap rows, kept for the numeric
  // workgroup pass: row i's span words at wgbits + wgslot[i] * wgstride
  // (wgslot[i] < 0: none); null when not kept
  uint32_t* wgbits;
  int32_t* wgslot;
  int32_t wgstride;
};
constexpr int kWordListCap = 64;  // word-list slots per output row (stride of RowArgs::wlist)

// Heavy-row bin: column windows of wc_sym / wc_num columns, two row-claim
// counters (work[0] symbolic, work[1] numeric), the B-sortedness flag of
// launch_rowstats, a cursor per A entry (nnz(A) int32) and a fault flag the
// numeric kernel sets (to 1) if a sparse window cannot advance.
struct HeavyArgs {
  int32_t wc_sym, wc_num, vcap, dcols;  // window columns (symbolic / numeric sparse), slots, dense width
  bool adapt;                           // numeric sparse windows shaped per row (no tuning cap)
  int32_t* work;
  const int32_t* bflag;
  int32_t* cur;
  int32_t* fault;
};

// ---- triangle-only Gram, op AAtS (gram.hip) ----
// Per A entry p = (i, k): vmeta[p] = the suffix of B row k (rows ascending)
// from its first row >= i (lower: the prefix up to its last row <= i), as a
// B-row descriptor; vac[p] = p.  Adds the full A*A^T product count (sum over
// B rows of nnz^2) to *full_products.
void launch_gram_operand(int32_t n, const int32_t* ap, const int32_t* ac, int32_t nb, const int4* bmeta,
                         const int32_t* bc, int lower, int4* vmeta, int32_t* vac,
                         unsigned long long* full_products, hipStream_t s);
// Expand T (one triangle of C's rows, diagonal included; offsets uoff, nu
// entries) into C: mir[j] += off-diagonal entries of column j (zeroed first) ...
void launch_gram_mirrored(int32_t n, int64_t nu, const int64_t* uoff, const int32_t* ucol, int32_t* mir,
                          hipStream_t s);
// ... strict[i] = T row i less its diagonal, total[i] = mir[i] + T row i ...
void launch_gram_counts(int32_t n, const int32_t* ucnt, const int32_t* mir, int32_t* strict, int32_t* total,
                        hipStream_t s);
// ... the off-diagonal entries in row-major order at sup_off (scan of strict) ...
void launch_gram_extract(int32_t n, int64_t nu, const int64_t* uoff, const int32_t* ucol, const double* uval,
                         const int64_t* sup_off, int lower, int32_t* keys, int32_t* rows, double* vals,
                         hipStream_t s);
// ... (stably sorted by column: launch_radix_pairs) placed with T's rows into
// C (mir_off: scan of mir).
void launch_gram_assemble(int32_t n, int64_t nc, const int64_t* coff, int lower_t, const int32_t* mir,
                          const int64_t* mir_off, const int32_t* srow, const double* sval, const int64_t* uoff,
                          const int32_t* ucol, const double* uval, int32_t* ccol, double* cval, hipStream_t s);
// Stable LSD radix sort of `count` (key < n, row, value) triples.  The input
// must sit in the buffers the first pass does not write: (rx.keys, rx.rows,
// rx.vals) when radix_passes(n) is odd, else (rx.keys2, rows2, vals2); the
// result is in (rx.keys2, rows2, vals2).
void launch_radix_pairs(int32_t n, int64_t count, const RadixScratch& rx, int32_t* rows2, double* vals2,
                        hipStream_t s);

// ---- row templates (templates.hip) ----
constexpr int kClassSlots = 4096;      // relative-pattern classes per call
constexpr int kTplSlots = 256;         // row templates per call
constexpr int kTplMaxLen = 64;         // entries of a template row's A row and of each of its B rows
constexpr int kTplMaxProducts = kTplMaxLen * kTplMaxLen;
struct TplArgs {
  unsigned long long* ckey;  // B-row class table: kClassSlots hash keys (0 = empty)
  int32_t* crep;             // class representative (a B row)
  int32_t* cpat;             // per B-row class: its length, then its representative's kTplMaxLen relative offsets
  unsigned long long* akey;  // A-side class table (unused when A's rows are B's rows)
  int32_t* arep;             // class representative (an A-side row)
  int32_t* bclass;           // per B row: class slot, or -1
  int32_t* aclass;           // per A-side row: class slot, or -1 (== bclass when A's rows are B's rows)
  unsigned long long* tkey;  // template table: kTplSlots keys
  int32_t* trep;             // template representative (A-side row)
  int32_t* tplid;            // per A-side row: template slot, or -1
  int4* info;                // per template: {valid, A entries m, slots nS, widest of the A row and its B rows}
  int32_t* posoff;           // per template: kTplMaxLen + 1 product offsets of the A entries
  uint16_t* pos;             // per template: kTplMaxProducts output slots of the products
  int32_t* soff;             // per template: kTplMaxProducts output offsets (column - row), ascending
  int32_t* tuple;            // per template: kTplMaxLen B classes of its representative's entries
  uint32_t* pos2;            // per template: [parity][entry u][pair l] slots of B entries 2l - par, 2l - par + 1
  int2* grid;                // per template: {line, plane} strides read off its A offsets (0: none)
  int32_t* matched;          // [0]: B rows whose pattern repeats another's; [1]: the same for A-side rows;
                             // [2], [3]: distinct patterns among sampled B rows, and the samples
  int32_t row0;              // global row index of A-side row 0 (row panels)
};
constexpr int kTplPairs = kTplMaxLen / 2 + 1;  // B-value pairs of one B row (either parity)
// LDS of one template row: its A entries' values and {B-row start, first
// product, products} (kTplMaxLen each), then nS f64 slots.
__host__ __device__ inline int64_t tpl_group_bytes(int64_t nsmax) {
  return (int64_t)kTplMaxLen * 24 + ((nsmax * 8 + 15) & ~(int64_t)15);
}
// The same for the pair-load kernel (rows of at most 31 entries): values and
// {B-row start, length} of 32 entries, then nS f64 slots.
__host__ __device__ inline int64_t tpl_pair_group_bytes(int64_t nsmax) {
  return (int64_t)32 * 16 + ((nsmax * 8 + 15) & ~(int64_t)15);
}
// Classes of B's rows (and of A's rows unless a_is_b), template candidates
// of A's rows (tplid, pending the pre-pass's check against the template's
// representative) and the template tables.
// bptr (optional): B's row pointers, when bmeta is not built yet -- the
// class pass then writes B's row metadata too (rows over kTplMaxLen entries
// by a pass of their own, launched only when *longrows was raised: longrows
// must be zero beforehand) and raises *bflag on a descending B row.
// apat: zero on the device when A's rows have B's patterns (A*A; A*A^T when
// the symmetric-pattern transpose gather succeeded): the A-side rows then
// take B's classes.
void launch_tpl_detect(const TplArgs& T, int32_t n, const int32_t* ap, const int32_t* ac, int32_t nb,
                       int4* bmeta, const int32_t* bc, bool a_is_b, const int32_t* bptr, int32_t* bflag,
                       int32_t* longrows, const int32_t* apat, hipStream_t s);
// Numeric pass of the template bin (maxb: longest B row, nsmax: most slots
// of a row of the bin).
// tile_n > 0: the bin holds every template row of an n = tile_n row matrix
// (no panel), so the pair kernel may walk the rows in grid tiles.
void launch_tpl_num(const RowArgs& a, const TplArgs& T, int32_t maxb, int32_t nsmax, int32_t tile_n,
                    const int64_t* rowoff,
                    int32_t* ccol, double* cval, int num_cus, int64_t lds_per_cu, hipStream_t s);

// Symbolic: rowcnt[i] = distinct output columns of row i.
// words_max: the bin's largest LDS statistic (span words or hash keys).
// Symbolic results of a bin's rows (panel-local ids, a.rows) copied from a
// row-pointer plan over all rows: rowcnt[i] = pcnt[r0 + i], rowwords[i] =
// pwords[r0 + i] without a word list.
void launch_take_plan(const RowArgs& a, int32_t r0, const int32_t* pcnt, const int32_t* pwords, int32_t* rowcnt,
                      hipStream_t s);
void launch_symbolic_bin(int bin, const RowArgs& a, int32_t words_max, int32_t* rowcnt, const HeavyArgs& h,
                         hipStream_t s);
// Numeric: writes sorted col/val of each row at rowoff[i].
// words_max / c_max: bin maxima.
void launch_numeric_bin(int bin, const RowArgs& a, int32_t words_max, int32_t c_max,
                        const int64_t* rowoff, int32_t* ccol, double* cval, const HeavyArgs& h, hipStream_t s);
// Bytes of LDS one row of a bin needs (0 for the heavy bin).
int64_t symbolic_row_lds_bytes(int bin, int32_t metric_max, int32_t words_max);
int64_t numeric_row_lds_bytes(int bin, int32_t c_max, int32_t words_max);

// AAtS: fill the lower entries of the complete, sorted C from row r's upper
// part (C's structure is symmetric; values were formed for j >= i only).
void launch_mirror_pull(int32_t n, int64_t nnz, const int64_t* off, const int32_t* ccol, double* cval,
                        hipStream_t s);

// ---- batched path (block-diagonal packing of many matrices) ----
struct BatchSrc {  // one separate device matrix to pack
  const int32_t* rowptr;
  const int32_t* col;
  const double* val;
  int32_t n;
  int64_t nnz;
};
void launch_batch_rows(int32_t N, int32_t K, const int32_t* blk_row, const uint8_t* blk_op, int32_t* rowbase,
                       int32_t* rowblk, uint8_t* rowop, hipStream_t s);
void launch_batch_acg(int32_t N, const int32_t* rp, const int32_t* col, const int32_t* rowbase,
                      const uint8_t* rowop, int32_t* acg, int32_t* tcnt, int32_t* rid, hipStream_t s);
void launch_batch_tscatter(int32_t N, const int32_t* rp, const int32_t* acg, const int32_t* rowbase,
                           const uint8_t* rowop, int32_t* cursor, int64_t nnz_a, int32_t* bcol, double* bval,
                           hipStream_t s);
// Largest A*A^T matrix (rows) the per-matrix workgroup transpose takes (its
// column counts live in LDS).
constexpr int32_t kBatchTBlockN = 16384;
void launch_batch_tblock(int32_t K, int32_t N, int32_t max_nt, int64_t nnz_t, const int32_t* blk_row,
                         const int64_t* blk_nnz, const uint8_t* blk_op, const int32_t* toff, const int32_t* rid,
                         const int32_t* col, const double* val, int32_t* tp, int32_t* bcol, double* bval,
                         hipStream_t s);
void launch_batch_local_rowptr(int32_t N, int32_t K, const int64_t* rowoff, const int32_t* rowbase,
                               const int32_t* rowblk, const int32_t* blk_row, int32_t* crp, int64_t* blkoff,
                               hipStream_t s);
void launch_batch_block_sum(int32_t K, const int32_t* blk_row, const int64_t* rowf, int64_t* out, hipStream_t s);
void launch_batch_gather(int32_t K, const BatchSrc* src, const int32_t* blk_row, const int64_t* blk_nnz, int32_t* rp,
                         int32_t* col, double* val, hipStream_t s);

// Numeric semantics: per-row counts of non-zero values, then their
// order-preserving compaction to offsets `off` (int64 exclusive scan).
void launch_count_nonzero(int32_t n, const void* rp, bool rp64, const double* val, int32_t* cnt, hipStream_t s);
void launch_compact_nonzero(int32_t n, const void* rp, bool rp64, const int32_t* col, const double* val,
                            const int64_t* off, int32_t* ocol, double* oval, hipStream_t s);

// Order-independent digest of a CSR result into out[0] (structure) and
// out[1] (values): sums of per-entry 64-bit mixes of (position, content).
void launch_digest(int32_t n, int64_t nnz, const void* rp, bool rp64, const int32_t* col, const double* val,
                   unsigned long long* out, hipStream_t s);

void set_num_cus(int n);
void set_lds_per_cu(int64_t b);
int num_cus();
int64_t lds_per_cu();
void launch_i64_to_i32(const int64_t* in, int64_t n, int32_t* out, hipStream_t s, const int32_t* gate = nullptr);
void launch_copy_i64(const int64_t* in, int64_t n, int64_t* out, hipStream_t s);

}  // namespace appt
