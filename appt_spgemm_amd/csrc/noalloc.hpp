// noalloc.hpp -- static device scratch of appt_spgemm_compute_noalloc.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace appt {

// Bytes of the per-device static scratch array.
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;

// Takes the device address of the scratch array on `device` (which must be
// the current device) once; later calls return at once.
hipError_t static_scratch_init(int device);
// Device address of the scratch array on `device`, or null before its init.
void* static_scratch(int device);

// Upper bound of the scratch bytes the pipeline takes for an n-row operand
// with nnz entries (AA, AAt or AAtS; no stored workgroup bitmaps, the
// binary-search mirror): per row, the pre-pass arrays, bins, offsets and the
// 64-entry word list; per entry, the transpose and its radix scratch and the
// heavy-row cursors; a fixed part for the bin statistics; 256-byte alignment
// of every buffer.
inline int64_t static_scratch_need(int64_t n, int64_t nnz) {
  const int64_t per_row = 4 + 4 + 8 + 8 + 4 + 16 + 8 + 4 + 4 + 1 + 4 + 4 + 4 + 512 + 4 + 1 + 4 + 8;
  const int64_t per_entry = 4 + 8 + 4 + 4 + 8 + 4 + 4 + 4 + 4;
  const int64_t fixed = (int64_t)3 << 20;
  return per_row * (n + 2) + per_entry * nnz + fixed;
}

}  // namespace appt
