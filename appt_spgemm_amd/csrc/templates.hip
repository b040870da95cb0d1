// templates.hip -- row templates: output rows whose structure is a shifted
// copy of a shared relative pattern (stencils, banded operators, structured
// grids), computed without a symbolic pass and without per-product index
// lookups.
//
// Relative pattern of a row r with entries at columns c_0 .. c_{m-1} (storage
// order): the sequence (c_0 - r, ..., c_{m-1} - r).  Output row i of
// C = A * B depends only on
//   * the relative pattern of A's row i,
//   * the relative patterns of the B rows k_u = c_u it references, in order.
// Rows that agree in all of these have the same relative output structure S
// (sorted, distinct offsets) and the same map from each product (u, v) (A
// entry u, B entry v) to its output slot.  Such rows share one template:
// nS = |S|, S, and pos[u][v].  Their counts need no symbolic pass (nS), and
// their values are accumulated straight into slot pos[u][v] in A-row storage
// order -- the order of the row-wise Gustavson loop -- so they are the same
// bit for bit as in the other wave-level bins.
//
// Detection is exact.  Relative patterns are grouped into classes through a
// table keyed by a 64-bit hash, and every row is compared entry by entry with
// its class representative: a row that differs (a hash collision) gets no
// class.  Output rows are grouped into templates by a hash of (A class,
// B classes in order), and every row is compared class by class with its
// template's representative row before it takes the template bin.  Rows
// longer than kTplMaxLen entries, with B rows longer than that, or whose
// template has more than kTplMaxProducts products, stay in the other bins;
// so do all rows of a matrix whose patterns do not repeat (the tables fill
// and stop admitting).
#include "kernels.hpp"
#include "device_common.hpp"

namespace appt {

namespace {

__device__ __forceinline__ unsigned long long tmix(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Open addressing with at most kProbes probes; returns the slot holding key h
// (inserted here if absent) or -1 when the neighbourhood is full, and sets
// *mine when this call inserted the key.  Key 0 marks an empty slot (hashes
// are made odd).  Probes are plain (cached) loads: most rows of a structured
// matrix find their key already there, and a stale empty slot only costs a
// compare-and-swap that returns the key.  Nothing here is a same-address
// atomic per row (those serialise in one L2 channel).
constexpr int kProbes = 8;
__device__ __forceinline__ int table_insert(unsigned long long* keys, int nslots, unsigned long long h, bool* mine) {
  int slot = (int)((h >> 17) & (unsigned long long)(nslots - 1));
  *mine = false;
  for (int p = 0; p < kProbes; p++) {
    // device-coherent read (the L2): a key inserted by another CU is seen
    const unsigned long long k = __hip_atomic_load(&keys[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == h) return slot;
    if (k == 0ull) {
      const unsigned long long old = atomicCAS(&keys[slot], 0ull, h);
      if (old == 0ull) {
        *mine = true;
        return slot;
      }
      if (old == h) return slot;
    }
    slot = (slot + 1) & (nslots - 1);
  }
  return -1;
}

// Block-local cache in front of table_insert: rows of one block (neighbouring
// rows of a structured matrix) mostly share their key, so only the first row
// of each key in a block probes the global table.  An entry packs the key's
// high 32 bits with slot + 1 in one 64-bit LDS word (written and read whole);
// a false hit between two keys that share their high bits only costs the row
// its class or template, which the exact checks that follow would drop.
constexpr int kCacheSlots = 64;
struct KeyCache {
  unsigned long long e[kCacheSlots];
};
__device__ __forceinline__ void cache_init(KeyCache& c) {
  for (int t = threadIdx.x; t < kCacheSlots; t += blockDim.x) c.e[t] = 0ull;
  __syncthreads();
}
__device__ __forceinline__ int cached_insert(KeyCache& c, unsigned long long* keys, int nslots, unsigned long long h,
                                             bool* mine) {
  const int idx = (int)((h >> 7) & (kCacheSlots - 1));
  const unsigned long long e = c.e[idx];
  *mine = false;
  if ((e >> 32) == (h >> 32) && (e & 0xffffffffull) != 0ull) return (int)(e & 0xffffffffull) - 1;
  const int slot = table_insert(keys, nslots, h, mine);
  if (slot >= 0) c.e[idx] = (h & 0xffffffff00000000ull) | (unsigned long long)(slot + 1);
  return slot;
}

// Entries of row r: [start, end) and its relative base.  B rows are read
// through bmeta (B row k is global row k); A-side rows through ap (panel row
// i is global row row0 + i).
struct Rows {
  const int32_t* ap;   // null: B rows through bmeta
  const int4* bmeta;
  int32_t row0;
  __device__ __forceinline__ int2 range(int r) const {
    if (ap) return make_int2(ap[r], ap[r + 1]);
    const int4 b = bmeta[r];
    return make_int2(b.x, b.y);
  }
  __device__ __forceinline__ int64_t base(int r) const { return ap ? (int64_t)row0 + r : (int64_t)r; }
};

// Insertion for the group leaders of a wavefront (lanes with want): leaders
// holding the same key are served by one probe (the first of them), so a
// wavefront of rows that share a pattern costs one cache lookup and at most
// one probe of the global table; the table's device-wide reads and
// compare-and-swaps go past the XCD's L2 and are kept rare.  *mine is set
// for the lane whose probe inserted the key.
__device__ __forceinline__ int wave_insert(KeyCache& kc, unsigned long long* keys, int nslots, unsigned long long h,
                                           bool want, bool* mine) {
  const int lane = threadIdx.x % kWave;
  unsigned long long pending = __ballot(want);
  int result = -1;
  *mine = false;
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    const unsigned long long hl = __shfl(h, leader, kWave);
    const unsigned long long same = __ballot(want && h == hl) & pending;
    int slot = -1;
    bool ins = false;
    if (lane == leader) slot = cached_insert(kc, keys, nslots, h, &ins);
    slot = __shfl(slot, leader, kWave);
    if ((same >> lane) & 1ull) {
      result = slot;
      *mine = lane == leader && ins;
    }
    pending &= ~same;
  }
  return result;
}

// The class and template kernels take kClsLanes lanes per row, each lane
// loading kClsUnroll entries per step (their loads in flight together), so a
// wavefront keeps many rows' loads outstanding.
constexpr int kClsLanes = 4;
constexpr int kClsUnroll = 4;
constexpr int64_t kClsMaxBlocks = 4096;

__device__ __forceinline__ unsigned long long entry_hash(int u, int64_t rel) {
  return tmix(((unsigned long long)(u + 1) << 32) ^ (unsigned long long)(uint32_t)(int32_t)rel);
}

// Whether B's row patterns repeat at all, from a sample: kSampleBlocks blocks
// hash the relative patterns of kSampleRows rows each (spread over B; one
// row per four lanes, so a block takes its rows in one step) into an LDS set and count the distinct ones into matched[2] (samples into
// matched[3]).  When more than three quarters of the sampled rows are
// distinct, the class and template passes skip their hashing and table
// probes: no template can pay (a quarter of B's rows must repeat another's
// pattern), and a table that fills with distinct keys costs more than
// anything else in the call.
constexpr int kSampleBlocks = 64, kSampleRows = 64, kSampleSlots = 256;
__global__ void __launch_bounds__(256) k_class_sample(int32_t nrows, Rows R, const int32_t* __restrict__ col,
                                                      int32_t* __restrict__ matched) {
  constexpr int G = kClsLanes;
  __shared__ unsigned long long set[kSampleSlots];
  __shared__ int32_t distinct, taken;
  for (int t = threadIdx.x; t < kSampleSlots; t += blockDim.x) set[t] = 0ull;
  if (threadIdx.x == 0) distinct = taken = 0;
  __syncthreads();
  const int lane = threadIdx.x % G;
  const int64_t total = (int64_t)kSampleBlocks * kSampleRows;
  for (int j0 = 0; j0 < kSampleRows; j0 += (int)blockDim.x / G) {
    const int64_t j = (int64_t)blockIdx.x * kSampleRows + j0 + threadIdx.x / G;   // sample index
    const int64_t r = nrows < total ? (j < nrows ? j : -1) : j * nrows / total;
    int2 se = make_int2(0, 0);
    if (r >= 0) se = R.range((int)r);
    const int m = se.y - se.x;
    const bool ok = r >= 0 && m > 0 && m <= kTplMaxLen;
    const int64_t b = ok ? R.base((int)r) : 0;
    unsigned long long h = 0;
    if (ok) {
      // a row has at most kTplMaxLen entries: all of a lane's loads at once
      constexpr int U = kTplMaxLen / G;
      int c[U];
#pragma unroll
      for (int x = 0; x < U; x++) c[x] = lane + x * G < m ? col[se.x + lane + x * G] : 0;
#pragma unroll
      for (int x = 0; x < U; x++)
        if (lane + x * G < m) h += entry_hash(lane + x * G, (int64_t)c[x] - b);
    }
    h = gsum<G, unsigned long long>(h, nullptr);
    h = (h + tmix((unsigned long long)m)) | 1ull;
    if (lane == 0 && r >= 0) {
      atomicAdd(&taken, 1);
      bool fresh = !ok;   // rows without a class count as distinct
      if (ok) {
        int slot = (int)(h & (kSampleSlots - 1));
        for (int p = 0; p < kSampleSlots; p++) {
          const unsigned long long old = atomicCAS(&set[slot], 0ull, h);
          if (old == 0ull) { fresh = true; break; }
          if (old == h) break;
          slot = (slot + 1) & (kSampleSlots - 1);
        }
      }
      if (fresh) atomicAdd(&distinct, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&matched[2], distinct);
    atomicAdd(&matched[3], taken);
  }
}
__device__ __forceinline__ bool patterns_distinct(const int32_t* matched) {
  return (int64_t)matched[2] * 4 > (int64_t)matched[3] * 3;
}

// Pattern class of each row: hash of its relative pattern, inserted into the
// class table; the row that inserts a key is the class representative (any
// row of the class would do: every other row is compared with it exactly).
// META (B rows through their row pointers, R.ap): also each row's metadata
// {start, end, min column, max column} for rows of at most kTplMaxLen
// entries, and *flag = 1 on a descending row, from the same loads.
template <bool META>
__global__ void __launch_bounds__(256) k_class_insert(int32_t nrows, Rows R, const int32_t* __restrict__ col,
                                                      unsigned long long* __restrict__ keys,
                                                      int32_t* __restrict__ rep, int32_t* __restrict__ cls,
                                                      int4* __restrict__ meta, int32_t* __restrict__ flag,
                                                      const int32_t* __restrict__ matched,
                                                      int32_t* __restrict__ longrows,
                                                      const int32_t* __restrict__ gate, int32_t* __restrict__ pat) {
  constexpr int G = kClsLanes;
  if (gate && *gate == 0) return;   // A-side rows whose classes are copied from B's
  __shared__ KeyCache kc;
  cache_init(kc);
  // sampled patterns mostly distinct: no classes (the metadata still)
  const bool off = patterns_distinct(matched);
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  const int64_t stride = (int64_t)gridDim.x * gpb;
  // every wave runs the same number of iterations (group shuffles stay converged)
  const int64_t wave_first = (int64_t)blockIdx.x * gpb + (threadIdx.x / kWave) * (kWave / G);
  for (int64_t w0 = wave_first; w0 < nrows; w0 += stride) {
    const int64_t r = w0 + (threadIdx.x % kWave) / G;
    int2 se = make_int2(0, 0);
    if (r < nrows) se = R.range((int)r);
    const int m = se.y - se.x;
    const bool ok = r < nrows && m > 0 && m <= kTplMaxLen && !off;
    const bool scan = r < nrows && m > 0 && m <= kTplMaxLen && (META || !off);
    const int64_t b = ok ? R.base((int)r) : 0;
    unsigned long long h = 0;
    int mn = INT32_MAX, mx = -1;
    bool desc = false;
    if (scan)
      for (int u0 = lane; u0 < m; u0 += G * kClsUnroll) {
        int c[kClsUnroll], nx[kClsUnroll];
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++) c[x] = u0 + x * G < m ? col[se.x + u0 + x * G] : 0;
        if (META) {
          // each entry's successor: the next lane's entry of the same step
          // (a quad rotation: G = 4 lanes per row); lane G - 1 takes lane 0's
          // entry of the next step, and the last step's from memory
          static_assert(G == 4, "successors come from a quad rotation");
          const int last = u0 + (kClsUnroll - 1) * G + 1;   // lane G - 1's last successor, own index
          const int tail = col[se.x + (last < m ? last : m - 1)];
#pragma unroll
          for (int x = 0; x < kClsUnroll; x++) {
            const int r0 = __builtin_amdgcn_update_dpp(0, c[x], 0x39, 0xf, 0xf, false);
            const int r1 = x + 1 < kClsUnroll ? __builtin_amdgcn_update_dpp(0, c[x + 1], 0x39, 0xf, 0xf, false) : tail;
            nx[x] = u0 + x * G + 1 < m ? (lane < G - 1 ? r0 : r1) : INT32_MAX;
          }
        }
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++)
          if (u0 + x * G < m) {
            if (!off) h += entry_hash(u0 + x * G, (int64_t)c[x] - b);
            if (META) {
              mn = c[x] < mn ? c[x] : mn;
              mx = c[x] > mx ? c[x] : mx;
              desc |= nx[x] < c[x];
            }
          }
      }
    if (META) {
      mn = gmin<G>(mn);
      mx = gmax<G>(mx);
      if (lane == 0 && r < nrows && m <= kTplMaxLen) meta[r] = make_int4(se.x, se.y, mn, mx);
      if (__ballot(desc) != 0ull && threadIdx.x % kWave == 0) flag[0] = 1;
      // rows past kTplMaxLen are left to the metadata pass, run only then
      if (__ballot(r < nrows && m > kTplMaxLen) != 0ull && threadIdx.x % kWave == 0) longrows[0] = 1;
    }
    if (off) {
      if (lane == 0 && r < nrows) cls[r] = -1;
      continue;
    }
    h = gsum<G, unsigned long long>(h, nullptr);
    h = (h + tmix((unsigned long long)m)) | 1ull;
    bool mine;
    const int slot = wave_insert(kc, keys, kClassSlots, h, lane == 0 && ok, &mine);
    if (lane == 0 && r < nrows) {
      if (mine) rep[slot] = (int32_t)r;
      cls[r] = ok ? slot : -1;
    }
    if (pat) {
      // the inserting row records the class's pattern (length, then its
      // relative offsets) for the checks that follow
      const int leader = (threadIdx.x % kWave) & ~(G - 1);
      const int gslot = __shfl(slot, leader, kWave);
      const bool gmine = __shfl(mine ? 1 : 0, leader, kWave) != 0;
      if (gmine) {
        int32_t* pp = pat + (size_t)gslot * (kTplMaxLen + 1);
        if (lane == 0) pp[0] = m;
        for (int u = lane; u < m; u += G) pp[1 + u] = (int32_t)((int64_t)col[se.x + u] - b);
      }
    }
  }
}

// Entry-by-entry comparison of each classed row with its class
// representative; a row that differs loses its class.
// Rows that matched a class other than their own (repeated patterns) are
// counted into *matched (one atomic per block).
__global__ void __launch_bounds__(256) k_class_verify(int32_t nrows, Rows R, const int32_t* __restrict__ col,
                                                      const int32_t* __restrict__ rep, int32_t* __restrict__ cls,
                                                      int32_t* __restrict__ matched, const int32_t* __restrict__ gate,
                                                      const int32_t* __restrict__ apat) {
  constexpr int G = kClsLanes;
  if (patterns_distinct(gate)) return;   // every row is classless already
  if (apat && *apat == 0) return;        // A-side rows whose classes are copied from B's
  __shared__ int32_t block_matched;
  if (threadIdx.x == 0) block_matched = 0;
  __syncthreads();
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  const int64_t stride = (int64_t)gridDim.x * gpb;
  const int64_t wave_first = (int64_t)blockIdx.x * gpb + (threadIdx.x / kWave) * (kWave / G);
  for (int64_t w0 = wave_first; w0 < nrows; w0 += stride) {
    const int64_t r = w0 + (threadIdx.x % kWave) / G;
    const int c = r < nrows ? cls[r] : -1;
    const int q = c >= 0 ? rep[c] : -1;
    int same = 1;
    if (c >= 0 && q != (int)r) {
      const int2 se = R.range((int)r), sq = R.range(q);
      const int64_t b = R.base((int)r), bq = R.base(q);
      const int m = se.y - se.x;
      if (m != sq.y - sq.x) {
        same = 0;
      } else {
        for (int u0 = lane; u0 < m; u0 += G * kClsUnroll) {
          int x0[kClsUnroll], x1[kClsUnroll];
#pragma unroll
          for (int x = 0; x < kClsUnroll; x++) {
            const bool in = u0 + x * G < m;
            x0[x] = in ? col[se.x + u0 + x * G] : 0;
            x1[x] = in ? col[sq.x + u0 + x * G] : 0;
          }
#pragma unroll
          for (int x = 0; x < kClsUnroll; x++)
            if (u0 + x * G < m && (int64_t)x0[x] - b != (int64_t)x1[x] - bq) same = 0;
        }
      }
    }
    same = gmin<G>(same);
    if (lane == 0 && c >= 0 && !same) cls[r] = -1;
    const int hits = __popcll(__ballot(lane == 0 && c >= 0 && same && q != (int)r));
    if (threadIdx.x % kWave == 0 && hits) atomicAdd(&block_matched, hits);
  }
  __syncthreads();
  if (threadIdx.x == 0 && block_matched) atomicAdd(matched, block_matched);
}

// Template candidate of each A-side row: A class and the classes of its B
// rows, in order.  Writes the template slot (or -1) to tplid; the row that
// inserts a template's key is its representative.
// Templates only pay when B's row patterns repeat: unless at least a quarter
// of B's rows matched another row's pattern (*matched of nb), every row gets
// tplid -1 and no table is built.
__device__ __forceinline__ bool templates_worth(const int32_t* matched, int32_t nb) {
  return (int64_t)matched[0] * 4 >= (int64_t)nb;
}

__global__ void __launch_bounds__(256) k_tpl_insert(int32_t n, const int32_t* __restrict__ ap,
                                                    const int32_t* __restrict__ ac, const int32_t* __restrict__ aclass,
                                                    const int32_t* __restrict__ bclass,
                                                    unsigned long long* __restrict__ keys, int32_t* __restrict__ rep,
                                                    int32_t* __restrict__ tplid, const int32_t* __restrict__ matched,
                                                    int32_t nb) {
  constexpr int G = kClsLanes;
  if (!templates_worth(matched, nb)) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
      tplid[i] = -1;
    return;
  }
  __shared__ KeyCache kc;
  cache_init(kc);
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  const int64_t stride = (int64_t)gridDim.x * gpb;
  const int64_t wave_first = (int64_t)blockIdx.x * gpb + (threadIdx.x / kWave) * (kWave / G);
  for (int64_t w0 = wave_first; w0 < n; w0 += stride) {
    const int64_t i = w0 + (threadIdx.x % kWave) / G;
    const int a = i < n ? aclass[i] : -1;
    unsigned long long h = 0;
    int ok = a >= 0 ? 1 : 0;
    if (a >= 0) {
      const int s = ap[i], e = ap[i + 1];
      for (int u0 = lane; u0 < e - s; u0 += G * kClsUnroll) {
        int k[kClsUnroll];
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++) k[x] = u0 + x * G < e - s ? ac[s + u0 + x * G] : -1;
        int bc[kClsUnroll];
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++) bc[x] = k[x] >= 0 ? bclass[k[x]] : 0;
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++)
          if (k[x] >= 0) {
            ok &= bc[x] >= 0 ? 1 : 0;
            h += tmix(((unsigned long long)(u0 + x * G + 1) << 32) ^ (unsigned long long)(uint32_t)bc[x]);
          }
      }
    }
    h = gsum<G, unsigned long long>(h, nullptr);
    ok = gmin<G>(ok);
    h = (h + tmix(0x7a3ull ^ ((unsigned long long)(uint32_t)a << 8))) | 1ull;
    bool mine;
    const int slot = wave_insert(kc, keys, kTplSlots, h, lane == 0 && ok, &mine);
    if (lane == 0 && i < n) {
      if (mine) rep[slot] = (int32_t)i;
      tplid[i] = ok ? slot : -1;
    }
  }
}

// A*A (A's rows are B's rows): the class check of each row and its template
// candidate in one pass.  Row r is compared entry by entry with its class's
// pattern, which the class pass recorded from the representative (so the
// check reads it by class id instead of through the representative's row
// bounds), and, as an output row, hashed with its own class and
// its entries' B classes (as k_tpl_insert).  The B classes are read while
// other rows' checks may still clear theirs: a candidate whose B class is
// cleared afterwards fails the pre-pass check, which requires every class
// of the row to be set and equal to the template's.  Templates go ahead
// unless the sampled patterns were mostly distinct.
__global__ void __launch_bounds__(256) k_verify_tpl(int32_t n, const int32_t* __restrict__ ap,
                                                    const int32_t* __restrict__ ac, const int32_t* __restrict__ cpat,
                                                    int32_t* cls, unsigned long long* __restrict__ tkeys,
                                                    int32_t* __restrict__ trep, int32_t* __restrict__ tplid,
                                                    const int32_t* __restrict__ gate) {
  constexpr int G = kClsLanes;
  if (patterns_distinct(gate)) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
      tplid[i] = -1;
    return;
  }
  __shared__ KeyCache kc;
  cache_init(kc);
  const int lane = threadIdx.x % G;
  const int64_t gpb = blockDim.x / G;
  const int64_t stride = (int64_t)gridDim.x * gpb;
  const int64_t wave_first = (int64_t)blockIdx.x * gpb + (threadIdx.x / kWave) * (kWave / G);
  for (int64_t w0 = wave_first; w0 < n; w0 += stride) {
    const int64_t r = w0 + (threadIdx.x % kWave) / G;
    const int c = r < n ? cls[r] : -1;
    const int s = r < n ? ap[r] : 0, e = r < n ? ap[r + 1] : 0;
    const int m = e - s;
    // the class's pattern as its inserting row recorded it: length, then
    // relative offsets (a classed row has at most kTplMaxLen entries)
    const int32_t* cp = cpat + (size_t)(c >= 0 ? c : 0) * (kTplMaxLen + 1);
    const int cm = c >= 0 ? cp[0] : 0;
    // one pass over the row's entries: the comparison with the class's
    // pattern and the template hash of the entries' B classes, their loads
    // in flight together.  The hash is used only if the row keeps its class.
    int same = 1;
    unsigned long long h = 0;
    int ok = c >= 0 ? 1 : 0;
    if (c >= 0) {
      for (int u0 = lane; u0 < m; u0 += G * kClsUnroll) {
        int k[kClsUnroll], x1[kClsUnroll];
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++) {
          const bool in = u0 + x * G < m;
          k[x] = in ? ac[s + u0 + x * G] : -1;
          x1[x] = in ? cp[1 + u0 + x * G] : 0;
        }
        int bc[kClsUnroll];
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++) bc[x] = k[x] >= 0 ? __hip_atomic_load(&cls[k[x]], __ATOMIC_RELAXED,
                                                                                   __HIP_MEMORY_SCOPE_AGENT) : 0;
#pragma unroll
        for (int x = 0; x < kClsUnroll; x++)
          if (k[x] >= 0) {
            if ((int64_t)k[x] - r != (int64_t)x1[x]) same = 0;
            ok &= bc[x] >= 0 ? 1 : 0;
            h += tmix(((unsigned long long)(u0 + x * G + 1) << 32) ^ (unsigned long long)(uint32_t)bc[x]);
          }
      }
    }
    if (cm != m) same = 0;
    same = gmin<G>(same);
    if (lane == 0 && c >= 0 && !same) cls[r] = -1;
    const int a = (c >= 0 && same) ? c : -1;
    if (a < 0) ok = 0;
    h = gsum<G, unsigned long long>(h, nullptr);
    ok = gmin<G>(ok);
    h = (h + tmix(0x7a3ull ^ ((unsigned long long)(uint32_t)a << 8))) | 1ull;
    bool mine;
    const int slot = wave_insert(kc, tkeys, kTplSlots, h, lane == 0 && ok, &mine);
    if (lane == 0 && r < n) {
      if (mine) trep[slot] = (int32_t)r;
      tplid[r] = ok ? slot : -1;
    }
  }
}

// Line and plane strides of a structured grid read off a row's column
// offsets (a row-ordering hint only): the positive offsets sorted, runs of
// consecutive values are clusters; the line stride is the centre of the
// first cluster past the immediate neighbours (offsets above 2), the plane
// stride the centre of the next cluster beyond twice that.  A 27-point 96^3
// stencil row gives {1}, {95, 96, 97}, {9215, 9216, 9217}, ...: 96 and 9216.
// {0, 0} when the offsets show no such clusters.
__device__ inline int2 grid_strides(const int32_t* cols, int m, int64_t row, int32_t* d /* kTplMaxLen, LDS */) {
  int np = 0;
  for (int u = 0; u < m; u++) {
    const int64_t x = (int64_t)cols[u] - row;
    if (x > 0 && x < INT32_MAX) {
      int k = np++;
      while (k > 0 && d[k - 1] > (int32_t)x) { d[k] = d[k - 1]; k--; }
      d[k] = (int32_t)x;
    }
  }
  int line = 0, plane = 0;
  int k = 0;
  while (k < np) {
    int e = k;
    while (e + 1 < np && d[e + 1] <= d[e] + 1) e++;   // run [k, e] (repeats included)
    const int a = d[k], b = d[e];
    if (a > 2) {
      const int c = a + (b - a) / 2;
      if (line == 0) line = c;
      else if (c > 2 * line) { plane = c; break; }
    }
    k = e + 1;
  }
  return make_int2(line, plane);
}

// Template tables, one block per template slot, from its representative row:
// the products' output offsets (B column - row), sorted; the distinct ones
// are S, and pos[(u, v)] is each product's rank among them.  The template is
// marked invalid when it has more than kTplMaxProducts products.
constexpr int kBuildT = 1024;
__global__ void __launch_bounds__(kBuildT) k_tpl_build(const int32_t* __restrict__ ap, const int32_t* __restrict__ ac,
                                                       const int4* __restrict__ bmeta, const int32_t* __restrict__ bc,
                                                       int32_t row0, const unsigned long long* __restrict__ keys,
                                                       const int32_t* __restrict__ rep, const int32_t* __restrict__ bclass,
                                                       int4* __restrict__ info, int32_t* __restrict__ posoff,
                                                       uint16_t* __restrict__ pos, int32_t* __restrict__ soff,
                                                       int32_t* __restrict__ tuple, uint32_t* __restrict__ pos2,
                                                       int2* __restrict__ grid) {
  __shared__ unsigned long long key[kTplMaxProducts];
  __shared__ uint16_t rank_of[kTplMaxProducts];
  __shared__ int32_t off[kTplMaxLen + 1];
  __shared__ int32_t bstart[kTplMaxLen];
  __shared__ int32_t scan[kBuildT / kWave + 1];
  __shared__ int32_t maxb_s;
  __shared__ int32_t dsort[kTplMaxLen];
  __shared__ int32_t cols[kTplMaxLen];
  __shared__ int32_t blen[kTplMaxLen];
  const int t = blockIdx.x, tid = threadIdx.x;
  if (keys[t] == 0ull) {
    if (tid == 0) {
      info[t] = make_int4(0, 0, 0, 0);
      grid[t] = make_int2(0, 0);
    }
    return;
  }
  const int r = rep[t];
  const int s = ap[r], m = ap[r + 1] - s;
  const int64_t gi = (int64_t)row0 + r;
  // the representative's entries and their B rows, one thread each (m <=
  // kTplMaxLen <= kBuildT: template rows are at most kTplMaxLen long)
  if (tid < m) {
    const int k = ac[s + tid];
    const int4 b = bmeta[k];
    cols[tid] = k;
    bstart[tid] = b.x;
    blen[tid] = b.y - b.x;
  }
  __syncthreads();
  if (tid == 0) {
    int run = 0, maxb = 0;
    for (int u = 0; u < m; u++) {
      off[u] = run;
      run += blen[u];
      maxb = blen[u] > maxb ? blen[u] : maxb;
    }
    off[m] = run;
    maxb_s = maxb;
    grid[t] = grid_strides(cols, m, gi, dsort);
  }
  __syncthreads();
  const int P = off[m];
  if (P > kTplMaxProducts || P <= 0) {
    if (tid == 0) info[t] = make_int4(0, m, 0, 0);
    return;
  }
  int32_t* po = posoff + (size_t)t * (kTplMaxLen + 1);
  for (int u = tid; u <= m; u += kBuildT) po[u] = off[u];
  // the representative's B classes, entry by entry: the pre-pass compares
  // every candidate row against them
  for (int u = tid; u < m; u += kBuildT) tuple[(size_t)t * kTplMaxLen + u] = bclass[cols[u]];
  // keys: (offset + 2^31) << 32 | product index; padding sorts last
  int np2 = 1;
  while (np2 < P) np2 <<= 1;
  for (int p = tid; p < np2; p += kBuildT) {
    unsigned long long k = ~0ull;
    if (p < P) {
      int u = 0;
      while (off[u + 1] <= p) u++;
      const int v = p - off[u];
      const int64_t d = (int64_t)bc[bstart[u] + v] - gi;
      k = ((unsigned long long)(d + 2147483648LL) << 32) | (unsigned long long)p;
    }
    key[p] = k;
  }
  __syncthreads();
  // bitonic sort of np2 keys in LDS
  for (int kk = 2; kk <= np2; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int p = tid; p < np2; p += kBuildT) {
        const int q = p ^ j;
        if (q > p) {
          const bool up = (p & kk) == 0;
          const unsigned long long x = key[p], y = key[q];
          if ((x > y) == up) {
            key[p] = y;
            key[q] = x;
          }
        }
      }
      __syncthreads();
    }
  // rank of each distinct offset: block scan of "first of its offset" flags,
  // kBuildT elements per round
  int base = 0;
  uint16_t* pp = pos + (size_t)t * kTplMaxProducts;
  int32_t* so = soff + (size_t)t * kTplMaxProducts;
  for (int p0 = 0; p0 < P; p0 += kBuildT) {
    const int p = p0 + tid;
    const bool head = p < P && (p == 0 || (key[p] >> 32) != (key[p - 1] >> 32));
    int tot;
    const int ex = block_scan_excl<int>(head ? 1 : 0, scan, &tot);
    if (p < P) {
      const int rank = base + ex + (head ? 1 : 0) - 1;
      pp[key[p] & 0xffffffffull] = (uint16_t)rank;
      rank_of[key[p] & 0xffffffffull] = (uint16_t)rank;
      if (head) so[rank] = (int32_t)((int64_t)(key[p] >> 32) - 2147483648LL);
    }
    base += tot;
    __syncthreads();
  }
  // pair tables: for parity par and entry u, pair l holds the slots of
  // entries 2l - par and 2l - par + 1 of B row k_u (0xffff: none)
  uint32_t* p2 = pos2 + (size_t)t * (2 * kTplMaxLen * kTplPairs);
  __syncthreads();
  for (int x = tid; x < 2 * m * kTplPairs; x += kBuildT) {
    const int par = x / (m * kTplPairs), u = (x / kTplPairs) % m, l = x % kTplPairs;
    const int len = off[u + 1] - off[u];
    const int v0 = 2 * l - par, v1 = v0 + 1;
    const uint32_t s0 = (v0 >= 0 && v0 < len) ? rank_of[off[u] + v0] : 0xffffu;
    const uint32_t s1 = (v1 >= 0 && v1 < len) ? rank_of[off[u] + v1] : 0xffffu;
    p2[((size_t)par * kTplMaxLen + u) * kTplPairs + l] = s0 | (s1 << 16);
  }
  // .w: the widest of the A row and its B rows (sizes the numeric kernel's
  // lanes per row and its staged entries)
  if (tid == 0) info[t] = make_int4(1, m, base, maxb_s > m ? maxb_s : m);
}

}  // namespace

// A-side rows that are B's rows by pattern (A*A, and A*A^T of a symmetric
// pattern, which *apat == 0 reports: the transpose gather succeeded): their
// classes are B's classes of the same global rows, copied instead of hashed
// and checked again.
__global__ void k_class_copy(int32_t n, int32_t row0, const int32_t* __restrict__ bclass,
                             int32_t* __restrict__ aclass, const int32_t* __restrict__ apat) {
  if (*apat != 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    aclass[i] = bclass[(int64_t)row0 + i];
}

void launch_tpl_detect(const TplArgs& T, int32_t n, const int32_t* ap, const int32_t* ac, int32_t nb,
                       int4* bmeta, const int32_t* bc, bool a_is_b, const int32_t* bptr, int32_t* bflag,
                       int32_t* longrows, const int32_t* apat, hipStream_t s) {
  // tables empty (a representative is written by the row that inserts its key)
  launch_fill3_i32(reinterpret_cast<int32_t*>(T.ckey), 2 * kClassSlots, reinterpret_cast<int32_t*>(T.tkey),
                   2 * kTplSlots, T.matched, 4, s);
  auto grid = [](int64_t rows) {
    int64_t b = (rows * kClsLanes + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > kClsMaxBlocks ? kClsMaxBlocks : b));
  };
  const Rows RB{nullptr, bmeta, 0};
  if (nb > 0) {
    hipLaunchKernelGGL(k_class_sample, dim3(kSampleBlocks), dim3(256), 0, s, nb, bptr ? Rows{bptr, nullptr, 0} : RB,
                       bc, T.matched);
    if (bptr) {
      // B's row metadata from the class pass (rows of at most kTplMaxLen
      // entries) and a pass of its own for the longer rows
      const Rows RP{bptr, nullptr, 0};
      hipLaunchKernelGGL(k_class_insert<true>, dim3(grid(nb)), dim3(256), 0, s, nb, RP, bc, T.ckey, T.crep,
                         T.bclass, bmeta, bflag, (const int32_t*)T.matched, longrows, nullptr, T.cpat);
      launch_rowstats(nb, bptr, nullptr, nullptr, 0, bc, bmeta, bflag, s, kTplMaxLen, longrows);
    } else {
      hipLaunchKernelGGL(k_class_insert<false>, dim3(grid(nb)), dim3(256), 0, s, nb, RB, bc, T.ckey, T.crep,
                         T.bclass, nullptr, nullptr, (const int32_t*)T.matched, nullptr, nullptr, T.cpat);
    }
    // A*A: the check runs inside the template pass below
    if (!a_is_b)
      hipLaunchKernelGGL(k_class_verify, dim3(grid(nb)), dim3(256), 0, s, nb, RB, bc, T.crep, T.bclass, T.matched,
                         (const int32_t*)T.matched, nullptr);
  }
  if (n <= 0) return;
  if (a_is_b) {
    hipLaunchKernelGGL(k_verify_tpl, dim3(grid(n)), dim3(256), 0, s, n, ap, ac, (const int32_t*)T.cpat, T.bclass,
                       T.tkey, T.trep, T.tplid, (const int32_t*)T.matched);
    hipLaunchKernelGGL(k_tpl_build, dim3(kTplSlots), dim3(kBuildT), 0, s, ap, ac, bmeta, bc, T.row0, T.tkey, T.trep,
                       T.bclass, T.info, T.posoff, T.pos, T.soff, T.tuple, T.pos2, T.grid);
    return;
  }
  if (!a_is_b) {
    // A-side rows (a row panel, or A next to B = A^T): a table of their own,
    // whose representatives are A-side rows
    // (rows whose pattern is B's take B's classes instead, decided on the
    // device by *apat)
    launch_fill_i32(reinterpret_cast<int32_t*>(T.akey), 2 * kClassSlots, 0, s);
    const Rows RA{ap, nullptr, T.row0};
    int64_t cb = ((int64_t)n + 255) / 256;
    if (cb > 4096) cb = 4096;
    hipLaunchKernelGGL(k_class_copy, dim3((unsigned)cb), dim3(256), 0, s, n, T.row0, (const int32_t*)T.bclass,
                       T.aclass, apat);
    hipLaunchKernelGGL(k_class_insert<false>, dim3(grid(n)), dim3(256), 0, s, n, RA, ac, T.akey, T.arep, T.aclass,
                       nullptr, nullptr, (const int32_t*)T.matched, nullptr, apat, nullptr);
    hipLaunchKernelGGL(k_class_verify, dim3(grid(n)), dim3(256), 0, s, n, RA, ac, T.arep, T.aclass, T.matched + 1,
                       (const int32_t*)T.matched, apat);
  }
  hipLaunchKernelGGL(k_tpl_insert, dim3(grid(n)), dim3(256), 0, s, n, ap, ac, T.aclass, T.bclass, T.tkey, T.trep,
                     T.tplid, T.matched, nb);
  hipLaunchKernelGGL(k_tpl_build, dim3(kTplSlots), dim3(kBuildT), 0, s, ap, ac, bmeta, bc, T.row0, T.tkey, T.trep,
                     T.bclass, T.info, T.posoff, T.pos, T.soff, T.tuple, T.pos2, T.grid);
}

// ---------------------------------------------------------------------------
// Numeric pass of template rows.  G = 32 lanes per row (two rows per
// wavefront), rows walked XCD-locally by persistent blocks and
// software-pipelined four rows deep, so that no load waits on another load
// issued in the same iteration: rows[r] -> (template, row start and length)
// -> (a_iu, B-row id, template product offsets of entry u) -> B-row start ->
// products.  For a row, entry u's values are staged in LDS; then every
// product's gathers (b_kv, and its slot pos[u][v] in the template) are issued
// at once -- lane v takes entry v of B row k_u, for all u -- and the LDS f64
// atomics follow in u order, so every slot receives its products in A-row
// storage order (bit-identical to the row-wise Gustavson loop).  The row is
// written out in slot order (S is sorted): columns row + S[s].
// Rows of up to kTplFast A entries; longer template rows take the same
// product loop kTplFast entries at a time.
// ---------------------------------------------------------------------------
#ifndef APPT_TPL_FAST
#define APPT_TPL_FAST 6
#endif
constexpr int kTplFast = APPT_TPL_FAST;  // A entries whose gathers are in flight together

// Gather with a 32-bit byte offset from a uniform base when the array is
// small enough (one address register per load instead of two).
template <typename V, bool W64>
__device__ __forceinline__ V gather(const V* base, uint32_t idx) {
  if constexpr (W64) return base[idx];
  else return *(const V*)((const char*)base + (size_t)(idx * (uint32_t)sizeof(V)));
}

struct TplRow {
  int i, t, s, len;   // row, template, first entry, entries
  int k, o, e, bs;    // this lane's entry (lane < len): B row, products [o, e), B-row start
  double a;           // a_iu of this lane's entry
};

template <int G, bool W64>
__global__ void __launch_bounds__(256) k_tpl_num(RowArgs A, TplArgs T, int32_t nsmax,
                                                 const int64_t* __restrict__ rowoff, int32_t* __restrict__ ccol,
                                                 double* __restrict__ cval) {
  // G covers the longest B row of the launch (lane v takes entry v)
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  unsigned char* base = lds_raw + (size_t)gi * (size_t)tpl_group_bytes(nsmax);
  double* sa = (double*)base;                         // a_iu
  int4* se = (int4*)(base + kTplMaxLen * 8);          // {B-row start, first product, products, -}
  double* vals = (double*)(base + kTplMaxLen * 24);

  const RowWalk walk = row_walk(A.nrows, R, gi);
  const int S = walk.step, r1 = walk.end;
  auto hop1 = [&](TplRow& q) {
    q.t = T.tplid[q.i];
    q.s = A.ap[q.i];
    q.len = A.ap[q.i + 1] - q.s;
  };
  auto hop2 = [&](TplRow& q) {
    q.k = -1;
    if (lane < q.len) {
      const int32_t* po = T.posoff + (size_t)q.t * (kTplMaxLen + 1);
      q.a = A.av[q.s + lane];
      q.k = A.ac[q.s + lane];
      q.o = po[lane];
      q.e = po[lane + 1];
    }
  };
  auto hop3 = [&](TplRow& q) {
    if (q.k >= 0) q.bs = A.bmeta[q.k].x;
  };
  int r = walk.first;
  TplRow q0{}, q1{}, q2{}, q3{};
  q0.len = q1.len = q2.len = q3.len = 0;
  q0.k = q1.k = q2.k = q3.k = -1;
  if (r < r1) { q0.i = A.rows[r]; hop1(q0); hop2(q0); hop3(q0); }
  if (r + S < r1) { q1.i = A.rows[r + S]; hop1(q1); hop2(q1); }
  if (r + 2 * S < r1) { q2.i = A.rows[r + 2 * S]; hop1(q2); }
  if (r + 3 * S < r1) q3.i = A.rows[r + 3 * S];
  for (; r < r1; r += S) {
    TplRow q4{};
    q4.len = 0;
    q4.k = -1;
    if (r + 4 * S < r1) q4.i = A.rows[r + 4 * S];
    if (r + 3 * S < r1) hop1(q3);
    if (r + 2 * S < r1) hop2(q2);
    if (r + S < r1) hop3(q1);
    const int i = q0.i, t = q0.t, m = q0.len;
    const int4 inf = T.info[t];
    const int nS = inf.z;
    const int64_t off = rowoff[i];
    const uint16_t* pp = T.pos + (size_t)t * kTplMaxProducts;
    for (int q = lane; q < nS; q += G) vals[q] = 0.0;
    if (lane < m) {
      sa[lane] = q0.a;
      se[lane] = make_int4(q0.bs, q0.o, q0.e - q0.o, 0);
    }
    for (int u = lane + G; u < m; u += G) {   // rows longer than G entries
      const int k = A.ac[q0.s + u];
      const int32_t* po = T.posoff + (size_t)t * (kTplMaxLen + 1);
      sa[u] = A.av[q0.s + u];
      se[u] = make_int4(A.bmeta[k].x, po[u], po[u + 1] - po[u], 0);
    }
    gsync<G>();
    for (int u0 = 0; u0 < m; u0 += kTplFast) {
      double pv[kTplFast];
      uint32_t slot[kTplFast];
#pragma unroll
      for (int d = 0; d < kTplFast; d++) {
        const int u = u0 + d < m ? u0 + d : u0;
        const int4 x = se[u];
        const bool ok = u0 + d < m && lane < x.z;
        const int v = ok ? lane : 0;
        slot[d] = ok ? (uint32_t)gather<uint16_t, false>(pp, (uint32_t)(x.y + v)) : 0xffffffffu;
        pv[d] = sa[u] * gather<double, W64>(A.bv, (uint32_t)(x.x + v));
      }
#pragma unroll
      for (int d = 0; d < kTplFast; d++)
        if (slot[d] != 0xffffffffu) atomicAdd(&vals[slot[d]], pv[d]);
    }
    gsync<G>();
    const int64_t row = (int64_t)T.row0 + i;
    const int32_t* so = T.soff + (size_t)t * kTplMaxProducts;
    for (int q = lane; q < nS; q += G) {
      ccol[off + q] = (int32_t)(row + so[q]);
      cval[off + q] = vals[q];
    }
    gsync<G>();
    q0 = q1; q1 = q2; q2 = q3; q3 = q4;
  }
}

// Template rows no wider than kTplPairMax (A row and every B row): G = 16
// lanes per row (four rows per wavefront), each lane gathering an aligned
// pair of B values with one 16-byte load.  B row k_u starting at position bs
// is read as pairs from bs & ~1, so lane l holds entries v = 2l - par and
// 2l - par + 1 (par = bs & 1); the template keeps the slots of both for each
// parity (pos2).  Entries u are taken in storage order, their gathers issued
// kTplFast at a time with no branch around them (dead lanes load entry 0),
// and the LDS f64 atomics follow in u order.  A lane holds two of the row's
// A entries (lane, lane + G) through the four-deep row pipeline, which also
// brings the row's output offset and slot count.  The write-out takes four
// slots per lane (16-byte stores) and clears them for the next row.
constexpr int kTplPairG = 16;
constexpr int kTplPairMax = 2 * kTplPairG - 1;
constexpr int kTplPairLen = 2 * kTplPairG;   // A entries a group stages (<= kTplPairMax)

// Two consecutive B values from entry idx on (8-byte aligned).
typedef double dpair __attribute__((ext_vector_type(2), aligned(8)));
template <bool W64>
__device__ __forceinline__ dpair load_pair(const double* base, uint32_t idx) {
  if constexpr (W64) return *(const dpair*)(base + idx);
  else return *(const dpair*)((const char*)base + (size_t)(idx * 8u));
}

// 16-byte access to int32 data that is only 4-byte aligned.
typedef int32_t i32x4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ int4 load_i4(const int32_t* p) {
  const i32x4u v = *(const i32x4u*)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_i4(int32_t* p, int4 v) {
  *(i32x4u*)p = i32x4u{v.x, v.y, v.z, v.w};
}

// A row in the pair kernel's pipeline.  Every hop stores what it loaded as it
// is and derives nothing from it: a value computed from a load at once would
// make the hop wait for that load, and the next rows' loads are meant to be
// in flight while this row's products are formed.
struct TplRow2 {
  int i, t, s, e;    // row, template, first entry, end
  int k0, k1;        // B rows of entries lane and lane + G (-1: none)
  int bs0, bs1;      // B-row starts
  int be0, be1;      // B-row ends
  double a0, a1;     // a_iu of those entries
  int64_t off;       // first output position of the row
  int nS;            // the template's slot count
};

template <bool W64>
__global__ void __launch_bounds__(256) k_tpl_num2(RowArgs A, TplArgs T, int32_t nsmax, int32_t tile_n,
                                                  const int64_t* __restrict__ rowoff, int32_t* __restrict__ ccol,
                                                  double* __restrict__ cval) {
  constexpr int G = kTplPairG;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int R = blockDim.x / G;
  const int gi = threadIdx.x / G, lane = threadIdx.x % G;
  unsigned char* base = lds_raw + (size_t)gi * (size_t)tpl_pair_group_bytes(nsmax);
  double* sa = (double*)base;                      // a_iu
  int2* se = (int2*)(base + kTplPairLen * 8);      // {B-row start, B-row length}
  double* vals = (double*)(base + kTplPairLen * 16);
  // Row order.  When the bin holds every template row of the matrix and the
  // middle row's template shows grid strides (line TL, plane TP), rows are
  // walked in tiles: blocks of TY lines through all planes, (tile, plane,
  // line, column) order, rows of no template skipped.  An XCD's rows in
  // flight then reach into a few planes of one tile instead of whole planes,
  // so the B rows they share stay in its L2 (C2s numeric 0.91 -> 0.86 ms).
  // The walk covers each row once whatever the strides; otherwise rows come
  // from the bin's list.
  int TL = 0, TP = 0, TY = 1, Ly = 1, Zs = 1;
  if (tile_n > 0) {
    const int th = T.tplid[tile_n / 2];
    if (th >= 0) {
      const int2 g = T.grid[th];
      if (g.x >= 2 && g.y >= 4 * g.x && g.y <= tile_n) {
        TL = g.x;
        TP = g.y;
        Ly = (TP + TL - 1) / TL;
        TY = 2304 / TL;
        TY = TY < 1 ? 1 : (TY > Ly ? Ly : TY);
        Zs = (int)(((int64_t)tile_n + TP - 1) / TP);
      }
    }
  }
  const int64_t tiles_n = TL > 0 ? (int64_t)((Ly + TY - 1) / TY) * Zs * TY * TL : 0;
  const bool tiled = TL > 0 && tiles_n < INT32_MAX;
  const RowWalk walk = row_walk(tiled ? (int)tiles_n : A.nrows, R, gi);
  // a tile position's row, or -1 off the grid; whether the row has a
  // template is known once hop1 has read it (rows without are skipped).
  // The walk order is a template argument of the row loop (tl), so the list
  // walk's row load and the tiled walk's arithmetic do not share a register:
  // a shared one makes the tiled walk wait for all loads in flight (C2s
  // numeric 0.825 -> 0.813 ms).
  auto row_at = [&](int rr, auto tl) -> int {
    if constexpr (!decltype(tl)::value) return A.rows[rr];
    const uint32_t ur = (uint32_t)rr;
    const uint32_t x = ur % (uint32_t)TL;
    uint32_t t = ur / (uint32_t)TL;
    const uint32_t yl = t % (uint32_t)TY;
    t /= (uint32_t)TY;
    const uint32_t z = t % (uint32_t)Zs, y = (t / (uint32_t)Zs) * (uint32_t)TY + yl;
    if (y >= (uint32_t)Ly || y * (uint32_t)TL + x >= (uint32_t)TP) return -1;
    const int64_t i = (int64_t)z * TP + (int64_t)y * TL + x;
    return i < tile_n ? (int)i : -1;
  };
  const int S = walk.step, r1 = walk.end;
  const uint32_t bnnz = (uint32_t)A.bnnz;
  for (int q = lane; q < nsmax; q += G) vals[q] = 0.0;
  gsync<G>();

  auto hop1 = [&](TplRow2& q) {
    if (q.i < 0) {   // a tile position off the grid
      q.t = -1;
      q.s = q.e = 0;
      return;
    }
    q.t = T.tplid[q.i];   // -1: not a template row (tiled walk), skipped
    q.s = A.ap[q.i];
    q.e = A.ap[q.i + 1];
  };
  auto hop2 = [&](TplRow2& q) {
    q.k0 = q.k1 = -1;
    if (q.t < 0) {
      q.s = q.e = 0;
      q.off = 0;
      q.nS = 0;
      return;
    }
    if (q.s + lane < q.e) {
      q.a0 = A.av[q.s + lane];
      q.k0 = A.ac[q.s + lane];
    }
    if (q.s + lane + G < q.e) {
      q.a1 = A.av[q.s + lane + G];
      q.k1 = A.ac[q.s + lane + G];
    }
    q.off = rowoff[q.i];
    q.nS = T.info[q.t].z;
  };
  auto hop3 = [&](TplRow2& q) {
    if (q.k0 >= 0) {
      const int4 b = A.bmeta[q.k0];
      q.bs0 = b.x;
      q.be0 = b.y;
    }
    if (q.k1 >= 0) {
      const int4 b = A.bmeta[q.k1];
      q.bs1 = b.x;
      q.be1 = b.y;
    }
  };
  auto run = [&](auto tl) {
    int r = walk.first;
    TplRow2 q0{}, q1{}, q2{}, q3{};
    q0.s = q0.e = q1.s = q1.e = q2.s = q2.e = q3.s = q3.e = 0;
    q0.k0 = q1.k0 = q2.k0 = q3.k0 = -1;
    q0.k1 = q1.k1 = q2.k1 = q3.k1 = -1;
    if (r < r1) { q0.i = row_at(r, tl); hop1(q0); hop2(q0); hop3(q0); }
    if (r + S < r1) { q1.i = row_at(r + S, tl); hop1(q1); hop2(q1); }
    if (r + 2 * S < r1) { q2.i = row_at(r + 2 * S, tl); hop1(q2); }
    if (r + 3 * S < r1) q3.i = row_at(r + 3 * S, tl);
    for (; r < r1; r += S) {
      TplRow2 q4{};
      q4.s = q4.e = 0;
      q4.k0 = q4.k1 = -1;
      if (r + 4 * S < r1) q4.i = row_at(r + 4 * S, tl);
      if (r + 3 * S < r1) hop1(q3);
      if (r + 2 * S < r1) hop2(q2);
      if (r + S < r1) hop3(q1);
      const int t = q0.t < 0 ? 0 : q0.t, m = q0.e - q0.s, nS = q0.nS;   // skipped rows: m = nS = 0
      if (lane < m) {
        sa[lane] = q0.a0;
        se[lane] = make_int2(q0.bs0, q0.be0 - q0.bs0);
      }
      if (lane + G < m) {
        sa[lane + G] = q0.a1;
        se[lane + G] = make_int2(q0.bs1, q0.be1 - q0.bs1);
      }
      gsync<G>();
      const uint32_t* p2 = T.pos2 + (size_t)t * (2 * kTplMaxLen * kTplPairs);
      const int64_t row = (int64_t)T.row0 + q0.i;
      const int64_t off = q0.off;
      const int32_t* so = T.soff + (size_t)t * kTplMaxProducts;
      // one batch of W entries from u0: their gathers issued together, no
      // branch around them (dead lanes and entries past m load entry 0)
      auto batch = [&](int u0, auto wc) {
        constexpr int W = decltype(wc)::value;
        double pv0[W], pv1[W];
        uint32_t sl[W];
        int2 xs[W];
        double as[W];
        // the batch's staged entries first (one LDS wait for all of them)
#pragma unroll
        for (int d = 0; d < W; d++) {
          const int u = u0 + d < m ? u0 + d : u0;
          xs[d] = se[u];
          as[d] = sa[u];
        }
#pragma unroll
        for (int d = 0; d < W; d++) {
          const int u = u0 + d < m ? u0 + d : u0;
          const int2 x = xs[d];
          const int par = x.x & 1;
          const bool live = (u0 + d < m) & (2 * lane < x.y + par);
          const uint32_t q = live ? (uint32_t)((x.x & ~1) + 2 * lane) : 0u;   // even: the pair (q, q + 1)
          // the last entry of an odd-length B: load the pair ending there
          const bool tail = q + 1 >= bnnz;
          const uint32_t v = gather<uint32_t, false>(p2, (uint32_t)(((par * kTplMaxLen) + u) * kTplPairs + lane));
          sl[d] = live ? v : 0xffffffffu;
          const dpair raw = load_pair<W64>(A.bv, tail ? q - 1 : q);
          const double bx = tail ? raw.y : raw.x, by = tail ? 0.0 : raw.y;
          pv0[d] = as[d] * bx;
          pv1[d] = as[d] * by;
        }
#pragma unroll
        for (int d = 0; d < W; d++) {
          const uint32_t s0 = sl[d] & 0xffffu, s1 = sl[d] >> 16;
          if (s0 != 0xffffu) atomicAdd(&vals[s0], pv0[d]);
          if (s1 != 0xffffu) atomicAdd(&vals[s1], pv1[d]);
        }
      };
      // full batches, then the rest in a batch of kTplFast or kTplFast / 2
      // entries: a 27-entry row ends in 3 entries instead of 3 live and 3
      // dead ones (the gathers' address processing bounds the kernel, and
      // dead entries cost it as much as live ones)
      int u0 = 0;
      for (; u0 + kTplFast <= m; u0 += kTplFast) batch(u0, std::integral_constant<int, kTplFast>{});
      if (m - u0 > kTplFast / 2) batch(u0, std::integral_constant<int, kTplFast>{});
      else if (u0 < m) batch(u0, std::integral_constant<int, kTplFast / 2>{});
      gsync<G>();
      // write-out: four consecutive slots per lane (16-byte stores, so a
      // row takes a third of the store instructions), clearing them
      const int32_t rowc = (int32_t)row;
      for (int q = 4 * lane; q < nS; q += 4 * G) {
        if (q + 4 <= nS) {
          const int4 o = load_i4(so + q);
          const dpair v0 = *(const dpair*)&vals[q], v1 = *(const dpair*)&vals[q + 2];
          *(dpair*)&vals[q] = dpair{0.0, 0.0};
          *(dpair*)&vals[q + 2] = dpair{0.0, 0.0};
          store_i4(ccol + off + q, make_int4(rowc + o.x, rowc + o.y, rowc + o.z, rowc + o.w));
          *(dpair*)(cval + off + q) = v0;
          *(dpair*)(cval + off + q + 2) = v1;
        } else {
          for (int e = q; e < nS; e++) {
            const int32_t o = so[e];
            const double v = vals[e];
            vals[e] = 0.0;
            ccol[off + e] = rowc + o;
            cval[off + e] = v;
          }
        }
      }
      gsync<G>();
      q0 = q1; q1 = q2; q2 = q3; q3 = q4;
    }
  };
  if (tiled) run(std::true_type{});
  else run(std::false_type{});
}

void launch_tpl_num(const RowArgs& a, const TplArgs& T, int32_t maxb, int32_t nsmax, int32_t tile_n,
                    const int64_t* rowoff, int32_t* ccol, double* cval, int num_cus, int64_t lds_per_cu,
                    hipStream_t s) {
  if (a.nrows <= 0) return;
  if (maxb <= kTplPairMax && a.bnnz >= 2) {
    // maxb: the widest A or B row of any template row (pairs of B values
    // need at least two of them)
    const int64_t gb = tpl_pair_group_bytes(nsmax);
    int R = 256 / kTplPairG;
    while (R > 4 && (int64_t)R * gb * 4 > lds_per_cu) R >>= 1;
    const size_t lds = (size_t)((int64_t)R * gb);
    const bool w64 = a.bnnz >= ((int64_t)1 << 29);
    const void* fn = w64 ? (const void*)k_tpl_num2<true> : (const void*)k_tpl_num2<false>;
    set_dyn_lds(fn, lds);
    // persistent blocks: as many as are resident at once (rows are split
    // among the blocks up front, so a block that waits for a slot is a tail)
    // (asked once per thread, device and LDS size)
    thread_local int fit_dev = -1, fit = 0;
    thread_local size_t fit_lds = 0;
    thread_local const void* fit_fn = nullptr;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (fit_dev != dev || fit_lds != lds || fit_fn != fn) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, fn, R * kTplPairG, lds) != hipSuccess || fit < 1) fit = 4;
      if (fit > 8) fit = 8;
      fit_dev = dev;
      fit_lds = lds;
      fit_fn = fn;
    }
    int64_t blocks = ((int64_t)a.nrows + R - 1) / R;
    const int64_t cap = (int64_t)num_cus * fit;
    if (blocks > cap) blocks = cap;
    if (w64)
      hipLaunchKernelGGL(k_tpl_num2<true>, dim3((unsigned)blocks), dim3(R * kTplPairG), lds, s, a, T, nsmax, tile_n,
                         rowoff, ccol, cval);
    else
      hipLaunchKernelGGL(k_tpl_num2<false>, dim3((unsigned)blocks), dim3(R * kTplPairG), lds, s, a, T, nsmax,
                         tile_n, rowoff, ccol, cval);
    return;
  }
  const int64_t gb = tpl_group_bytes(nsmax);
  const int G = maxb <= 32 ? 32 : 64;
  int R = 256 / G;
  while (R > 2 && (int64_t)R * gb * 4 > lds_per_cu) R >>= 1;
  const size_t lds = (size_t)R * (size_t)gb;
  int64_t blocks = ((int64_t)a.nrows + R - 1) / R;
  const int64_t cap = (int64_t)num_cus * 8;
  if (blocks > cap) blocks = cap;
  const bool w64 = a.bnnz >= ((int64_t)1 << 29);
  const void* fn = G == 32 ? (w64 ? (const void*)k_tpl_num<32, true> : (const void*)k_tpl_num<32, false>)
                           : (w64 ? (const void*)k_tpl_num<64, true> : (const void*)k_tpl_num<64, false>);
  set_dyn_lds(fn, lds);
  const dim3 grid((unsigned)blocks), block((unsigned)(R * G));
  if (G == 32) {
    if (w64) hipLaunchKernelGGL((k_tpl_num<32, true>), grid, block, lds, s, a, T, nsmax, rowoff, ccol, cval);
    else hipLaunchKernelGGL((k_tpl_num<32, false>), grid, block, lds, s, a, T, nsmax, rowoff, ccol, cval);
  } else {
    if (w64) hipLaunchKernelGGL((k_tpl_num<64, true>), grid, block, lds, s, a, T, nsmax, rowoff, ccol, cval);
    else hipLaunchKernelGGL((k_tpl_num<64, false>), grid, block, lds, s, a, T, nsmax, rowoff, ccol, cval);
  }
}

}  // namespace appt
