/*
 * oracle.c -- CPU restatement of the reference SpGEMM semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * check and the CPU-baseline leg of the benchmark load it; the product
 * library never links or calls it.
 *
 * What it restates (Wendy-Hamlet/appt_spgemm):
 *   - row-wise Gustavson C = A*B with B = A ("AA") or B = A^T ("AAt",
 *     "AAtS"): src/spgemm_engine.cuh:71-115 (symbolic/numeric), task
 *     selection :134-135; the CPU correctness path is scipy's A @ A and
 *     A @ A.T in src/export_bin.py:46-47, but with the *structural*
 *     semantics the engine and cuSPARSE use (report/CORRECTNESS_NOTES.md:3-16):
 *     entries that cancel to 0.0 are kept, columns strictly ascending.
 *   - "AAtS": only j >= i is formed (src/spgemm_engine.cuh:82,104), then
 *     mirrored, diagonal once (k_mirror, :121-127).
 *   - products per row k_flops (:47-50), table sizing np2l/k_caps (:33,51-57).
 *   - checksum (s, as, sq) summed sequentially over the row-major sorted
 *     values (:116-118, :194-196).
 *   - descriptors deg_cv / fill_aa / fill_aat (bench/dispatch.cu:26-37),
 *     decision rule pick_ours (:14-17), psym (src/fetch_and_prep.py:37-48).
 *
 * Value order (fixed so GPU kernels can be bit-compared): for row i, A's
 * entries p are visited in storage order, and for each, B row k's entries in
 * storage order; an output entry starts at +0.0 on first touch and each
 * contribution is added as acc = acc + (a*b), product rounded first, no fused
 * multiply-add (built with -ffp-contract=off).  B = A^T is built by a stable
 * counting sort (row k of A^T lists source rows ascending).  Alongside each
 * value the oracle can emit scale = sum |a*b| of the same contributions, the
 * denominator of the per-entry tolerance |c - c_ref| <= tol * scale.
 *
 * Pinning: tests check this file against closed forms (2D Laplacian values
 * and counts, 27-point stencil counts), dense brute force on small random
 * matrices (tests/golden fixtures), scipy pattern products, and invariants
 * (A symmetric => AA == AAt; AAtS == AAt structurally; row sums).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OP_AA 0
#define OP_AAT 1
#define OP_AATS 2

/* Stable counting-sort transpose (restates the k_colcount/scan/k_scatter
 * build of B = A^T, src/spgemm_engine.cuh:37-44,140-146, with a fixed order). */
int oracle_transpose(int32_t n, const int32_t* ap, const int32_t* ac, const double* av,
                     int32_t* bp, int32_t* bc, double* bv) {
  int64_t nnz = ap[n];
  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  if (!cnt) return 1;
  for (int64_t p = 0; p < nnz; p++) cnt[ac[p] + 1]++;
  for (int32_t k = 0; k < n; k++) cnt[k + 1] += cnt[k];
  for (int32_t k = 0; k <= n; k++) bp[k] = (int32_t)cnt[k];
  for (int32_t i = 0; i < n; i++)
    for (int32_t p = ap[i]; p < ap[i + 1]; p++) {
      int64_t pos = cnt[ac[p]]++;
      bc[pos] = i;
      if (bv && av) bv[pos] = av[p];
    }
  free(cnt);
  return 0;
}

/* k_flops: f_i = sum over row i of nnz(B row k) (src/spgemm_engine.cuh:47-50). */
void oracle_row_products(int32_t n, const int32_t* ap, const int32_t* ac, const int32_t* bp,
                         int64_t* f) {
  for (int32_t i = 0; i < n; i++) {
    int64_t s = 0;
    for (int32_t p = ap[i]; p < ap[i + 1]; p++) s += bp[ac[p] + 1] - bp[ac[p]];
    f[i] = s;
  }
}

/* np2l + k_caps sizing rule (src/spgemm_engine.cuh:33,51-57): shared-memory
 * capacity (0 if global) and global capacity per row. */
static int64_t np2l(int64_t x) { if (x < 2) return 2; int64_t p = 2; while (p < x) p <<= 1; return p; }
void oracle_caps(int32_t n, const int64_t* f, int32_t sh_cap, int32_t* rcap, int64_t* gcap) {
  for (int32_t i = 0; i < n; i++) {
    if (f[i] == 0) { rcap[i] = 0; gcap[i] = 0; continue; }
    int64_t want = np2l((f[i] * 4) / 3 + 1);
    if (want <= sh_cap) { rcap[i] = (int32_t)want; gcap[i] = 0; }
    else { rcap[i] = 0; gcap[i] = want; }
  }
}

typedef struct {
  int32_t* mark;   /* last row that touched column j, or -1 */
  double* acc;
  double* scl;
  int32_t* list;
  int32_t nlist;
} spa_t;

static int spa_init(spa_t* s, int32_t n, int want_scale) {
  s->mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  s->acc = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  s->scl = want_scale ? (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1)) : NULL;
  s->list = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!s->mark || !s->acc || !s->list || (want_scale && !s->scl)) return 1;
  for (int32_t j = 0; j < n; j++) s->mark[j] = -1;
  s->nlist = 0;
  return 0;
}
static void spa_free(spa_t* s) { free(s->mark); free(s->acc); free(s->scl); free(s->list); }

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* One output row: accumulate into the SPA, return distinct column count. */
static int32_t row_accumulate(spa_t* s, int32_t i, int upper_only, const int32_t* ap,
                              const int32_t* ac, const double* av, const int32_t* bp,
                              const int32_t* bc, const double* bv, int want_vals) {
  s->nlist = 0;
  for (int32_t p = ap[i]; p < ap[i + 1]; p++) {
    int32_t k = ac[p];
    double a = want_vals ? av[p] : 0.0;
    for (int32_t q = bp[k]; q < bp[k + 1]; q++) {
      int32_t j = bc[q];
      if (upper_only && j < i) continue;
      if (s->mark[j] != i) {
        s->mark[j] = i;
        s->list[s->nlist++] = j;
        if (want_vals) { s->acc[j] = 0.0; if (s->scl) s->scl[j] = 0.0; }
      }
      if (want_vals) {
        double prod = a * bv[q];
        s->acc[j] = s->acc[j] + prod;
        if (s->scl) s->scl[j] = s->scl[j] + fabs(prod);
      }
    }
  }
  return s->nlist;
}

/* Row-range Gustavson product C[r0:r1) = A[r0:r1) * B.  Caller passes B
 * explicitly (A or its transpose).  Two passes: counts into cnt[i - r0], then
 * (if col != NULL) fill at off[i - r0].  upper_only keeps j >= i.  Returns 0
 * on success, 1 on allocation failure.  threads <= 0 means all. */
int oracle_gustavson_range(int32_t n, const int32_t* ap, const int32_t* ac, const double* av,
                           const int32_t* bp, const int32_t* bc, const double* bv, int upper_only,
                           int32_t r0, int32_t r1, int64_t* cnt_or_off, int32_t* col, double* val,
                           double* scale, int threads) {
  int fail = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel reduction(| : fail)
#endif
  {
    spa_t s;
    if (spa_init(&s, n, scale != NULL)) fail = 1;
    else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
      for (int32_t i = r0; i < r1; i++) {
        if (!col) {
          cnt_or_off[i - r0] = row_accumulate(&s, i, upper_only, ap, ac, av, bp, bc, bv, 0);
          continue;
        }
        int32_t m = row_accumulate(&s, i, upper_only, ap, ac, av, bp, bc, bv, val != NULL);
        qsort(s.list, (size_t)m, sizeof(int32_t), cmp_i32);
        int64_t o = cnt_or_off[i - r0];
        for (int32_t t = 0; t < m; t++) {
          int32_t j = s.list[t];
          col[o + t] = j;
          if (val) val[o + t] = s.acc[j];
          if (scale) scale[o + t] = s.scl[j];
        }
      }
    }
    spa_free(&s);
  }
  return fail;
}

/* Full product for an op.  Outputs are malloc'd: *cp (n+1 int64), *cc, *cv,
 * *cs (scale; only if want_scale).  Free with oracle_free.  Returns 0 on
 * success. */
int oracle_spgemm(int32_t n, const int32_t* ap, const int32_t* ac, const double* av, int op,
                  int threads, int want_scale, int64_t** cp_out, int32_t** cc_out,
                  double** cv_out, double** cs_out) {
  int64_t nnz = ap[n];
  const int32_t* bp = ap; const int32_t* bc = ac; const double* bv = av;
  int32_t *tp = NULL, *tc = NULL; double* tv = NULL;
  if (op != OP_AA) {
    tp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    tc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    tv = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    if (!tp || !tc || !tv || oracle_transpose(n, ap, ac, av, tp, tc, tv)) return 1;
    bp = tp; bc = tc; bv = tv;
  }
  int upper = (op == OP_AATS);
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
  if (!off) return 1;
  if (oracle_gustavson_range(n, ap, ac, av, bp, bc, bv, upper, 0, n, off, NULL, NULL, NULL, threads))
    return 1;
  /* exclusive scan */
  int64_t run = 0;
  for (int32_t i = 0; i < n; i++) { int64_t c = off[i]; off[i] = run; run += c; }
  off[n] = run;
  int32_t* uc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(run > 0 ? run : 1));
  double* uv = (double*)malloc(sizeof(double) * (size_t)(run > 0 ? run : 1));
  double* us = want_scale ? (double*)malloc(sizeof(double) * (size_t)(run > 0 ? run : 1)) : NULL;
  if (!uc || !uv || (want_scale && !us)) return 1;
  if (oracle_gustavson_range(n, ap, ac, av, bp, bc, bv, upper, 0, n, off, uc, uv, us, threads))
    return 1;
  free(tp); free(tc); free(tv);
  if (!upper) { *cp_out = off; *cc_out = uc; *cv_out = uv; if (cs_out) *cs_out = us; else free(us); return 0; }

  /* Mirror the upper triangle (k_mirror, src/spgemm_engine.cuh:121-127):
   * (i,j) with i<j also yields (j,i) with the same value; diagonal once.
   * Row r of the full matrix: lower part (columns j<r from U rows j, in
   * ascending j) followed by U row r. */
  int64_t* low = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  if (!low) return 1;
  for (int32_t i = 0; i < n; i++)
    for (int64_t t = off[i]; t < off[i + 1]; t++) if (uc[t] > i) low[uc[t]]++;
  int64_t* fp = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
  if (!fp) return 1;
  fp[0] = 0;
  for (int32_t i = 0; i < n; i++) fp[i + 1] = fp[i] + low[i] + (off[i + 1] - off[i]);
  int64_t tot = fp[n];
  int32_t* fc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tot > 0 ? tot : 1));
  double* fv = (double*)malloc(sizeof(double) * (size_t)(tot > 0 ? tot : 1));
  double* fs = want_scale ? (double*)malloc(sizeof(double) * (size_t)(tot > 0 ? tot : 1)) : NULL;
  int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
  if (!fc || !fv || !cur || (want_scale && !fs)) return 1;
  for (int32_t i = 0; i < n; i++) cur[i] = fp[i];
  /* rows visited ascending, so each lower segment fills in ascending j */
  for (int32_t i = 0; i < n; i++)
    for (int64_t t = off[i]; t < off[i + 1]; t++) {
      int32_t j = uc[t];
      if (j > i) { int64_t d = cur[j]++; fc[d] = i; fv[d] = uv[t]; if (fs) fs[d] = us[t]; }
    }
  for (int32_t i = 0; i < n; i++) {
    int64_t d = fp[i] + low[i];
    for (int64_t t = off[i]; t < off[i + 1]; t++, d++) { fc[d] = uc[t]; fv[d] = uv[t]; if (fs) fs[d] = us[t]; }
  }
  free(low); free(cur); free(off); free(uc); free(uv); free(us);
  *cp_out = fp; *cc_out = fc; *cv_out = fv;
  if (cs_out) *cs_out = fs; else free(fs);
  return 0;
}

void oracle_free(void* p) { free(p); }

/* checksum(): s, as, sq sequentially over the values (src/spgemm_engine.cuh:116-118). */
void oracle_checksum(const double* v, int64_t m, double* out3) {
  double s = 0, as = 0, sq = 0;
  for (int64_t t = 0; t < m; t++) { double x = v[t]; s += x; as += x < 0 ? -x : x; sq += x * x; }
  out3[0] = s; out3[1] = as; out3[2] = sq;
}

/* pick_ours (bench/dispatch.cu:14-17). */
int oracle_pick_ours(double dcv, double fill, int64_t nnz) {
  if (dcv <= 0.63) return fill <= 53.23;
  return (nnz > 20164 && fill <= 47.17);
}

/* Descriptors (bench/dispatch.cu:26-37; src/fetch_and_prep.py:37-48).
 * out[0]=deg_mean out[1]=deg_cv out[2]=deg_max out[3]=fill_aa out[4]=fill_aat
 * out[5]=psym out[6]=products_aa out[7]=products_aat out[8]=bandwidth
 * Degrees are the stored row lengths (bench/dispatch.cu:27-30, two-pass
 * variance, ddof=0).  psym follows src/fetch_and_prep.py:39-46: P is the
 * numeric pattern after merging duplicates -- the positions whose stored
 * values, summed in storage order, are not zero (NaN counts as nonzero) --
 * and psym = |P & P^T| / max(1, |P|). */
int oracle_descriptors(int32_t n, const int32_t* ap, const int32_t* ac, const double* av, double* out) {
  int64_t nnz = ap[n];
  double dmean = n > 0 ? (double)nnz / n : 0.0, var = 0;
  int32_t dmax = 0;
  for (int32_t i = 0; i < n; i++) {
    int32_t d = ap[i + 1] - ap[i];
    double x = d - dmean; var += x * x;
    if (d > dmax) dmax = d;
  }
  double dcv = n > 0 ? sqrt(var / n) / (dmean > 1e-9 ? dmean : 1e-9) : 0.0;
  int64_t acc_aa = 0;
  for (int64_t p = 0; p < nnz; p++) { int32_t k = ac[p]; acc_aa += ap[k + 1] - ap[k]; }
  int64_t* cc = (int64_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t));
  if (!cc) return 1;
  for (int64_t p = 0; p < nnz; p++) cc[ac[p]]++;
  int64_t acc_aat = 0;
  for (int32_t k = 0; k < n; k++) acc_aat += cc[k] * cc[k];
  free(cc);
  /* P, row by row: merge each row's duplicates into a dense accumulator */
  int32_t* prp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  int32_t* pcol = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  double* accv = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (!prp || !pcol || !order || !accv || !mark) return 1;
  for (int32_t j = 0; j < n; j++) mark[j] = -1;
  int32_t bw = 0;
  int64_t np = 0;
  prp[0] = 0;
  for (int32_t i = 0; i < n; i++) {
    int64_t nd = 0;
    for (int32_t p = ap[i]; p < ap[i + 1]; p++) {
      int32_t j = ac[p];
      int32_t d = i > j ? i - j : j - i;
      if (d > bw) bw = d;
      if (mark[j] != i) { mark[j] = i; accv[j] = 0.0; order[nd++] = j; }
      accv[j] += av[p];
    }
    for (int64_t t = 0; t < nd; t++)
      if (accv[order[t]] != 0.0) pcol[np++] = order[t];
    prp[i + 1] = (int32_t)np;
  }
  /* |P & P^T|: row i of P against row i of P^T */
  int32_t* tp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  int32_t* tc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
  if (!tp || !tc) return 1;
  if (oracle_transpose(n, prp, pcol, NULL, tp, tc, NULL)) return 1;
  for (int32_t j = 0; j < n; j++) mark[j] = -1;
  int64_t inter = 0;
  for (int32_t i = 0; i < n; i++) {
    for (int32_t q = tp[i]; q < tp[i + 1]; q++) mark[tc[q]] = i;
    for (int32_t p = prp[i]; p < prp[i + 1]; p++)
      if (mark[pcol[p]] == i) inter++;
  }
  free(prp); free(pcol); free(order); free(accv); free(mark); free(tp); free(tc);
  int64_t distinct = np;
  out[0] = dmean; out[1] = dcv; out[2] = dmax;
  out[3] = nnz > 0 ? (double)acc_aa / nnz : 0.0;
  out[4] = nnz > 0 ? (double)acc_aat / nnz : 0.0;
  out[5] = (double)inter / (double)(distinct > 1 ? distinct : 1);
  out[6] = (double)acc_aa; out[7] = (double)acc_aat; out[8] = bw;
  return 0;
}
