/*
 * appt_spgemm.h -- C ABI of the MI355X (gfx950) SpGEMM engine.
 *
 * Computes C = A*A, C = A*A^T, or the symmetric C = A*A^T of the reference's
 * Gram task (full rows whose values are bit-symmetric by default; optionally
 * only one triangle's products, mirrored), for square CSR matrices with int32
 * indices and fp64 values,
 * using structural semantics: an entry whose contributions cancel to 0.0 is
 * kept, columns of every output row are strictly ascending.
 *
 * Plain pointers and sizes only; every function returns an appt_status and
 * never exits the process.  All "device" pointers are HIP device memory on the
 * engine's device.  One engine per host thread (thread-compatible, not
 * thread-safe); engines on different threads are independent.
 *
 * Reference interface each entry point replaces is cited as file:line of
 * Wendy-Hamlet/appt_spgemm.
 */
#ifndef APPT_SPGEMM_H
#define APPT_SPGEMM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference prints and exit(1)s on every error
 * (src/spgemm_engine.cuh:18 CK macro); this ABI returns them instead. */
typedef enum appt_status {
  APPT_OK = 0,
  APPT_ERR_INVALID = 1,      /* bad argument / malformed CSR / bad task string */
  APPT_ERR_HIP = 2,          /* a HIP runtime call failed */
  APPT_ERR_OOM = 3,          /* device allocation failed or budget exceeded */
  APPT_ERR_ADMISSION = 4,    /* job rejected by the HBM budget admission */
  APPT_ERR_IO = 5,           /* CSR1 file could not be read/written */
  APPT_ERR_UNSUPPORTED = 6,  /* shape outside what the kernels support */
  APPT_ERR_OVERFLOW = 7,     /* a size does not fit the requested index width */
  APPT_ERR_NO_DEVICE = 8     /* no HIP device visible */
} appt_status;

/* Operation.  Task strings "AA" | "AAt" | "AAtS" of src/spgemm_engine.cuh:134. */
typedef enum appt_op {
  APPT_OP_AA = 0,    /* C = A*A */
  APPT_OP_AAT = 1,   /* C = A*A^T, B = A^T materialised */
  APPT_OP_AATS = 2   /* C = A*A^T, symmetric: full rows, or (tuning.gram) one triangle's products, mirrored */
} appt_op;

/* Device CSR operand (src/spgemm_engine.cuh:132 dOff/dCol/dVal, n, nnz).
 * Rows need not be sorted; duplicate columns inside a row are merged by
 * summation.  The engine only reads these arrays. */
typedef struct appt_csr_view {
  int32_t n;                /* rows == columns */
  int64_t nnz;              /* < 2^31 */
  const int32_t* rowptr;    /* device, n+1 */
  const int32_t* col;       /* device, nnz */
  const double* val;        /* device, nnz */
} appt_csr_view;

/* Device CSR result.  rowptr is int32 when nnz < 2^31, else int64
 * (rowptr_is_64 = 1).  The buffers belong to the engine until
 * appt_csr_release(). */
typedef struct appt_csr_result {
  int32_t n;
  int64_t nnz;
  int32_t rowptr_is_64;
  void* rowptr;             /* device, n+1 entries of int32 or int64 */
  int32_t* col;             /* device, nnz */
  double* val;              /* device, nnz */
  void* _priv;              /* engine bookkeeping; do not touch */
} appt_csr_result;

/* struct EngResult { float ms; int64_t nnz; double s, as, sq; }
 * (src/spgemm_engine.cuh:129), same field order and types. */
typedef struct appt_eng_result {
  float ms;
  int64_t nnz;
  double s, as, sq;
} appt_eng_result;

/* Bin thresholds and launch shapes.  Zero-filled means "built-in defaults". */
#define APPT_MAX_CLASSES 8
typedef struct appt_tuning {
  int32_t f_class_max[APPT_MAX_CLASSES]; /* products upper bound of LDS classes, ascending, 0-terminated */
  int32_t wg_lds_bytes;       /* LDS cap of the workgroup-per-row bins; rows needing more take the
                                 windowed heavy-row bin (0 = the device's LDS per block) */
  int32_t lds_budget_bytes;   /* LDS bytes a workgroup may take */
  int32_t bitmap_words_factor;/* numeric dense window when words(span) <= factor * distinct words */
  int32_t group_for_len[4];   /* lanes per row for mean B-row length <= 8,16,32,inf (0 = default) */
  int32_t heavy_window_cols;  /* fixed column window of the heavy-row bin (0 = shaped per row within the LDS) */
  int32_t sort_span_factor;   /* numeric pass: rows of at most 512 products whose columns are
                                 scattered (fewer than 4 columns per 3 distinct 32-column words) take
                                 the register-sort bin when their output span exceeds factor * products
                                 words, even where a window fits (0 = default, < 0 = only rows no
                                 window fits) */
  int32_t sort_span_factor_symbolic; /* symbolic pass: the same by span alone (0 = default, < 0 = off) */
  int32_t transpose_radix_min_nnz; /* asymmetric A^T of at least this many entries by radix sort (0 = default
                                      262,144; < 0 = never) */
  int32_t mirror_radix_min_nnz;    /* AAtS mirror by radix sort when nnz(C) reaches this (0 = default 1,048,576;
                                      < 0 = never) */
  int32_t class_lds_bytes[APPT_MAX_CLASSES]; /* LDS bytes one row of products class c may use in the
                                      wave-level bins (0 = default 16 KB, 24 KB, 32 KB) */
  int32_t sort_max_products;  /* largest register-sort bin (products and A entries of a row, rounded up to
                                 a power of two in [16, 2048]; 0 = 2048; < 0 = no register-sort bins) */
  int32_t rowptr64_min_nnz;   /* results with at least this many entries get int64 row pointers
                                 (0 = only when nnz(C) >= 2^31, where int32 cannot hold them) */
  int32_t row_templates;      /* rows whose own relative pattern and whose B rows' patterns repeat
                                 share a template: no symbolic pass, products straight to their slots.
                                 Detected for operands of at least this many entries (0 = 262,144;
                                 < 0 = never) */
  int32_t gram;               /* AAtS: 1 / 2 = Gram over the upper / lower triangle (only j >= i /
                                 j <= i products are formed, then mirrored); 0 or -1 = full rows
                                 (the default: the mirror's transposition outweighs the saved
                                 products on the measured inputs) */
} appt_tuning;

typedef struct appt_options {
  int32_t device;             /* HIP device ordinal */
  void* stream;               /* hipStream_t to run on; NULL = engine creates its own default-priority stream */
  int64_t hbm_budget_bytes;   /* cap on engine-allocated device memory; 0 = 90% of free memory at create */
  appt_tuning tuning;
} appt_options;

typedef struct appt_engine appt_engine;

/* Per-phase device times of the last compute call, in milliseconds
 * (measured with HIP events when timing is enabled). */
typedef struct appt_phase_times {
  float transpose_ms, prepass_ms, symbolic_ms, scan_ms, numeric_ms, mirror_ms, total_ms;
  int64_t products;            /* sum over A's nonzeros of nnz(B row k): 2*products = FLOPs */
  int64_t nnz_c;
  int32_t n_bins_used;
  int32_t sync_points;         /* device->host synchronisations in the call */
} appt_phase_times;

/* Structural descriptors of src/fetch_and_prep.py:37-48 and bench/dispatch.cu:26-37
 * plus products and bandwidth.  Degrees are the stored row lengths (as
 * bench/dispatch.cu:27-30 reads them); psym is taken over the numeric pattern
 * after merging duplicates (src/fetch_and_prep.py:39-46): positions whose
 * stored values, summed in storage order, are not zero. */
typedef struct appt_descriptors {
  int32_t n;
  int64_t nnz;
  double deg_mean, deg_cv;     /* row-degree mean and std/mean (ddof=0, variance exact) */
  int32_t deg_max;
  double fill_aa;              /* sum_p deg(col_p) / nnz       (bench/dispatch.cu:31-33) */
  double fill_aat;             /* sum_k colcount_k^2 / nnz     (bench/dispatch.cu:34-36) */
  double psym;                 /* nnz(P & P^T) / nnz(P)        (src/fetch_and_prep.py:42-46) */
  int64_t products_aa, products_aat;
  int32_t max_row_products_aa; /* clamped to int32 range */
  int32_t bandwidth;           /* max |i - j| over stored entries */
  int32_t pick_ours;           /* decision rule of bench/dispatch.cu:14-17 on (deg_cv, fill, nnz) */
} appt_descriptors;

/* ---- engine lifetime ------------------------------------------------------ */
appt_status appt_engine_create(const appt_options* opt, appt_engine** out);
appt_status appt_engine_destroy(appt_engine* eng);
/* Enable per-phase HIP-event timing (adds event records, no extra syncs). */
appt_status appt_engine_set_timing(appt_engine* eng, int enable);
appt_status appt_engine_phase_times(const appt_engine* eng, appt_phase_times* out);

/* Row bins of the last compute call: how many output rows each symbolic and
 * numeric bin received (the engine's internal dispatch, the role of the
 * reference dispatcher's ours/cuSPARSE verdict, bench/dispatch.cu:39-47, at
 * row granularity).  Bin names come from appt_bin_name. */
#define APPT_MAX_BINS 64
typedef struct appt_bin_report {
  int32_t n_bins;                       /* bins defined: ids 0 .. n_bins - 1 */
  int32_t sym_rows[APPT_MAX_BINS];      /* rows per symbolic bin */
  int32_t num_rows[APPT_MAX_BINS];      /* rows per numeric bin */
  int32_t sym_metric_max[APPT_MAX_BINS];/* largest products count in the bin */
  int32_t num_metric_max[APPT_MAX_BINS];/* largest output row in the bin */
} appt_bin_report;
appt_status appt_engine_bin_report(const appt_engine* eng, appt_bin_report* out);
/* Static name of a bin id, e.g. "window/c1/G16", "hash/c0/G8", "list/c0/G32",
 * "sort/N256", "wg-window", "wg-hash", "heavy", "template"; "empty" for 0, NULL past the end. */
const char* appt_bin_name(int32_t bin);
/* Release cached device memory held by the engine's pool. */
appt_status appt_engine_trim(appt_engine* eng);
/* Bytes currently allocated from HIP by the engine (pool + live results). */
int64_t appt_engine_device_bytes(const appt_engine* eng);
/* Last error text of this engine (or of the thread's last engine-less call). */
const char* appt_last_error(void);

/* ---- the hot path ---------------------------------------------------------- */
/* C = op(A).  Replaces run_ours()'s pipeline (src/spgemm_engine.cuh:132-210)
 * and, unlike it, returns C (the reference frees C, :203-209).  Synchronous
 * with respect to the host on return (C is complete). */
appt_status appt_spgemm_compute(appt_engine* eng, const appt_csr_view* A, appt_op op,
                                appt_csr_result* C);
appt_status appt_csr_release(appt_engine* eng, appt_csr_result* C);

/* C = op(A) with no device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  The caller passes the output buffers: rowptr (n+1
 * int32), col and val of `capacity` entries, all device memory on the
 * engine's device.  Every scratch buffer comes from one static device array
 * of 64 MiB per device (taken when an engine is created there), so n and nnz
 * are limited to what fits in it: about 606 bytes per row plus 44 per entry
 * plus 3 MiB; larger inputs are refused with APPT_ERR_UNSUPPORTED before any
 * work.  *nnz_out receives nnz(C); when it exceeds capacity (or needs 64-bit
 * row pointers) the call returns APPT_ERR_OVERFLOW with col/val untouched, so
 * the caller can size the buffers and call again.  Same kernels and results
 * as appt_spgemm_compute.  The array is shared by every call on one device
 * and keeps the last call's contents: calls on one device that overlap in
 * time (from several host threads) race; calls from one thread are
 * synchronous and never overlap. */
appt_status appt_spgemm_compute_noalloc(appt_engine* eng, const appt_csr_view* A, appt_op op, int32_t* rowptr,
                                        int32_t* col, double* val, int64_t capacity, int64_t* nnz_out);

/* Numeric semantics on request: drop the entries of an engine-owned result
 * whose value is exactly zero (scipy's eliminate_zeros; the default is the
 * reference's structural semantics, report/CORRECTNESS_NOTES.md:3-16).  C is
 * replaced by the compacted result. */
appt_status appt_csr_eliminate_zeros(appt_engine* eng, appt_csr_result* C);

/* Row panel: rows [row_begin, row_end) of op(A) as a CSR with row_end -
 * row_begin rows and global column indices.  A panel of AATS holds complete
 * rows of the symmetric product, computed as A*A^T's rows (no triangle to
 * mirror from when only some rows are formed).  For products whose C
 * does not fit the device at once; no counterpart in the reference, whose
 * int32 arena overflows first (src/spgemm_engine.cuh:156-162). */
appt_status appt_spgemm_compute_rows(appt_engine* eng, const appt_csr_view* A, appt_op op, int32_t row_begin,
                                     int32_t row_end, appt_csr_result* C);
/* Exact row pointers of op(A) (n+1 int64, host memory) from the symbolic
 * phase alone: sizes row panels before any numeric work.  The engine keeps
 * the rows' symbolic results (8 bytes per row) until its next call of this
 * function: appt_spgemm_compute_rows on the same view (rowptr, col, val, n,
 * nnz) and op takes its heavy rows' counts from them instead of recomputing
 * them, so A's contents must not change in between, nor may another matrix
 * take over the same device pointers (plan again first).  appt_spgemm_compute
 * never uses the plan. */
appt_status appt_spgemm_symbolic_rowptr(appt_engine* eng, const appt_csr_view* A, appt_op op, int64_t* rowptr);

/* run_ours(task, dOff, dCol, dVal, n, nnz, reps) (src/spgemm_engine.cuh:132-133):
 * one checked call (checksums from the sorted values on the host, :194-196),
 * then `reps` timed calls; ms is the mean per call over the timed loop
 * (:199-202).  Timed scope: transpose, pre-pass, symbolic, scans, numeric,
 * mirror. */
appt_status appt_spgemm_run_timed(const char* task, const int32_t* dOff, const int32_t* dCol,
                                  const double* dVal, int32_t n, int64_t nnz, int32_t reps,
                                  appt_eng_result* out);
/* Same on a caller-owned engine (keeps its pool and tuning). */
appt_status appt_engine_run_timed(appt_engine* eng, const char* task, const appt_csr_view* A,
                                  int32_t reps, appt_eng_result* out);

/* Descriptors and the decision rule (bench/dispatch.cu:26-38), computed on
 * the GPU from A as it lies (nothing staged), with one host synchronisation. */
appt_status appt_dispatch_descriptors(appt_engine* eng, const appt_csr_view* A, appt_op op,
                                      appt_descriptors* out);

/* Parse "AA" | "AAt" | "AAtS" (src/spgemm_engine.cuh:134). */
appt_status appt_op_from_task(const char* task, appt_op* out);

/* ---- batched path --------------------------------------------------------------
 * Many independent matrices (each C_k = A_k*A_k or A_k*A_k^T) run as one
 * block-diagonal product: one pipeline over all packed rows, planned once (two
 * host synchronisations), then replayed as a captured HIP graph with no host
 * synchronisation inside.  The structure of the inputs is fixed at plan time;
 * values may change between executions.  No counterpart in the reference,
 * which runs one matrix per call (src/spgemm_engine.cuh:132). */
typedef struct appt_batch appt_batch;

/* Packed device input of a batch, written by the caller (or appt_batch_gather):
 * matrix k owns rows [row0_k, row0_k + n_k) with row0 the prefix sum of n_each,
 * rowptr holds global entry offsets (n_total + 1 entries), col holds indices
 * local to each matrix (0 .. n_k-1), val the values. */
typedef struct appt_batch_buffers {
  int32_t n_total;
  int64_t nnz;
  int32_t* rowptr;
  int32_t* col;
  double* val;
} appt_batch_buffers;

typedef struct appt_batch_info {
  int32_t count;
  int32_t n_total;
  int64_t nnz_a;             /* packed input entries */
  int64_t nnz_c;             /* sum of the results' nnz */
  int64_t products;          /* sum of the matrices' product counts (2*products = FLOPs) */
  int32_t graph;             /* 1 when execute replays a captured HIP graph */
  int32_t bins_used;
  appt_tuning tuning;        /* bin configuration in use */
} appt_batch_info;

/* ops[k] is APPT_OP_AA or APPT_OP_AAT; sum of nnz_each < 2^31, sum of n_each + count < 2^31. */
appt_status appt_batch_create(appt_engine* eng, int32_t count, const int32_t* n_each, const int64_t* nnz_each,
                              const appt_op* ops, appt_batch** out);
appt_status appt_batch_destroy(appt_batch* b);
appt_status appt_batch_inputs(appt_batch* b, appt_batch_buffers* out);
/* Copy a packed host batch (same layout, n = n_total, nnz = sum of nnz_each) into the packed input. */
struct appt_host_csr;
appt_status appt_batch_upload(appt_batch* b, const struct appt_host_csr* packed);
/* Copy `count` separate device matrices (each rowptr starting at 0) into the packed input. */
appt_status appt_batch_gather(appt_batch* b, const appt_csr_view* each);
/* Planning run (sizes, C buffers); use_graph != 0 then captures the replay graph. */
appt_status appt_batch_plan(appt_batch* b, int32_t use_graph);
/* Time candidate bin configurations (reps executions each) on this batch,
 * keep the fastest and re-plan with it. */
appt_status appt_batch_autotune(appt_batch* b, int32_t reps, appt_tuning* chosen);
/* Run the planned batch; synchronous on return. */
appt_status appt_batch_execute(appt_batch* b);
/* View of C_k inside the batch's buffers (int32 rowptr, local columns); valid
 * until the batch is re-planned or destroyed; not to be passed to appt_csr_release. */
appt_status appt_batch_result(const appt_batch* b, int32_t k, appt_csr_result* out);
appt_status appt_batch_info_get(const appt_batch* b, appt_batch_info* out);

/* ---- serving -------------------------------------------------------------------
 * A request loop over device-resident jobs (C = A*A or A*A^T).  A job is
 * admitted by memory: once its symbolic phase has fixed nnz(C), the exact
 * result size plus a working-set estimate must fit the HBM budget beside what
 * admitted jobs hold; otherwise it waits for memory to free up, or is refused
 * (APPT_ERR_ADMISSION) when it could never fit.  Up to `slots` jobs run at
 * once, each on its own engine and default-priority stream, so their kernels
 * overlap on the device; small jobs waiting together are run as one
 * block-diagonal batch.  No counterpart in the reference. */
typedef struct appt_server appt_server;

typedef struct appt_server_options {
  int32_t device;
  int32_t slots;              /* jobs (or job groups) in flight at once; 0 = 4 */
  int64_t hbm_budget_bytes;   /* admission budget; 0 = 200 GB */
  int32_t batch_max_jobs;     /* most small jobs run as one group; 0 = 64, 1 = no grouping */
  int32_t batch_row_limit;    /* a job is small when n <= this; 0 = 65536 */
  int32_t admission_wait_ms;  /* longest wait for memory before refusing; 0 = 10000 */
  appt_tuning tuning;
} appt_server_options;

typedef enum appt_job_state {
  APPT_JOB_QUEUED = 0,
  APPT_JOB_RUNNING = 1,
  APPT_JOB_DONE = 2,          /* finished; status says whether it succeeded */
  APPT_JOB_RELEASED = 3
} appt_job_state;

typedef struct appt_job_info {
  int32_t state;              /* appt_job_state */
  int32_t status;             /* appt_status of a finished job */
  int32_t group;              /* jobs it ran with (1 = alone) */
  double queue_ms, run_ms, latency_ms;  /* submit->start, start->done, submit->done */
  int64_t products;
  int64_t reserved_bytes;     /* admission bytes held by the result */
  appt_csr_result C;          /* device result of a successful job until appt_server_release */
} appt_job_info;

typedef struct appt_server_stats {
  int64_t submitted, completed, refused, failed;
  double p50_ms, p99_ms, mean_ms;  /* latency of completed jobs */
  double gflops;                   /* 2 * sum(products) of completed jobs / (last done - first submit) */
  double jobs_per_s;
  /* budget bytes charged now and at peak: every byte the slots' pools hold
   * (cached working sets, free blocks, live results) plus what running jobs
   * were admitted for and have not allocated yet */
  int64_t bytes_reserved, peak_bytes_reserved;
  int64_t groups_run;              /* launches that ran more than one job */
  int64_t device_bytes, peak_device_bytes;  /* device memory the slots' pools hold, now and at peak */
  int64_t overruns;                /* runs that allocated more than they were admitted for (0 expected) */
} appt_server_stats;

appt_status appt_server_create(const appt_server_options* opt, appt_server** out);
/* Finishes running jobs, fails queued ones, frees every result. */
appt_status appt_server_destroy(appt_server* s);
/* Queue C = op(A) (op AA or AAT); A is read until the job finishes. */
appt_status appt_server_submit(appt_server* s, const appt_csr_view* A, appt_op op, uint64_t* job);
appt_status appt_server_poll(appt_server* s, uint64_t job, appt_job_info* out);
appt_status appt_server_wait(appt_server* s, uint64_t job, appt_job_info* out);
/* Free a finished job's result and its admission bytes. */
appt_status appt_server_release(appt_server* s, uint64_t job);
appt_status appt_server_stats_get(appt_server* s, appt_server_stats* out);

/* ---- host CSR, CSR1 files and synthetic generators --------------------------- */
/* Host CSR owned by the library (malloc'd); free with appt_host_csr_free. */
typedef struct appt_host_csr {
  int32_t n;
  int64_t nnz;
  int32_t* rowptr;   /* n+1 */
  int32_t* col;      /* nnz */
  double* val;       /* nnz */
} appt_host_csr;

void appt_host_csr_free(appt_host_csr* m);

/* CSR1 layout of src/export_bin.py:6-8 (reader src/spgemm_engine.cuh:22-32):
 * "CSR1", int32 n, int32 pad, int64 nnz, int32 indptr[n+1], int32 indices[nnz],
 * f64 data[nnz], little-endian. */
appt_status appt_csr1_read(const char* path, appt_host_csr* out);
appt_status appt_csr1_write(const char* path, const appt_host_csr* m);

/* 2D 5-point Laplacian on a g x g grid, 4 on the diagonal, -1 off it. */
appt_status appt_gen_laplace2d(int32_t g, appt_host_csr* out);
/* 27-point stencil on a g^3 grid, index x + g*y + g*g*z, values U(-1,1) from
 * seed; drop_p > 0 drops each off-diagonal independently with probability
 * drop_p (second random stream of the same seed). */
appt_status appt_gen_stencil27(int32_t g, double drop_p, uint64_t seed, appt_host_csr* out);
/* R-MAT: 2^scale rows, edge_factor * 2^scale sampled edges with quadrant
 * probabilities (a,b,c, d = 1-a-b-c), duplicates merged, values U(0,1). */
appt_status appt_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                          uint64_t seed, appt_host_csr* out);
/* Random banded: half-bandwidth bw, each in-band off-diagonal kept with
 * probability fill, diagonal always kept, values U(-1,1). */
appt_status appt_gen_banded(int32_t n, int32_t bw, double fill, uint64_t seed, appt_host_csr* out);
/* Erdos-Renyi: round(mean_deg * n) uniform (i,j) pairs, duplicates merged,
 * values U(-1,1). */
appt_status appt_gen_erdos_renyi(int32_t n, double mean_deg, uint64_t seed, appt_host_csr* out);

/* ---- host <-> device helpers (for callers without their own allocator) ------- */
/* Copies a host CSR to new device buffers; APPT_ERR_INVALID (nothing
 * allocated) unless rowptr[0] = 0, rowptr[n] = nnz, row pointers are
 * non-decreasing and every column is in [0, n). */
appt_status appt_device_upload(appt_engine* eng, const appt_host_csr* h, appt_csr_view* out_dev);
/* The same structure check for a caller's device CSR, on the GPU (one pass
 * over rowptr and col, one synchronisation): APPT_OK, or APPT_ERR_INVALID
 * with the first failed condition in appt_last_error().  The compute calls
 * take their operands as valid, as the reference does (an out-of-range
 * column is undefined behaviour there); this is the check for untrusted
 * input.  No reference counterpart. */
appt_status appt_csr_check(appt_engine* eng, const appt_csr_view* A);
appt_status appt_device_free_view(appt_engine* eng, appt_csr_view* v);
/* Rows [row_begin, row_end) of a device result to host arrays: rowptr gets
 * row_end - row_begin + 1 int64 offsets relative to the first row, col / val
 * that range's entries (sized by the caller from the row pointers). */
appt_status appt_csr_download_rows(appt_engine* eng, const appt_csr_result* C, int32_t row_begin, int32_t row_end,
                                   int64_t* rowptr, int32_t* col, double* val);
/* Digests of a device result computed on the GPU (one pass, one
 * synchronisation): out[0] hashes the structure (row pointers as int64 with
 * their row index, columns with their position), out[1] the value bits with
 * their position; each term is a 64-bit mix, summed modulo 2^64.  Equal
 * results give equal digests; compares results too large to download.  No
 * reference counterpart (the reference compares host checksums,
 * src/spgemm_engine.cuh:116-118). */
appt_status appt_csr_digest(appt_engine* eng, const appt_csr_result* C, uint64_t* out);
/* Copy a device result to caller-provided host arrays (rowptr as int64);
 * eng may be NULL (the calling thread's current device is used). */
appt_status appt_csr_download(appt_engine* eng, const appt_csr_result* C, int64_t* rowptr,
                              int32_t* col, double* val);

#ifdef __cplusplus
}
#endif
#endif /* APPT_SPGEMM_H */
