// appt_rocsparse <tag> <bin.csr | gen:...> <AA|AAt|AAtS> [reps=20]
//
// Vendor-library context timing: the same product through rocSPARSE's
// generic SpGEMM, timed the way the reference times its vendor path
// (src/cusparse_engine.cuh:45-68): per call, the transpose for A*A^T
// (csr2csc), the buffer-size / nnz / compute stages and every allocation are
// inside the timed region; one checked call, two warmups, then `reps` timed
// calls.  Prints the reference's stdout lines ("  __chk ..." and
// "RESULT,tag,task,n,nnz,C_nnz,ms").  AAtS runs as A*A^T (the library has no
// upper-triangle variant).
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <vector>

#include "common.hpp"

#define HIPCK(x)                                                                            \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) {                                                                 \
      std::fprintf(stderr, "%s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      std::exit(1);                                                                         \
    }                                                                                       \
  } while (0)
#define SPCK(x)                                                                             \
  do {                                                                                      \
    rocsparse_status s_ = (x);                                                              \
    if (s_ != rocsparse_status_success) {                                                   \
      std::fprintf(stderr, "%s failed: rocsparse status %d (%s:%d)\n", #x, (int)s_, __FILE__, __LINE__); \
      std::exit(1);                                                                         \
    }                                                                                       \
  } while (0)

struct Dev {
  int n;
  int64_t nnz;
  int* rp;
  int* col;
  double* val;
};

// B = A^T through csr2csc (all buffers allocated here; caller frees)
static Dev transpose(rocsparse_handle h, const Dev& A) {
  Dev T{A.n, A.nnz, nullptr, nullptr, nullptr};
  HIPCK(hipMalloc(&T.rp, sizeof(int) * (A.n + 1)));
  HIPCK(hipMalloc(&T.col, sizeof(int) * (A.nnz > 0 ? A.nnz : 1)));
  HIPCK(hipMalloc(&T.val, sizeof(double) * (A.nnz > 0 ? A.nnz : 1)));
  size_t bs = 0;
  SPCK(rocsparse_csr2csc_buffer_size(h, A.n, A.n, (rocsparse_int)A.nnz, A.rp, A.col, rocsparse_action_numeric, &bs));
  void* buf = nullptr;
  HIPCK(hipMalloc(&buf, bs > 0 ? bs : 8));
  SPCK(rocsparse_dcsr2csc(h, A.n, A.n, (rocsparse_int)A.nnz, A.val, A.rp, A.col, T.val, T.col, T.rp,
                          rocsparse_action_numeric, rocsparse_index_base_zero, buf));
  HIPCK(hipFree(buf));
  return T;
}

// one full C = A*B; returns nnz(C); copies C's values to the host when asked
static int64_t spgemm(rocsparse_handle h, const Dev& A, const Dev& B, std::vector<double>* out) {
  rocsparse_spmat_descr mA, mB, mC, mD;
  SPCK(rocsparse_create_csr_descr(&mA, A.n, A.n, A.nnz, A.rp, A.col, A.val, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, rocsparse_datatype_f64_r));
  SPCK(rocsparse_create_csr_descr(&mB, B.n, B.n, B.nnz, B.rp, B.col, B.val, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, rocsparse_datatype_f64_r));
  int* crp = nullptr;
  HIPCK(hipMalloc(&crp, sizeof(int) * (A.n + 1)));
  SPCK(rocsparse_create_csr_descr(&mC, A.n, B.n, 0, crp, nullptr, nullptr, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, rocsparse_datatype_f64_r));
  SPCK(rocsparse_create_csr_descr(&mD, A.n, B.n, 0, crp, nullptr, nullptr, rocsparse_indextype_i32,
                                  rocsparse_indextype_i32, rocsparse_index_base_zero, rocsparse_datatype_f64_r));
  const double alpha = 1.0;
  const rocsparse_operation op = rocsparse_operation_none;
  size_t bs = 0;
  SPCK(rocsparse_spgemm(h, op, op, &alpha, mA, mB, nullptr, mD, mC, rocsparse_datatype_f64_r,
                        rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_buffer_size, &bs, nullptr));
  void* buf = nullptr;
  HIPCK(hipMalloc(&buf, bs > 0 ? bs : 8));
  SPCK(rocsparse_spgemm(h, op, op, &alpha, mA, mB, nullptr, mD, mC, rocsparse_datatype_f64_r,
                        rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_nnz, &bs, buf));
  int64_t rows, cols, cnnz;
  SPCK(rocsparse_spmat_get_size(mC, &rows, &cols, &cnnz));
  int* ccol = nullptr;
  double* cval = nullptr;
  HIPCK(hipMalloc(&ccol, sizeof(int) * (cnnz > 0 ? cnnz : 1)));
  HIPCK(hipMalloc(&cval, sizeof(double) * (cnnz > 0 ? cnnz : 1)));
  SPCK(rocsparse_csr_set_pointers(mC, crp, ccol, cval));
  SPCK(rocsparse_spgemm(h, op, op, &alpha, mA, mB, nullptr, mD, mC, rocsparse_datatype_f64_r,
                        rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_compute, &bs, buf));
  if (out) {
    out->resize((size_t)cnnz);
    if (cnnz) HIPCK(hipMemcpy(out->data(), cval, sizeof(double) * cnnz, hipMemcpyDeviceToHost));
  }
  rocsparse_destroy_spmat_descr(mA);
  rocsparse_destroy_spmat_descr(mB);
  rocsparse_destroy_spmat_descr(mC);
  rocsparse_destroy_spmat_descr(mD);
  HIPCK(hipFree(buf));
  HIPCK(hipFree(crp));
  HIPCK(hipFree(ccol));
  HIPCK(hipFree(cval));
  return cnnz;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <tag> <bin.csr|gen:...> <AA|AAt|AAtS> [reps]\n", argv[0]);
    return 1;
  }
  const char* tag = argv[1];
  const char* task = argv[3];
  const int reps = argc > 4 ? std::atoi(argv[4]) : 20;
  appt_op opk;
  CLI_CHECK(appt_op_from_task(task, &opk));
  const bool aat = opk != APPT_OP_AA;
  appt_host_csr hm{};
  CLI_CHECK(load_matrix(argv[2], &hm));
  Dev A{hm.n, hm.nnz, nullptr, nullptr, nullptr};
  HIPCK(hipMalloc(&A.rp, sizeof(int) * (hm.n + 1)));
  HIPCK(hipMalloc(&A.col, sizeof(int) * (hm.nnz > 0 ? hm.nnz : 1)));
  HIPCK(hipMalloc(&A.val, sizeof(double) * (hm.nnz > 0 ? hm.nnz : 1)));
  HIPCK(hipMemcpy(A.rp, hm.rowptr, sizeof(int) * (hm.n + 1), hipMemcpyHostToDevice));
  if (hm.nnz) {
    HIPCK(hipMemcpy(A.col, hm.col, sizeof(int) * hm.nnz, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(A.val, hm.val, sizeof(double) * hm.nnz, hipMemcpyHostToDevice));
  }
  rocsparse_handle h;
  SPCK(rocsparse_create_handle(&h));
  auto once = [&](std::vector<double>* out) {
    if (!aat) return spgemm(h, A, A, out);
    Dev T = transpose(h, A);
    const int64_t r = spgemm(h, A, T, out);
    HIPCK(hipFree(T.rp));
    HIPCK(hipFree(T.col));
    HIPCK(hipFree(T.val));
    return r;
  };
  std::vector<double> C;
  const int64_t cnnz = once(&C);
  double s = 0, as = 0, sq = 0;
  for (double x : C) { s += x; as += x < 0 ? -x : x; sq += x * x; }
  for (int w = 0; w < 2; w++) once(nullptr);
  HIPCK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  HIPCK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++) once(nullptr);
  HIPCK(hipEventRecord(e1, nullptr));
  HIPCK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCK(hipEventElapsedTime(&ms, e0, e1));
  ms /= (float)(reps > 0 ? reps : 1);
  std::printf("  __chk s=%.6e as=%.6e sq=%.6e\n", s, as, sq);
  std::printf("RESULT,%s,%s,%d,%lld,%lld,%.4f\n", tag, task, hm.n, (long long)hm.nnz, (long long)cnnz, ms);
  rocsparse_destroy_handle(h);
  hipFree(A.rp);
  hipFree(A.col);
  hipFree(A.val);
  appt_host_csr_free(&hm);
  return 0;
}
