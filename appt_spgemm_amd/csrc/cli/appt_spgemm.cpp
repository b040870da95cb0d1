// appt_spgemm <tag> <bin.csr | gen:...> <AA|AAt|AAtS> [reps=20]
//
// Same contract and stdout lines as the reference CLI (src/spgemm.cu:128-226):
//   "  __chk s=%.6e as=%.6e sq=%.6e"   checksums of the sorted values
//   "RESULT,tag,task,n,nnz,C_nnz,ms"    mean ms per call over reps
#include "common.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <tag> <bin.csr|gen:...> <AA|AAt|AAtS> [reps]\n", argv[0]);
    return 1;
  }
  const char* tag = argv[1];
  const char* task = argv[3];
  const int reps = argc > 4 ? std::atoi(argv[4]) : 20;
  appt_op op;
  CLI_CHECK(appt_op_from_task(task, &op));
  appt_host_csr h{};
  CLI_CHECK(load_matrix(argv[2], &h));
  appt_engine* eng = nullptr;
  appt_options o{};
  CLI_CHECK(appt_engine_create(&o, &eng));
  appt_csr_view A{};
  CLI_CHECK(appt_device_upload(eng, &h, &A));
  appt_eng_result r{};
  CLI_CHECK(appt_engine_run_timed(eng, task, &A, reps, &r));
  std::printf("  __chk s=%.6e as=%.6e sq=%.6e\n", r.s, r.as, r.sq);
  std::printf("RESULT,%s,%s,%d,%lld,%lld,%.4f\n", tag, task, h.n, (long long)h.nnz, (long long)r.nnz, r.ms);
  appt_device_free_view(eng, &A);
  appt_engine_destroy(eng);
  appt_host_csr_free(&h);
  return 0;
}
