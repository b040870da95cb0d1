// appt_dispatch <tag> <bin.csr | gen:...> <AA|AAt> [reps=20]
//
// The reference's online dispatcher (bench/dispatch.cu:19-53) computes
// descriptors on the host, applies pick_ours, and runs either its engine or
// cuSPARSE.  Here the descriptors are computed on the GPU and the choice is
// among the engine's internal row bins (there is no vendor path); the
// reference rule's verdict is still reported as rule=ours|cusparse.
//   DISPATCH,tag,task,ours,dcv=..,fill=..,C_nnz=..,desc_ms=..,run_ms=..,total_ms=..,rule=..,bins=..
#include <chrono>

#include "common.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <tag> <bin.csr|gen:...> <AA|AAt> [reps]\n", argv[0]);
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
  appt_descriptors d{};
  CLI_CHECK(appt_dispatch_descriptors(eng, &A, op, &d));  // warm
  auto t0 = std::chrono::high_resolution_clock::now();
  CLI_CHECK(appt_dispatch_descriptors(eng, &A, op, &d));
  auto t1 = std::chrono::high_resolution_clock::now();
  const double desc_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  appt_eng_result r{};
  CLI_CHECK(appt_engine_run_timed(eng, task, &A, reps, &r));
  const double fill = op == APPT_OP_AA ? d.fill_aa : d.fill_aat;
  appt_phase_times pt{};
  CLI_CHECK(appt_engine_phase_times(eng, &pt));
  std::printf("DISPATCH,%s,%s,ours,dcv=%.3f,fill=%.1f,C_nnz=%lld,desc_ms=%.4f,run_ms=%.4f,total_ms=%.4f,rule=%s,bins=%d\n",
              tag, task, d.deg_cv, fill, (long long)r.nnz, desc_ms, r.ms, desc_ms + r.ms,
              d.pick_ours ? "ours" : "cusparse", pt.n_bins_used);
  appt_device_free_view(eng, &A);
  appt_engine_destroy(eng);
  appt_host_csr_free(&h);
  return 0;
}
