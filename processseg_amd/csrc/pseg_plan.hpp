// Host-side view of a compiled pipeline plan, shared by the C ABI
// (pseg_host.cpp) and the serving loop (pseg_serve.cpp).
#pragma once
#include <string>

#include "../../include/pseg.h"
#include "pseg_internal.hpp"

struct pseg_plan {
  using DevPlan = pseg::DevPlan;
  static constexpr int kMaxSteps = pseg::kMaxSteps;

  int dtype;
  int threads;           // workgroup size of the fused kernel (pipeline_threads_for)
  int full;              // the kernel form its steps need (step_form, JobLaunch::full)
  double dt_ms;
  int32_t samples, record, margin;
  int n_steps;
  int n_parts;
  int part_a[kMaxSteps + 1], part_b[kMaxSteps + 1];
  uint64_t generation;
  DevPlan host;          // with device pointers
  DevPlan *dev;          // the device copy (start of the allocation)
  char *dev_base;
  size_t step_lds[kMaxSteps];
  int32_t step_kind[kMaxSteps];
  int32_t step_nfft[kMaxSteps], step_pad[kMaxSteps], step_margin[kMaxSteps],
      step_reflect[kMaxSteps];
  int32_t norm_stat[kMaxSteps];
  double norm_target[kMaxSteps], norm_pct[kMaxSteps];
};

namespace pseg {

// Sets this thread's last error (pseg_last_error) and returns st.
pseg_status set_error(pseg_status st, const std::string &msg);
// PSEG_E_HIP with "what: <HIP error string>".
pseg_status hip_error(hipError_t e, const char *what);

}  // namespace pseg
