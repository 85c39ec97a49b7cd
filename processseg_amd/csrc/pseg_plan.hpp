// Host-side view of a compiled pipeline plan, shared by the C ABI
// (pseg_host.cpp) and the serving loop (pseg_serve.cpp).
#pragma once
#include <memory>
#include <string>

#include "../../include/pseg.h"
#include "pseg_internal.hpp"

namespace pseg {
struct CalibPlan;
// initCalibrationPlan (core/calibrate.nim:316-360): "" or the error text
// ("rule i: ...").
std::string calib_compile(const pseg_calib_rule *rules, int n, int target, int file_factors,
                          int scale_aux, CalibPlan &out);
std::shared_ptr<CalibPlan> calib_new();
// resolveRow and count (core/calibrate.nim:430-512).
pseg_row_calibration calib_resolve_row(const CalibPlan &p, const pseg_calib_row &row,
                                       bool has_file_scales);
void calib_count(pseg_calib_summary &s, const pseg_row_calibration &r, const CalibPlan &p);
}  // namespace pseg

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
  int amplitude_units;   // amplitudeUnits: 0 recorded, 1 calibrated, 2 relative
  // calibration steps in plan order: their rules and whether they invert
  int n_calib;
  std::shared_ptr<const pseg::CalibPlan> calib[2];
  bool calib_inverse[2];
};

namespace pseg {

// Sets this thread's last error (pseg_last_error) and returns st.
pseg_status set_error(pseg_status st, const std::string &msg);
// PSEG_E_HIP with "what: <HIP error string>".
pseg_status hip_error(hipError_t e, const char *what);

}  // namespace pseg
