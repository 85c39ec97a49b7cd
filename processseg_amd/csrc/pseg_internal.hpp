// Internal structures shared by the host plan compiler (pseg_host.cpp) and the
// gfx950 kernels (pseg_kernels.hip).  Everything a kernel reads of a plan is
// passed by value in a kernel argument (DevPlan), with pointers to the plan's
// device tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pseg {

constexpr int kMaxSteps = 16;
constexpr int kMaxRadices = 24;

enum StepKind : int32_t {
  kDemean = 0,
  kDetrend = 1,
  kScale = 2,
  kClip = 3,
  kPolarity = 4,
  kFilter = 5,
  kAgc = 6,
  kNormTrace = 7,
  kNormPanel = 8,
};

// Divide by a runtime constant with a multiply and a shift (Granlund-
// Montgomery, 32-bit): q = umulhi(n, mul) >> shift, exact for n < 2^31.
struct FastDiv {
  uint32_t d, mul, shift;
};

// One mixed-radix FFT of m complex points as Stockham stages: radix r,
// stride s = product of the earlier radices, sub = m / (s r).  Twiddles
// w_cur^{pk} of stage i start at tw_off[i] (complex, interleaved re/im).
struct DevFft {
  int32_t n;          // real length (even)
  int32_t m;          // complex length n / 2
  int32_t nrad;
  int32_t rad[kMaxRadices];
  int32_t sub[kMaxRadices];
  int32_t tw_off[kMaxRadices];
  FastDiv sdiv[kMaxRadices];
  const void *tw;     // T2 twiddles of every stage
  const void *split;  // T2 (cos, sin)(-2 pi k / n), k = 0 .. m
};

struct DevFilter {
  DevFft fft;
  int32_t pad;
  int32_t reflect;
  int32_t zero_dc;
  int32_t trend;      // 0 none, 1 mean, 2 linear
  double dc_gain;
  const void *gain;   // T, m + 1 bins
  const void *fall;   // T, pad values
};

struct DevAgc {
  int32_t half;
  int32_t mode;       // 0 rms, 1 mean |x|
  double target;
  double epsilon;
  const double *inv;  // 1 / c, c = 0 .. 2h + 1
  const double *rev;  // 1 / (2h + 1 - m), m = 0 .. 2h
};

struct DevStep {
  int32_t kind;
  int32_t stat;       // normalize
  double factor;      // scale
  double lo, hi;      // clip
  double target;      // normalize
  double percentile;  // normalize
  DevFilter filter;
  DevAgc agc;
};

struct DevPlan {
  int32_t n_steps;
  DevStep steps[kMaxSteps];
};

// What one launch of the pipeline kernel runs: steps [first, last) of the
// plan on rows `cols` wide, after scaling the live rows by *factor when
// has_factor (the start of a part after a per-panel normalization).
struct PipeLaunch {
  int32_t first, last, cols, has_factor;
  size_t lds_bytes;
  const double *factor;
};

struct DevCorr {
  DevFft fft;
  const void *mul;    // T2 multiplier per bin, m + 1 bins
  int32_t shift;
  int32_t pilot;
  int32_t record;
  int32_t out;
};

// Launchers (pseg_kernels.hip).  Return hipSuccess or the launch error.
hipError_t launch_pipeline(int dtype, const DevPlan *dev_plan, const PipeLaunch &a,
                           const void *src, void *dst, int32_t rows, const int32_t *row_len,
                           hipStream_t stream);
// Per-panel normalization factor (into *factor, NaN: leave the panel alone)
// of the live rows of x; scratch: rows doubles, 2 rows uint32 and, for a
// percentile, 65536 uint32 of histogram (cleared here).
hipError_t launch_panel_factor(int dtype, const void *x, int32_t rows, int32_t cols,
                               const int32_t *row_len, int32_t stat, double target,
                               double percentile, double *row_sq, uint32_t *row_max,
                               uint32_t *hist, double *factor, hipStream_t stream);
hipError_t launch_correlate(int dtype, const DevCorr *dev_plan, size_t lds_bytes,
                            const void *src, int32_t rows, int32_t src_cols,
                            const int32_t *row_len, void *dst, hipStream_t stream);
int pipeline_max_cols();

}  // namespace pseg
