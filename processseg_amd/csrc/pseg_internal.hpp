// Internal structures shared by the host plan compiler (pseg_host.cpp) and the
// gfx950 kernels (pseg_kernels.hip).  Everything a kernel reads of a plan is
// passed by value in a kernel argument (DevPlan), with pointers to the plan's
// device tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pseg.h"

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
  kMuteTop = 9,
  kMuteBottom = 10,
  kGainCurve = 11,
  kIir = 12,
  kCalibrate = 13,    // calibrateOp and uncalibrateOp: a per-row multiplier
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
  int32_t n;          // real length
  int32_t m;          // complex length: n / 2 for an even n; n for an odd n (the
                      // n-point complex transform of (x, 0), no split)
  int32_t nrad;
  int32_t rad[kMaxRadices];
  int32_t sub[kMaxRadices];
  int32_t tw_off[kMaxRadices];
  FastDiv sdiv[kMaxRadices];
  // The first pass (butterflies of `pad` points, stride 1) writes its
  // output with one spare element after every `pad`, so that the lanes of
  // a store hit different LDS banks; the second pass reads that layout.
  // 0: no padding (a single pass).  `stride`: complex elements of the
  // transform's LDS buffer (m + m / pad); the passes run in place in it.
  int32_t pad;
  int32_t stride;
  const void *tw;     // T2 twiddles of every stage
  const void *split;  // T2 (cos, sin)(-2 pi k / n), k = 0 .. m
};

// Two consecutive Stockham stages whose butterfly holds at most `fuse`
// points run as one pass through LDS (host and device share the rule: the
// host sizes the first pass's padded layout and checks the passes' registers
// from it).  Both kernels fuse up to 16 points.
constexpr int kFusePipeline = 16;
constexpr int kFuseCorr = 16;
__host__ __device__ constexpr bool fusable(int r1, int r2, int fuse) { return r1 * r2 <= fuse; }
// Whether stage radix r2 can follow r1: a plan's radices come as 4s, at most
// one 2, 3s, then 5s (the device instantiates only such pairs).
__host__ __device__ constexpr int radix_rank(int r) { return r == 4 ? 0 : r == 2 ? 1 : r == 3 ? 2 : 3; }
__host__ __device__ constexpr bool may_follow(int r1, int r2) {
  return radix_rank(r1) < radix_rank(r2) || (r1 == r2 && r1 != 2);
}

// A pass of butterflies of r points runs in place with each thread holding
// the inputs of at most this many butterflies: `budget` complex values.  The
// fused row kernel holds 24 (its passes' registers beside the rest of the
// row's state decide how many rows share a CU); the one-transform kernels
// (correlation, spectra, Hilbert) hold 32.
constexpr int kBudgetPipeline = 24;
constexpr int kBudgetCorr = 32;
__host__ __device__ constexpr int pass_iters_max(int r, int budget) {
  return r >= budget ? 1 : budget / r;
}

struct DevFilter {
  DevFft fft;
  int32_t pad;
  int32_t reflect;
  int32_t zero_dc;
  int32_t trend;      // 0 none, 1 mean, 2 linear
  double dc_gain;
  const void *coef;   // T4, m / 2 + 1: the split, gain and unsplit per bin pair
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

// A mute (core/mute.nim): the curve giving each trace's mute time and the
// taper, both in ms; the row's column of that time needs its keys.
struct DevMute {
  int32_t kind;        // 0 time, 1 table, 2 LMO
  int32_t knots;
  double t_ms;         // time; LMO: t0Ms
  double ms_per_dist;  // LMO
  double pad_ms;       // LMO
  double taper_ms;
  double dt_ms;
  const double *offsets;  // table knots
  const double *times;
};

// A GainCurvePlan (gain/curves.nim:86-93): the curve, and its factors as T
// at t_min_ms + k dt_ms for k < len.
struct DevGain {
  int32_t kind;       // 0 knots, 1 power, 2 exponential
  int32_t knots;
  int64_t len;
  double power, ref_ms, min_ms;
  double db_per_s;
  double t_min_ms, dt_ms;
  const void *table;
  const double *knot_ms, *knot_db;
};

// An IirPlan (filter/iir.nim:86-101): the cascade's sections as (b0, b1, b2,
// a1, a2, g) float64 sextets, the pads before (lead) and after (tail) the
// data, steady-state starts with odd-reflection pads or starts from rest,
// forward only (minimum phase) or forward then backward (zero phase), and
// the trend's response G(1)·(a + b·i) − b·G'(1).
struct DevIir {
  int32_t n_sections;
  int32_t zero_phase;
  int32_t reflect;
  int32_t trend;
  int32_t lead, tail;
  int32_t span;         // lead + samples + tail: the float64 row buffer in LDS
  int32_t reserved;
  double dc_gain, dc_slope;
  const double *sec;
};

struct DevStep {
  int32_t kind;
  int32_t stat;       // normalize; calibration: which row of the job's multipliers
  double factor;      // scale
  double lo, hi;      // clip
  double target;      // normalize
  double percentile;  // normalize
  DevFilter filter;
  DevAgc agc;
  DevMute mute;
  DevGain gain;
  DevIir iir;
};

struct DevPlan {
  int32_t n_steps;
  DevStep steps[kMaxSteps];
};

// One job of the pipeline kernel: steps [step_first, step_last) of a plan on
// `rows` rows of a window `cols` wide.  Row r of the window is
// src[r * src_pitch .. + cols); its trace holds num[r] samples from trace
// sample 0 (-1: the row is not live), the window starts at trace sample
// `first`, so its live part is clamp(num[r] - first, 0, cols).  Window columns
// [crop0, crop1) are stored to dst[r * dst_pitch + 0 ..).  When has_factor,
// the live part is first scaled by *factor (a part after a per-panel
// normalization).  Jobs of a batch are sorted by item0, the first work item
// (row) of the job.  Mutes and gain curves read the rows' trace keys (t0Ms,
// offset per row; NULL: 0 and NaN) and col0, the trace sample of window
// column 0.
struct Job {
  const DevPlan *plan;
  const void *src;
  void *dst;
  const int32_t *num;
  const double *factor;
  const double *keys;
  int64_t src_pitch, dst_pitch;
  int64_t col0;
  int64_t src_off;    // raw traces: the byte of each trace's sample 0
  int32_t rows, cols, first, crop0, crop1, item0;
  int32_t step_first, step_last, has_factor;
  int32_t src_fmt;    // 0: src holds T; else raw SEG-Y traces of this format,
                      // src_pitch bytes apart (the kernel form with every step)
  // Calibration steps: every row's live length whatever its status (the
  // same convention as num) and the multipliers, [calibration step][rows]
  // (NULL when the plan has none).
  const int32_t *cal_len;
  const double *cal_k;
};

// One launch: a single job by value (n_jobs == 0) or n_jobs jobs in device
// memory; `total` work items; direct (one block per item) or persistent
// (counter != NULL, `grid` blocks take items from *counter, cleared first).
// `full`: the kernel form (step_form): 0 when every step is demean / filter /
// AGC (the smallest kernel, for the instruction cache), 1 for the form with
// every step but the IIR cascades, 2 for the form with every step.
struct JobLaunch {
  Job single;
  const Job *jobs;
  int32_t n_jobs, total, grid, max_cols;
  uint32_t *counter;
  size_t lds_bytes;
  int32_t full;
};

struct DevCorr {
  DevFft fft;
  const void *coef;   // T4, m + 2: the split, multiplier and unsplit per bin pair
  int32_t shift;
  int32_t pilot;
  int32_t record;
  int32_t out;
};

// Launchers.  Return hipSuccess or the launch error.
// The pipeline kernel with `threads` (128, 256 or 512) threads per row
// (pseg_kernels.hip, which dispatches to pipeline_nt<threads>.hip).
hipError_t launch_pipeline(int dtype, int threads, const JobLaunch &a, hipStream_t stream);
hipError_t launch_pipeline_nt128(int dtype, const JobLaunch &a, hipStream_t stream);
hipError_t launch_pipeline_nt256(int dtype, const JobLaunch &a, hipStream_t stream);
hipError_t launch_pipeline_nt512(int dtype, const JobLaunch &a, hipStream_t stream);
// Workgroups of the kernel launch_pipeline would pick for (dtype, threads,
// max_cols, full) that fit on the device at once with `lds` bytes each.
hipError_t pipeline_resident_blocks(int dtype, int threads, int max_cols, size_t lds, int full,
                                    int *blocks);
hipError_t resident_pipeline_nt128(int dtype, int max_cols, size_t lds, int full, int *blocks);
hipError_t resident_pipeline_nt256(int dtype, int max_cols, size_t lds, int full, int *blocks);
hipError_t resident_pipeline_nt512(int dtype, int max_cols, size_t lds, int full, int *blocks);
// Launch classes of a batch at most: 2 dtypes x 3 thread counts x 2 forms.
constexpr int kMaxBatchClasses = 12;

// The steps the smaller kernel runs (see JobLaunch::full); a per-panel
// normalization's factor is applied by the job itself, not by a step.
constexpr bool core_step(int kind) {
  return kind == kDemean || kind == kFilter || kind == kAgc || kind == kNormPanel;
}
// The kernel form a step needs (JobLaunch::full): 0 core, 1 every step but
// IIR, 2 every step.
constexpr int step_form(int kind) { return core_step(kind) ? 0 : kind == kIir ? 2 : 1; }
// Per-panel normalization factor (into *factor, NaN: leave the panel alone)
// of the live rows of x; scratch: rows doubles, 2 rows uint32 and, for a
// percentile, 65536 uint32 of histogram (cleared here).
hipError_t launch_panel_factor(int dtype, const void *x, int32_t rows, int32_t cols,
                               const int32_t *row_len, int32_t stat, double target,
                               double percentile, double *row_sq, uint32_t *row_max,
                               uint32_t *hist, double *factor, hipStream_t stream);
// The correlation kernel with `threads` (512 or 1024) threads per record.
// A SpectrumPlan (fft/spectrum.nim:61-101): the real FFT of nfft points, the
// window's Tukey taper (T, `window` values) and its float64 sum.
// An f-k plan (fft/fk.nim:60-68): the time transform (nTime real points),
// the spatial one (nK real points, run twice per frequency), the tapers
// (T) and where the spatial pass keeps the first half spectrum in LDS.
struct DevFk {
  DevFft time;
  DevFft space;
  const void *time_w;   // T, samples
  const void *trace_w;  // T, traces
  int32_t traces, samples, n_freq, n_k;
  int64_t space_lds;    // bytes of the spatial transform's buffer
  int64_t im_lds;       // where the spatial pass keeps a column's imaginary parts (0: re-read)
};
// fkAmplitudeInto (fft/fk.nim:438-504): spec (workspace, traces x n_freq T
// pairs) and flag (one uint32, cleared first) between the two passes; amp
// (n_k x n_freq float64, frequency fastest).
hipError_t launch_fk(int dtype, int threads, const DevFk *dev, size_t lds_time, size_t lds_space,
                     int traces, int n_freq, const void *src, int32_t cols,
                     const int32_t *row_len, const int32_t *live_rows, int32_t first_col,
                     void *spec, uint32_t *flag, double *amp, hipStream_t s);

struct DevSpec {
  DevFft fft;
  const void *taper;
  double taper_sum;
  int32_t window;
  int32_t demean;
};
// averageAmplitudeInto (fft/spectrum.nim:334-433) of rows [rows, cols] from
// column first_col: acc (bins + 1 float64: the power sums, then the rows
// used) is cleared, filled, and turned into amp (bins float64).
hipError_t launch_spectrum(int dtype, int threads, const DevSpec *dev, size_t lds, const void *src,
                           int32_t rows, int32_t cols, const int32_t *row_len, int32_t first_col,
                           double *parts, double *amp, double *used, int32_t bins, int32_t nfft,
                           hipStream_t stream);
hipError_t launch_spectrum_noalloc(int dtype, int threads, const DevSpec *dev, size_t lds,
                                   const void *src, int32_t rows, int32_t cols,
                                   const int32_t *row_len, int32_t first_col, double *amp,
                                   double *used, int32_t bins, int32_t nfft, hipStream_t stream);

// amplitudeInto (fft/spectrum.nim:295-312) of each row's window from
// first_col: amp and phase, [rows, bins] of dtype.
hipError_t launch_spectrum_rows(int dtype, int threads, const DevSpec *dev, size_t lds,
                                const void *src, int32_t rows, int32_t cols,
                                const int32_t *row_len, int32_t first_col, void *amp, void *phase,
                                hipStream_t s);

// hilbertInto / envelopeInto (fft/hilbert.nim:127-150) of each row's live
// samples (src and dst [rows, cols] of dtype; dst may be src).
hipError_t launch_hilbert(int dtype, int threads, const DevFft *dev, size_t lds, const void *src,
                          int32_t rows, int32_t cols, const int32_t *row_len, int32_t envelope,
                          void *dst, hipStream_t s);
hipError_t launch_analytic(int dtype, int threads, const DevFft *dev, size_t lds, const void *src,
                           int32_t rows, int32_t cols, const int32_t *row_len, double dt_ms,
                           double thr_scale, double *env, double *phase, double *hz,
                           hipStream_t stream);

// spectrogramInto (fft/spectrogram.nim) of one trace of `len` samples:
// frames x bins float64 amplitudes (the DevSpec's taper holds the Hann
// weights).
hipError_t launch_spectrogram(int dtype, int threads, const DevSpec *dev, size_t lds,
                              const void *x, int32_t len, int32_t hop, int32_t frames,
                              double *amp, hipStream_t s);
// A first-break plan on the device (qc/firstbreak.nim:172-183): the
// parameters, the windows in samples, and for a band-pass the minimum-phase
// biquads (b0, b1, b2, a1, a2, g each) with their pads and trend.
constexpr int kFbMaxSections = 32;
struct DevFirstBreak {
  int32_t method, filtered, whole, n_sections, reflect, trend;
  int64_t samples, w, lta, snr, before, after, lead;
  double dt, stabilization, search_start, search_end, guide_v, guide_t0, before_ms, after_ms;
  double low_db, high_db, dc_gain, dc_slope;
  double sec[kFbMaxSections][6];
};
// pickInto (qc/firstbreak.nim:492-541), one thread per row; the scratch
// arrays are [index][rows] (the lanes of a wave read neighbouring values).
hipError_t launch_firstbreak(int dtype, const DevFirstBreak *plan, const void *src, int32_t rows,
                             int32_t cols, const int32_t *row_len, const double *keys,
                             int64_t first_sample, double *p, double *a, double *buf, void *copy,
                             double *time_ms, int64_t *col, double *snr_db, double *conf,
                             hipStream_t stream);

// addFold / clear / foldDifference (qc/fold.nim) on the device.
hipError_t launch_fold_add(const pseg_binning &b, const double *bounds, int32_t nb,
                           int32_t sectors, double start_deg, int32_t offset_range,
                           const double *keys, const int32_t *row_len, int32_t rows, int32_t *fold,
                           int32_t *ofold, int32_t *afold, float *omin, float *omax,
                           unsigned long long *counters, hipStream_t stream);
hipError_t launch_fold_clear(int64_t bins, int32_t classes, int32_t sectors, int32_t offset_range,
                             int32_t *fold, int32_t *ofold, int32_t *afold, float *omin,
                             float *omax, unsigned long long *counters, hipStream_t stream);
hipError_t launch_fold_diff(const int32_t *pre, const int32_t *act, int32_t *out, int64_t n,
                            hipStream_t stream);

// n float64 at p set to 0 by a kernel (graph-capturable, unlike a memset).
hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t stream);
// averageSpectrogramInto (fft/spectrogram.nim:370-430) over `rows` rows
// [rows, cols]: status (rows int32), parts (groups·frames·bins float64) and
// used (groups float64) are the workspace; amp: frames·bins float64.
hipError_t launch_spectrogram_average(int dtype, int threads, const DevSpec *dev, size_t lds,
                                      int32_t bins, const void *src, int32_t rows, int32_t cols,
                                      const int32_t *row_len, int32_t hop, int32_t frames,
                                      int32_t groups, int32_t *status, double *parts,
                                      double *used, double *amp, double *used_out,
                                      hipStream_t stream);
// Workgroup rows of an averaged spectrogram: about 8 waves per SIMD of
// workgroups over the frames (2048 of 256 threads, 8192 of 64), at least one
// row each; a function of the counts and the plan only (the sum's order is
// fixed by it).
inline int32_t sgram_groups(int32_t rows, int32_t frames, int threads) {
  const int64_t target = (int64_t)2048 * 256 / (threads > 0 ? threads : 256);
  int64_t g = (target + (int64_t)frames - 1) / (frames > 0 ? frames : 1);
  if (g > rows) g = rows;
  if (g > 65535) g = 65535;
  return (int32_t)(g < 1 ? 1 : g);
}

// SEG-Y sample decode (format 1, 2, 3, 5 or 8, big-endian) of `rows` rows
// `pitch` bytes apart, samples from byte `offset` of each, to a [rows, cols]
// panel of dtype.
hipError_t launch_decode(int dtype, int format, const void *raw, int64_t pitch, int64_t offset,
                         int32_t rows, int32_t cols, void *dst, hipStream_t s);
hipError_t launch_correlate(int dtype, int threads, const DevCorr *dev_plan, size_t lds_bytes,
                            const void *src, int32_t rows, int32_t src_cols,
                            const int32_t *row_len, void *dst, hipStream_t stream);
// Dynamic LDS one row of the pipeline kernel may hold (the workgroup's
// 160 KiB less room for the kernel's static LDS).
constexpr size_t kMaxRowLds = 160 * 1024 - 2048;

// An averaged spectrum's partial sums: one per workgroup, bins + 1 float64
// each (the power sums, then the rows used).  Workgroups: at most 1024, one
// per row below that, and at most kSpecPartials / (bins + 1) so the partials
// stay within kSpecPartials float64 (16 MiB).  A function of the row count
// and the plan only, so the sum's order (and its bits) is too.
constexpr int64_t kSpecPartials = (int64_t)1 << 21;
inline int32_t spectrum_groups(int32_t rows, int32_t bins) {
  int64_t g = rows < 1024 ? rows : 1024;
  const int64_t cap = kSpecPartials / ((int64_t)bins + 1);
  if (g > cap) g = cap;
  return (int32_t)(g < 1 ? 1 : g);
}
// The widest row each thread count serves (samples per thread x threads).
constexpr int pipeline_max_cols(int threads) {
  return threads == 128 ? 128 * 24 : threads == 256 ? 256 * 24 : threads == 512 ? 512 * 24 : 0;
}
// The fewest threads whose widest row holds rows of `samples` samples (0:
// wider than any).  A plan takes this count, or a larger one when its
// transforms need more threads to run in place, and keeps it for every
// apply, so a row's arithmetic (the order of its block reductions) depends
// on its plan alone, never on the batch around it.
constexpr int pipeline_threads_for(int samples) {
  return samples <= pipeline_max_cols(128) ? 128
         : samples <= pipeline_max_cols(256) ? 256
         : samples <= pipeline_max_cols(512) ? 512 : 0;
}

}  // namespace pseg
