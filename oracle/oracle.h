/*
 * Oracle: a CPU restatement of processSEG's compiled-pipeline hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (processseg_amd/, the
 * C-ABI library) links, loads or calls this code.  Only tests/, the smoke
 * check in __graft_entry__.py and bench.py's cpu_baseline leg use it, as the
 * checker and as the CPU baseline ("port").
 *
 * The reference is Nim (processSEG 0.14.1) and cannot be compiled here: no Nim
 * toolchain, and its dependencies (parseSEG, Arraymancer) are not vendored.
 * This file restates its algorithm in C, following the reference file by file
 * (each function cites the file:line it follows).  Every kernel exists twice,
 * with the sample type T = float ("FP32 mirror": float64 tables and sums, T
 * butterflies, exactly the reference's precisions) and T = double ("FP64
 * oracle").  Built with -ffp-contract=off, so no multiply-add is fused.
 *
 * Parity pins (see tests/test_oracle_*.py): nextFastLen table and FFT vs a
 * naive DFT and numpy.fft (tfftplan.nim); response closed forms and
 * impulse-length known answers L=110 / nfft=2250, L=881 / nfft=3840
 * (bench/README.md:116-117); AGC vs brute-force window sums (tgain.nim:
 * 105-121), h=125 / window 251 (tgain.nim:160); percentile vs sort
 * (tstats.nim:15); correlation vs a direct sum and nfft=16200
 * (tcorrelate.nim:15-64, bench/README.md:369); decodeWindow table
 * (tpipeline.nim:320-324); filter invariants (tfilterapply.nim).
 */
#ifndef PSEG_ORACLE_H
#define PSEG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_NOTCH 8
#define OR_MAX_KNOTS 32

/* FilterSpec (filter/design.nim:78-90).  kind 0 = rkTrapezoid, 1 = rkButterworth. */
typedef struct or_spec {
  int32_t kind;
  int32_t zero_dc;
  int32_t n_notch;
  int32_t reserved;
  double f1, f2, f3, f4;                  /* trapezoid corners, Hz (Inf: none) */
  double low_hz, low_slope;               /* Butterworth low cut (0: none) */
  double high_hz, high_slope;             /* Butterworth high cut (Inf: none) */
  double notch_hz[OR_MAX_NOTCH];
  double notch_width[OR_MAX_NOTCH];
} or_spec;

/* Op (pipeline.nim:158-174), as plain data.  kind follows OpKind
 * (pipeline.nim:92-108): 0 demean, 1 detrend, 2 scale, 3 gainDb, 4 clip,
 * 5 polarity, 6 muteTop, 7 muteBottom, 8 filter, 9 iir (spec, trend and
 * iir_phase), 10 agc, 11 normalize, 12 gainCurve.
 * Other kinds are refused. */
typedef struct or_op {
  int32_t kind;
  int32_t trend;      /* filter: 0 trNone, 1 trMean, 2 trLinear */
  int32_t mode;       /* agc: 0 agcRms, 1 agcMeanAbs */
  int32_t stat;       /* normalize: 0 normMax, 1 normRms, 2 normPercentile */
  int32_t scope;      /* normalize: 0 perTrace, 1 perPanel */
  int32_t reserved;
  double value;       /* scale: factor; gainDb: dB */
  double lo, hi;      /* clip */
  double window_ms;   /* agc */
  double target;      /* agc, normalize */
  double epsilon;     /* agc */
  double percentile;  /* normalize */
  or_spec spec;       /* filter */
  /* muteTop / muteBottom (core/mute.nim:31-98): mute_kind 0 timeMute
   * (mute_t_ms), 1 tableMute (mute_knots |offset| -> time knots), 2 lmoMute
   * (mute_t_ms = t0Ms, mute_ms_per_dist = 1000 / velocity, mute_pad_ms) */
  int32_t mute_kind;
  int32_t mute_knots;
  double mute_t_ms;
  double mute_ms_per_dist;
  double mute_pad_ms;
  double taper_ms;
  double mute_offsets[OR_MAX_KNOTS];
  double mute_times[OR_MAX_KNOTS];
  /* gainCurve (gain/curves.nim:61-153, pipeline.nim:318-322): gain_kind 0
   * knotGain (gain_knots knots, times gain_knot_ms strictly increasing, gains
   * gain_knot_db), 1 powerGain (gain_power, gain_ref_ms, gain_min_ms), 2
   * exponentialGain (gain_db_per_s); tabulated over gain_tmin_ms ..
   * gain_tmax_ms (NaN: the end of the record).  The knot arrays are read
   * while the plan is built and used. */
  int32_t gain_kind;
  int32_t gain_knots;
  const double *gain_knot_ms;
  const double *gain_knot_db;
  double gain_power, gain_ref_ms, gain_min_ms;
  double gain_db_per_s;
  double gain_tmin_ms, gain_tmax_ms;
  int32_t iir_phase;  /* iir: 0 ipZero, 1 ipMinimum (filter/iir.nim:77-80) */
  int32_t reserved2;
} or_op;

/* ---------------------------------------------------------------- IIR -- */
/* IirPlan (filter/iir.nim:82-101, 193-239): a Butterworth biquad cascade
 * (transposed direct form II, a0 = 1), minimum phase (one causal pass) or
 * zero phase (forward-backward). */
#define OR_MAX_SECTIONS 32
typedef struct or_biquad {
  double b0, b1, b2, a1, a2;
  double g;               /* gain at 0 Hz */
} or_biquad;
typedef struct or_iir {
  int32_t n_sections;
  int32_t phase;          /* 0 ipZero, 1 ipMinimum */
  int32_t trend;          /* 0 trNone, 1 trMean, 2 trLinear */
  int32_t reflect;        /* odd-reflection pads and steady-state starts */
  int64_t pad, lead, tail, samples;
  double dt_ms, dc_gain, dc_slope;
  or_biquad sec[OR_MAX_SECTIONS];
} or_iir;
/* initIirPlan; 0 ok, -1 invalid (message in err). */
int or_iir_init(or_iir *p, const or_spec *spec, double dt_ms, int64_t samples, int32_t phase,
                int32_t trend, char *err, int errlen);
/* magnitudeAt (filter/iir.nim:268-278); order (:260-262). */
double or_iir_magnitude_at(const or_iir *p, double f);
int32_t or_iir_order(const or_iir *p);
/* applyFilter on one trace x[0 ..< n] (n <= samples), in place (filter/iir.nim:
 * 418-433): trend removed, padded, the passes in float64, the trend's
 * steady-state response added back. */
void or_iir_apply_f32(const or_iir *p, float *x, int64_t n);
void or_iir_apply_f64(const or_iir *p, double *x, int64_t n);

/* ------------------------------------------------------------- lengths -- */
int or_is_fast_len(int64_t n);
int64_t or_next_fast_len(int64_t n, int even);
int or_factorize(int64_t n, int32_t *radices); /* returns the count */

/* ----------------------------------------------------------------- FFT -- */
/* Real forward transform (unscaled) and inverse (scaled 1/n) of one trace. */
int or_rfft_f32(const float *x, int64_t n, float *re, float *im);
int or_rfft_f64(const double *x, int64_t n, double *re, double *im);
int or_irfft_f32(const float *re, const float *im, int64_t n, float *x);
int or_irfft_f64(const double *re, const double *im, int64_t n, double *x);

/* -------------------------------------------------------------- design -- */
double or_response(const or_spec *spec, double f, double dt_ms);
double or_narrowest_ramp_hz(const or_spec *spec);
/* 0 ok; -1 invalid spec (message in err). */
int or_impulse_length(const or_spec *spec, double dt_ms, int64_t *L, char *err, int errlen);

/* Filter plan decisions (filter/apply.nim:118-151). */
typedef struct or_filter_info {
  int64_t margin, pad, nfft;
  int32_t reflect, zero_dc;
  double dc_gain;
} or_filter_info;
int or_filter_info_of(const or_spec *spec, double dt_ms, int64_t samples, int32_t trend,
                      or_filter_info *out, char *err, int errlen);

/* ---------------------------------------------------------------- meta -- */
/* isLive / liveSamples (core/types.nim:212-230) flattened to one int32 per
 * row: the live length, or -1 for a row that is not live.  status: 0 tsLive,
 * 1 tsUnknown, 2 tsDead, 3 tsAuxiliary.  used_for_corr != 0 gives the rows
 * correlation uses instead (vibroseis/correlate.nim:442-443: not dead and not
 * failed). */
int or_meta_flatten(int32_t rows, int32_t cols, int64_t first_sample,
                    const int64_t *num_samples, const uint8_t *status,
                    const int64_t *failed, int32_t n_failed, int used_for_corr,
                    int32_t *row_len);

/* ------------------------------------------------------------ pipeline -- */
typedef struct or_plan_info {
  int64_t margin;
  int32_t parts;
  int32_t n_steps;
  int64_t nfft[16];   /* per step: the FFT length of a filter step, else 0 */
  int64_t half[16];   /* per step: the AGC half window h, else 0 */
} or_plan_info;

int or_plan_info_of(const or_op *ops, int32_t n_ops, int32_t dtype, double dt_ms,
                    int64_t samples, int64_t record_samples, or_plan_info *out,
                    char *err, int errlen);
/* compile + apply (pipeline.nim:466-523, 725-772).  src may equal dst (in
 * place).  row_len as or_meta_flatten.  Returns 0, or -1 with a message. */
int or_pipeline_apply_f32(const or_op *ops, int32_t n_ops, double dt_ms, int64_t samples,
                          int64_t record_samples, const float *src, float *dst,
                          int32_t rows, int32_t cols, const int32_t *row_len,
                          char *err, int errlen);
int or_pipeline_apply_f64(const or_op *ops, int32_t n_ops, double dt_ms, int64_t samples,
                          int64_t record_samples, const double *src, double *dst,
                          int32_t rows, int32_t cols, const int32_t *row_len,
                          char *err, int errlen);
/* The same with the rows' trace keys (mutes read them): keys holds
 * (t0Ms, offset) per row, or is NULL (t0Ms 0, offset NaN); first_sample is
 * ChunkMeta.firstSample (the column origin of the mute times). */
int or_pipeline_apply_keys_f32(const or_op *ops, int32_t n_ops, double dt_ms, int64_t samples,
                               int64_t record_samples, const float *src, float *dst,
                               int32_t rows, int32_t cols, const int32_t *row_len,
                               const double *keys, int64_t first_sample, char *err, int errlen);
int or_pipeline_apply_keys_f64(const or_op *ops, int32_t n_ops, double dt_ms, int64_t samples,
                               int64_t record_samples, const double *src, double *dst,
                               int32_t rows, int32_t cols, const int32_t *row_len,
                               const double *keys, int64_t first_sample, char *err, int errlen);
/* muteTimeMs (core/mute.nim:100-118) of a mute op at `offset`. */
double or_mute_time_ms(const or_op *op, double offset);
/* gainDbAt / factorAt (gain/curves.nim:140-160) of a gainCurve op's curve. */
double or_gain_db_at(const or_op *op, double t_ms);
double or_gain_factor_at(const or_op *op, double t_ms);
/* A gainCurve op's plan (initGainCurvePlan, gain/curves.nim:214-249) for
 * traces of `samples` samples at dt_ms, applied to one trace x[0 ..< n] whose
 * sample 0 is at start_ms, as a pipeline step applies it (the workspace form,
 * gain/curves.nim:380-416: off the table the factor is rounded to the sample
 * type first; a zero stays zero).  table_len (may be NULL) receives the
 * table's length.  0 ok, -1 invalid (message in err). */
int or_gain_curve_apply_f32(const or_op *op, double dt_ms, int64_t samples, float *x, int64_t n,
                            double start_ms, int64_t *table_len, char *err, int errlen);
int or_gain_curve_apply_f64(const or_op *op, double dt_ms, int64_t samples, double *x, int64_t n,
                            double start_ms, int64_t *table_len, char *err, int errlen);
/* muteTop / muteBottom of one row in column units (core/mute.nim:136-180). */
void or_mute_top_f32(float *x, int64_t n, double start_col, double taper_cols);
void or_mute_top_f64(double *x, int64_t n, double start_col, double taper_cols);
void or_mute_bottom_f32(float *x, int64_t n, double end_col, double taper_cols);
void or_mute_bottom_f64(double *x, int64_t n, double end_col, double taper_cols);

/* The same through a ProcessPool clone (core/pool.nim:145-308): `threads`
 * persistent workers taking runs of ~256 KiB of rows (multiples of 8) from an
 * atomic counter, one workspace each, every step on a tile while it is in
 * cache.  Bit-identical to the single-thread form.  Used as the CPU baseline. */
typedef struct or_pool or_pool;
or_pool *or_pool_create(int32_t threads);
void or_pool_destroy(or_pool *pool);
int or_pool_apply_f32(or_pool *pool, const or_op *ops, int32_t n_ops, double dt_ms,
                      int64_t samples, const float *src, float *dst, int32_t rows,
                      int32_t cols, const int32_t *row_len, int32_t repeats,
                      char *err, int errlen);

/* ------------------------------------------------------- single kernels -- */
int or_demean_f32(float *x, int64_t n);
int or_demean_f64(double *x, int64_t n);
int or_agc_f32(float *x, int64_t n, double window_ms, double dt_ms, int32_t mode,
               double target, double epsilon);
int or_agc_f64(double *x, int64_t n, double window_ms, double dt_ms, int32_t mode,
               double target, double epsilon);
double or_percentile_abs_f32(const float *x, int64_t n, double p);
double or_percentile_abs_f64(const double *x, int64_t n, double p);
/* panelStats (gain/stats.nim:253-304): out = {rms, maxAbs, clip, rows, samples} */
int or_panel_stats_f32(const float *x, int32_t rows, int32_t cols, const int32_t *row_len,
                       double clip_percentile, double *out);
int or_panel_stats_f64(const double *x, int32_t rows, int32_t cols, const int32_t *row_len,
                       double clip_percentile, double *out);
/* Single-trace applyFilter (filter/apply.nim:340-363). */
int or_filter_f32(float *x, int64_t n, const or_spec *spec, double dt_ms, int64_t samples,
                  int32_t trend, char *err, int errlen);
int or_filter_f64(double *x, int64_t n, const or_spec *spec, double dt_ms, int64_t samples,
                  int32_t trend, char *err, int errlen);

/* --------------------------------------------------------- correlation -- */
typedef struct or_corr_info {
  int64_t nfft, shift, out_samples, pilot_samples, record_samples;
  double scale;
} or_corr_info;
int or_corr_info_of(const double *pilot, int64_t pilot_len, double pilot_dt_ms,
                    double pilot_t0_ms, double dt_ms, int64_t record_samples,
                    int64_t out_samples, int32_t mode, double eps_percent,
                    int32_t normalize, or_corr_info *out, char *err, int errlen);
int or_correlate_f32(const double *pilot, int64_t pilot_len, double pilot_dt_ms,
                     double pilot_t0_ms, double dt_ms, int64_t record_samples,
                     int64_t out_samples, int32_t mode, double eps_percent,
                     int32_t normalize, const float *src, int32_t rows, int32_t src_cols,
                     const int32_t *row_len, float *dst, int32_t dst_cols,
                     char *err, int errlen);
int or_correlate_f64(const double *pilot, int64_t pilot_len, double pilot_dt_ms,
                     double pilot_t0_ms, double dt_ms, int64_t record_samples,
                     int64_t out_samples, int32_t mode, double eps_percent,
                     int32_t normalize, const double *src, int32_t rows, int32_t src_cols,
                     const int32_t *row_len, double *dst, int32_t dst_cols,
                     char *err, int errlen);

/* ---------------------------------------------------------- first breaks -- */
/* FirstBreakParams (qc/firstbreak.nim:146-170).  method: 0 pmStaLta,
 * 1 pmEnergyRatio, 2 pmCoppens. */
typedef struct or_fb_params {
  int32_t method;
  int32_t reserved;
  double window_ms, lta_ms, stabilization;
  double low_cut_hz, high_cut_hz, slope_db_per_octave;
  double search_start_ms, search_end_ms;
  double guide_velocity, guide_t0_ms, before_ms, after_ms;
  double snr_window_ms, low_db, high_db;
} or_fb_params;
/* FirstBreakPlan (qc/firstbreak.nim:172-183). */
typedef struct or_fb_plan {
  or_fb_params p;
  double dt_ms;
  int64_t samples, w, lta, snr, before, after;
  int32_t filtered, whole;
  or_iir filter;
} or_fb_plan;
/* initFirstBreakPlan (qc/firstbreak.nim:195-251); 0 ok, -1 invalid. */
int or_fb_init(or_fb_plan *plan, const or_fb_params *p, double dt_ms, int64_t samples, char *err,
               int errlen);
/* requiredMargin (qc/firstbreak.nim:265-275). */
int64_t or_fb_required_margin(const or_fb_plan *plan);
/* pickInto (qc/firstbreak.nim:492-541) over rows [rows, cols]: row_len[r] the
 * live samples (< 0: not live, no pick); keys (t0Ms, offset) per row, or
 * NULL (0, NaN); first_sample the block's first column.  Per row: time_ms,
 * col (-1: no pick), snr_db, confidence. */
void or_fb_pick_f32(const or_fb_plan *plan, const float *src, int32_t rows, int32_t cols,
                    const int32_t *row_len, const double *keys, int64_t first_sample,
                    double *time_ms, int64_t *col, double *snr_db, double *confidence);
void or_fb_pick_f64(const or_fb_plan *plan, const double *src, int32_t rows, int32_t cols,
                    const int32_t *row_len, const double *keys, int64_t first_sample,
                    double *time_ms, int64_t *col, double *snr_db, double *confidence);

/* ------------------------------------------------------------ fold maps -- */
/* Binning (stack/binning.nim:48-66): kind 0 bnCdp, 1 bnLine, 2 bnGrid,
 * 3 bnInlineCrossline; 2D binnings have inlines = 1. */
typedef struct or_binning {
  int32_t kind;
  int32_t reserved;
  int64_t inlines, crosslines;
  double first0, step0, first1, step1;
  double x0, y0, ux, uy, vx, vy;
  double size0, size1, max_off;
} or_binning;
/* The trace keys fold counting reads, [rows][OR_FOLD_KEYS] float64. */
enum { OR_K_SRC_X, OR_K_SRC_Y, OR_K_RCV_X, OR_K_RCV_Y, OR_K_CMP_X, OR_K_CMP_Y, OR_K_OFFSET,
       OR_K_ENSEMBLE, OR_K_INLINE, OR_K_CROSSLINE, OR_FOLD_KEYS };
/* binAt / binOf (stack/binning.nim:208-237); -1 for no bin. */
int64_t or_bin_at(const or_binning *b, double x, double y);
int64_t or_bin_of(const or_binning *b, const double *keys);
/* addFold(meta) (qc/fold.nim:429-588): the live rows (row_len >= 0) of
 * keys counted into fold [bins], offset_fold [classes][bins] (n_bounds
 * edges, NULL for none), azimuth_fold [sectors][bins], offset_min/max
 * [bins] (offset_range); counters: binned, outside, no coordinates, not
 * live. */
void or_fold_add(const or_binning *b, const double *bounds, int32_t n_bounds, int32_t sectors,
                 double sector_start_deg, int32_t offset_range, const double *keys,
                 const int32_t *row_len, int32_t rows, int32_t *fold, int32_t *offset_fold,
                 int32_t *azimuth_fold, float *offset_min, float *offset_max, int64_t *counters);

/* ------------------------------------------------------ sweep / pilot -- */
int64_t or_samples_in_length(double length_ms, double dt_ms);
/* linearSweep + sweepInto + applyTaper(tpCos2) (vibroseis/sweep.nim:185-342,
 * taper.nim:125-179), as float64 samples (generatePilot's default type). */
int or_linear_sweep(double f1_hz, double f2_hz, double length_ms, double dt_ms,
                    double taper_start_ms, double taper_end_ms, double start_phase_deg,
                    double amplitude, double *out, int64_t n, char *err, int errlen);

#ifdef __cplusplus
}
#endif
#endif
