/*
 * pseg.h — C ABI of the MI355X (gfx950) panel hot path of processSEG.
 *
 * The library (processseg_amd/libpseg.so) compiles processSEG pipelines and
 * correlation plans on the host (float64 tables, as the reference does) and
 * applies them to [rows, cols] panels in HBM with hand-written HIP kernels.
 * Every entry point replaces one reference interface; the citation is given
 * beside each declaration (paths relative to the reference repository).
 *
 * Conventions.
 *  - Every function returns a pseg_status; on failure a thread-local message
 *    is kept for pseg_last_error.  Argument errors are found before any sample
 *    changes (pipeline.nim:619-641), and data never raises: a NaN or Inf stays
 *    in its own row.
 *  - Panels are row-major, C-contiguous, float32 or float64, in device memory.
 *    src and dst are the same panel (in place) or do not overlap.
 *  - Rows are described by one int32 per row, `row_len`, in device memory:
 *    the live length (ChunkMeta.liveSamples, core/types.nim:226-230) of a row
 *    the op processes, or -1 for a row it leaves alone (dead, auxiliary or
 *    failed: core/types.nim:221-224).  pseg_meta_flatten builds it on the host.
 *  - `stream` is a hipStream_t (NULL: the default stream).  Apply calls are
 *    asynchronous on it and allocate nothing.
 *  - Plans are immutable after compile and may be shared across threads and
 *    streams (pipeline.nim:78-82).
 */
#ifndef PSEG_H
#define PSEG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pseg_status {
  PSEG_OK = 0,
  PSEG_E_INVALID = 1,      /* an argument the reference rejects with ValueError */
  PSEG_E_UNSUPPORTED = 2,  /* an op outside the GPU path: keep the CPU plan */
  PSEG_E_HIP = 3,          /* a HIP runtime error */
  PSEG_E_WORKSPACE = 4,    /* workspace missing or smaller than the plan needs */
  PSEG_E_BUSY = 5          /* a queue is full (serving) */
} pseg_status;

typedef enum pseg_dtype { PSEG_F32 = 0, PSEG_F64 = 1 } pseg_dtype;

/* OpKind (pipeline.nim:92-108), same order. */
typedef enum pseg_op_kind {
  PSEG_OP_DEMEAN = 0, PSEG_OP_DETREND = 1, PSEG_OP_SCALE = 2, PSEG_OP_GAIN_DB = 3,
  PSEG_OP_CLIP = 4, PSEG_OP_POLARITY = 5, PSEG_OP_MUTE_TOP = 6, PSEG_OP_MUTE_BOTTOM = 7,
  PSEG_OP_FILTER = 8, PSEG_OP_IIR = 9, PSEG_OP_AGC = 10, PSEG_OP_NORMALIZE = 11,
  PSEG_OP_GAIN_CURVE = 12, PSEG_OP_CALIBRATE = 13, PSEG_OP_UNCALIBRATE = 14
} pseg_op_kind;

#define PSEG_MAX_NOTCH 8
#define PSEG_MAX_OPS 16
#define PSEG_MAX_KNOTS 32
#define PSEG_MAX_GAIN_KNOTS (1 << 20)

/* FilterSpec (filter/design.nim:78-90): kind 0 = cos² trapezoid (corners
 * f1 < f2 <= f3 < f4; f2 = 0: no low cut, f3 = f4 = Inf: no high cut),
 * kind 1 = Butterworth magnitude (low_hz = 0: no low cut, high_hz = Inf: no
 * high cut; slopes in dB/octave, effective). */
typedef struct pseg_filter_spec {
  int32_t kind;
  int32_t zero_dc;
  int32_t n_notch;
  int32_t reserved;
  double f1, f2, f3, f4;
  double low_hz, low_slope, high_hz, high_slope;
  double notch_hz[PSEG_MAX_NOTCH];
  double notch_width[PSEG_MAX_NOTCH];
} pseg_filter_spec;

/* Op (pipeline.nim:158-174) as plain data; the builders of pipeline.nim:
 * 258-316 fill it.  Fields a kind does not use are ignored. */
typedef struct pseg_op {
  int32_t kind;        /* pseg_op_kind */
  int32_t trend;       /* filter: 0 trNone, 1 trMean, 2 trLinear (filter/edges.nim:8-14) */
  int32_t mode;        /* agc: 0 agcRms, 1 agcMeanAbs (gain/gain.nim:92-95) */
  int32_t stat;        /* normalize: 0 normMax, 1 normRms, 2 normPercentile */
  int32_t scope;       /* normalize: 0 perTrace, 1 perPanel */
  int32_t reserved;
  double value;        /* scale: factor; gainDb: dB */
  double lo, hi;       /* clip */
  double window_ms;    /* agc */
  double target;       /* agc, normalize */
  double epsilon;      /* agc */
  double percentile;   /* normalize */
  pseg_filter_spec spec; /* filter */
  /* muteTop / muteBottom (core/mute.nim:31-98, pipeline.nim:283-289): curve
   * mute_kind 0 timeMute (mute_t_ms), 1 tableMute (mute_knots knots of
   * |offset| -> time, offsets strictly ascending), 2 lmoMute (mute_t_ms =
   * t0Ms, mute_ms_per_dist = 1000 / velocity, mute_pad_ms); the taper in ms. */
  int32_t mute_kind;
  int32_t mute_knots;
  double mute_t_ms;
  double mute_ms_per_dist;
  double mute_pad_ms;
  double taper_ms;
  double mute_offsets[PSEG_MAX_KNOTS];
  double mute_times[PSEG_MAX_KNOTS];
  /* gainCurveOp (pipeline.nim:318-322, gain/curves.nim:61-153): gain_kind 0
   * knotGain (gain_knots knots: times gain_knot_ms, strictly increasing, and
   * gains gain_knot_db, read during pseg_plan_compile only; at most
   * PSEG_MAX_GAIN_KNOTS), 1 powerGain (gain_power, gain_ref_ms, gain_min_ms),
   * 2 exponentialGain (gain_db_per_s); the table covers gain_tmin_ms ..
   * gain_tmax_ms (NaN: the end of the record, recordSamples). */
  int32_t gain_kind;
  int32_t gain_knots;
  const double *gain_knot_ms;
  const double *gain_knot_db;
  double gain_power, gain_ref_ms, gain_min_ms;
  double gain_db_per_s;
  double gain_tmin_ms, gain_tmax_ms;
  /* iirOp (pipeline.nim:300-303, filter/iir.nim:193-239): a Butterworth
   * spec in `spec` (kind 1, no notches, no zero_dc), `trend`, and the phase:
   * 0 ipZero (forward-backward, slopes multiples of 12 dB/octave), 1
   * ipMinimum (one causal pass, multiples of 6). */
  int32_t iir_phase;
  int32_t reserved2;
  /* calibrateOp / uncalibrateOp (pipeline.nim:324-334, 444-447): the
   * Calibration's rules (calib_n_rules of them, read during
   * pseg_plan_compile only), its target (pseg_calib_target), how the file's
   * own factors rank (pseg_file_factors) and whether auxiliary rows are
   * converted to volts.  The rules' errors read "op i (opCalibrate): rule
   * j: ..." as the reference's compile raises them. */
  const struct pseg_calib_rule *calib_rules;
  int32_t calib_n_rules;
  int32_t calib_target;
  int32_t calib_file_factors;
  int32_t calib_scale_aux;
} pseg_op;

typedef struct pseg_plan pseg_plan;
typedef struct pseg_corr_plan pseg_corr_plan;

/* What a compiled plan reports (pipeline.nim:562-602, filter/apply.nim:153-176). */
typedef struct pseg_plan_info {
  uint64_t generation;     /* process-unique, increasing (pipeline.nim:213,523) */
  int32_t dtype;
  int32_t n_steps;
  int32_t parts;           /* partCount (pipeline.nim:643-647) */
  int32_t samples;
  int32_t record_samples;
  int32_t margin;          /* requiredMargin: the sum of the steps' margins */
  double dt_ms;
  int32_t step_kind[PSEG_MAX_OPS];   /* the op kind of each step */
  int32_t step_nfft[PSEG_MAX_OPS];   /* FFT length of a filter step, else 0 */
  int32_t step_pad[PSEG_MAX_OPS];    /* pad of a filter step */
  int32_t step_margin[PSEG_MAX_OPS]; /* impulse length L, or AGC half window h */
  int32_t step_reflect[PSEG_MAX_OPS];
  size_t lds_bytes;        /* LDS one row of the fused kernel holds */
  int32_t threads_per_row; /* workgroup size of the fused kernel for this plan
                              (128, 256 or 512: rows up to 3072, 6144, 12288) */
  int32_t amplitude_units; /* amplitudeUnits (pipeline.nim:571-577): 0 paRecorded,
                              1 paCalibrated, 2 paRelative */
  int32_t n_calib;         /* calibration steps (calibrateOp, uncalibrateOp) */
  int32_t reserved;
} pseg_plan_info;

/* ---------------------------------------------------------- sample decode -- */

/* SEG-Y data sample formats (SEG-Y rev 2, binary header bytes 3225-3226)
 * and SEG-D demultiplexed formats, all big-endian: the decode parseSEG runs
 * before the prefetcher's transform hands a chunk to the plan
 * (batch.nim:267-309, docs/integration.md:294-298). */
typedef enum pseg_sample_format {
  PSEG_FMT_IBM32 = 1,   /* IBM System/360 float: (-1)^s * f/2^24 * 16^(e-64) */
  PSEG_FMT_INT32 = 2,
  PSEG_FMT_INT16 = 3,
  PSEG_FMT_IEEE32 = 5,
  PSEG_FMT_INT8 = 8,
  /* SEG-D demultiplexed data (general header 1 format code, bytes 3-4):
   * the formats of the reference's SEG-D test files (tests/utils/
   * calibfiles.nim: 8058, 8036) and their companions.  Big-endian. */
  PSEG_FMT_SEGD_INT24 = 8036,   /* 24-bit two's complement integer */
  PSEG_FMT_SEGD_INT32 = 8038,   /* 32-bit two's complement integer */
  PSEG_FMT_SEGD_IEEE32 = 8058,  /* IEEE float32 */
  PSEG_FMT_SEGD_IEEE64 = 8080   /* IEEE float64 */
} pseg_sample_format;

/* --------------------------------------------------------------- spectra -- */

/* A SpectrumPlan (fft/spectrum.nim:77-101): windows of `window` samples,
 * demeaned (demean != 0), Tukey-tapered (sin² ramps of round(taper_fraction *
 * window) samples at each end), zero-padded to nfft (0: the next power of
 * two; else a 2·3·5-smooth length >= the window; odd lengths give
 * PSEG_E_UNSUPPORTED here). */
typedef struct pseg_spectrum_plan pseg_spectrum_plan;
pseg_status pseg_spectrum_compile(int32_t dtype, double dt_ms, int32_t window,
                                  double taper_fraction, int32_t nfft, int32_t demean,
                                  pseg_spectrum_plan **plan);
pseg_status pseg_spectrum_destroy(pseg_spectrum_plan *plan);
pseg_status pseg_spectrum_info_get(const pseg_spectrum_plan *plan, int32_t *nfft, int32_t *bins);
/* The workspace pseg_spectrum_average needs for `rows` rows, in bytes (the
 * partial sums of its workgroups; 0 for no rows).  Replaces the pool's
 * partial buffer, ensureAvgScratch (core/pool.nim:448-460). */
pseg_status pseg_spectrum_workspace_size(const pseg_spectrum_plan *plan, int32_t rows,
                                         size_t *bytes);
/* averageAmplitudeInto (fft/spectrum.nim:334-433; the pool form
 * core/pool.nim:776-800): the RMS amplitude spectrum, sqrt(mean over rows
 * of |X[k]|² / Σw²) · sides(k), of the windows from column first_col of the
 * live rows of src ([rows, cols], C-contiguous) with a live sample there and
 * no NaN or Inf in the window; row_len from pseg_meta_flatten.  ws: device,
 * at least pseg_spectrum_workspace_size bytes (PSEG_E_WORKSPACE otherwise),
 * nothing to clear; amp: device, bins float64; used: device, one float64 (the
 * rows averaged), or NULL.  The rows' powers are summed in an order fixed by
 * the row count and the plan, so repeated calls give the same bits.  Calls on
 * different streams need their own ws. */
pseg_status pseg_spectrum_average(const pseg_spectrum_plan *plan, const void *src, int32_t rows,
                                  int32_t cols, const int32_t *row_len, int32_t first_col,
                                  void *ws, size_t ws_bytes, double *amp, double *used,
                                  void *stream);

/* pseg_spectrum_average without a workspace and without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums go to a static device array of 2^21
 * float64 (16 MiB) per GPU, held until the process ends; every row count and
 * plan fits (the workgroups are capped to it), a larger need would be
 * PSEG_E_INVALID.  The same bits as pseg_spectrum_average on the same input.
 * All calls on one GPU share the array: calls that overlap on one GPU (two
 * streams, two threads) race and give wrong averages; serialize them. */
pseg_status pseg_spectrum_average_noalloc(const pseg_spectrum_plan *plan, const void *src,
                                          int32_t rows, int32_t cols, const int32_t *row_len,
                                          int32_t first_col, double *amp, double *used,
                                          void *stream);

/* amplitudeInto (fft/spectrum.nim:295-312) of each row's window from
 * first_col: amp = |X[k]|·sides(k)/Σw and phase = atan2(Im, Re) (radians,
 * from the window's first sample), [rows, bins] of the plan's dtype on the
 * device.  A window holding NaN or Inf gives NaN in every bin; rows with no
 * live sample in the window are left as they are. */
pseg_status pseg_spectrum_amplitude(const pseg_spectrum_plan *plan, const void *src, int32_t rows,
                                    int32_t cols, const int32_t *row_len, int32_t first_col,
                                    void *amp, void *phase, void *stream);

/* A SpectrogramPlan (fft/spectrogram.nim): frames i = 0 .. (samples-1)/hop
 * centred on column i*hop, windowLen samples weighted by a Hann window
 * sin²(pi (j+1)/(windowLen+1)), zero-padded to nfft (0: the next power of
 * two); samples outside the trace count as zero. */
typedef struct pseg_spectrogram_plan pseg_spectrogram_plan;
pseg_status pseg_spectrogram_compile(int32_t dtype, int32_t window, int32_t hop, double dt_ms,
                                     int32_t samples, int32_t nfft, pseg_spectrogram_plan **plan);
pseg_status pseg_spectrogram_destroy(pseg_spectrogram_plan *plan);
pseg_status pseg_spectrogram_info_get(const pseg_spectrogram_plan *plan, int32_t *frames,
                                      int32_t *bins, int32_t *nfft);
/* spectrogramInto of one device trace x[0 ..< len] (len <= samples): amp
 * (device, frames * bins float64) = |X[k]|·sides(k)/Σw per frame, Σw the
 * weights of the frame's samples inside the trace; a frame with none is 0. */
pseg_status pseg_spectrogram_apply(const pseg_spectrogram_plan *plan, const void *x, int32_t len,
                                   double *amp, void *stream);
/* The workspace pseg_spectrogram_average needs for `rows` rows, in bytes. */
pseg_status pseg_spectrogram_average_workspace_size(const pseg_spectrogram_plan *plan,
                                                    int32_t rows, size_t *bytes);
/* averageSpectrogramInto (fft/spectrogram.nim:370-430; the pool form
 * core/pool.nim:569-631): the power mean sqrt(Σ amp² / used) of the
 * spectrograms of the live rows of src ([rows, cols], cols <= the plan's
 * samples; row_len from pseg_meta_flatten), each over its live samples,
 * into amp (device, frames·bins float64).  Left out: rows not live or with
 * no live sample, float64 rows whose live samples leave the safe range (NaN,
 * infinity, subnormal, a nonzero magnitude below 1e-100 or above 1e140),
 * rows whose spectrogram holds a NaN or an infinity.  used: device, one
 * float64 (the rows averaged) or NULL; no row: amp all zero.  ws: device, at
 * least pseg_spectrogram_average_workspace_size bytes (PSEG_E_WORKSPACE
 * otherwise), one per stream.  One launch over rows × frames; the rows'
 * powers are summed in an order fixed by the counts, so repeated calls give
 * the same bits. */
pseg_status pseg_spectrogram_average(const pseg_spectrogram_plan *plan, const void *src,
                                     int32_t rows, int32_t cols, const int32_t *row_len,
                                     void *ws, size_t ws_bytes, double *amp, double *used,
                                     void *stream);

/* A HilbertPlan (fft/hilbert.nim:44-53): traces of up to `samples` samples,
 * zero-padded to nfft >= 2 * samples (0: nextFastLen(2 * samples)) so the
 * transform does not wrap. */
typedef struct pseg_hilbert_plan pseg_hilbert_plan;
pseg_status pseg_hilbert_compile(int32_t dtype, int32_t samples, int32_t nfft,
                                 pseg_hilbert_plan **plan);
pseg_status pseg_hilbert_destroy(pseg_hilbert_plan *plan);
pseg_status pseg_hilbert_nfft(const pseg_hilbert_plan *plan, int32_t *nfft);
/* hilbertInto (envelope = 0) or envelopeInto (envelope != 0,
 * sqrt(x² + H[x]²)) (fft/hilbert.nim:127-150) of each row's live samples
 * (row_len from pseg_meta_flatten; cols <= samples) into dst ([rows, cols],
 * the same block as src or disjoint).  Rows not live and samples past a
 * live end are left as they are in dst. */
pseg_status pseg_hilbert_apply(const pseg_hilbert_plan *plan, const void *src, void *dst,
                               int32_t rows, int32_t cols, const int32_t *row_len,
                               int32_t envelope, void *stream);

/* analyticInto (fft/hilbert.nim:160-215) of every live row of src ([rows,
 * cols], C-contiguous, cols <= the plan's samples; row_len from
 * pseg_meta_flatten), over its live length: env = |x + i·H[x]|, phase =
 * arg(x + i·H[x]) in degrees in (-180, 180] (NaN where env is 0), hz = the
 * circular mean of the two wrapped one-sample phase steps about each sample
 * (NaN at the row's first and last live samples and where the sample or a
 * neighbour is 0 or below peak·10^(floor_db/20), peak the row's largest
 * envelope).  env, phase, hz: device, [rows, cols] float64; rows not live and
 * columns past the live end are left as they are.  dt_ms finite and
 * positive, floor_db not NaN (PSEG_E_INVALID). */
pseg_status pseg_hilbert_analytic(const pseg_hilbert_plan *plan, const void *src, int32_t rows,
                                  int32_t cols, const int32_t *row_len, double dt_ms,
                                  double floor_db, double *env, double *phase, double *hz,
                                  void *stream);

/* ------------------------------------------------------------------- f-k -- */

/* An FkPlan (fft/fk.nim:60-106): the f-k amplitude of `traces` live rows by
 * windows of `samples` columns at dt_ms, tapered over taper_fraction (0 ..
 * 0.5) of each axis; nTime = nextFastLen(samples), nK = nextFastLen(2·traces).
 * PSEG_E_INVALID for traces < 4, samples < 8 and the reference's other
 * refusals; PSEG_E_UNSUPPORTED when a transform exceeds one workgroup. */
typedef struct pseg_fk_plan pseg_fk_plan;
pseg_status pseg_fk_compile(int32_t dtype, double dt_ms, int32_t traces, int32_t samples,
                            double taper_fraction, pseg_fk_plan **plan);
pseg_status pseg_fk_destroy(pseg_fk_plan *plan);
/* FkAxes (fft/fk.nim:50-58): frequency bins, wavenumber bins, Hz per bin,
 * cycles per trace per wavenumber bin. */
pseg_status pseg_fk_axes(const pseg_fk_plan *plan, int32_t *n_freq, int32_t *n_k, double *df_hz,
                         double *dk_per_trace);
/* The device workspace pseg_fk_amplitude needs (the half spectra of the
 * traces between the two passes, and a flag). */
pseg_status pseg_fk_workspace_size(const pseg_fk_plan *plan, size_t *bytes);
/* fkAmplitudeInto (fft/fk.nim:438-504): the unscaled f-k modulus of the
 * window first_col ..< first_col + samples of src ([rows, cols]) into amp
 * (device, n_k·n_freq float64, frequency fastest, zero wavenumber at n_k/2,
 * later-on-later events at positive k).  live_rows: device int32, the row
 * index of each of the plan's `traces` rows that enter (fkRows, fft/fk.nim:
 * 204-216: live rows with a sample at first_col or later, in order);
 * row_len: device, pseg_meta_flatten's.  The modulus is sqrt(re² + im²) in
 * float64, or hypot when a float64 sample of the window leaves [1e-100,
 * 1e140] (fft/fk.nim:164-202).  Asynchronous on `stream`. */
pseg_status pseg_fk_amplitude(const pseg_fk_plan *plan, const void *src, int32_t rows,
                              int32_t cols, const int32_t *row_len, const int32_t *live_rows,
                              int32_t first_col, double *amp, void *ws, size_t ws_bytes,
                              void *stream);

/* --------------------------------------------------------- first breaks -- */

/* FirstBreakParams (qc/firstbreak.nim:146-170): method 0 pmStaLta,
 * 1 pmEnergyRatio, 2 pmCoppens; NaN search bounds and guide velocity for
 * none; high_cut_hz = Inf and low_cut_hz = 0 for no band-pass. */
typedef struct pseg_fb_params {
  int32_t method;
  int32_t reserved;
  double window_ms, lta_ms, stabilization;
  double low_cut_hz, high_cut_hz, slope_db_per_octave;
  double search_start_ms, search_end_ms;
  double guide_velocity, guide_t0_ms, before_ms, after_ms;
  double snr_window_ms, low_db, high_db;
} pseg_fb_params;

typedef struct pseg_firstbreak_plan pseg_firstbreak_plan;

/* initFirstBreakPlan (qc/firstbreak.nim:195-251) for traces of up to
 * `samples` samples at dt_ms; PSEG_E_INVALID with the reference's
 * ValueError text for bad parameters. */
pseg_status pseg_firstbreak_compile(int32_t dtype, const pseg_fb_params *params, double dt_ms,
                                    int64_t samples, pseg_firstbreak_plan **out);
pseg_status pseg_firstbreak_destroy(pseg_firstbreak_plan *plan);
/* requiredMargin (qc/firstbreak.nim:269-281): the samples a pick reads
 * before its search window (the whole trace when it reads all of it). */
pseg_status pseg_firstbreak_required_margin(const pseg_firstbreak_plan *plan, int64_t *out);
/* workspaceSize (qc/firstbreak.nim:291-300) for a [rows, cols] block: the
 * bytes of ws pseg_firstbreak_pick takes. */
pseg_status pseg_firstbreak_workspace_size(const pseg_firstbreak_plan *plan, int32_t rows,
                                           int32_t cols, size_t *out);
/* pickInto (qc/firstbreak.nim:492-541) of every row of src ([rows, cols],
 * C-contiguous, cols <= samples; row_len from pseg_meta_flatten, -1 for
 * rows not live), each to its live length: keys NULL or [rows, 2] float64
 * (t0Ms, offset) on the device, first_sample the meta's.  ws: device,
 * 256-byte aligned, at least pseg_firstbreak_workspace_size bytes.  Per row:
 * time_ms (NaN: no pick), col (-1: no pick), snr_db (NaN: none), conf
 * (0 .. 1), all device arrays of rows. */
pseg_status pseg_firstbreak_pick(const pseg_firstbreak_plan *plan, const void *src, int32_t rows,
                                 int32_t cols, const int32_t *row_len, const double *keys,
                                 int64_t first_sample, void *ws, size_t ws_bytes, double *time_ms,
                                 int64_t *col, double *snr_db, double *conf, void *stream);

/* ------------------------------------------------------------ fold maps -- */

/* BinningKind (stack/binning.nim:48-53). */
enum { PSEG_BIN_CDP = 0, PSEG_BIN_LINE = 1, PSEG_BIN_GRID = 2, PSEG_BIN_INLINE_CROSSLINE = 3 };

/* Binning (stack/binning.nim:55-66) as plain data, built and validated by
 * cdpBinning / lineBinning / gridBinning / inlineCrosslineBinning
 * (:80-175); 2D binnings have inlines = 1.  Bins are numbered
 * inline * crosslines + crossline. */
typedef struct pseg_binning {
  int32_t kind;
  int32_t reserved;
  int64_t inlines, crosslines;
  double first0, step0, first1, step1;
  double x0, y0, ux, uy, vx, vy;
  double size0, size1, max_off;
} pseg_binning;

/* The trace keys fold counting reads (TraceKeys fields of parseSEG), one
 * row of PSEG_FOLD_KEYS float64 per trace, NaN where a header is absent. */
enum { PSEG_KEY_SRC_X, PSEG_KEY_SRC_Y, PSEG_KEY_RCV_X, PSEG_KEY_RCV_Y, PSEG_KEY_CMP_X,
       PSEG_KEY_CMP_Y, PSEG_KEY_OFFSET, PSEG_KEY_ENSEMBLE, PSEG_KEY_INLINE, PSEG_KEY_CROSSLINE,
       PSEG_FOLD_KEYS };

/* addFold(m, meta) (qc/fold.nim:583-588, with addTrace :561-581 and
 * addToBin :429-451): every row of keys ([rows][PSEG_FOLD_KEYS], device)
 * with row_len[r] >= 0 (pseg_meta_flatten: live) is counted into the map's
 * device arrays: fold [bins]; offset_fold [n_bounds - 1][bins] when
 * n_bounds > 1 (bounds: device, ascending); azimuth_fold [sectors][bins]
 * when sectors > 0; offset_min / offset_max [bins] when offset_range;
 * counters [4] += binned, outside, no coordinates, not live.  Integer
 * atomics and compare-and-swap extremes: the map does not depend on the
 * order the rows are counted in. */
pseg_status pseg_fold_add(const pseg_binning *binning, const double *bounds, int32_t n_bounds,
                          int32_t sectors, double sector_start_deg, int32_t offset_range,
                          const double *keys, const int32_t *row_len, int32_t rows,
                          int32_t *fold, int32_t *offset_fold, int32_t *azimuth_fold,
                          float *offset_min, float *offset_max, uint64_t *counters,
                          void *stream);
/* clear (qc/fold.nim:416-427): counts and counters to 0, the offset range
 * to NaN, on the device (a kernel, no memset). */
pseg_status pseg_fold_clear(int64_t bins, int32_t classes, int32_t sectors, int32_t offset_range,
                            int32_t *fold, int32_t *offset_fold, int32_t *azimuth_fold,
                            float *offset_min, float *offset_max, uint64_t *counters,
                            void *stream);
/* foldDifference (qc/fold.nim:612-622): out = pre - actual per bin
 * (wrapping int32), device arrays of bins. */
pseg_status pseg_fold_difference(const int32_t *pre, const int32_t *actual, int32_t *out,
                                 int64_t bins, void *stream);

/* --------------------------------------------------------------- errors -- */

/* The last error of this thread (the ValueError text of the reference,
 * "op i (kind): ..." for an op: pipeline.nim:384-386).  Returns the length. */
size_t pseg_last_error(char *buf, size_t n);

/* --------------------------------------------------------------- device -- */

/* Number of visible devices; 0 without a GPU.  Never fails. */
int32_t pseg_device_count(void);

/* ----------------------------------------------------------- calibration -- */

/* AmplitudeUnit (core/units.nim:57-64), CalibrationTarget and FileFactors
 * (core/calibrate.nim:104-115), CalibrationStatus and FactorSource
 * (core/calibrate.nim:155-165), in the reference's order. */
typedef enum pseg_amplitude_unit {
  PSEG_AU_RECORDED = 0, PSEG_AU_VOLT = 1, PSEG_AU_VELOCITY = 2, PSEG_AU_ACCELERATION = 3,
  PSEG_AU_PRESSURE = 4, PSEG_AU_FORCE = 5
} pseg_amplitude_unit;
typedef enum pseg_calib_target { PSEG_CAL_PHYSICAL = 0, PSEG_CAL_VOLTS = 1 } pseg_calib_target;
typedef enum pseg_file_factors {
  PSEG_FF_FALLBACK = 0, PSEG_FF_FIRST = 1, PSEG_FF_IGNORE = 2
} pseg_file_factors;
typedef enum pseg_calib_status {
  PSEG_CS_CONVERTED = 0, PSEG_CS_UNRESOLVED = 1, PSEG_CS_KEPT = 2
} pseg_calib_status;
typedef enum pseg_factor_source {
  PSEG_FS_NONE = 0, PSEG_FS_RULE = 1, PSEG_FS_FILE = 2
} pseg_factor_source;

/* CalibrationRule (core/calibrate.nim:133-153) with RuleMatch's criteria
 * (-1 / NaN / has_station 0: any) and its factors in SI, as Descale.
 * voltsPerCount and Sensitivity.voltsPerUnit give them (core/units.nim:
 * 265-268, 347-349): NaN where the rule gives none. */
typedef struct pseg_calib_rule {
  int32_t sensor_type;      /* SEG-D sensor type 0 .. 255, -1: any */
  int32_t component;        /* 0 .. 255, -1: any */
  int32_t channel_set;      /* >= 0, -1: any */
  int32_t has_station;      /* 0: any station */
  double channel;           /* NaN: any */
  double station_line, station_point;
  int32_t station_index;    /* 0: any index */
  int32_t sensitivity_unit; /* the quantity of the sensitivity (pseg_amplitude_unit) */
  double descale;           /* volts per count */
  double sensitivity;       /* volts per m/s, m/s², Pa or N */
  double gain_db;           /* a recorder gain the descale leaves out */
} pseg_calib_rule;

/* What resolveRow reads of one row (core/calibrate.nim:430-497): its status,
 * whether it failed, TraceKeys' sensorType, component, channel and receiver
 * station, TraceRows.channelSetIndex, and the row's FileScale (core/units.nim:
 * 105-117; NaN where the headers say nothing). */
typedef struct pseg_calib_row {
  uint8_t status;           /* 0 tsLive, 1 tsUnknown, 2 tsDead, 3 tsAuxiliary */
  uint8_t failed;
  uint8_t sensor_type;
  uint8_t file_physical;    /* FileScale.physical */
  int32_t component;
  int32_t channel_set;      /* -1: none */
  int32_t rcv_index;        /* 0: none */
  double channel, rcv_line, rcv_point;
  double file_descale, file_sensitivity;
  int32_t file_unit;        /* pseg_amplitude_unit */
  int32_t reserved;
} pseg_calib_row;

/* RowCalibration (core/calibrate.nim:167-182). */
typedef struct pseg_row_calibration {
  double factor;            /* 1 for a kept row, NaN for an unresolved one */
  int32_t unit, status;     /* pseg_amplitude_unit, pseg_calib_status */
  int32_t descale_from, sensitivity_from;     /* pseg_factor_source */
  int32_t descale_rule, sensitivity_rule, gain_rule;  /* -1: none */
  int32_t conflict, mismatch;
} pseg_row_calibration;

/* CalibrationSummary (core/calibrate.nim:184-193). */
typedef struct pseg_calib_summary {
  int32_t units[6];         /* rows by unit after calibration */
  int32_t converted, unresolved, kept;
  int32_t no_descale, no_sensitivity;
  int32_t conflicts, mismatches;
  int32_t reserved;
} pseg_calib_summary;

typedef struct pseg_calib pseg_calib;
/* initCalibrationPlan (core/calibrate.nim:316-360): validates the rules
 * (PSEG_E_INVALID "rule i: ..." with the reference's text) and indexes them. */
pseg_status pseg_calib_compile(const pseg_calib_rule *rules, int32_t n_rules, int32_t target,
                               int32_t file_factors, int32_t scale_auxiliary, pseg_calib **out);
pseg_status pseg_calib_destroy(pseg_calib *cal);
/* resolveInto (core/calibrate.nim:514-527): each row's RowCalibration into
 * out (n_rows of them) and the counts into summary (may be NULL);
 * has_file_scales: the meta carries a FileScale column (ChunkMeta.
 * hasFileScales).  Host memory; allocates nothing. */
pseg_status pseg_calib_resolve(const pseg_calib *cal, const pseg_calib_row *rows, int32_t n_rows,
                               int32_t has_file_scales, pseg_row_calibration *out,
                               pseg_calib_summary *summary);

/* The multipliers of a plan's calibration steps for the rows of a block
 * (calibrateImpl's k, core/calibrate.nim:615-624, from each step's
 * resolveInto): factors[s * n_rows + r] for the s-th calibration step in
 * the plan's order: the row's factor for calibrateOp, its inverse for
 * uncalibrateOp, NaN for an unresolved row, 1 for a kept one.  Host arrays;
 * factors may be NULL to ask for n_calib only; summaries NULL or one per
 * calibration step. */
pseg_status pseg_plan_calib_factors(const pseg_plan *plan, const pseg_calib_row *rows,
                                    int32_t n_rows, int32_t has_file_scales, double *factors,
                                    pseg_calib_summary *summaries, int32_t *n_calib);

/* ------------------------------------------------------------------ meta -- */

/* Replaces ChunkMeta's isLive / liveSamples / isFailed (core/types.nim:
 * 212-230) and correlation's row test (vibroseis/correlate.nim:442-443) with
 * one host array: row_len[r] = clamp(num_samples[r] - first_sample, 0, cols)
 * for a row the op uses, else -1.  status: 0 tsLive, 1 tsUnknown, 2 tsDead,
 * 3 tsAuxiliary.  failed: ascending row indices (may be NULL when n_failed is
 * 0).  for_correlation == 1: rows not dead and not failed (auxiliary rows
 * included); 2: every row (the live length a calibration step scales); 0:
 * live or unknown rows that did not fail. */
pseg_status pseg_meta_flatten(int32_t rows, int32_t cols, int64_t first_sample,
                              const int64_t *num_samples, const uint8_t *status,
                              const int64_t *failed, int32_t n_failed,
                              int32_t for_correlation, int32_t *row_len);

/* -------------------------------------------------------------- lengths -- */

/* nextFastLen (fft/plan.nim:113-118): the smallest 2·3·5-smooth length
 * >= max(n, 1), even unless even == 0. */
int64_t pseg_next_fast_len(int64_t n, int32_t even);

/* impulseLength (filter/design.nim:353-366), on the host in float64. */
pseg_status pseg_impulse_length(const pseg_filter_spec *spec, double dt_ms, int64_t *length);

/* response (filter/design.nim:236-255). */
double pseg_response(const pseg_filter_spec *spec, double f_hz, double dt_ms);

/* ------------------------------------------------------------- pipeline -- */

/* Pipeline.compile (pipeline.nim:466-523): validates every op for blocks of up
 * to `samples` samples at dt_ms and builds the device tables (twiddles, gains,
 * pad tapers, AGC tables) on the current device.  record_samples: 0, or the
 * trace length a window of `samples` is cut from.  Every op kind runs on the
 * GPU; PSEG_E_UNSUPPORTED is left for transforms beyond one workgroup's LDS
 * (the caller keeps the CPU plan for such a flow). */
pseg_status pseg_plan_compile(const pseg_op *ops, int32_t n_ops, int32_t dtype, double dt_ms,
                              int32_t samples, int32_t record_samples, pseg_plan **out);
pseg_status pseg_plan_destroy(pseg_plan *plan);
pseg_status pseg_plan_info_get(const pseg_plan *plan, pseg_plan_info *out);

/* responseModuli (pipeline.nim:774-838): gain[i] = |Σ_j x[j] exp(-2πi
 * hz[i] j dt_s)| for the n host samples x at m arbitrary frequencies, one
 * sine and cosine per frequency.  Host arrays. */
pseg_status pseg_response_moduli(const double *x, int32_t n, const double *hz, int32_t m,
                                 double dt_s, double *gain);
/* appliedResponseInto (pipeline.nim:840-880): the response of the flow as
 * applied — the flow with every FFT and IIR filter's trend removal off,
 * compiled for `samples` samples at dt_ms, applied on the current device
 * to a unit impulse at sample samples/2 of one live row (t0Ms 0, offset 0,
 * no file scales), and pseg_response_moduli of that output at hz (host,
 * m values) into gain (host).  PSEG_E_INVALID as pseg_plan_compile. */
pseg_status pseg_applied_response(const pseg_op *ops, int32_t n_ops, int32_t dtype, double dt_ms,
                                  int32_t samples, const double *hz, int32_t m, double *gain);

/* Pipeline.requiredMargin(dtMs) without compiling (pipeline.nim:525-548). */
pseg_status pseg_required_margin(const pseg_op *ops, int32_t n_ops, double dt_ms,
                                 int32_t *margin);

/* decodeWindow (pipeline.nim:550-560): the samples to decode for a zoom that
 * shows shown_first .. shown_last (inclusive) of traces trace_len long. */
void pseg_decode_window(int32_t shown_first, int32_t shown_last, int64_t margin,
                        int32_t trace_len, int32_t *first, int32_t *last);

/* The device workspace apply needs for `rows` x `cols` (ws.reserve(plan),
 * pipeline.nim:599-602): 0 unless the plan has a per-panel normalization. */
pseg_status pseg_plan_workspace_size(const pseg_plan *plan, int32_t rows, int32_t cols,
                                     size_t *bytes);

/* Plan.apply(src, dst, meta, ws) (pipeline.nim:736-772) and
 * ProcessPool.apply (core/pool.nim:277-308): runs the plan on every row of
 * src into dst; rows with row_len < 0 are copied unchanged, samples past a
 * row's live length too.  src == dst runs in place.  cols <= plan samples.
 * ws: device memory of pseg_plan_workspace_size bytes (NULL when 0). */
pseg_status pseg_plan_apply(const pseg_plan *plan, const void *src, void *dst, int32_t rows,
                            int32_t cols, const int32_t *row_len, void *ws, size_t ws_bytes,
                            void *stream);

/* The same with the rows' trace keys, which mutes and gain curves read
 * (muteCol, core/mute.nim:189-194: the mute time's column (t - t0Ms) / dtMs -
 * firstSample; sampleTimeMs, core/types.nim:237-240: sample j of a row at
 * t0Ms + (firstSample + j)·dtMs): row_keys holds (t0Ms, offset) per row as
 * float64 in device memory (ChunkMeta.keys), or is NULL (t0Ms 0, offset NaN:
 * a timeMute then applies, an offset-dependent mute leaves the rows alone); a
 * NaN t0Ms leaves the row to mutes and gain curves; first_sample is
 * ChunkMeta.firstSample.  pseg_plan_apply is this with NULL and 0. */
pseg_status pseg_plan_apply_keys(const pseg_plan *plan, const void *src, void *dst, int32_t rows,
                                 int32_t cols, const int32_t *row_len, const double *row_keys,
                                 int64_t first_sample, void *ws, size_t ws_bytes,
                                 void *stream);

/* pseg_plan_apply_keys for a plan with calibration steps (calibrateOp,
 * uncalibrateOp; pipeline.nim:686-688): calib_len holds every row's live
 * length whatever its status (liveSamples, core/types.nim:226-230;
 * pseg_meta_flatten with for_correlation = 2) and calib_factors the
 * multipliers of pseg_plan_calib_factors, [n_calib][rows] float64, both in
 * device memory.  A calibration step scales a row's live samples by its
 * multiplier in T, dead and auxiliary rows included (pipeline.nim:44-45);
 * a multiplier of 1 leaves the row as it is.  Every other step leaves rows
 * with row_len < 0 alone.  The other apply entry points refuse a plan with
 * calibration steps (PSEG_E_INVALID). */
pseg_status pseg_plan_apply_calib(const pseg_plan *plan, const void *src, void *dst, int32_t rows,
                                  int32_t cols, const int32_t *row_len, const double *row_keys,
                                  int64_t first_sample, const int32_t *calib_len,
                                  const double *calib_factors, void *ws, size_t ws_bytes,
                                  void *stream);

/* Decodes `rows` traces of big-endian samples on the device into a [rows,
 * cols] panel of `dtype` (C-contiguous): trace r's sample j is the word at
 * raw + r*pitch + offset + j*bytes (a SEG-Y file's traces: offset 240, pitch
 * 240 + bytes * samples; a SEG-D record's demultiplexed traces: offset the
 * trace header and its extensions).  IBM, int8, int16, int24 and IEEE words
 * decode exactly (IEEE float64 to float32 rounds to nearest), int32 to
 * float32 rounds to nearest.  For 4- and 8-byte formats raw, pitch and
 * offset must be multiples of 4.  Asynchronous on `stream`.  No counterpart
 * in processSEG itself: parseSEG decodes on the host (its decode workers). */
pseg_status pseg_decode_samples(int32_t format, const void *raw, int64_t pitch, int64_t offset,
                                int32_t rows, int32_t cols, int32_t dtype, void *dst,
                                void *stream);

/* apply(plan, src: Tensor[T], dst: var Tensor[T], meta, ws) and the
 * in-place apply(plan, samples: var Tensor[T], meta, ws) (pipeline.nim:
 * 758-772) for a caller whose panels live in host memory: src, dst, row_len
 * (pseg_meta_flatten's) and row_keys (NULL or [rows, 2] float64) are HOST
 * arrays; src == dst runs in place.  The block is copied to device memory
 * the library keeps per thread and device (grown on demand; the plan's
 * workspace is carved from it too), the plan runs there on `stream`, and
 * the result is copied back: the call returns when dst holds it.  The same
 * checks and errors as pseg_plan_apply_keys, before any copy. */
pseg_status pseg_plan_apply_host(const pseg_plan *plan, const void *src, void *dst, int32_t rows,
                                 int32_t cols, const int32_t *row_len, const double *row_keys,
                                 int64_t first_sample, void *stream);
/* Frees this thread's staging memory of pseg_plan_apply_host (on every
 * device it used); the next call allocates it again. */
pseg_status pseg_host_staging_release(void);

/* pseg_plan_apply_keys with the source as raw SEG-Y traces (as
 * pseg_decode_samples reads them) instead of a panel: each row is decoded
 * as it is loaded and the plan runs on it, so the traces cross HBM once,
 * into dst ([rows, cols] of the plan's dtype; samples past a live end hold
 * the decoded values).  Plans with a per-panel normalization are refused
 * (PSEG_E_UNSUPPORTED: their statistics read a decoded panel).  No
 * counterpart in processSEG: the prefetcher's transform (batch.nim:267-309)
 * runs the plan on parseSEG's decoded chunk, a second pass. */
pseg_status pseg_plan_apply_raw(const pseg_plan *plan, int32_t format, const void *raw,
                                int64_t pitch, int64_t offset, void *dst, int32_t rows,
                                int32_t cols, const int32_t *row_len, const double *row_keys,
                                int64_t first_sample, void *stream);

/* ---------------------------------------------------------------- batch -- */

/* Many blocks, each with its own plan (different sample intervals, widths
 * and row counts), applied as one unit of work: what a caller otherwise does
 * as one ProcessPool.apply(plan_i, src_i, dst_i, meta_i) per record
 * (core/pool.nim:277-308, in a loop over records as bench/viewersim.nim
 * does).  Each item's result equals pseg_plan_apply of that item alone, bit
 * for bit.  The batch runs as one persistent launch per (dtype, threads per
 * row) class present, rows handed out heaviest first from a device counter. */
typedef struct pseg_batch_item {
  const pseg_plan *plan;   /* one part (no per-panel normalization) */
  const void *src;         /* device [rows, cols], C-contiguous */
  void *dst;               /* device, same shape; == src runs in place */
  int32_t rows, cols;      /* cols <= the plan's samples */
  const int32_t *row_len;  /* device, rows int32 from pseg_meta_flatten */
  const double *row_keys;  /* device (t0Ms, offset) per row, or NULL (as in
                              pseg_plan_apply_keys): mutes and gain curves */
  int64_t first_sample;    /* ChunkMeta.firstSample, >= 0 */
} pseg_batch_item;
typedef struct pseg_batch pseg_batch;

typedef struct pseg_batch_info {
  int32_t items;
  int32_t launches;        /* persistent launches per pseg_batch_launch */
  int64_t rows;            /* rows of all items */
  int64_t samples;         /* rows x cols of all items */
  int32_t threads[12];     /* threads per row of each launch */
  int32_t grid[12];        /* workgroups of each persistent launch; 0: one
                              workgroup per row (pseg_batch_tune chose it) */
} pseg_batch_info;

/* What pseg_batch_tune measured and chose per launch class. */
typedef struct pseg_batch_tuning {
  int32_t launches;
  int32_t grid[12];            /* the workgroups chosen */
  int32_t direct[12];          /* 1: one workgroup per row, no work counter */
  float best_ms[12];           /* the chosen shape's median launch time */
  float candidate_ms[12][4];   /* 1/2, 3/4, all of the resident workgroups, direct (0: skipped) */
} pseg_batch_tuning;

/* Validates every item as pseg_plan_apply would (before any sample changes;
 * the first bad item is named: "item i: ...") and keeps the work table on the
 * device.  The plans and buffers must outlive the batch. */
pseg_status pseg_batch_create(const pseg_batch_item *items, int32_t n_items, pseg_batch **out);
/* Runs the batch on `stream`.  No allocation, no host-device copy and no
 * memset node: the call can be captured in a graph and replayed. */
pseg_status pseg_batch_launch(const pseg_batch *batch, void *stream);
pseg_status pseg_batch_info_get(const pseg_batch *batch, pseg_batch_info *out);
pseg_status pseg_batch_destroy(pseg_batch *batch);
/* No counterpart in the reference (its CPU pool has no launch shape): times
 * each launch class of the batch `repeats` times (after one warm-up run)
 * with a persistent grid of 1/2, 3/4 and all of the workgroups the device
 * holds at once, and with one workgroup per row, and keeps the fastest
 * shape in the batch.  A row's arithmetic does not depend on the launch
 * shape, so later launches give the same bits, faster.  The batch runs
 * 4·(repeats + 1) times per class: items must not run in place
 * (PSEG_E_INVALID).  Synchronizes `stream`. */
pseg_status pseg_batch_tune(pseg_batch *batch, int32_t repeats, void *stream,
                            pseg_batch_tuning *out);

/* -------------------------------------------------------------- serving -- */

/* A viewer's zoom windows served from a survey resident in device memory.
 * Each request is the reference's zoomWindow (tests/utils/viewersim.nim:
 * 167-191): margin = the flow's requiredMargin, window = decodeWindow(shown,
 * margin, samples), the plan for (flow, window width, samples) compiled on
 * first use (planFor, viewersim.nim:53-62), apply on the window's rows, crop
 * to the shown samples.  Requests queue and are served by continuous
 * batching: a worker thread packs every admitted request (admission: room
 * for the result in the server's device arena) into persistent launches, two
 * batches in flight on their own streams.  A request's samples equal
 * pseg_plan_apply of its window alone, bit for bit, whatever else shares its
 * launch. */
typedef struct pseg_server pseg_server;

typedef struct pseg_server_config {
  int32_t dtype;               /* PSEG_F32 or PSEG_F64 */
  int32_t traces;              /* rows of the survey */
  int32_t samples;             /* samples per trace (the record length) */
  int32_t max_batch_requests;  /* requests per launch; 1 = one launch per request */
  int64_t pitch;               /* elements between trace starts, >= samples */
  const void *survey;          /* device [traces, pitch] */
  const int32_t *trace_len;    /* device, traces int32: the trace's numSamples,
                                  -1 for a trace that is not live */
  double dt_ms;
  size_t arena_bytes;          /* device memory for results */
  int32_t max_batch_rows;      /* rows per launch (0: no limit) */
  int32_t max_inflight;        /* launches in flight, 1 or 2 */
  const double *trace_keys;    /* device, traces x (t0Ms, offset) float64, or
                                  NULL (t0Ms 0, offset NaN): what mutes and
                                  gain curves read */
} pseg_server_config;

typedef struct pseg_serve_result {
  int32_t status;              /* PSEG_OK or the request's error code */
  int32_t rows, cols;          /* traces x shown samples */
  int32_t first, window;       /* the decoded window: samples first .. first+window-1 */
  int32_t batch_requests;      /* requests in the launch that served it */
  int32_t plan_compiled;       /* 1: its plan was compiled for this request */
  int32_t reserved;
  const void *samples;         /* device [rows, cols]; valid until release */
  uint64_t generation;         /* the plan's generation (staleness tag) */
  int64_t submit_ns, start_ns, done_ns;  /* steady clock of the host */
} pseg_serve_result;

typedef struct pseg_server_info {
  int64_t requests, batches, launches, plans_compiled;
  int64_t arena_peak_bytes;
} pseg_server_info;

/* The survey and trace_len must outlive the server. */
pseg_status pseg_server_create(const pseg_server_config *cfg, pseg_server **out);
/* Registers a flow (the ops of a Pipeline); *flow is its index.  A flow with
 * a per-panel normalization is PSEG_E_UNSUPPORTED here. */
pseg_status pseg_server_add_flow(pseg_server *srv, const pseg_op *ops, int32_t n_ops,
                                 int32_t *flow);
/* Queues the zoom window of traces trace0 .. trace0+ntraces-1, samples
 * shown_first .. shown_last (inclusive).  Thread-safe. */
pseg_status pseg_server_submit(pseg_server *srv, int32_t flow, int32_t trace0, int32_t ntraces,
                               int32_t shown_first, int32_t shown_last, uint64_t *ticket);
/* Waits up to timeout_ms (< 0: no limit) for a request; PSEG_E_BUSY when it
 * is not done yet.  out->status holds the request's own result. */
pseg_status pseg_server_wait(pseg_server *srv, uint64_t ticket, int32_t timeout_ms,
                             pseg_serve_result *out);
/* Copies a finished request's samples into dst (device, rows x cols). */
pseg_status pseg_server_read(pseg_server *srv, uint64_t ticket, void *dst, void *stream);
/* Frees a finished request's result memory and forgets the ticket. */
pseg_status pseg_server_release(pseg_server *srv, uint64_t ticket);
pseg_status pseg_server_info_get(pseg_server *srv, pseg_server_info *out);
/* Finishes the queued requests, then stops the worker and frees the server's
 * memory (plans compiled by the server included). */
pseg_status pseg_server_destroy(pseg_server *srv);

/* ---------------------------------------------------------- correlation -- */

/* What a correlation plan reports (vibroseis/correlate.nim:162-188). */
typedef struct pseg_corr_info {
  int32_t dtype;
  int32_t nfft;
  int32_t shift;           /* whole samples of the pilot's start time */
  int32_t out_samples;
  int32_t pilot_samples;
  int32_t record_samples;
  double dt_ms;
  double scale;            /* normalization factor folded into the multiplier */
  size_t lds_bytes;
  int32_t threads_per_record;  /* workgroup size of the correlation kernel */
  int32_t reserved;
} pseg_corr_info;

/* initCorrelationPlan (vibroseis/correlate.nim:82-160): pilot samples in
 * float64 (Pilot[T].samples converted), the pilot's own interval and start
 * time, records of up to record_samples at dt_ms, out_samples columns (0: the
 * listen time), mode 0 cmCorrelate / 1 cmDeconvolve, eps_percent, and
 * normalize.  The pilot spectrum is computed in float64 on the host; the
 * per-bin multiplier stays on the device. */
pseg_status pseg_corr_compile(int32_t dtype, const double *pilot, int32_t pilot_len,
                              double pilot_dt_ms, double pilot_t0_ms, double dt_ms,
                              int32_t record_samples, int32_t out_samples, int32_t mode,
                              double eps_percent, int32_t normalize, pseg_corr_plan **out);
pseg_status pseg_corr_destroy(pseg_corr_plan *plan);
pseg_status pseg_corr_info_get(const pseg_corr_plan *plan, pseg_corr_info *out);

/* correlateInto(plan, src, dst, meta, ws) (vibroseis/correlate.nim:421-464)
 * and ProcessPool.correlateInto (core/pool.nim:329-347): src is
 * [rows, src_cols <= record_samples], dst [rows, out_samples]; row_len from
 * pseg_meta_flatten(for_correlation = 1); rows with -1 give zeros. */
pseg_status pseg_corr_apply(const pseg_corr_plan *plan, const void *src, int32_t rows,
                            int32_t src_cols, const int32_t *row_len, void *dst,
                            int32_t dst_cols, void *stream);

/* ------------------------------------------------------ sweep and pilot -- */

/* linearSweep + generatePilot (vibroseis/sweep.nim:185-192, 292-349;
 * pilot.nim:115-119) with tpCos2 end tapers (vibroseis/taper.nim:125-179):
 * n must be samplesInLength(length_ms, dt_ms) (core/types.nim:147-166). */
pseg_status pseg_linear_sweep(double f1_hz, double f2_hz, double length_ms, double dt_ms,
                              double taper_start_ms, double taper_end_ms,
                              double start_phase_deg, double amplitude, double *out,
                              int32_t n);

#ifdef __cplusplus
}
#endif
#endif
