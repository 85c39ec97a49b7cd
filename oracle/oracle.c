/*
 * Oracle: CPU restatement of processSEG's panel hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the product never links this.
 *
 * This file holds the type-independent parts (lengths, filter design and
 * impulse length, meta flattening, sweep generation, the ProcessPool clone)
 * and instantiates or_tmpl.h for float64 and float32.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.141592653589793
#define OR_MAX_STEPS 32
#define OR_HIST_BINS 65536
#define OR_HIST_SHIFT 15
#define OR_MAX_KEY 0xFEFF

/* Step kinds of a compiled plan (pipeline.nim:180-183), the subset restated. */
enum { ST_DEMEAN, ST_DETREND, ST_SCALE, ST_CLIP, ST_POLARITY, ST_FILTER, ST_AGC,
       ST_NORM_TRACE, ST_NORM_PANEL, ST_MUTE_TOP, ST_MUTE_BOTTOM, ST_GAIN_CURVE, ST_IIR };

static int or_seterr(char *err, int errlen, const char *msg) {
  if (err && errlen > 0) {
    strncpy(err, msg, (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
  return -1;
}

/* finitePositive (gain/gain.nim:112-113): normal or subnormal, and > 0. */
static int or_finite_positive(double x) { return isfinite(x) && x > 0; }

static const char *or_op_name(int kind) {
  static const char *names[] = {"opDemean", "opDetrend", "opScale", "opGainDb", "opClip",
                                "opPolarity", "opMuteTop", "opMuteBottom", "opFilter",
                                "opIir", "opAgc", "opNormalize", "opGainCurve",
                                "opCalibrate", "opUncalibrate"};
  if (kind < 0 || kind > 14) return "op?";
  return names[kind];
}

/* ---------------------------------------------------------------- mutes -- */

/* muteTimeMs (core/mute.nim:100-118): the mute time of a trace at `offset`
 * (|offset| used; NaN when the curve needs the offset and it is NaN). */
double or_mute_time_ms(const or_op *c, double offset) {
  if (c->mute_kind == 0) return c->mute_t_ms;
  if (c->mute_kind == 2) return c->mute_t_ms + fabs(offset) * c->mute_ms_per_dist + c->mute_pad_ms;
  const double x = fabs(offset);
  if (isnan(x)) return NAN;
  const int last = c->mute_knots - 1;
  if (x <= c->mute_offsets[0]) return c->mute_times[0];
  if (x >= c->mute_offsets[last]) return c->mute_times[last];
  int lo = 0, hi = last;  /* the last knot <= x, by bisection */
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (c->mute_offsets[mid] <= x) lo = mid;
    else hi = mid;
  }
  const double u = (x - c->mute_offsets[lo]) / (c->mute_offsets[hi] - c->mute_offsets[lo]);
  return c->mute_times[lo] + u * (c->mute_times[hi] - c->mute_times[lo]);
}

/* firstAtOrAfter / firstAfter (core/mute.nim:122-132): the first column
 * >= c (> c), clamped to 0 .. n (0 for NaN). */
static int64_t or_first_at_or_after(double c, int64_t n) {
  if (!(c > 0)) return 0;
  if (c >= (double)n) return n;
  return (int64_t)ceil(c);
}
static int64_t or_first_after(double c, int64_t n) {
  if (!(c >= 0)) return 0;
  if (c >= (double)n) return n;
  return (int64_t)floor(c) + 1;
}

/* A mute op's validation (timeMute / tableMute / lmoMute, core/mute.nim:
 * 58-98, and the taper check of pipeline.nim:417-421). */
static int or_mute_check(const or_op *op, char *msg, int msglen) {
  if (!isfinite(op->taper_ms) || op->taper_ms < 0) {
    snprintf(msg, msglen, "a mute taper must be finite and >= 0 ms, got %g", op->taper_ms);
    return -1;
  }
  if (op->mute_kind == 0) {
    if (!isfinite(op->mute_t_ms)) { snprintf(msg, msglen, "a mute time must be finite"); return -1; }
    return 0;
  }
  if (op->mute_kind == 2) {
    if (!or_finite_positive(op->mute_ms_per_dist)) { snprintf(msg, msglen, "an LMO mute velocity must be finite and positive"); return -1; }
    if (!isfinite(op->mute_t_ms) || !isfinite(op->mute_pad_ms)) { snprintf(msg, msglen, "LMO mute times must be finite"); return -1; }
    return 0;
  }
  if (op->mute_kind != 1) { snprintf(msg, msglen, "unknown mute curve %d", op->mute_kind); return -1; }
  if (op->mute_knots < 1 || op->mute_knots > OR_MAX_KNOTS) { snprintf(msg, msglen, "a mute table needs 1 .. %d knots", OR_MAX_KNOTS); return -1; }
  for (int i = 0; i < op->mute_knots; i++) {
    if (!isfinite(op->mute_offsets[i]) || op->mute_offsets[i] < 0) { snprintf(msg, msglen, "mute table offsets must be finite and >= 0"); return -1; }
    if (i > 0 && op->mute_offsets[i] <= op->mute_offsets[i - 1]) { snprintf(msg, msglen, "mute table offsets must be strictly ascending"); return -1; }
    if (!isfinite(op->mute_times[i])) { snprintf(msg, msglen, "mute table times must be finite"); return -1; }
  }
  return 0;
}

/* ---------------------------------------------------------- gain curves -- */

/* gainDbAt (gain/curves.nim:134-153): linear in dB between knots and flat
 * outside them; 20·power·log10(max(t, minMs)/refMs); dbPerS·t/1000.  NaN for
 * a NaN time or a knot curve without knots. */
double or_gain_db_at(const or_op *c, double t) {
  if (isnan(t)) return NAN;
  if (c->gain_kind == 1) return 20 * c->gain_power * log10(fmax(t, c->gain_min_ms) / c->gain_ref_ms);
  if (c->gain_kind == 2) return c->gain_db_per_s * t / 1000;
  if (c->gain_knots < 1) return NAN;
  const double *kt = c->gain_knot_ms, *kd = c->gain_knot_db;
  const int last = c->gain_knots - 1;
  if (t <= kt[0]) return kd[0];
  if (t >= kt[last]) return kd[last];
  int lo = 0, hi = last;  /* kt[lo] < t < kt[hi] */
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (kt[mid] <= t) lo = mid;
    else hi = mid;
  }
  const double f = (t - kt[lo]) / (kt[hi] - kt[lo]);
  return kd[lo] + f * (kd[hi] - kd[lo]);
}

/* factorAt (gain/curves.nim:154-160): 10^(dB/20), and (max(t, minMs)/refMs)^power
 * directly for a power curve. */
double or_gain_factor_at(const or_op *c, double t) {
  if (isnan(t)) return NAN;
  if (c->gain_kind == 1) return pow(fmax(t, c->gain_min_ms) / c->gain_ref_ms, c->gain_power);
  return pow(10.0, or_gain_db_at(c, t) / 20);
}

/* The curve checks of knotGain / powerGain / exponentialGain
 * (gain/curves.nim:100-132) and initGainCurvePlan's "no knots"
 * (gain/curves.nim:225-227). */
static int or_gain_check(const or_op *op, char *msg, int msglen) {
  if (op->gain_kind == 1) {
    if (!isfinite(op->gain_power)) { snprintf(msg, msglen, "a power must be finite, got %g", op->gain_power); return -1; }
    if (!or_finite_positive(op->gain_ref_ms)) { snprintf(msg, msglen, "refMs must be finite and positive, got %g", op->gain_ref_ms); return -1; }
    if (!or_finite_positive(op->gain_min_ms)) { snprintf(msg, msglen, "minMs must be finite and positive, got %g", op->gain_min_ms); return -1; }
    return 0;
  }
  if (op->gain_kind == 2) {
    if (!isfinite(op->gain_db_per_s)) { snprintf(msg, msglen, "dbPerS must be finite, got %g", op->gain_db_per_s); return -1; }
    return 0;
  }
  if (op->gain_kind != 0) { snprintf(msg, msglen, "unknown gain curve %d", op->gain_kind); return -1; }
  if (op->gain_knots < 1 || !op->gain_knot_ms || !op->gain_knot_db) {
    snprintf(msg, msglen, "the curve has no knots: build it with knotGain, powerGain or exponentialGain");
    return -1;
  }
  for (int i = 0; i < op->gain_knots; i++) {
    if (!isfinite(op->gain_knot_ms[i]) || !isfinite(op->gain_knot_db[i])) { snprintf(msg, msglen, "knot %d is not finite", i); return -1; }
    if (i > 0 && !(op->gain_knot_ms[i] > op->gain_knot_ms[i - 1])) {
      snprintf(msg, msglen, "knot times must strictly increase: %g ms is followed by %g ms",
               op->gain_knot_ms[i - 1], op->gain_knot_ms[i]);
      return -1;
    }
  }
  return 0;
}

/* spanOf (gain/curves.nim:340-355): where a row of n samples whose sample 0
 * is at `start` lies on a table of len factors from tmin every dt.
 * x[j0 ..< j1] takes table entry s + j (interpolated by f unless on_grid);
 * the rest of the row is off the table. */
typedef struct or_gain_span {
  int64_t s, j0, j1;
  double f;
  int on_grid;
} or_gain_span;

static int64_t or_clamp64(int64_t x, int64_t a, int64_t b) { return x < a ? a : (x > b ? b : x); }

static or_gain_span or_gain_span_of(int64_t len, double tmin, double dt, int64_t n, double start) {
  or_gain_span sp = {0, n, n, 0.0, 0};
  const double u = (start - tmin) / dt; /* the table position of x[0] */
  if (!isfinite(u) || fabs(u) > 1e15) return sp;
  const double r = round(u);
  sp.on_grid = fabs(u - r) <= 1e-9 * fmax(1.0, fabs(u));
  sp.s = (int64_t)(sp.on_grid ? r : floor(u));
  sp.f = u - (double)sp.s;
  const int64_t last = len - (sp.on_grid ? 1 : 2); /* the last j with s + j on the table */
  sp.j0 = or_clamp64(-sp.s, 0, n);
  sp.j1 = or_clamp64(last - sp.s + 1, sp.j0, n);
  return sp;
}

/* ------------------------------------------------------------- lengths -- */

/* isFastLen (fft/plan.nim:105-111). */
int or_is_fast_len(int64_t n) {
  if (n < 1) return 0;
  int64_t k = n;
  static const int ps[3] = {2, 3, 5};
  for (int i = 0; i < 3; i++)
    while (k % ps[i] == 0) k /= ps[i];
  return k == 1;
}

/* nextFastLen (fft/plan.nim:113-118). */
int64_t or_next_fast_len(int64_t n, int even) {
  int64_t r = n > 1 ? n : 1;
  while (!or_is_fast_len(r) || (even && r % 2 == 1)) r++;
  return r;
}

/* factorize (fft/plan.nim:120-132): 4s first, then a 2, 3s and 5s. */
int or_factorize(int64_t n, int32_t *radices) {
  int c = 0;
  int64_t k = n;
  while (k % 4 == 0) { radices[c++] = 4; k /= 4; }
  if (k % 2 == 0) { radices[c++] = 2; k /= 2; }
  while (k % 3 == 0) { radices[c++] = 3; k /= 3; }
  while (k % 5 == 0) { radices[c++] = 5; k /= 5; }
  return c;
}

/* selectK (gain/stats.nim:135-157): quickselect, median of three. */
static void or_select_k(double *a, int64_t lo0, int64_t hi0, int64_t k) {
  int64_t lo = lo0, hi = hi0;
  while (hi > lo) {
    int64_t mid = (lo + hi) >> 1;
    double t;
    if (a[mid] < a[lo]) { t = a[mid]; a[mid] = a[lo]; a[lo] = t; }
    if (a[hi] < a[lo]) { t = a[hi]; a[hi] = a[lo]; a[lo] = t; }
    if (a[hi] < a[mid]) { t = a[hi]; a[hi] = a[mid]; a[mid] = t; }
    double pivot = a[mid];
    int64_t i = lo, j = hi;
    while (i <= j) {
      while (a[i] < pivot) i++;
      while (a[j] > pivot) j--;
      if (i <= j) { t = a[i]; a[i] = a[j]; a[j] = t; i++; j--; }
    }
    if (k <= j) hi = j;
    else if (k >= i) lo = i;
    else return;
  }
}

/* keyOf / binLow (gain/stats.nim:208-213). */
static int or_key_of(float a) {
  uint32_t b;
  memcpy(&b, &a, 4);
  uint32_t k = b >> OR_HIST_SHIFT;
  return (int)(k < OR_MAX_KEY ? k : OR_MAX_KEY);
}

static double or_bin_low(int64_t k) {
  uint32_t b = (uint32_t)k << OR_HIST_SHIFT;
  float f;
  memcpy(&f, &b, 4);
  return (double)f;
}

/* orderStat (gain/stats.nim:237-251). */
static double or_order_stat(const uint32_t *h, int64_t j, double mx) {
  int64_t below = 0;
  for (int64_t k = 0; k <= OR_MAX_KEY; k++) {
    int64_t c = h[k];
    if (below + c > j) {
      if (k == 0) return 0;
      double lo = or_bin_low(k), hi = or_bin_low(k + 1);
      hi = hi < mx ? hi : mx;
      double v = lo + ((double)(j - below) + 0.5) / (double)c * (hi - lo);
      v = v > lo ? v : lo;
      return v < mx ? v : mx;
    }
    below += c;
  }
  return mx;
}

/* runSingle (filter/iir.nim:393-396), below with the IIR design. */
static void or_iir_run(const or_iir *p, double *buf, int64_t m);

/* ------------------------------------------------- instantiate float64 -- */

#define T double
#define SFX(x) x##_f64
#define UINT_T uint64_t
#define SINT_T int64_t
#define ABS_MASK 0x7FFFFFFFFFFFFFFFull
#define SQRT_T sqrt
#include "or_tmpl.h"
#undef T
#undef SFX
#undef UINT_T
#undef SINT_T
#undef ABS_MASK
#undef SQRT_T

/* -------------------------------------------------------------- design -- */

static int has_low(const or_spec *s) { return s->f2 > 0; }
static int has_high(const or_spec *s) { return s->f3 < INFINITY; }

/* warp and ramp (filter/design.nim:227-234). */
static double or_warp(double f, double dt_ms) {
  double x = OR_PI * f * dt_ms / 1000;
  return x >= OR_PI / 2 ? INFINITY : tan(x);
}

static double or_ramp(double x) {
  if (x <= 0) return 0.0;
  if (x >= 1) return 1.0;
  double s = sin(OR_PI / 2 * x);
  return s * s;
}

/* response (filter/design.nim:236-255). */
double or_response(const or_spec *spec, double f, double dt_ms) {
  double r = 1.0;
  if (spec->kind == 0) {
    if (has_low(spec)) r *= or_ramp((f - spec->f1) / (spec->f2 - spec->f1));
    if (has_high(spec)) r *= 1 - or_ramp((f - spec->f3) / (spec->f4 - spec->f3));
  } else {
    double wf = or_warp(f, dt_ms);
    if (spec->low_hz > 0)
      r *= f <= 0 ? 0.0 : 1 / (1 + pow(or_warp(spec->low_hz, dt_ms) / wf, spec->low_slope / 6));
    if (spec->high_hz < INFINITY)
      r *= wf == INFINITY ? 0.0
                          : 1 / (1 + pow(wf / or_warp(spec->high_hz, dt_ms), spec->high_slope / 6));
  }
  for (int i = 0; i < spec->n_notch; i++)
    r *= or_ramp(fabs(f - spec->notch_hz[i]) / (spec->notch_width[i] / 2));
  return r;
}

/* narrowestRampHz (filter/design.nim:267-282). */
double or_narrowest_ramp_hz(const or_spec *spec) {
  double r = INFINITY;
  if (spec->kind == 0) {
    if (has_low(spec)) r = fmin(r, spec->f2 - spec->f1);
    if (has_high(spec)) r = fmin(r, spec->f4 - spec->f3);
  } else {
    if (spec->low_hz > 0) r = fmin(r, spec->low_hz * fmin(1.0, 24 / spec->low_slope) / 2);
    if (spec->high_hz < INFINITY) r = fmin(r, spec->high_hz * fmin(1.0, 24 / spec->high_slope) / 2);
  }
  for (int i = 0; i < spec->n_notch; i++) r = fmin(r, spec->notch_width[i] / 2);
  return r;
}

/* validate (filter/design.nim:100-173). */
static int or_validate(const or_spec *s, double dt_ms, char *err, int errlen) {
  if (!(dt_ms > 0 && dt_ms < INFINITY)) return or_seterr(err, errlen, "dtMs must be finite and > 0");
  double nyq = 500 / dt_ms;
  if (s->n_notch < 0 || s->n_notch > OR_MAX_NOTCH) return or_seterr(err, errlen, "too many notches");
  if (s->kind == 0) {
    double c[4] = {s->f1, s->f2, s->f3, s->f4};
    for (int i = 0; i < 4; i++) if (isnan(c[i])) return or_seterr(err, errlen, "a corner is NaN");
    if (has_low(s) && !(s->f1 >= 0 && s->f1 < s->f2)) return or_seterr(err, errlen, "a low cut needs 0 <= f1 < f2");
    if (!has_low(s) && (s->f1 != 0 || s->f2 != 0)) return or_seterr(err, errlen, "no low cut needs f1 = f2 = 0");
    if (has_high(s) && !(s->f3 < s->f4 && s->f4 < INFINITY)) return or_seterr(err, errlen, "a high cut needs f3 < f4 < Inf");
    if (!has_high(s) && s->f4 != INFINITY) return or_seterr(err, errlen, "no high cut needs f3 = f4 = Inf");
    if (has_low(s) && has_high(s) && !(s->f2 <= s->f3)) return or_seterr(err, errlen, "a band-pass needs f2 <= f3");
    if (has_high(s) && s->f4 > nyq) return or_seterr(err, errlen, "f4 is above Nyquist");
    if (has_low(s) && s->f2 >= nyq) return or_seterr(err, errlen, "f2 is not below Nyquist");
  } else {
    if (!(s->low_hz >= 0 && s->low_hz < INFINITY)) return or_seterr(err, errlen, "a low-cut corner must be finite and >= 0");
    if (!(s->high_hz > 0)) return or_seterr(err, errlen, "a high-cut corner must be > 0");
    if (s->low_hz > 0 && !(s->low_slope > 0 && s->low_slope < INFINITY)) return or_seterr(err, errlen, "a slope must be finite and > 0");
    if (s->high_hz < INFINITY && !(s->high_slope > 0 && s->high_slope < INFINITY)) return or_seterr(err, errlen, "a slope must be finite and > 0");
    if (!(s->low_hz < s->high_hz)) return or_seterr(err, errlen, "the low-cut corner must be below the high-cut one");
    if (s->high_hz < INFINITY && s->high_hz >= nyq) return or_seterr(err, errlen, "the high-cut corner is not below Nyquist");
    if (s->low_hz >= nyq) return or_seterr(err, errlen, "the low-cut corner is not below Nyquist");
  }
  for (int i = 0; i < s->n_notch; i++) {
    if (!(s->notch_hz[i] > 0 && s->notch_hz[i] < INFINITY)) return or_seterr(err, errlen, "a notch frequency must be finite and > 0");
    if (!(s->notch_width[i] > 0 && s->notch_width[i] < INFINITY)) return or_seterr(err, errlen, "a notch width must be finite and > 0");
    if (s->notch_hz[i] >= nyq) return or_seterr(err, errlen, "a notch is not below Nyquist");
  }
  return 0;
}

/* ---------------------------------------------------------------- IIR -- */

static or_biquad or_biquad_of(double b0, double b1, double b2, double a1, double a2) {
  or_biquad q = {b0, b1, b2, a1, a2, (b0 + b1 + b2) / (1 + a1 + a2)};
  return q;
}

/* butterSections (filter/iir.nim:113-148): `order` poles with the -3 dB
 * corner at fc Hz by the prewarped bilinear transform; high: a high-pass
 * (gain 1 at Nyquist), else a low-pass (gain 1 at 0 Hz).  Appends to sec. */
static int or_butter_sections(double fc, double dt_ms, int order, int high, or_biquad *sec) {
  const double fs = 1000 / dt_ms;
  const double wc = 2 * fs * tan(OR_PI * fc / fs);
  const double k2 = 2 * fs;
  int n = 0;
  for (int k = 0; k < order / 2; k++) {
    const double th = OR_PI * (double)(2 * k + 1) / (double)(2 * order);
    const double sr = -wc * sin(th), si = wc * cos(th);   /* analog pole s */
    /* z = (k2 + s) / (k2 - s) */
    const double nr = k2 + sr, ni = si, dr = k2 - sr, di = -si;
    const double den = dr * dr + di * di;
    const double zr = (nr * dr + ni * di) / den, zi = (ni * dr - nr * di) / den;
    const double a1 = -2 * zr, a2 = zr * zr + zi * zi;
    if (high) {
      const double g = (1 - a1 + a2) / 4;
      sec[n++] = or_biquad_of(g, -2 * g, g, a1, a2);
    } else {
      const double g = (1 + a1 + a2) / 4;
      sec[n++] = or_biquad_of(g, 2 * g, g, a1, a2);
    }
  }
  if (order % 2 == 1) {
    const double z = (k2 - wc) / (k2 + wc);
    const double a1 = -z;
    if (high) {
      const double g = (1 - a1) / 2;
      sec[n++] = or_biquad_of(g, -g, 0, a1, 0);
    } else {
      const double g = (1 + a1) / 2;
      sec[n++] = or_biquad_of(g, g, 0, a1, 0);
    }
  }
  return n;
}

/* causalLength (filter/iir.nim:150-182): the samples after a spike that hold
 * 99.99 % of the energy of the causal impulse response, or of its complement,
 * whichever is longer. */
static int64_t or_causal_length(const or_biquad *sec, int ns) {
  int64_t n = 1024;
  for (;;) {
    double *h = (double *)calloc((size_t)n, sizeof(double));
    h[0] = 1;
    for (int c = 0; c < ns; c++) {
      double s1 = 0, s2 = 0;
      const or_biquad q = sec[c];
      for (int64_t j = 0; j < n; j++) {
        const double x = h[j];
        const double y = q.b0 * x + s1;
        s1 = q.b1 * x - q.a1 * y + s2;
        s2 = q.b2 * x - q.a2 * y;
        h[j] = y;
      }
    }
    double rest = 0, tail = 0;
    for (int64_t j = 1; j < n; j++) {
      rest += h[j] * h[j];
      if (j >= n - n / 4) tail += h[j] * h[j];
    }
    if (tail <= 1e-12 * fmax(rest, 1e-300) || n >= ((int64_t)1 << 24)) {
      int64_t best = 0;
      const double cs[2] = {h[0], 1 - h[0]};
      for (int i = 0; i < 2; i++) {
        const double total = cs[i] * cs[i] + rest;
        double e = cs[i] * cs[i];
        int64_t L = 0;
        while (e < 0.9999 * total && L < n - 1) {
          L++;
          e += h[L] * h[L];
        }
        if (L > best) best = L;
      }
      free(h);
      return best;
    }
    free(h);
    n *= 2;
  }
}

/* poles (filter/iir.nim:184-191). */
static int or_iir_poles(double slope, double per_pole, const char *side, char *err, int errlen) {
  const double n = slope / per_pole;
  if (!(n >= 1 && n <= 32 && n == round(n))) {
    char msg[200];
    snprintf(msg, sizeof msg, "the %s slope of an IIR filter must be a multiple of %g dB/octave, "
             "1 .. 32 poles; got %g", side, per_pole, slope);
    or_seterr(err, errlen, msg);
    return -1;
  }
  return (int)n;
}

/* initIirPlan (filter/iir.nim:193-239). */
int or_iir_init(or_iir *p, const or_spec *spec, double dt_ms, int64_t samples, int32_t phase,
                int32_t trend, char *err, int errlen) {
  memset(p, 0, sizeof(*p));
  if (samples < 1) return or_seterr(err, errlen, "an IIR plan needs samples >= 1");
  if (or_validate(spec, dt_ms, err, errlen)) return -1;
  if (spec->kind != 1)
    return or_seterr(err, errlen, "an IIR filter needs a Butterworth spec (butterLowCut, butterHighCut, butterBandPass)");
  if (spec->n_notch > 0 || spec->zero_dc)
    return or_seterr(err, errlen, "an IIR filter takes no notches and no 0 Hz suppression");
  const double per_pole = phase == 0 ? 12.0 : 6.0;
  or_biquad low[OR_MAX_SECTIONS], high[OR_MAX_SECTIONS];
  int nl = 0, nh = 0;
  if (spec->low_hz > 0) {
    int np = or_iir_poles(spec->low_slope, per_pole, "low-cut", err, errlen);
    if (np < 0) return -1;
    nl = or_butter_sections(spec->low_hz, dt_ms, np, 1, low);
  }
  if (spec->high_hz < INFINITY) {
    int np = or_iir_poles(spec->high_slope, per_pole, "high-cut", err, errlen);
    if (np < 0) return -1;
    nh = or_butter_sections(spec->high_hz, dt_ms, np, 0, high);
  }
  for (int i = 0; i < nl; i++) p->sec[p->n_sections++] = low[i];
  for (int i = 0; i < nh; i++) p->sec[p->n_sections++] = high[i];
  p->phase = phase;
  if (phase == 0) {
    if (or_impulse_length(spec, dt_ms, &p->pad, err, errlen)) return -1;
  } else {
    int64_t a = or_causal_length(low, nl), b = or_causal_length(high, nh);
    p->pad = a > b ? a : b;
  }
  p->samples = samples;
  p->dt_ms = dt_ms;
  /* G(1) and G'(1) of the cascade from each section's N(q)/D(q) */
  double g = 1, dg = 0;
  for (int i = 0; i < p->n_sections; i++) {
    const or_biquad c = p->sec[i];
    const double n1 = c.b0 + c.b1 + c.b2, d1 = 1 + c.a1 + c.a2;
    const double gi = n1 / d1;
    const double dgi = ((c.b1 + 2 * c.b2) * d1 - n1 * (c.a1 + 2 * c.a2)) / (d1 * d1);
    dg = dg * gi + g * dgi;
    g *= gi;
  }
  p->dc_gain = phase == 0 ? g * g : g;
  p->dc_slope = phase == 0 ? 0.0 : dg;
  p->trend = trend;
  p->reflect = fabs(p->dc_gain) >= 0.5 || trend != 2;
  p->lead = p->reflect ? p->pad : 0;
  p->tail = phase == 0 ? p->pad : 0;
  return 0;
}

double or_iir_magnitude_at(const or_iir *p, double f) {
  const double w = 2 * OR_PI * f * p->dt_ms / 1000;
  const double zr = cos(w), zi = -sin(w);        /* z^-1 */
  const double z2r = zr * zr - zi * zi, z2i = 2 * zr * zi;
  double hr = 1, hi = 0;
  for (int i = 0; i < p->n_sections; i++) {
    const or_biquad s = p->sec[i];
    const double nr = s.b0 + s.b1 * zr + s.b2 * z2r, ni = s.b1 * zi + s.b2 * z2i;
    const double dr = 1 + s.a1 * zr + s.a2 * z2r, di = s.a1 * zi + s.a2 * z2i;
    const double den = dr * dr + di * di;
    const double qr = (nr * dr + ni * di) / den, qi = (ni * dr - nr * di) / den;
    const double tr = hr * qr - hi * qi;
    hi = hr * qi + hi * qr;
    hr = tr;
  }
  const double m = hypot(hr, hi);
  return p->phase == 0 ? m * m : m;
}

int32_t or_iir_order(const or_iir *p) {
  int32_t o = 0;
  for (int i = 0; i < p->n_sections; i++) o += p->sec[i].a2 == 0 ? 1 : 2;
  return o;
}

/* sosKernel (filter/iir.nim:319-351) on one trace of m samples in place: a
 * steady-state start on the pass's first input or a start from rest, and
 * states below 1e-250 set to 0 every 256 steps. */
static void or_sos_pass(double *buf, int64_t m, const or_biquad *c, int steady, int backward) {
  double s1 = 0, s2 = 0;
  if (steady && m > 0) {
    const double u = buf[backward ? m - 1 : 0];
    const double y = c->g * u;
    s2 = c->b2 * u - c->a2 * y;
    s1 = y - c->b0 * u;
  }
  int64_t done = 0;
  while (done < m) {
    const int64_t run = m - done < 256 ? m - done : 256;
    for (int64_t k = 0; k < run; k++) {
      const int64_t j = backward ? m - 1 - done - k : done + k;
      const double x = buf[j];
      const double y = c->b0 * x + s1;
      s1 = c->b1 * x - c->a1 * y + s2;
      s2 = c->b2 * x - c->a2 * y;
      buf[j] = y;
    }
    s1 = fabs(s1) < 1e-250 ? 0.0 : s1;
    s2 = fabs(s2) < 1e-250 ? 0.0 : s2;
    done += run;
  }
}

/* runSingle (filter/iir.nim:393-396). */
static void or_iir_run(const or_iir *p, double *buf, int64_t m) {
  for (int i = 0; i < p->n_sections; i++) or_sos_pass(buf, m, &p->sec[i], p->reflect, 0);
  if (p->phase == 0)
    for (int i = 0; i < p->n_sections; i++) or_sos_pass(buf, m, &p->sec[i], p->reflect, 1);
}

/* energyLength (filter/design.nim:288-298). */
static int64_t or_energy_length(const double *h, int64_t n, double centre) {
  double total = centre * centre;
  for (int64_t j = 1; j < n; j++) total += h[j] * h[j];
  double e = centre * centre;
  int64_t L = 0;
  while (e < 0.999 * total && L < n / 2) {
    L++;
    e += h[L] * h[L] + (n - L != L ? h[n - L] * h[n - L] : 0.0);
  }
  return L;
}

#define OR_MAX_GRID (1 << 24)

/* featureLength (filter/design.nim:300-325). */
static int or_feature_length(const or_spec *spec, double dt_ms, int64_t *L, char *err, int errlen) {
  double feature = or_narrowest_ramp_hz(spec);
  if (feature == INFINITY) { *L = 0; return 0; }
  double fs = 1000 / dt_ms;
  double start = ceil(16 * fs / feature);
  int64_t n = start > (double)OR_MAX_GRID ? (int64_t)OR_MAX_GRID + 1
                                          : or_next_fast_len((int64_t)fmax(64, start), 1);
  for (;;) {
    if (n > OR_MAX_GRID) return or_seterr(err, errlen, "the filter's impulse response is too long");
    fftplan_f64 p;
    fft_init_f64(&p, n);
    int64_t bins = n / 2 + 1;
    double *re = (double *)malloc(sizeof(double) * (size_t)bins);
    double *im = (double *)calloc((size_t)bins, sizeof(double));
    double *h = (double *)malloc(sizeof(double) * (size_t)n);
    double *work = (double *)malloc(sizeof(double) * (size_t)(4 * p.m));
    double df = 1000 / ((double)n * dt_ms);
    for (int64_t k = 0; k < bins; k++) re[k] = or_response(spec, (double)k * df, dt_ms);
    if (spec->zero_dc) re[0] = 0;
    fft_inverse_f64(&p, re, im, h, work);
    int64_t a = or_energy_length(h, n, h[0]), b = or_energy_length(h, n, 1 - h[0]);
    free(re); free(im); free(h); free(work);
    fft_free_f64(&p);
    int64_t l = a > b ? a : b;
    if (l <= n / 8) { *L = l; return 0; }
    n = or_next_fast_len(2 * n, 1);
  }
}

/* impulseLength + features (filter/design.nim:327-366). */
int or_impulse_length(const or_spec *spec, double dt_ms, int64_t *L, char *err, int errlen) {
  if (or_validate(spec, dt_ms, err, errlen)) return -1;
  or_spec feats[2 + OR_MAX_NOTCH];
  int nf = 0;
  or_spec flat;
  memset(&flat, 0, sizeof(flat));
  flat.kind = 0; flat.f1 = 0; flat.f2 = 0; flat.f3 = INFINITY; flat.f4 = INFINITY;
  if (spec->kind == 0) {
    if (has_low(spec)) { feats[nf] = flat; feats[nf].f1 = spec->f1; feats[nf].f2 = spec->f2; nf++; }
    if (has_high(spec)) { feats[nf] = flat; feats[nf].f3 = spec->f3; feats[nf].f4 = spec->f4; nf++; }
  } else {
    or_spec b;
    memset(&b, 0, sizeof(b));
    b.kind = 1;
    if (spec->low_hz > 0) {
      feats[nf] = b; feats[nf].low_hz = spec->low_hz; feats[nf].low_slope = spec->low_slope;
      feats[nf].high_hz = INFINITY; feats[nf].high_slope = spec->high_slope; nf++;
    }
    if (spec->high_hz < INFINITY) {
      feats[nf] = b; feats[nf].low_hz = 0; feats[nf].low_slope = spec->low_slope;
      feats[nf].high_hz = spec->high_hz; feats[nf].high_slope = spec->high_slope; nf++;
    }
  }
  for (int i = 0; i < spec->n_notch; i++) {
    feats[nf] = flat;
    feats[nf].n_notch = 1;
    feats[nf].notch_hz[0] = spec->notch_hz[i];
    feats[nf].notch_width[0] = spec->notch_width[i];
    nf++;
  }
  int64_t best = 0;
  for (int i = 0; i < nf; i++) {
    int64_t l;
    if (or_feature_length(&feats[i], dt_ms, &l, err, errlen)) return -1;
    best = l > best ? l : best;
  }
  *L = best;
  return 0;
}

/* initFilterPlan's integer decisions (filter/apply.nim:118-137). */
int or_filter_info_of(const or_spec *spec, double dt_ms, int64_t samples, int32_t trend,
                      or_filter_info *out, char *err, int errlen) {
  if (samples < 1) return or_seterr(err, errlen, "a filter plan needs samples >= 1");
  int64_t margin;
  if (or_impulse_length(spec, dt_ms, &margin, err, errlen)) return -1;
  double dc = or_response(spec, 0, dt_ms);
  int reflect = dc >= 0.5 || trend != 2;
  int64_t pad = reflect ? margin : (margin < samples ? margin : samples);
  int64_t n = samples + (reflect ? 4 : 2) * pad;
  double ramp = or_narrowest_ramp_hz(spec);
  if (ramp < INFINITY && pad == margin) {
    int64_t b = (int64_t)ceil(2000 / (ramp * dt_ms));
    n = n > b ? n : b;
  }
  n = or_next_fast_len(n, 1);
  out->margin = margin;
  out->pad = pad;
  out->nfft = n;
  out->reflect = reflect;
  out->zero_dc = spec->zero_dc;
  out->dc_gain = dc;
  return 0;
}

/* ------------------------------------------------- instantiate float32 -- */

#define T float
#define SFX(x) x##_f32
#define UINT_T uint32_t
#define SINT_T int32_t
#define ABS_MASK 0x7FFFFFFFu
#define SQRT_T sqrtf
#include "or_tmpl.h"
#undef T
#undef SFX
#undef UINT_T
#undef SINT_T
#undef ABS_MASK
#undef SQRT_T

/* ---------------------------------------------------------------- meta -- */

int or_meta_flatten(int32_t rows, int32_t cols, int64_t first_sample,
                    const int64_t *num_samples, const uint8_t *status,
                    const int64_t *failed, int32_t n_failed, int used_for_corr,
                    int32_t *row_len) {
  for (int32_t r = 0; r < rows; r++) {
    /* isFailed: binary search in the ascending list (core/types.nim:212-219) */
    int32_t lo = 0, hi = n_failed;
    while (lo < hi) {
      int32_t mid = (lo + hi) >> 1;
      if (failed[mid] < r) lo = mid + 1; else hi = mid;
    }
    int is_failed = lo < n_failed && failed[lo] == r;
    int ok = used_for_corr ? (!is_failed && status[r] != 2)
                           : (!is_failed && (status[r] == 0 || status[r] == 1));
    int64_t live = num_samples[r] - first_sample;
    live = live < 0 ? 0 : (live > cols ? cols : live);
    row_len[r] = ok ? (int32_t)live : -1;
  }
  return 0;
}

/* ------------------------------------------------------------ plan info -- */

int or_plan_info_of(const or_op *ops, int32_t n_ops, int32_t dtype, double dt_ms,
                    int64_t samples, int64_t record_samples, or_plan_info *out,
                    char *err, int errlen) {
  memset(out, 0, sizeof(*out));
  if (dtype == 0) {
    plan_f32 *P = (plan_f32 *)malloc(sizeof(plan_f32));
    if (plan_compile_f32(P, ops, n_ops, dt_ms, samples, record_samples, err, errlen)) { free(P); return -1; }
    out->margin = P->margin; out->parts = P->n_parts; out->n_steps = P->n_steps;
    for (int i = 0; i < P->n_steps && i < 16; i++) {
      if (P->steps[i].kind == ST_FILTER) out->nfft[i] = P->steps[i].filter.fft.n;
      if (P->steps[i].kind == ST_AGC) out->half[i] = P->steps[i].agc.half;
    }
    plan_free_f32(P);
    free(P);
  } else {
    plan_f64 *P = (plan_f64 *)malloc(sizeof(plan_f64));
    if (plan_compile_f64(P, ops, n_ops, dt_ms, samples, record_samples, err, errlen)) { free(P); return -1; }
    out->margin = P->margin; out->parts = P->n_parts; out->n_steps = P->n_steps;
    for (int i = 0; i < P->n_steps && i < 16; i++) {
      if (P->steps[i].kind == ST_FILTER) out->nfft[i] = P->steps[i].filter.fft.n;
      if (P->steps[i].kind == ST_AGC) out->half[i] = P->steps[i].agc.half;
    }
    plan_free_f64(P);
    free(P);
  }
  return 0;
}

int or_corr_info_of(const double *pilot, int64_t pilot_len, double pilot_dt_ms,
                    double pilot_t0_ms, double dt_ms, int64_t record_samples,
                    int64_t out_samples, int32_t mode, double eps_percent,
                    int32_t normalize, or_corr_info *out, char *err, int errlen) {
  corrplan_f64 c;
  if (corr_init_f64(&c, pilot, pilot_len, pilot_dt_ms, pilot_t0_ms, dt_ms, record_samples,
                    out_samples, mode, eps_percent, normalize, err, errlen))
    return -1;
  out->nfft = c.fft.n;
  out->shift = c.shift;
  out->out_samples = c.out;
  out->pilot_samples = c.pilot;
  out->record_samples = c.record;
  out->scale = c.scale;
  corr_free_f64(&c);
  return 0;
}

/* ------------------------------------------------------ sweep / pilot -- */

/* samplesInLength (core/types.nim:147-166). */
int64_t or_samples_in_length(double length_ms, double dt_ms) {
  if (!(length_ms > 0 && length_ms < INFINITY && dt_ms > 0 && dt_ms < INFINITY)) return 0;
  double n = length_ms / dt_ms;
  if (n > 2147483647.0) return 0;
  double r = round(n);
  return fabs(n - r) <= 1e-6 ? (int64_t)r : (int64_t)floor(n);
}

/* tpCos2 taper weight (vibroseis/taper.nim:125-145). */
static double or_cos2_weight(double x) {
  x = x < 0 ? 0 : (x > 1 ? 1 : x);
  return 0.5 - 0.5 * cos(OR_PI * x);
}

int or_linear_sweep(double f1, double f2, double length_ms, double dt_ms, double ts,
                    double te, double phase_deg, double amp, double *x, int64_t n,
                    char *err, int errlen) {
  /* validate (vibroseis/sweep.nim:132-183), the parts a linear sweep uses */
  if (!(dt_ms > 0 && dt_ms < INFINITY)) return or_seterr(err, errlen, "dtMs must be finite and positive");
  if (!(length_ms >= dt_ms && length_ms < INFINITY)) return or_seterr(err, errlen, "lengthMs must be at least one sample interval");
  double steps = length_ms / dt_ms;
  if (fabs(steps - round(steps)) > 1e-6) return or_seterr(err, errlen, "lengthMs must be a whole number of samples");
  double nyq = 500 / dt_ms, limit = 0.8 * nyq;
  double fs[2] = {f1, f2};
  for (int i = 0; i < 2; i++) {
    if (!(fs[i] > 0 && fs[i] < INFINITY)) return or_seterr(err, errlen, "a sweep frequency must be > 0");
    if (fs[i] > nyq) return or_seterr(err, errlen, "a sweep frequency is above Nyquist");
    if (fs[i] > limit * (1 + 1e-12)) return or_seterr(err, errlen, "a sweep frequency is above the anti-alias limit");
  }
  double tsl = isnan(ts) ? 0 : ts, tel = isnan(te) ? 0 : te;
  if (tsl < 0 || tel < 0) return or_seterr(err, errlen, "a taper length must be >= 0");
  if (tsl + tel > length_ms * (1 + 1e-9)) return or_seterr(err, errlen, "the sweep tapers overlap");
  int64_t len = or_samples_in_length(length_ms, dt_ms);
  if (n != len) return or_seterr(err, errlen, "the buffer is not the sweep's length");
  /* phasesInto (slLinear) + sweepInto (vibroseis/sweep.nim:292-342) */
  double dt = dt_ms / 1000, Tl = length_ms / 1000, rate = (f2 - f1) / Tl;
  double p0 = phase_deg * OR_PI / 180;
  for (int64_t j = 0; j < n; j++) {
    double t = (double)j * dt;
    double ph = 2 * OR_PI * (f1 * t + 0.5 * rate * t * t);
    x[j] = amp * sin(p0 + ph);
  }
  /* applyTaper (vibroseis/taper.nim:160-179) */
  double sweep_ms = (double)n * dt_ms;
  if (tsl > 0) {
    int64_t j = 0;
    while (j < n && (double)j * dt_ms < tsl) {
      x[j] = x[j] * or_cos2_weight((double)j * dt_ms / tsl);
      j++;
    }
  }
  if (tel > 0) {
    int64_t j = n - 1;
    while (j >= 0 && sweep_ms - (double)j * dt_ms < tel) {
      x[j] = x[j] * or_cos2_weight((sweep_ms - (double)j * dt_ms) / tel);
      j--;
    }
  }
  return 0;
}

/* ------------------------------------------------------- ProcessPool -- */

/* A clone of core/pool.nim's scheduling (:71-308) for the CPU baseline:
 * persistent workers sleeping on a condition variable, one job at a time,
 * rows handed out in runs of about 256 KiB (a multiple of 8 rows) from an
 * atomic counter, one workspace per worker. */
struct or_pool {
  int32_t n;
  pthread_t *threads;
  pthread_mutex_t mu;
  pthread_cond_t wake, done;
  int64_t job_id;
  int32_t running;
  int quit;
  void (*job)(struct or_pool *, int32_t worker);
  /* the apply job */
  const plan_f32 *plan;
  const float *src;
  float *dst;
  int32_t rows, cols;
  const int32_t *row_len;
  int part;
  double factor;
  atomic_long next;
  int64_t grain;
  void **scratch;
};

typedef struct { or_pool *p; int32_t id; } or_worker_arg;

static void *or_worker_main(void *arg) {
  or_worker_arg *wa = (or_worker_arg *)arg;
  or_pool *p = wa->p;
  int32_t id = wa->id;
  free(wa);
  int64_t seen = 0;
  pthread_mutex_lock(&p->mu);
  for (;;) {
    while (!p->quit && p->job_id == seen) pthread_cond_wait(&p->wake, &p->mu);
    if (p->quit) break;
    seen = p->job_id;
    pthread_mutex_unlock(&p->mu);
    p->job(p, id);
    pthread_mutex_lock(&p->mu);
    if (--p->running == 0) pthread_cond_signal(&p->done);
  }
  pthread_mutex_unlock(&p->mu);
  return NULL;
}

or_pool *or_pool_create(int32_t threads) {
  if (threads < 1) threads = 1;
  or_pool *p = (or_pool *)calloc(1, sizeof(or_pool));
  p->n = threads;
  pthread_mutex_init(&p->mu, NULL);
  pthread_cond_init(&p->wake, NULL);
  pthread_cond_init(&p->done, NULL);
  p->threads = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  p->scratch = (void **)calloc((size_t)threads, sizeof(void *));
  for (int32_t i = 0; i < threads; i++) {
    or_worker_arg *wa = (or_worker_arg *)malloc(sizeof(or_worker_arg));
    wa->p = p;
    wa->id = i;
    pthread_create(&p->threads[i], NULL, or_worker_main, wa);
  }
  return p;
}

void or_pool_destroy(or_pool *p) {
  if (!p) return;
  pthread_mutex_lock(&p->mu);
  p->quit = 1;
  pthread_cond_broadcast(&p->wake);
  pthread_mutex_unlock(&p->mu);
  for (int32_t i = 0; i < p->n; i++) pthread_join(p->threads[i], NULL);
  for (int32_t i = 0; i < p->n; i++) free(p->scratch[i]);
  free(p->scratch);
  free(p->threads);
  pthread_mutex_destroy(&p->mu);
  pthread_cond_destroy(&p->wake);
  pthread_cond_destroy(&p->done);
  free(p);
}

/* run (core/pool.nim:172-183): wake every worker, wait for all. */
static void or_pool_run(or_pool *p, void (*job)(or_pool *, int32_t)) {
  pthread_mutex_lock(&p->mu);
  p->job = job;
  p->running = p->n;
  p->job_id++;
  pthread_cond_broadcast(&p->wake);
  while (p->running > 0) pthread_cond_wait(&p->done, &p->mu);
  pthread_mutex_unlock(&p->mu);
}

/* partJob (core/pool.nim:239-249). */
static void or_part_job(or_pool *p, int32_t worker) {
  void *scratch = p->scratch[worker];
  for (;;) {
    long first = atomic_fetch_add(&p->next, (long)p->grain);
    if (first >= p->rows) break;
    int32_t last = (int32_t)(first + p->grain < p->rows ? first + p->grain : p->rows);
    run_part_rows_f32(p->plan, p->part, p->part == 0 ? p->src : NULL, p->dst, (int32_t)first,
                      last, p->cols, p->row_len, p->factor, p->grain, scratch, NULL, 0);
  }
}

/* panelJob (core/pool.nim:251-260): worker 0 computes the factor. */
static void or_panel_job(or_pool *p, int32_t worker) {
  if (worker != 0) return;
  const plan_f32 *P = p->plan;
  p->factor = panel_factor_f32(&P->steps[P->part_a[p->part]], p->part == 0 ? p->src : p->dst,
                               p->rows, p->cols, p->row_len);
}

int or_pool_apply_f32(or_pool *pool, const or_op *ops, int32_t n_ops, double dt_ms,
                      int64_t samples, const float *src, float *dst, int32_t rows,
                      int32_t cols, const int32_t *row_len, int32_t repeats,
                      char *err, int errlen) {
  plan_f32 *P = (plan_f32 *)malloc(sizeof(plan_f32));
  if (plan_compile_f32(P, ops, n_ops, dt_ms, samples, 0, err, errlen)) { free(P); return -1; }
  if (check_apply_f32(P, cols, err, errlen)) { plan_free_f32(P); free(P); return -1; }
  for (int32_t i = 0; i < pool->n; i++) {
    free(pool->scratch[i]);
    pool->scratch[i] = malloc(sizeof(double) * (size_t)(P->scratch + 8));
  }
  pool->plan = P;
  pool->src = src;
  pool->dst = dst;
  pool->rows = rows;
  pool->cols = cols;
  pool->row_len = row_len;
  /* reset (core/pool.nim:210-216): runs of runBytes, whole batches of 8 */
  int64_t g = (256 * 1024) / ((int64_t)(cols > 0 ? cols : 1) * 4) / 8 * 8;
  pool->grain = g > 8 ? g : 8;
  for (int32_t rep = 0; rep < (repeats > 0 ? repeats : 1); rep++) {
    for (int part = 0; part < P->n_parts; part++) {
      pool->part = part;
      pool->factor = NAN;
      int a = P->part_a[part];
      if (P->part_b[part] > a && P->steps[a].kind == ST_NORM_PANEL) or_pool_run(pool, or_panel_job);
      atomic_store(&pool->next, 0);
      or_pool_run(pool, or_part_job);
    }
  }
  plan_free_f32(P);
  free(P);
  return 0;
}

/* ---------------------------------------------------------- first breaks -- */

static int or_fb_positive(double x) { return x > 0 && x < INFINITY; }

/* samplesIn (qc/firstbreak.nim:189-192): a window in samples, 1 .. samples + 1. */
static int64_t or_fb_samples_in(double ms, double dt, int64_t samples) {
  double v = round(ms / dt);
  if (v < 1) v = 1;
  if (v > (double)(samples + 1)) v = (double)(samples + 1);
  return (int64_t)v;
}

int or_fb_init(or_fb_plan *P, const or_fb_params *p, double dt, int64_t samples, char *err,
               int errlen) {
  memset(P, 0, sizeof(*P));
  if (!or_fb_positive(dt)) return or_seterr(err, errlen, "dtMs must be finite and positive");
  if (samples < 1) return or_seterr(err, errlen, "a first-break plan needs samples >= 1");
  if (!(or_fb_positive(p->window_ms) && or_fb_positive(p->lta_ms) && or_fb_positive(p->snr_window_ms)))
    return or_seterr(err, errlen, "first-break windows must be finite and positive");
  if (p->lta_ms < p->window_ms) return or_seterr(err, errlen, "ltaMs must be >= windowMs");
  if (!(p->stabilization >= 0 && p->stabilization < INFINITY))
    return or_seterr(err, errlen, "stabilization must be finite and >= 0");
  if (!(p->low_db < p->high_db) || p->low_db == -INFINITY || p->high_db == INFINITY)
    return or_seterr(err, errlen, "need finite lowDb < highDb");
  if (p->search_end_ms < p->search_start_ms) return or_seterr(err, errlen, "searchEndMs < searchStartMs");
  if (p->search_start_ms == INFINITY || p->search_end_ms == -INFINITY)
    return or_seterr(err, errlen, "the search window is empty");
  if (p->method < 0 || p->method > 2) return or_seterr(err, errlen, "unknown pick method");
  if (!isnan(p->guide_velocity)) {
    if (!or_fb_positive(p->guide_velocity))
      return or_seterr(err, errlen, "guideVelocity must be finite and positive (or NaN)");
    if (!(p->before_ms >= 0 && p->after_ms >= 0 && p->before_ms < INFINITY && p->after_ms < INFINITY))
      return or_seterr(err, errlen, "beforeMs and afterMs must be finite and >= 0");
    if (!isfinite(p->guide_t0_ms)) return or_seterr(err, errlen, "guideT0Ms must be finite");
  }
  P->p = *p;
  P->dt_ms = dt;
  P->samples = samples;
  P->w = or_fb_samples_in(p->window_ms, dt, samples);
  P->lta = or_fb_samples_in(p->lta_ms, dt, samples);
  P->snr = or_fb_samples_in(p->snr_window_ms, dt, samples);
  if (!(p->low_cut_hz >= 0 && p->high_cut_hz > p->low_cut_hz))
    return or_seterr(err, errlen, "need 0 <= lowCutHz < highCutHz");
  P->filtered = p->low_cut_hz > 0 || p->high_cut_hz < INFINITY;
  P->whole = p->method == 2 ||
             (isnan(p->guide_velocity) && isnan(p->search_start_ms) && isnan(p->search_end_ms));
  if (P->filtered) {
    or_spec spec;
    memset(&spec, 0, sizeof spec);
    spec.kind = 1;
    spec.f3 = spec.f4 = INFINITY;
    spec.low_hz = p->low_cut_hz;
    spec.low_slope = p->slope_db_per_octave;
    spec.high_hz = p->high_cut_hz;
    spec.high_slope = p->slope_db_per_octave;
    if (or_iir_init(&P->filter, &spec, dt, samples, 1, P->whole ? 2 : 0, err, errlen)) return -1;
  }
  const int64_t w = P->w, snr = P->snr;
  const int64_t need = p->method == 0 ? w + P->lta : w;
  const int64_t b0 = snr + w / 2 > need ? snr + w / 2 : need;
  P->before = b0 + (P->filtered ? P->filter.pad : 0);
  P->after = w > snr ? w : snr;
  return 0;
}

int64_t or_fb_required_margin(const or_fb_plan *P) { return P->whole ? P->samples : P->before; }


/* ------------------------------------------------------------ fold maps -- */

/* nearest (stack/binning.nim:197-206): bin i holds [i - 0.5, i + 0.5); -1
 * outside [-0.5, n - 0.5) or NaN; the half-way test on u - floor(u), which
 * is exact. */
static int64_t or_nearest(double u, int64_t n) {
  if (!(u >= -0.5 && u < (double)n - 0.5)) return -1;
  const double whole = floor(u);
  const int64_t k = (int64_t)whole + (u - whole >= 0.5 ? 1 : 0);
  return k < n - 1 ? k : n - 1;
}

static int64_t or_grid_bin(const or_binning *b, double u, double v) {
  const int64_t i = or_nearest(v, b->inlines), j = or_nearest(u, b->crosslines);
  return i < 0 || j < 0 ? -1 : i * b->crosslines + j;
}

int64_t or_bin_at(const or_binning *b, double x, double y) {
  const double dx = x - b->x0, dy = y - b->y0;
  if (b->kind == 1) {
    if (!(fabs(dx * b->uy - dy * b->ux) <= b->max_off)) return -1;
    return or_nearest((dx * b->ux + dy * b->uy) / b->size0, b->crosslines);
  }
  if (b->kind == 2)
    return or_grid_bin(b, (dx * b->ux + dy * b->uy) / b->size0, (dx * b->vx + dy * b->vy) / b->size1);
  return -1;
}

int64_t or_bin_of(const or_binning *b, const double *k) {
  if (b->kind == 0) return or_nearest((k[OR_K_ENSEMBLE] - b->first0) / b->step0, b->crosslines);
  if (b->kind == 3)
    return or_grid_bin(b, (k[OR_K_CROSSLINE] - b->first1) / b->step1,
                       (k[OR_K_INLINE] - b->first0) / b->step0);
  return or_bin_at(b, k[OR_K_CMP_X], k[OR_K_CMP_Y]);
}

static int or_finite(double x) { return isfinite(x); }

/* floorMod for floats (Nim std/math): fmod, then moved to the sign of y. */
static double or_floor_mod(double x, double y) {
  double r = fmod(x, y);
  if ((r > 0 && y < 0) || (r < 0 && y > 0)) r += y;
  return r;
}

/* addToBin (qc/fold.nim:429-451). */
static void or_add_to_bin(int64_t bin, int64_t size, double offset, double bearing,
                          const double *bounds, int32_t nb, int32_t sectors, double start,
                          int32_t offset_range, int32_t *fold, int32_t *ofold, int32_t *afold,
                          float *omin, float *omax, int64_t *counters) {
  counters[0]++;
  fold[bin]++;
  if (nb > 1 && offset >= bounds[0] && offset < bounds[nb - 1]) {
    int32_t k = 0; /* upperBound - 1: bounds[k] <= offset < bounds[k + 1] */
    while (k + 1 < nb && bounds[k + 1] <= offset) k++;
    ofold[(int64_t)k * size + bin]++;
  }
  if (offset_range && !isnan(offset)) {
    const float f = (float)offset;
    if (!(omin[bin] <= f)) omin[bin] = f;
    if (!(omax[bin] >= f)) omax[bin] = f;
  }
  if (sectors > 0 && !isnan(bearing)) {
    int64_t k = (int64_t)(or_floor_mod(bearing - start, 360.0) * (double)sectors / 360.0);
    if (k > sectors - 1) k = sectors - 1;
    afold[k * size + bin]++;
  }
}

/* radToDeg (Nim std/math): d / (PI / 180). */
static double or_rad_to_deg(double d) { return d / (OR_PI / 180.0); }

void or_fold_add(const or_binning *b, const double *bounds, int32_t n_bounds, int32_t sectors,
                 double sector_start_deg, int32_t offset_range, const double *keys,
                 const int32_t *row_len, int32_t rows, int32_t *fold, int32_t *offset_fold,
                 int32_t *azimuth_fold, float *offset_min, float *offset_max, int64_t *counters) {
  const int64_t size = b->inlines * b->crosslines;
  for (int32_t r = 0; r < rows; r++) {
    if (row_len[r] < 0) { counters[3]++; continue; }
    /* addTrace (qc/fold.nim:561-581) */
    const double *k = keys + (int64_t)r * OR_FOLD_KEYS;
    const int coords = or_finite(k[OR_K_SRC_X]) && or_finite(k[OR_K_SRC_Y]) &&
                       or_finite(k[OR_K_RCV_X]) && or_finite(k[OR_K_RCV_Y]);
    int64_t bin;
    if (b->kind == 1 || b->kind == 2) {
      double x = k[OR_K_CMP_X], y = k[OR_K_CMP_Y];
      if (!(or_finite(x) && or_finite(y))) {
        if (!coords) { counters[2]++; continue; }
        x = (k[OR_K_SRC_X] + k[OR_K_RCV_X]) * 0.5;
        y = (k[OR_K_SRC_Y] + k[OR_K_RCV_Y]) * 0.5;
      }
      bin = or_bin_at(b, x, y);
    } else {
      bin = or_bin_of(b, k);
    }
    if (bin < 0) { counters[1]++; continue; }
    const double offset = coords ? hypot(k[OR_K_RCV_X] - k[OR_K_SRC_X], k[OR_K_RCV_Y] - k[OR_K_SRC_Y])
                                 : fabs(k[OR_K_OFFSET]);
    const double bearing = coords && sectors > 0
                               ? or_rad_to_deg(atan2(k[OR_K_RCV_X] - k[OR_K_SRC_X],
                                                     k[OR_K_RCV_Y] - k[OR_K_SRC_Y]))
                               : NAN;
    or_add_to_bin(bin, size, offset, bearing, bounds, n_bounds, sectors, sector_start_deg,
                  offset_range, fold, offset_fold, azimuth_fold, offset_min, offset_max, counters);
  }
}
