// First-break picking (qc/firstbreak.nim) on the GPU, one thread per row.
// A pick is an integer column chosen by an argmax over float64 attributes;
// to give the same column as the serial reference, every row runs the
// reference's serial arithmetic in the same order: the minimum-phase
// band-pass's recursion (filter/iir.nim, with its trend, pads and flush),
// the prefix energies, the picker's attribute and its argmax, the
// confidence.  No multiply-add is contracted in this file, so each operation
// rounds as the serial code's does.  The scratch arrays are laid out
// [index][rows]: the lanes of a wave step through neighbouring rows'
// values together, so their loads and stores coalesce.  Every per-sample
// loop reads its inputs 8 or 16 samples at a time before it computes on them: a
// load that follows a store to a scratch array the compiler cannot tell
// apart would otherwise wait out the memory latency once per sample, and
// with one lane per row those waits are the kernel's time.  The arithmetic
// and its order are the serial code's.
#pragma clang fp contract(off)

#include "pseg_internal.hpp"

#include <math.h>

namespace pseg {

namespace {

constexpr int CH = 8;    // samples loaded at once per lane, loops over several arrays
constexpr int CH1 = 16;  // the same for loops over one array

struct Col {
  double *base;
  int64_t rows, r;
  __device__ __forceinline__ double &operator[](int64_t i) const { return base[i * rows + r]; }
};
template <typename T> struct ColT {
  T *base;
  int64_t rows, r;
  __device__ __forceinline__ T &operator[](int64_t i) const { return base[i * rows + r]; }
};

// sum64 / sumIndexed64 (core/sums.nim:12-40): eight accumulators, then the
// rest, in float64.
template <typename X>
__device__ double sum64(const X &x, int64_t n) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, r = 0;
  const int64_t n8 = n / 8 * 8;
  for (int64_t i = 0; i < n8; i += 8)
    for (int j = 0; j < 8; j++) acc[j] += (double)x[i + j];
  for (int j = 0; j < 8; j++) r += acc[j];
  for (int64_t k = n8; k < n; k++) r += (double)x[k];
  return r;
}
template <typename X>
__device__ double sum_indexed64(const X &x, int64_t n) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, idx[8], r = 0;
  for (int j = 0; j < 8; j++) idx[j] = (double)j;
  const int64_t n8 = n / 8 * 8;
  for (int64_t i = 0; i < n8; i += 8)
    for (int j = 0; j < 8; j++) {
      acc[j] += idx[j] * (double)x[i + j];
      idx[j] += 8;
    }
  for (int j = 0; j < 8; j++) r += acc[j];
  for (int64_t k = n8; k < n; k++) r += (double)k * (double)x[k];
  return r;
}

// applyFilter on the stretch copy x[0 ..< n] (filter/iir.nim:363-433,
// minimum phase: one causal pass per biquad): trend removed, the odd
// reflection before (a steady-state start), the passes in float64 with the
// states flushed below 1e-250 every 256 samples, the trend's steady-state
// response added back.
template <typename T>
__device__ void iir_row(const DevFirstBreak &P, const ColT<T> &x, int64_t n, const Col &buf) {
  if (n <= 0) return;
  double a = 0, b = 0;
  if (P.trend == 1) {
    a = sum64(x, n) / (double)n;
  } else if (P.trend == 2) {
    const double sx = sum64(x, n), nf = (double)n, mean = sx / nf;
    if (n == 1) {
      a = mean;
    } else {
      const double c = (nf - 1) / 2;
      const double slope = (sum_indexed64(x, n) - c * sx) / (nf * (nf * nf - 1) / 12);
      a = mean - slope * c;
      b = slope;
    }
  }
  const int64_t L = P.lead;
  for (int64_t i0 = 0; i0 < n; i0 += CH1) {
    T xv[CH1];
#pragma unroll
    for (int q = 0; q < CH1; q++)
      if (i0 + q < n) xv[q] = x[i0 + q];
#pragma unroll
    for (int q = 0; q < CH1; q++) {
      const int64_t i = i0 + q;
      if (i < n) buf[L + i] = (double)xv[q] - (a + b * (double)i);
    }
  }
  if (P.reflect) {
    const double d0 = buf[L];
    for (int64_t k = 1; k <= L; k++) buf[L - k] = 2 * d0 - buf[L + (k < n - 1 ? k : n - 1)];
  }
  const int64_t m = L + n;
  for (int s = 0; s < P.n_sections; s++) {
    const double b0 = P.sec[s][0], b1 = P.sec[s][1], b2 = P.sec[s][2];
    const double a1 = P.sec[s][3], a2 = P.sec[s][4], g = P.sec[s][5];
    double s1 = 0, s2 = 0;
    if (P.reflect && m > 0) {
      const double u = buf[0];
      const double y = g * u;
      s2 = b2 * u - a2 * y;
      s1 = y - b0 * u;
    }
    int64_t done = 0;
    while (done < m) {
      const int64_t run = m - done < 256 ? m - done : 256;
      for (int64_t k0 = 0; k0 < run; k0 += CH1) {
        // each index is read and written once per pass, so the chunk's
        // inputs can be read before its outputs are stored
        double xv[CH1];
#pragma unroll
        for (int q = 0; q < CH1; q++)
          if (k0 + q < run) xv[q] = buf[done + k0 + q];
#pragma unroll
        for (int q = 0; q < CH1; q++) {
          if (k0 + q < run) {
            const double y = b0 * xv[q] + s1;
            s1 = b1 * xv[q] - a1 * y + s2;
            s2 = b2 * xv[q] - a2 * y;
            buf[done + k0 + q] = y;
          }
        }
      }
      s1 = fabs(s1) < 1e-250 ? 0.0 : s1;
      s2 = fabs(s2) < 1e-250 ? 0.0 : s2;
      done += run;
    }
  }
  const double ga = a * P.dc_gain - b * P.dc_slope, gb = b * P.dc_gain;
  for (int64_t i0 = 0; i0 < n; i0 += CH1) {
    double yv[CH1];
#pragma unroll
    for (int q = 0; q < CH1; q++)
      if (i0 + q < n) yv[q] = buf[L + i0 + q];
#pragma unroll
    for (int q = 0; q < CH1; q++) {
      const int64_t i = i0 + q;
      if (i < n) x[i] = (T)(yv[q] + (ga + gb * (double)i));
    }
  }
}

// iir_row for cascades of at most kFusedSections biquads, in ONE pass over
// the row and no padded copy: CH1 padded samples at a time (the odd
// reflection before the row, then the trend-removed row, formed as the
// serial code forms them) run through every biquad in turn, the biquads'
// states kept per lane in LDS between chunks, and the last biquad's output
// goes back to x with the trend's response.  Each biquad sees the same
// inputs in the same order and does the same operations as in the
// section-by-section passes (its steady-state start from its own first
// input, its states flushed at the same 256-sample boundaries, which are
// chunk ends), so every value is the same; the row crosses memory once
// instead of once per biquad.
constexpr int kFusedSections = 8;
// The trend iir_row removes (filter/edges.nim through iir_row): the mean,
// or the least-squares line, from sum64 / sum_indexed64 of x.
__device__ __forceinline__ void trend_of(int trend, int64_t n, double sx, double sxi, double &a,
                                         double &b) {
  a = 0;
  b = 0;
  if (trend == 1) {
    a = sx / (double)n;
  } else if (trend == 2) {
    const double nf = (double)n, mean = sx / nf;
    if (n == 1) {
      a = mean;
    } else {
      const double c = (nf - 1) / 2;
      const double slope = (sxi - c * sx) / (nf * (nf * nf - 1) / 12);
      a = mean - slope * c;
      b = slope;
    }
  }
}

// sx, sxi: sum64 and sum_indexed64 of x (taken as x was copied).  p (when
// not NULL) receives the prefix energies of the output as pick_stretch
// makes them (p[0] = 0, p[i + 1] = p[i] + y[i]²), their total returned.
template <typename T>
__device__ double iir_row_fused(const DevFirstBreak &P, const ColT<T> &x, int64_t n,
                                double sx, double sxi, double (*st)[2][64], const Col *p) {
  if (n <= 0) return 0;
  double a, b;
  trend_of(P.trend, n, sx, sxi, a, b);
  double energy = 0;
  if (p) (*p)[0] = 0;
  const int64_t L = P.lead, m = L + n;
  const int S = P.n_sections;
  const int lane = threadIdx.x;
  // buf[L] of the serial code: the first trend-removed sample
  const double d0 = (double)x[0] - (a + b * (double)0);
  const double ga = a * P.dc_gain - b * P.dc_slope, gb = b * P.dc_gain;
  for (int s = 0; s < S; s++) st[s][0][lane] = st[s][1][lane] = 0;
  static_assert(256 % CH1 == 0, "flush boundaries must be chunk ends");
  for (int64_t j0 = 0; j0 < m; j0 += CH1) {
    // the chunk's inputs: padded sample j reads x at min(L - j, n - 1)
    // (before the row, the reflection) or j - L
    T xv[CH1];
#pragma unroll
    for (int q = 0; q < CH1; q++) {
      const int64_t j = j0 + q;
      if (j < m) {
        const int64_t k = L - j;
        xv[q] = x[j < L ? (k < n - 1 ? k : n - 1) : j - L];
      }
    }
    double u[CH1];
#pragma unroll
    for (int q = 0; q < CH1; q++) {
      const int64_t j = j0 + q;
      if (j < L) {
        const int64_t k = L - j, i = k < n - 1 ? k : n - 1;
        u[q] = 2 * d0 - ((double)xv[q] - (a + b * (double)i));
      } else {
        u[q] = (double)xv[q] - (a + b * (double)(j - L));
      }
    }
    const int cnt = m - j0 < CH1 ? (int)(m - j0) : CH1;
    const bool flush = (j0 + CH1) % 256 == 0 || j0 + cnt == m;
    for (int s = 0; s < S; s++) {
      const double b0 = P.sec[s][0], b1 = P.sec[s][1], b2 = P.sec[s][2];
      const double a1 = P.sec[s][3], a2 = P.sec[s][4];
      double s1 = st[s][0][lane], s2 = st[s][1][lane];
      if (j0 == 0 && P.reflect) {
        // the steady-state start on this biquad's first input
        const double y = P.sec[s][5] * u[0];
        s2 = b2 * u[0] - a2 * y;
        s1 = y - b0 * u[0];
      }
#pragma unroll
      for (int q = 0; q < CH1; q++) {
        if (q < cnt) {
          const double y = b0 * u[q] + s1;
          s1 = b1 * u[q] - a1 * y + s2;
          s2 = b2 * u[q] - a2 * y;
          u[q] = y;
        }
      }
      if (flush) {
        s1 = fabs(s1) < 1e-250 ? 0.0 : s1;
        s2 = fabs(s2) < 1e-250 ? 0.0 : s2;
      }
      st[s][0][lane] = s1;
      st[s][1][lane] = s2;
    }
#pragma unroll
    for (int q = 0; q < CH1; q++) {
      const int64_t j = j0 + q;
      if (q < cnt && j >= L) {
        const T y = (T)(u[q] + (ga + gb * (double)(j - L)));
        x[j - L] = y;
        if (p) {
          const double v = (double)y;
          energy += v * v;
          (*p)[j - L + 1] = energy;
        }
      }
    }
  }
  return energy;
}

// meanPower (qc/snr.nim:151-161) over the columns c0 .. c1 cut to 0 .. n - 1.
template <typename X>
__device__ double mean_power(const X &x, int64_t n, int64_t c0, int64_t c1) {
  const int64_t a = c0 > 0 ? c0 : 0, b = c1 < n - 1 ? c1 : n - 1;
  if (a > b) return NAN;
  double s = 0;
  for (int64_t j0 = a; j0 <= b; j0 += CH1) {
    double v[CH1];
#pragma unroll
    for (int q = 0; q < CH1; q++)
      if (j0 + q <= b) v[q] = (double)x[j0 + q];
#pragma unroll
    for (int q = 0; q < CH1; q++)
      if (j0 + q <= b) s += v[q] * v[q];
  }
  return s / (double)(b - a + 1);
}

// A strided view of a row of src (the lanes' rows are cols apart).
template <typename T> struct RowT {
  const T *p;
  __device__ __forceinline__ T operator[](int64_t i) const { return p[i]; }
};

template <typename T, typename X>
__device__ void pick_stretch(const DevFirstBreak &P, const X &x, int64_t first, int64_t m,
                             int64_t lo, int64_t hi, double origin, const Col &p, const Col &a,
                             double &time_ms, int64_t &col, double &snr_db, double &conf,
                             bool have_p = false, double p_total = 0) {
  double s = have_p ? p_total : 0;
  if (!have_p) p[0] = 0;
  for (int64_t i0 = 0; i0 < (have_p ? 0 : m); i0 += CH1) {
    double v[CH1];
#pragma unroll
    for (int q = 0; q < CH1; q++)
      if (i0 + q < m) v[q] = (double)x[i0 + q];
#pragma unroll
    for (int q = 0; q < CH1; q++) {
      if (i0 + q < m) {
        s += v[q] * v[q];
        p[i0 + q + 1] = s;
      }
    }
  }
  const double total = s;  // p[m]
  if (!(total > 0) || total == INFINITY) return;
  const double mean = total / (double)m;
  const double eps = 1e-6 * mean;
  const double beta = P.stabilization * mean * (double)P.w;
  const int64_t w = P.w, lta = P.lta, n = m;
  int64_t best_i = -1;
  double best = 0.0;
  const int64_t sw = w / 2 > 1 ? w / 2 : 1;
  const int64_t last_read = n - 1 < hi - 2 + sw ? n - 1 : hi - 2 + sw;
  int64_t i0 = 0;
  if (P.method == 0) {
    i0 = lo > w + sw ? lo : w + sw;
    for (int64_t b0 = i0 - 1; b0 <= last_read; b0 += CH) {
      double pe[CH], ps[CH], pm[CH];
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int64_t i = b0 + q;
        if (i <= last_read) {
          const int64_t s0 = i + 1 - w > 0 ? i + 1 - w : 0;
          const int64_t m0 = s0 - lta > 0 ? s0 - lta : 0;
          pe[q] = p[i + 1];
          ps[q] = p[s0];
          pm[q] = p[m0];
        }
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int64_t i = b0 + q;
        if (i <= last_read) {
          const int64_t s0 = i + 1 - w > 0 ? i + 1 - w : 0;
          const int64_t m0 = s0 - lta > 0 ? s0 - lta : 0;
          a[i] = s0 - m0 < sw ? 0.0
                              : ((pe[q] - ps[q]) / (double)(i + 1 - s0)) /
                                    ((ps[q] - pm[q]) / (double)(s0 - m0) + eps);
        }
      }
    }
  } else if (P.method == 1) {
    const int64_t q = w / 4 > 1 ? w / 4 : 1;
    for (int64_t b0 = lo; b0 < hi; b0 += CH) {
      double pa[CH], pi[CH], pb[CH], pe[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int64_t i = b0 + u;
        int64_t k = i < n - i ? i : n - i;
        k = w < k ? w : k;
        if (i < hi && k >= sw) {
          const int64_t e = n < i + q ? n : i + q;
          pa[u] = p[i + k];
          pi[u] = p[i];
          pb[u] = p[i - k];
          pe[u] = p[e];
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int64_t i = b0 + u;
        int64_t k = i < n - i ? i : n - i;
        k = w < k ? w : k;
        if (i < hi && k >= sw) {
          const double r = (pa[u] - pi[u]) / ((pi[u] - pb[u]) + eps * (double)k) *
                           __dsqrt_rn(pe[u] - pi[u]);
          if (r > best) {
            best = r;
            best_i = i;
          }
        }
      }
    }
  } else {
    i0 = lo > w ? lo : w;
    for (int64_t b0 = i0 - 1; b0 <= last_read; b0 += CH) {
      double pe[CH], ps[CH];
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int64_t i = b0 + q;
        if (i <= last_read) {
          const int64_t s0 = i + 1 - w > 0 ? i + 1 - w : 0;
          pe[q] = p[i + 1];
          ps[q] = p[s0];
        }
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int64_t i = b0 + q;
        if (i <= last_read) a[i] = (pe[q] - ps[q]) / (pe[q] + beta + eps);
      }
    }
  }
  if (P.method != 1) {
    for (int64_t b0 = i0; b0 < hi; b0 += CH) {
      double aj[CH], ai[CH];
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int64_t i = b0 + q;
        const int64_t j = n - 1 < i - 1 + sw ? n - 1 : i - 1 + sw;
        if (i < hi && j > i - 1) {
          aj[q] = a[j];
          ai[q] = a[i - 1];
        }
      }
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int64_t i = b0 + q;
        const int64_t j = n - 1 < i - 1 + sw ? n - 1 : i - 1 + sw;
        if (i < hi && j > i - 1) {
          const double r = (aj[q] - ai[q]) / (double)(j - i + 1);
          if (r > best) {
            best = r;
            best_i = i;
          }
        }
      }
    }
  }
  if (best_i < 0) return;
  const int64_t c = best_i;
  const int64_t gap_end = c - w / 2 > 0 ? c - w / 2 : 0;
  col = first + c;
  time_ms = origin + (double)(first + c) * P.dt;
  const int64_t nz0 = gap_end - P.snr > 0 ? gap_end - P.snr : 0;
  const double noise = mean_power(x, m, nz0, gap_end - 1);
  if (noise > 0) {
    const int64_t sig1 = m < c + P.snr ? m : c + P.snr;
    const double sg = mean_power(x, m, c, sig1 - 1);
    const double ss = sg + eps, nn = noise + eps;
    double snr = NAN;
    if (!isnan(ss) && nn > 0) {
      const double r = ss / nn;
      snr = isnan(r) ? (double)NAN : (r > 0 ? 10 * log10(r) : -INFINITY);
    }
    if (!isnan(snr)) {
      snr_db = snr;
      const double cf = (snr - P.low_db) / (P.high_db - P.low_db);
      conf = cf < 0 ? 0.0 : (cf > 1 ? 1.0 : cf);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(64) void firstbreak_kernel(const DevFirstBreak *__restrict__ plan,
                                                        const T *__restrict__ src, int32_t rows,
                                                        int32_t cols,
                                                        const int32_t *__restrict__ row_len,
                                                        const double *__restrict__ keys,
                                                        double first_sample, double *p_, double *a_,
                                                        double *buf_, T *copy_,
                                                        double *__restrict__ time_ms,
                                                        int64_t *__restrict__ col,
                                                        double *__restrict__ snr_db,
                                                        double *__restrict__ conf) {
  __shared__ double iir_st[kFusedSections][2][64];  // the biquads' states, per lane
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const DevFirstBreak &P = *plan;
  double t_out = NAN, s_out = NAN, c_out = 0;
  int64_t col_out = -1;
  const int live = row_len[r];
  if (live >= 0) {
    const int64_t n = live < cols ? live : cols;
    const RowT<T> x{src + r * (int64_t)cols};
    const double t0 = keys ? keys[2 * r] : 0.0, offset = keys ? keys[2 * r + 1] : (double)NAN;
    // stretchOf (qc/firstbreak.nim:394-430)
    int64_t first = 0, len = 0, lo = 0, hi = 0;
    double origin = NAN;
    bool ok = n >= 2;
    double t_lo = isnan(P.search_start) ? -INFINITY : P.search_start;
    double t_hi = isnan(P.search_end) ? INFINITY : P.search_end;
    if (ok && !isnan(P.guide_v)) {
      if (isnan(offset)) {
        ok = false;
      } else {
        const double tl = P.guide_t0 + 1000 * fabs(offset) / P.guide_v;
        t_lo = fmax(t_lo, tl - P.before_ms);
        t_hi = fmin(t_hi, tl + P.after_ms);
      }
    }
    if (ok) {
      const double t0c = t0 + first_sample * P.dt;
      origin = isinf(t0c) ? (double)NAN : t0c;
      if (isnan(origin) && (t_lo > -INFINITY || t_hi < INFINITY)) ok = false;
    }
    int64_t la = 0, lb = -1;
    if (ok) {
      int64_t i = 0;
      while (i < n && x[i] == (T)0) i++;
      if (i < n) {
        int64_t j = n - 1;
        while (x[j] == (T)0) j--;
        la = i;
        lb = j;
      }
      if (lb - la + 1 < 2) ok = false;
    }
    if (ok) {
      const double c_lo = isnan(origin) ? -INFINITY : (t_lo - origin) / P.dt;
      const double c_hi = isnan(origin) ? INFINITY : (t_hi - origin) / P.dt;
      const int64_t lo_c = c_lo <= 0 ? 0 : c_lo >= (double)n ? n : (int64_t)ceil(c_lo);
      const int64_t hi_c = c_hi < 0 ? 0 : c_hi >= (double)(n - 1) ? n : (int64_t)floor(c_hi) + 1;
      const int64_t wlo = la > lo_c ? la : lo_c;
      const int64_t whi = lb + 1 < hi_c ? lb + 1 : hi_c;
      if (wlo >= whi) {
        ok = false;
      } else {
        const int64_t a = P.whole ? la : (la > wlo - P.before ? la : wlo - P.before);
        const int64_t b = P.whole ? lb + 1 : (lb + 1 < whi + P.after ? lb + 1 : whi + P.after);
        first = a;
        len = b - a;
        lo = wlo - a;
        hi = whi - a;
      }
    }
    if (ok && len > 0) {
      const Col p{p_, rows, r}, aa{a_, rows, r};
      if (P.filtered) {
        const ColT<T> c{copy_, rows, r};
        // the copy, with sum64 and sum_indexed64 of it in their order: eight
        // accumulators over the first len / 8 * 8 samples (sample i into
        // accumulator i mod 8, the chunks being multiples of 8), the rest
        // added after them (read back below)
        const int64_t n8 = len / 8 * 8;
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, acci[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t j0 = 0; j0 < len; j0 += CH1) {
          T v[CH1];
#pragma unroll
          for (int q = 0; q < CH1; q++)
            if (j0 + q < len) v[q] = x[first + j0 + q];
#pragma unroll
          for (int q = 0; q < CH1; q++) {
            const int64_t j = j0 + q;
            if (j < len) c[j] = v[q];
            if (j < n8) {
              acc[q % 8] += (double)v[q];
              acci[q % 8] += (double)j * (double)v[q];
            }
          }
        }
        if (P.n_sections <= kFusedSections) {
          double sx = 0, sxi = 0;
          for (int k = 0; k < 8; k++) sx += acc[k];
          for (int k = 0; k < 8; k++) sxi += acci[k];
          for (int64_t k = n8; k < len; k++) {
            const double v = (double)c[k];
            sx += v;
            sxi += (double)k * v;
          }
          const double total = iir_row_fused<T>(P, c, len, sx, sxi, iir_st, &p);
          pick_stretch<T>(P, c, first, len, lo, hi, origin, p, aa, t_out, col_out, s_out, c_out,
                          true, total);
        } else {
          iir_row<T>(P, c, len, Col{buf_, rows, r});
          pick_stretch<T>(P, c, first, len, lo, hi, origin, p, aa, t_out, col_out, s_out, c_out);
        }
      } else {
        const RowT<T> y{src + r * (int64_t)cols + first};
        pick_stretch<T>(P, y, first, len, lo, hi, origin, p, aa, t_out, col_out, s_out, c_out);
      }
    }
  }
  time_ms[r] = t_out;
  col[r] = col_out;
  snr_db[r] = s_out;
  conf[r] = c_out;
}

}  // namespace

hipError_t launch_firstbreak(int dtype, const DevFirstBreak *plan, const void *src, int32_t rows,
                             int32_t cols, const int32_t *row_len, const double *keys,
                             int64_t first_sample, double *p, double *a, double *buf, void *copy,
                             double *time_ms, int64_t *col, double *snr_db, double *conf,
                             hipStream_t stream) {
  if (rows == 0) return hipSuccess;
  const int grid = (rows + 63) / 64;
  if (dtype == 0)
    hipLaunchKernelGGL(firstbreak_kernel<float>, dim3(grid), dim3(64), 0, stream, plan,
                       (const float *)src, rows, cols, row_len, keys, (double)first_sample, p, a,
                       buf, (float *)copy, time_ms, col, snr_db, conf);
  else
    hipLaunchKernelGGL(firstbreak_kernel<double>, dim3(grid), dim3(64), 0, stream, plan,
                       (const double *)src, rows, cols, row_len, keys, (double)first_sample, p, a,
                       buf, (double *)copy, time_ms, col, snr_db, conf);
  return hipGetLastError();
}

}  // namespace pseg
