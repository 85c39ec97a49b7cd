// gfx950 kernels beside the row pipeline: vibroseis correlation and the
// per-panel normalization statistics, and the launch dispatch.
//
// corr_kernel: one workgroup (1024 threads) per record: gather to the live
// length, rFFT, times the device-resident pilot multiplier, irFFT, lag spans
// (vibroseis/correlate.nim:241-290).
//
// panel kernels: per-row sums of squares and maxima, a 65536-bin histogram of
// float32 |x| bits >> 15 with integer atomics (bit-exact counts), and a
// one-block kernel for the per-panel factor (gain/stats.nim:253-304,
// gain/gain.nim:221-240).
#include "pseg_device.hpp"

#include <mutex>

namespace pseg {

// Per-row sum of squares (float64) and maximum |x| bits (as float32 for the
// histogram key and in T for the maximum), for the rows a panel statistic
// uses (gain/stats.nim:253-281).
template <typename T>
__global__ __launch_bounds__(256) void row_stats_kernel(const T *__restrict__ x, int32_t cols,
                                                        const int32_t *__restrict__ row_len,
                                                        double *__restrict__ row_sq,
                                                        uint32_t *__restrict__ row_max) {
  __shared__ double red[8];
  __shared__ typename Bits<T>::S smax[4];
  using S = typename Bits<T>::S;
  const int64_t row = blockIdx.x;
  const int n = min(row_len[row], cols);  // a length past the row reads the row only
  if (n <= 0) {
    if (threadIdx.x == 0) {
      row_sq[row] = n < 0 ? -1.0 : 0.0;
      row_max[2 * row] = 0;
      row_max[2 * row + 1] = 0;
    }
    return;
  }
  const T *r = x + row * (int64_t)cols;
  double sq = 0;
  S mb = 0;
  for (int j = threadIdx.x; j < n; j += 256) {
    T v = r[j];
    sq += (double)v * (double)v;
    S b = Bits<T>::abs_bits(v);
    mb = b > mb ? b : mb;
  }
  sq = block_sum<256>(sq, red);
  mb = wave_max_u<S>(mb);
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mb;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; i++) mb = smax[i] > mb ? smax[i] : mb;
    row_sq[row] = sq;
    unsigned long long b = (unsigned long long)mb;
    row_max[2 * row] = (uint32_t)b;
    row_max[2 * row + 1] = (uint32_t)(b >> 32);
  }
}

template <typename T>
__device__ __forceinline__ double row_max_value(const uint32_t *row_max, int64_t row) {
  unsigned long long b = (unsigned long long)row_max[2 * row] |
                         ((unsigned long long)row_max[2 * row + 1] << 32);
  return Bits<T>::value((typename Bits<T>::S)b);
}

// usable (gain/stats.nim:232-235): finite samples and a finite sum.
template <typename T>
__device__ __forceinline__ bool row_usable(const double *row_sq, const uint32_t *row_max,
                                           int64_t row) {
  double s = row_sq[row];
  if (s < 0) return false;
  double m = row_max_value<T>(row_max, row);
  return isfinite(m) && s != INFINITY;
}

// Histogram of float32(|x|) bits >> 15, min 0xFEFF (gain/stats.nim:208-230).
template <typename T>
__global__ __launch_bounds__(256) void hist_kernel(const T *__restrict__ x, int32_t cols,
                                                   const int32_t *__restrict__ row_len,
                                                   const double *__restrict__ row_sq,
                                                   const uint32_t *__restrict__ row_max,
                                                   uint32_t *__restrict__ hist) {
  const int64_t row = blockIdx.x;
  const int n = min(row_len[row], cols);
  if (n <= 0 || !row_usable<T>(row_sq, row_max, row)) return;
  const T *r = x + row * (int64_t)cols;
  for (int j = threadIdx.x; j < n; j += 256) {
    float a = (float)fabs((double)r[j]);
    uint32_t key = __float_as_uint(a) >> 15;
    key = key < 0xFEFFu ? key : 0xFEFFu;
    atomicAdd(&hist[key], 1u);
  }
}

// panelStats + panelFactor (gain/stats.nim:253-304, gain/gain.nim:221-240)
// in one block: sums in row order (deterministic), the percentile from the
// histogram by orderStat.
template <typename T>
__global__ __launch_bounds__(1024) void panel_factor_kernel(
    int32_t rows, const int32_t *__restrict__ row_len, const double *__restrict__ row_sq,
    const uint32_t *__restrict__ row_max, const uint32_t *__restrict__ hist, int32_t stat,
    double target, double percentile, double *__restrict__ factor) {
  __shared__ double s_sum[1024];
  __shared__ double s_max[1024];
  __shared__ long long s_cnt[1024];
  __shared__ long long s_scan[1024];
  __shared__ double s_out[2];
  const int t = threadIdx.x;
  // each thread a contiguous run of rows, then the runs in order
  const int per = (rows + 1023) / 1024;
  double sum = 0, mx = 0;
  long long samples = 0;
  for (int r = t * per; r < min(rows, (t + 1) * per); r++) {
    if (row_len[r] < 0 || !row_usable<T>(row_sq, row_max, r)) continue;
    sum += row_sq[r];
    double m = row_max_value<T>(row_max, r);
    mx = m > mx ? m : mx;
    samples += row_len[r];
  }
  s_sum[t] = sum;
  s_max[t] = mx;
  s_cnt[t] = samples;
  __syncthreads();
  if (t == 0) {
    double a = 0, b = 0;
    long long c = 0;
    for (int i = 0; i < 1024; i++) {
      a += s_sum[i];
      b = s_max[i] > b ? s_max[i] : b;
      c += s_cnt[i];
    }
    s_sum[0] = a;
    s_max[0] = b;
    s_cnt[0] = c;
  }
  __syncthreads();
  const double tsum = s_sum[0], tmax = s_max[0];
  const long long tn = s_cnt[0];
  double level = NAN;
  if (tn > 0) {
    if (stat == 0) level = tmax;
    else if (stat == 1) level = sqrt(tsum / (double)tn);
    else if (percentile == 100) level = tmax;
    else {
      // orderStat for ranks k and k + 1: a scan over the 65536 bins, 64 per
      // thread
      const double rank = (double)(tn - 1) * percentile / 100;
      long long k = (long long)floor(rank);
      if (k > tn - 1) k = tn - 1;
      long long local = 0;
      for (int i = 0; i < 64; i++) local += hist[t * 64 + i];
      s_scan[t] = local;
      __syncthreads();
      if (t == 0) {
        long long run = 0;
        for (int i = 0; i < 1024; i++) {
          long long c = s_scan[i];
          s_scan[i] = run;
          run += c;
        }
      }
      __syncthreads();
      for (int which = 0; which < 2; which++) {
        long long j = k + which;
        long long below = s_scan[t];
        for (int i = 0; i < 64; i++) {
          int key = t * 64 + i;
          if (key > 0xFEFF) break;
          long long c = hist[key];
          if (below <= j && below + c > j) {
            double val;
            if (key == 0) {
              val = 0;
            } else {
              double lo = (double)__uint_as_float((uint32_t)key << 15);
              double hi = (double)__uint_as_float((uint32_t)(key + 1) << 15);
              hi = hi < tmax ? hi : tmax;
              val = lo + ((double)(j - below) + 0.5) / (double)c * (hi - lo);
              val = val > lo ? val : lo;
              val = val < tmax ? val : tmax;
            }
            s_out[which] = val;
          }
          below += c;
        }
        __syncthreads();
      }
      if (t == 0) {
        // a rank past every count reads the maximum
        const double lower = k < tn ? s_out[0] : tmax;
        const double f = rank - (double)k;
        level = (f == 0 || k + 1 >= tn) ? lower : lower + f * (s_out[1] - lower);
        s_out[0] = level;
      }
      __syncthreads();
      level = s_out[0];
    }
  }
  if (t == 0) {
    double f = NAN;
    if (isfinite(level) && level > 0 && fits_in<T>(target / level)) f = target / level;
    *factor = f;
  }
}

// Real transforms of any 2·3·5-smooth length n.  Even n: the samples sit
// in the buffer as T values, the transform is the half-length complex one
// plus the split (combine).  Odd n (a plan's explicit nfft): the samples sit
// as complex (x, 0) and the transform is the n-point complex one (the plan's
// DevFft then holds m = n).  The bins of either go to the same hooks in
// combine's convention: 2·X[k], X[k] at 0 Hz (and at Nyquist for even n).
__device__ __forceinline__ bool fft_odd(const DevFft &f) { return (f.n & 1) != 0; }

// Sample j of the transform's input.
template <typename T>
__device__ __forceinline__ void put_sample(T *buf, int j, T y, bool odd) {
  if (odd) {
    buf[2 * j] = y;
    buf[2 * j + 1] = (T)0;
  } else {
    buf[j] = y;
  }
}

// The forward real transform of the samples in A (put by put_sample, the
// block synchronized), every bin k = 0 .. n/2 handed to hook.
template <typename T, int NTC, typename Hook>
__device__ __forceinline__ void real_forward(const DevFft &f, V2<T> *A, Hook hook) {
  complex_forward<T, NTC, kFuseCorr>(f, A);
  if (!fft_odd(f)) {
    combine<T, NTC>(f, A, A, hook);
    return;
  }
  const int half = f.n / 2;
  for (int k = threadIdx.x; k <= half; k += NTC) {
    V2<T> x = A[k];
    if (k != 0) x = mk<T>(x.x * (T)2, x.y * (T)2);
    hook(x, k);
  }
}

// The Hilbert transform of the real samples in A (put by put_sample, the
// block synchronized): -i·sign(f) on the spectrum, 0 Hz (and Nyquist) set to
// zero, inverse.  Afterwards H[x] at j is hilbert_at<T>(A, j, n).  Even n: the
// half-length transform, the split with HilbertMul and the unsplit, the
// inverse as the forward one on swapped pairs.  Odd n: the n-point complex
// transform, conj(-i·sign(f)·X) in place, the forward transform again (the
// inverse of Y is conj(FFT(conj Y)) / n; its real part is H[x]).
template <typename T> struct HilbertMul;
template <typename T, int NTC>
__device__ __forceinline__ void hilbert_transform(const DevFft &f, V2<T> *A) {
  const int n = f.n;
  complex_forward<T, NTC, kFuseCorr>(f, A);
  if (!fft_odd(f)) {
    combine<T, NTC>(f, A, A, HilbertMul<T>{(T)(0.5 / (double)n), f.m});
  } else {
    for (int k = threadIdx.x; k <= n / 2; k += NTC) {
      if (k == 0) {
        A[0] = mk<T>((T)0, (T)0);
      } else {
        const V2<T> a = A[k], b = A[n - k];
        A[k] = mk<T>(a.y, a.x);       // conj(-i·a)
        A[n - k] = mk<T>(-b.y, -b.x); // conj(+i·b)
      }
    }
  }
  __syncthreads();
  complex_forward<T, NTC, kFuseCorr>(f, A);
}

template <typename T>
__device__ __forceinline__ T hilbert_at(const V2<T> *A, int j, int n) {
  if (n & 1) return A[j].x * (T)(1.0 / (double)n);
  return ((const T *)A)[j ^ 1];
}

// Vibroseis correlation of one record per workgroup.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void corr_kernel(const DevCorr *__restrict__ plan,
                                                   const T *__restrict__ src, int32_t src_cols,
                                                   const int32_t *__restrict__ row_len,
                                                   T *__restrict__ dst) {
  extern __shared__ __align__(16) char lds[];
  const DevCorr &c = *plan;
  const int64_t row = blockIdx.x;
  const int live = min(row_len[row], src_cols);
  T *d = dst + row * (int64_t)c.out;
  if (live < 0) {
    for (int j = threadIdx.x; j < c.out; j += NTC) d[j] = (T)0;
    return;
  }
  const DevFft &f = c.fft;
  const int n = f.n;
  T *buf = (T *)lds;
  V2<T> *A = (V2<T> *)lds;  // the transform's buffer (f.stride complex)
  const T *s = src + row * (int64_t)src_cols;
  for (int j = threadIdx.x; j < n; j += NTC) buf[j] = j < live ? s[j] : (T)0;
  __syncthreads();
  complex_forward<T, NTC, kFuseCorr>(f, A);
  combine_cplx<T, NTC>(f, A, QTab<T>{as_global<T>(c.coef)});
  __syncthreads();
  complex_forward<T, NTC, kFuseCorr>(f, A);
  const T *out = (const T *)A;
  // lagSpans (vibroseis/correlate.nim:265-274)
  const int wrapped = min(max(1 - c.pilot - c.shift, 0), c.out);
  const int direct = min(max(-c.shift, wrapped), c.out);
  const int lastc = min(max(c.record - c.shift, direct), c.out);
  const int wrap = c.shift + n;
  for (int mcol = threadIdx.x; mcol < c.out; mcol += NTC) {
    T y = (T)0;
    if (mcol >= wrapped && mcol < direct) y = out[(mcol + wrap) ^ 1];
    else if (mcol >= direct && mcol < lastc) y = out[(mcol + c.shift) ^ 1];
    d[mcol] = y;
  }
}

// ---------------------------------------------------------------- spectrum

// |X[k]|² / Σw² of one row added to the workgroup's sums in LDS (each bin
// has one owner thread in combine, so plain adds).  combine hands over
// 2·X[k] except at 0 and Nyquist.
template <typename T> struct PowerAcc {
  double *acc;
  double inv_sw2;
  int m;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    const double re = (double)x.x, im = (double)x.y;
    double p = re * re + im * im;
    if (k != 0 && k != m) p *= 0.25;
    acc[k] += p * inv_sw2;
    return x;
  }
};

// Rows blockIdx.x, + gridDim.x, ... (a fixed assignment: the grid depends
// on the row count and the plan only): each row's window (windowInto,
// fft/spectrum.nim:271-293) through the real FFT into the workgroup's power
// sums, which it writes as its partial, parts[blockIdx.x] = (bins power
// sums, rows used); spectrum_finish_kernel adds the partials in a fixed
// order, so the result's bits do not depend on the order workgroups run
// (the reference's pool form does the same with fixed batches of rows,
// core/pool.nim:776-800).  A window is demeaned when the plan says so and
// its float64 sum is finite (a NaN or Inf leaves the row out), tapered (a
// window cut by the live end takes min(taper[j], taper[n-1-j]) and its own
// Σw), zero-padded to nfft.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void spectrum_kernel(const DevSpec *__restrict__ plan,
                                                       const T *__restrict__ src, int32_t rows,
                                                       int32_t cols,
                                                       const int32_t *__restrict__ row_len,
                                                       int32_t first_col, size_t acc_off,
                                                       double *__restrict__ parts) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red_lds[4 * (NTC / 64)];
  const DevSpec &P = *plan;
  const DevFft &f = P.fft;
  const int nf = f.n, bins = f.n / 2 + 1;
  const bool odd = fft_odd(f);
  double *lacc = (double *)(lds + acc_off);
  for (int k = threadIdx.x; k < bins; k += NTC) lacc[k] = 0;
  __syncthreads();
  RedSlots red{red_lds, 2 * (NTC / 64), 0};
  gptr<T> tp = as_global<T>(P.taper);
  const int W = P.window;
  double used = 0;
  for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
    const int live = min(row_len[row], cols);
    if (live <= first_col) continue;  // not live, or nothing live in the window
    const int n = min(live, first_col + W) - first_col;
    const T *s = src + row * (int64_t)cols + first_col;
    double sum = 0;
    for (int j = threadIdx.x; j < n; j += NTC) sum += (double)s[j];
    sum = block_sum<NTC>(sum, red.next());
    if (!(sum - sum == 0)) continue;  // a NaN or Inf in the window
    const T mean = P.demean ? (T)(sum / (double)n) : (T)0;
    T *buf = (T *)lds;
    const bool full = n == W;
    double swl = 0;
    for (int j = threadIdx.x; j < nf; j += NTC) {
      T y = (T)0;
      if (j < n) {
        T w = tp[j];
        if (!full) {
          const T w2 = tp[n - 1 - j];
          w = w2 < w ? w2 : w;
        }
        y = (s[j] - mean) * w;
        swl += (double)w;
      }
      put_sample(buf, j, y, odd);
    }
    const double sw = full ? P.taper_sum : block_sum<NTC>(swl, red.next());
    __syncthreads();
    V2<T> *A = (V2<T> *)lds;
    real_forward<T, NTC>(f, A, PowerAcc<T>{lacc, 1.0 / (sw * sw), f.m});
    __syncthreads();  // the buffer is free for the next row
    used += 1;
  }
  double *part = parts + (size_t)blockIdx.x * ((size_t)bins + 1);
  for (int k = threadIdx.x; k < bins; k += NTC) part[k] = lacc[k];
  if (threadIdx.x == 0) part[bins] = used;
}

// amplitudeInto (fft/spectrum.nim:295-312) per row: |X[k]|·sides(k)/Σw and
// atan2(Im, Re) as T into the row's bins.  combine hands over 2·X[k] except
// at 0 and Nyquist (the phase does not care).
template <typename T> struct AmpPhaseOut {
  T *amp, *phase;
  double inv_sw;
  int m, nfft;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    const double re = (double)x.x, im = (double)x.y;
    const double half = (k != 0 && k != m) ? 0.5 : 1.0;
    const double sides = (k == 0 || 2 * k == nfft) ? 1.0 : 2.0;
    amp[k] = (T)(hypot(re, im) * half * sides * inv_sw);
    phase[k] = (T)atan2(im, re);
    return x;
  }
};

// One workgroup per row: the row's window from first_col (as averaged
// above, but a NaN or Inf stays: the window's sum and Σw are NaN, and so is
// every bin), through the real FFT, into amp and phase [rows, bins].  Rows
// with no live sample in the window are left as they are.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void spectrum_rows_kernel(const DevSpec *__restrict__ plan,
                                                            const T *__restrict__ src, int32_t cols,
                                                            const int32_t *__restrict__ row_len,
                                                            int32_t first_col, T *__restrict__ amp,
                                                            T *__restrict__ phase) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red_lds[4 * (NTC / 64)];
  const DevSpec &P = *plan;
  const DevFft &f = P.fft;
  const int nf = f.n, bins = f.n / 2 + 1;
  const bool odd = fft_odd(f);
  const int64_t row = blockIdx.x;
  const int live = min(row_len[row], cols);
  if (live <= first_col) return;
  const int W = P.window;
  const int n = min(live, first_col + W) - first_col;
  const T *s = src + row * (int64_t)cols + first_col;
  RedSlots red{red_lds, 2 * (NTC / 64), 0};
  double sum = 0;
  for (int j = threadIdx.x; j < n; j += NTC) sum += (double)s[j];
  sum = block_sum<NTC>(sum, red.next());
  const bool finite = sum - sum == 0;
  const T mean = (P.demean && finite) ? (T)(sum / (double)n) : (T)0;
  gptr<T> tp = as_global<T>(P.taper);
  T *buf = (T *)lds;
  const bool full = n == W;
  double swl = 0;
  for (int j = threadIdx.x; j < nf; j += NTC) {
    T y = (T)0;
    if (j < n) {
      T w = tp[j];
      if (!full) {
        const T w2 = tp[n - 1 - j];
        w = w2 < w ? w2 : w;
      }
      y = (s[j] - mean) * w;
      swl += (double)w;
    }
    put_sample(buf, j, y, odd);
  }
  double sw = full ? P.taper_sum : block_sum<NTC>(swl, red.next());
  if (!finite) sw = NAN;
  __syncthreads();
  V2<T> *A = (V2<T> *)lds;
  real_forward<T, NTC>(f, A, AmpPhaseOut<T>{amp + row * (int64_t)bins, phase + row * (int64_t)bins,
                                          1.0 / sw, f.m, nf});
}

// The Hilbert transform's multiplier (fft/hilbert.nim:107-125): -i·sign(f)
// on the bins between 0 Hz and Nyquist, 0 on those two, with the split's 1/2
// and the inverse's 1/n folded in (combine's convention).
template <typename T> struct HilbertMul {
  T fold;
  int m;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    if (k == 0 || k == m) return mk<T>((T)0, (T)0);
    return mk<T>(x.y * fold, -x.x * fold);
  }
};

// One workgroup per row: the row's live samples zero-padded to nfft
// (>= twice the plan's samples: no wrap), forward, the multiplier, inverse;
// H[x] (or the envelope sqrt(x² + H²), envelope != 0) into dst over the live
// length.  Rows not live and samples past the live end are left as they are.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void hilbert_kernel(const DevFft *__restrict__ fp,
                                                      const T *__restrict__ src, int32_t cols,
                                                      const int32_t *__restrict__ row_len,
                                                      int32_t envelope, T *__restrict__ dst) {
  extern __shared__ __align__(16) char lds[];
  const DevFft &f = *fp;
  const int64_t row = blockIdx.x;
  const int live = min(row_len[row], cols);
  if (live <= 0) return;
  const int n = f.n;
  const T *s = src + row * (int64_t)cols;
  T *buf = (T *)lds;
  const bool odd = fft_odd(f);
  for (int j = threadIdx.x; j < n; j += NTC) put_sample(buf, j, j < live ? s[j] : (T)0, odd);
  __syncthreads();
  V2<T> *A = (V2<T> *)lds;
  hilbert_transform<T, NTC>(f, A);
  T *d = dst + row * (int64_t)cols;
  for (int j = threadIdx.x; j < live; j += NTC) {
    const T h = hilbert_at<T>(A, j, n);
    if (envelope) {
      const T x = s[j];
      d[j] = sqrt(x * x + h * h);
    } else {
      d[j] = h;
    }
  }
}

constexpr double kPi = 3.141592653589793;

// The instantaneous phase of x + i·h in degrees, in (-180, 180]
// (fft/hilbert.nim:190-194).
__device__ __forceinline__ double phase_deg(double x, double h) {
  double deg = atan2(h, x) * (180 / kPi);
  if (deg <= -180) deg += 360;
  return deg;
}

// wrapPi (fft/hilbert.nim:152-158): d brought into (-pi, pi].
__device__ __forceinline__ double wrap_pi(double d) {
  const double two = 2 * kPi;
  double r = d - two * floor((d + kPi) / two);
  if (r <= -kPi) r += two;
  else if (r > kPi) r -= two;
  return r;
}

// analyticInto (fft/hilbert.nim:160-215), one workgroup per row: the row's
// live samples through the Hilbert transform (as hilbert_kernel), then per
// sample in float64 the envelope hypot(x, h), the phase in degrees (NaN where
// the envelope is 0) and the frequency, the circular mean of the two wrapped
// one-sample phase steps about the sample, NaN at the row's ends and where
// the sample or a neighbour is 0 or below peak·thr_scale (peak: the row's
// largest envelope; thr_scale = 10^(floorDb/20)).  A neighbour's envelope and
// phase are recomputed from x and H[x] (in LDS), not read back from the
// outputs.  Rows not live and columns past the live end are left as they
// are.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void analytic_kernel(const DevFft *__restrict__ fp,
                                                       const T *__restrict__ src, int32_t cols,
                                                       const int32_t *__restrict__ row_len,
                                                       double dt_ms, double thr_scale,
                                                       double *__restrict__ env,
                                                       double *__restrict__ phase,
                                                       double *__restrict__ hz) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red[NTC / 64];
  const DevFft &f = *fp;
  const int64_t row = blockIdx.x;
  const int live = min(row_len[row], cols);
  if (live <= 0) return;
  const int n = f.n;
  const T *s = src + row * (int64_t)cols;
  T *buf = (T *)lds;
  const bool odd = fft_odd(f);
  for (int j = threadIdx.x; j < n; j += NTC) put_sample(buf, j, j < live ? s[j] : (T)0, odd);
  __syncthreads();
  V2<T> *A = (V2<T> *)lds;
  hilbert_transform<T, NTC>(f, A);
  double *e_row = env + row * (int64_t)cols;
  double *p_row = phase + row * (int64_t)cols;
  double *f_row = hz + row * (int64_t)cols;
  double mx = 0;
  for (int j = threadIdx.x; j < live; j += NTC) {
    const double xj = (double)s[j], hj = (double)hilbert_at<T>(A, j, n);
    const double e = hypot(xj, hj);
    e_row[j] = e;
    if (e > mx) mx = e;  // a NaN never raises the peak
    p_row[j] = e > 0 ? phase_deg(xj, hj) : (double)NAN;
  }
  // the row's peak envelope (lane maxima, then the waves')
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double other = __shfl_xor(mx, o, 64);
    if (other > mx) mx = other;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  double peak = 0;
  for (int w = 0; w < NTC / 64; w++)
    if (red[w] > peak) peak = red[w];
  const double thr = peak * thr_scale;
  const double scale = 1000 / (4 * kPi * dt_ms);
  for (int j = threadIdx.x; j < live; j += NTC) {
    double out = NAN;
    if (j > 0 && j < live - 1) {
      const double x0 = (double)s[j - 1], h0 = (double)hilbert_at<T>(A, j - 1, n);
      const double x1 = (double)s[j], h1 = (double)hilbert_at<T>(A, j, n);
      const double x2 = (double)s[j + 1], h2 = (double)hilbert_at<T>(A, j + 1, n);
      const double e0 = hypot(x0, h0), e1 = hypot(x1, h1), e2 = hypot(x2, h2);
      if (!(e0 <= 0 || e1 <= 0 || e2 <= 0 || e0 < thr || e1 < thr || e2 < thr)) {
        const double p0 = e0 > 0 ? phase_deg(x0, h0) : (double)NAN;
        const double p1 = e1 > 0 ? phase_deg(x1, h1) : (double)NAN;
        const double p2 = e2 > 0 ? phase_deg(x2, h2) : (double)NAN;
        const double back = wrap_pi((p1 - p0) * kPi / 180);
        const double fwd = wrap_pi((p2 - p1) * kPi / 180);
        out = wrap_pi(back + 0.5 * wrap_pi(fwd - back)) * (2 * scale);
      }
    }
    f_row[j] = out;
  }
}

// |X[k]|·sides(k)/Σw of a frame, float64, into the frame's bins (combine
// hands over 2·X[k] except at 0 and Nyquist).
template <typename T> struct FrameAmp {
  double *amp;
  double inv_sw;
  int m, nfft;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    const double re = (double)x.x, im = (double)x.y;
    const double half = (k != 0 && k != m) ? 0.5 : 1.0;
    const double sides = (k == 0 || 2 * k == nfft) ? 1.0 : 2.0;
    amp[k] = hypot(re, im) * half * sides * inv_sw;
    return x;
  }
};

// Frame `frame` of the trace x (len samples) through the real FFT: the
// window's samples x[frame·hop − W/2 + j] (outside the trace: zero) times the
// Hann weights, zero-padded to nfft; `hook` gets every bin (combine's
// convention, see FrameAmp) with 1/Σw, Σw the weights of the samples inside
// the trace.  Returns false (and calls nothing) for a frame with no sample
// inside the trace.
template <typename T, int NTC, typename Hook>
__device__ __forceinline__ bool spectrogram_frame(const DevSpec &P, const T *__restrict__ x,
                                                  int32_t len, int32_t hop, int64_t frame,
                                                  char *lds, RedSlots &red, Hook hook) {
  const DevFft &f = P.fft;
  const int nf = f.n, W = P.window;
  const int64_t start = frame * (int64_t)hop - W / 2;
  const int a = (int)max((int64_t)0, -start);
  const int b = (int)min((int64_t)W, (int64_t)len - start);
  if (b <= a) return false;
  gptr<T> wt = as_global<T>(P.taper);
  T *buf = (T *)lds;
  const bool odd = fft_odd(f);
  double swl = 0;
  for (int j = threadIdx.x; j < nf; j += NTC) {
    T y = (T)0;
    if (j >= a && j < b) {
      const T w = wt[j];
      y = x[start + j] * w;
      swl += (double)w;
    }
    put_sample(buf, j, y, odd);
  }
  const double sw = block_sum<NTC>(swl, red.next());
  __syncthreads();
  V2<T> *A = (V2<T> *)lds;
  hook.inv_sw = 1.0 / sw;
  real_forward<T, NTC>(f, A, hook);
  __syncthreads();  // the buffer is free for the next frame
  return true;
}

// spectrogramInto (fft/spectrogram.nim): one workgroup per frame i; a
// frame with no sample inside the trace reads 0.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void spectrogram_kernel(const DevSpec *__restrict__ plan,
                                                          const T *__restrict__ x, int32_t len,
                                                          int32_t hop, double *__restrict__ amp) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red_lds[4 * (NTC / 64)];
  const DevSpec &P = *plan;
  const int bins = P.fft.n / 2 + 1;
  const int64_t frame = blockIdx.x;
  double *out = amp + frame * (int64_t)bins;
  RedSlots red{red_lds, 2 * (NTC / 64), 0};
  if (!spectrogram_frame<T, NTC>(P, x, len, hop, frame, lds, red,
                                 FrameAmp<T>{out, 0.0, P.fft.m, P.fft.n}))
    for (int k = threadIdx.x; k < bins; k += NTC) out[k] = 0;
}

// A frame's amplitudes (as FrameAmp) squared into the workgroup's float64
// sums in LDS, and any NaN or infinity flagged (each bin has one owner
// thread in combine, so plain adds).
template <typename T> struct FrameAmpSq {
  double *acc;
  double inv_sw;
  int m, nfft;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    const double re = (double)x.x, im = (double)x.y;
    const double half = (k != 0 && k != m) ? 0.5 : 1.0;
    const double sides = (k == 0 || 2 * k == nfft) ? 1.0 : 2.0;
    const double a = hypot(re, im) * half * sides * inv_sw;
    acc[k] += a * a;
    return x;
  }
};

// Whether a frame's amplitudes are finite: hypot(re, im)·half·sides/Σw in
// float64 is finite exactly when re and im are (as T) finite, Σw > 0 for a
// frame with a sample; any other bin sets *bad.
template <typename T> struct FrameFinite {
  int *bad;
  double inv_sw;
  __device__ V2<T> operator()(V2<T> x, int) const {
    if (!(x.x - x.x == (T)0 && x.y - x.y == (T)0)) *bad = 1;
    return x;
  }
};

// Which rows averageSpectrogramInto (fft/spectrogram.nim:370-430) averages,
// one workgroup per row: status[r] = 0 for a row left out, 1 for a row
// averaged, 2 for a row whose spectrogram sgram_avg_kernel checks first.
// Out: rows not live or with no live sample, and a float64 row whose live
// samples leave the safe range (a NaN, an infinity, a subnormal, a nonzero
// magnitude below 1e-100 or above 1e140: averagedRow, leavesFloat64Range).
// A float32 row of finite samples within ±1e30 is averaged: no sum of a
// frame's weighted samples reaches FLT_MAX, so its spectrogram is finite;
// any other float32 row is checked (the reference's own test: transform,
// then amplitudesFinite).
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void sgram_status_kernel(const T *__restrict__ src,
                                                           int32_t cols,
                                                           const int32_t *__restrict__ row_len,
                                                           int32_t *__restrict__ status) {
  __shared__ int s_flag;
  const int64_t row = blockIdx.x;
  const int live = min(row_len[row], cols);
  if (live <= 0) {
    if (threadIdx.x == 0) status[row] = 0;
    return;
  }
  const T *x = src + row * (int64_t)cols;
  if (threadIdx.x == 0) s_flag = 0;
  __syncthreads();
  int flag = 0;
  for (int j = threadIdx.x; j < live; j += NTC) {
    const double v = fabs((double)x[j]);
    if (sizeof(T) == 8) {
      if (!(v <= 1e140) || (v != 0 && v < 1e-100)) flag = 1;
    } else {
      if (!(v <= 1e30)) flag = 2;
    }
  }
  if (flag) atomicOr(&s_flag, flag);
  __syncthreads();
  if (threadIdx.x == 0) status[row] = (s_flag & 1) ? 0 : (s_flag & 2) ? 2 : 1;
}

// Frame blockIdx.x of the rows g, g + G, ... (G = gridDim.y, fixed by the
// row and frame counts).  check != 0: the rows of status 2 only, each set to
// 0 when this frame's amplitudes hold a NaN or an infinity (every frame of a
// row is checked by one workgroup or another before any is summed: the
// check is a launch of its own).  check == 0: the squared amplitudes of the
// rows of status 1 or 2 summed in row order, written as the partial
// parts[g][frame][bins]; the (frame 0) workgroups count their rows into
// used[g].
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void sgram_avg_kernel(const DevSpec *__restrict__ plan,
                                                        const T *__restrict__ src,
                                                        int32_t rows, int32_t cols,
                                                        const int32_t *__restrict__ row_len,
                                                        int32_t *__restrict__ status,
                                                        int32_t hop, int32_t frames,
                                                        int32_t check, size_t acc_off,
                                                        double *__restrict__ parts,
                                                        double *__restrict__ used) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red_lds[4 * (NTC / 64)];
  __shared__ int s_bad;
  const DevSpec &P = *plan;
  const int bins = P.fft.n / 2 + 1;
  const int64_t frame = blockIdx.x;
  const int g = blockIdx.y, G = gridDim.y;
  double *acc = (double *)(lds + acc_off);
  if (!check)
    for (int k = threadIdx.x; k < bins; k += NTC) acc[k] = 0;
  __syncthreads();
  RedSlots red{red_lds, 2 * (NTC / 64), 0};
  double n_used = 0;
  for (int64_t r = g; r < rows; r += G) {
    const int st = __builtin_amdgcn_readfirstlane(status[r]);
    if (check ? st != 2 : st == 0) continue;
    const int live = min(row_len[r], cols);
    if (check) {
      if (threadIdx.x == 0) s_bad = 0;
      __syncthreads();
      spectrogram_frame<T, NTC>(P, src + r * (int64_t)cols, live, hop, frame, lds, red,
                                FrameFinite<T>{&s_bad, 0.0});
      if (threadIdx.x == 0 && s_bad) atomicExch(status + r, 0);
      __syncthreads();
    } else {
      spectrogram_frame<T, NTC>(P, src + r * (int64_t)cols, live, hop, frame, lds, red,
                                FrameAmpSq<T>{acc, 0.0, P.fft.m, P.fft.n});
      n_used += 1;
    }
  }
  if (check) return;
  __syncthreads();
  double *part = parts + ((size_t)g * (size_t)frames + (size_t)frame) * (size_t)bins;
  for (int k = threadIdx.x; k < bins; k += NTC) part[k] = acc[k];
  if (frame == 0 && threadIdx.x == 0) used[g] = n_used;
}

__global__ void zero_f64_kernel(double *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// amp[i] = sqrt(Σ_g parts[g][i] / used) over the G partials in order (i =
// frame·bins + k), 0 when no row was used; *used_out = the rows used.
__global__ void sgram_finish_kernel(const double *__restrict__ parts,
                                    const double *__restrict__ used, int32_t groups, int64_t n,
                                    double *__restrict__ amp, double *__restrict__ used_out) {
  double u = 0;
  for (int g = 0; g < groups; g++) u += used[g];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = 0;
    for (int g = 0; g < groups; g++) t += parts[(size_t)g * (size_t)n + (size_t)i];
    amp[i] = u > 0 ? sqrt(t / u) : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && used_out != nullptr) *used_out = u;
}

// The partials of `groups` workgroups, parts[g] = (bins power sums, rows
// used), summed in a fixed order.  A block of 1024 threads takes kFinishBins
// bins and has kFinishLanes = 1024 / kFinishBins lanes of groups per bin:
// lane l adds the groups l, l + kFinishLanes, ... of its bin (a wave reads
// 64 / kFinishBins groups' runs of kFinishBins contiguous bins), then the
// lane sums are added in lane order.  Every block adds the rows used the
// same way.  amp[k] = sqrt(Σ / used) · sides(k) (1 at 0 Hz and Nyquist), 0
// when no row was used; *used_out (when not NULL) = the rows used.
constexpr int kFinishBins = 8;
constexpr int kFinishLanes = 1024 / kFinishBins;
static_assert(1024 % kFinishBins == 0 && kFinishBins <= 64, "bins per finish block");
__global__ __launch_bounds__(1024) void spectrum_finish_kernel(const double *__restrict__ parts,
                                                               int32_t groups, int32_t bins,
                                                               int32_t nfft,
                                                               double *__restrict__ amp,
                                                               double *__restrict__ used_out) {
  __shared__ double s_sum[kFinishLanes][kFinishBins + 1];
  __shared__ double s_used[kFinishLanes];
  const int b = threadIdx.x % kFinishBins, gl = threadIdx.x / kFinishBins;
  const int k = blockIdx.x * kFinishBins + b;
  const size_t pitch = (size_t)bins + 1;
  const int kk = k < bins ? k : bins;  // lanes past the last bin read the used column
  double s = 0, u = 0;
#pragma unroll 4
  for (int g = gl; g < groups; g += kFinishLanes) {
    const double *p = parts + (size_t)g * pitch;
    s += p[kk];
    if (b == 0) u += p[bins];
  }
  s_sum[gl][b] = s;
  if (b == 0) s_used[gl] = u;
  __syncthreads();
  if (threadIdx.x < kFinishBins) {
    double t = 0, used = 0;
    for (int w = 0; w < kFinishLanes; w++) {
      t += s_sum[w][b];
      used += s_used[w];
    }
    if (k < bins) {
      const double sides = (k == 0 || 2 * k == nfft) ? 1.0 : 2.0;
      amp[k] = used > 0 ? sqrt(t / used) * sides : 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && used_out != nullptr) *used_out = used;
  }
}

// -------------------------------------------------------------- launchers

namespace {

template <typename T>
hipError_t panel_typed(const T *x, int32_t rows, int32_t cols, const int32_t *row_len,
                       int32_t stat, double target, double percentile, double *row_sq,
                       uint32_t *row_max, uint32_t *hist, double *factor, hipStream_t stream) {
  hipLaunchKernelGGL((row_stats_kernel<T>), dim3(rows), dim3(256), 0, stream, x, cols, row_len,
                     row_sq, row_max);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (stat == 2 && percentile != 100) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(64), dim3(256), 0, stream, hist, 65536);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((hist_kernel<T>), dim3(rows), dim3(256), 0, stream, x, cols, row_len,
                       row_sq, row_max, hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((panel_factor_kernel<T>), dim3(1), dim3(1024), 0, stream, rows, row_len,
                     row_sq, row_max, hist, stat, target, percentile, factor);
  return hipGetLastError();
}

template <typename T, int NTC>
hipError_t spectrum_rows_typed(const DevSpec *dev, size_t lds, const T *src, int32_t rows,
                               int32_t cols, const int32_t *row_len, int32_t first_col, T *amp,
                               T *phase, hipStream_t stream) {
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)spectrum_rows_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((spectrum_rows_kernel<T, NTC>), dim3(rows), dim3(NTC), lds, stream, dev, src,
                     cols, row_len, first_col, amp, phase);
  return hipGetLastError();
}

template <typename T, int NTC>
hipError_t hilbert_typed(const DevFft *dev, size_t lds, const T *src, int32_t rows, int32_t cols,
                         const int32_t *row_len, int32_t envelope, T *dst, hipStream_t stream) {
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)hilbert_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((hilbert_kernel<T, NTC>), dim3(rows), dim3(NTC), lds, stream, dev, src, cols,
                     row_len, envelope, dst);
  return hipGetLastError();
}

template <typename T, int NTC>
hipError_t analytic_typed(const DevFft *dev, size_t lds, const T *src, int32_t rows, int32_t cols,
                          const int32_t *row_len, double dt_ms, double thr_scale, double *env,
                          double *phase, double *hz, hipStream_t stream) {
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)analytic_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((analytic_kernel<T, NTC>), dim3(rows), dim3(NTC), lds, stream, dev, src, cols,
                     row_len, dt_ms, thr_scale, env, phase, hz);
  return hipGetLastError();
}

template <typename T, int NTC>
hipError_t spectrogram_typed(const DevSpec *dev, size_t lds, const T *x, int32_t len, int32_t hop,
                             int32_t frames, double *amp, hipStream_t stream) {
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)spectrogram_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((spectrogram_kernel<T, NTC>), dim3(frames), dim3(NTC), lds, stream, dev, x,
                     len, hop, amp);
  return hipGetLastError();
}

// The FFT buffer, then the workgroup's bins float64 power sums (16-byte
// aligned); one workgroup per CU-slot pass over the rows.
template <typename T, int NTC>
hipError_t spectrum_typed(const DevSpec *dev, size_t lds, int32_t bins, const T *src,
                          int32_t rows, int32_t cols, const int32_t *row_len, int32_t first_col,
                          double *parts, hipStream_t stream) {
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)spectrum_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (attr != hipSuccess) return attr;
  const size_t acc_off = (lds + 15) / 16 * 16;
  const size_t total = acc_off + (size_t)bins * sizeof(double);
  const int grid = spectrum_groups(rows, bins);
  hipLaunchKernelGGL((spectrum_kernel<T, NTC>), dim3(grid), dim3(NTC), total, stream, dev, src,
                     rows, cols, row_len, first_col, acc_off, parts);
  return hipGetLastError();
}

template <typename T, int NTC>
hipError_t corr_typed(const DevCorr *dplan, size_t lds, const T *src, int32_t rows,
                      int32_t src_cols, const int32_t *row_len, T *dst, hipStream_t stream) {
  // once per process (a thread-safe static), like the pipeline kernels
  static const hipError_t attr = hipFuncSetAttribute(
      (const void *)corr_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((corr_kernel<T, NTC>), dim3(rows), dim3(NTC), lds, stream, dplan, src,
                     src_cols, row_len, dst);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_pipeline(int dtype, int threads, const JobLaunch &a, hipStream_t stream) {
  if (a.total <= 0) return hipSuccess;
  switch (threads) {
    case 128: return launch_pipeline_nt128(dtype, a, stream);
    case 256: return launch_pipeline_nt256(dtype, a, stream);
    case 512: return launch_pipeline_nt512(dtype, a, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t pipeline_resident_blocks(int dtype, int threads, int max_cols, size_t lds, int full,
                                    int *blocks) {
  switch (threads) {
    case 128: return resident_pipeline_nt128(dtype, max_cols, lds, full, blocks);
    case 256: return resident_pipeline_nt256(dtype, max_cols, lds, full, blocks);
    case 512: return resident_pipeline_nt512(dtype, max_cols, lds, full, blocks);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_panel_factor(int dtype, const void *x, int32_t rows, int32_t cols,
                               const int32_t *row_len, int32_t stat, double target,
                               double percentile, double *row_sq, uint32_t *row_max,
                               uint32_t *hist, double *factor, hipStream_t stream) {
  if (dtype == 0)
    return panel_typed<float>((const float *)x, rows, cols, row_len, stat, target, percentile,
                              row_sq, row_max, hist, factor, stream);
  return panel_typed<double>((const double *)x, rows, cols, row_len, stat, target, percentile,
                             row_sq, row_max, hist, factor, stream);
}

hipError_t launch_correlate(int dtype, int threads, const DevCorr *dev_plan, size_t lds_bytes,
                            const void *src, int32_t rows, int32_t src_cols,
                            const int32_t *row_len, void *dst, hipStream_t stream) {
  if (rows <= 0) return hipSuccess;
  const float *sf = (const float *)src;
  const double *sd = (const double *)src;
  if (threads == 512) {
    if (dtype == 0)
      return corr_typed<float, 512>(dev_plan, lds_bytes, sf, rows, src_cols, row_len,
                                    (float *)dst, stream);
    return corr_typed<double, 512>(dev_plan, lds_bytes, sd, rows, src_cols, row_len,
                                   (double *)dst, stream);
  }
  if (dtype == 0)
    return corr_typed<float, 1024>(dev_plan, lds_bytes, sf, rows, src_cols, row_len,
                                   (float *)dst, stream);
  return corr_typed<double, 1024>(dev_plan, lds_bytes, sd, rows, src_cols, row_len,
                                  (double *)dst, stream);
}

}  // namespace pseg

namespace pseg {

hipError_t launch_spectrum(int dtype, int threads, const DevSpec *dev, size_t lds, const void *src,
                           int32_t rows, int32_t cols, const int32_t *row_len, int32_t first_col,
                           double *parts, double *amp, double *used, int32_t bins, int32_t nfft,
                           hipStream_t s) {
  hipError_t e = hipSuccess;
  const int32_t groups = rows > 0 && cols > 0 ? spectrum_groups(rows, bins) : 0;
  if (groups > 0) {
    if (dtype == 0) {
      const float *x = (const float *)src;
      e = threads == 256 ? spectrum_typed<float, 256>(dev, lds, bins, x, rows, cols, row_len, first_col, parts, s)
        : threads == 512 ? spectrum_typed<float, 512>(dev, lds, bins, x, rows, cols, row_len, first_col, parts, s)
                         : spectrum_typed<float, 1024>(dev, lds, bins, x, rows, cols, row_len, first_col, parts, s);
    } else {
      const double *x = (const double *)src;
      e = threads == 256 ? spectrum_typed<double, 256>(dev, lds, bins, x, rows, cols, row_len, first_col, parts, s)
        : threads == 512 ? spectrum_typed<double, 512>(dev, lds, bins, x, rows, cols, row_len, first_col, parts, s)
                         : spectrum_typed<double, 1024>(dev, lds, bins, x, rows, cols, row_len, first_col, parts, s);
    }
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(spectrum_finish_kernel, dim3((bins + kFinishBins - 1) / kFinishBins),
                     dim3(1024), 0, s, parts,
                     groups, bins, nfft, amp, used);
  return hipGetLastError();
}

}  // namespace pseg

namespace pseg {

hipError_t launch_spectrum_rows(int dtype, int threads, const DevSpec *dev, size_t lds,
                                const void *src, int32_t rows, int32_t cols,
                                const int32_t *row_len, int32_t first_col, void *amp, void *phase,
                                hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  if (dtype == 0) {
    const float *x = (const float *)src;
    float *a = (float *)amp, *p = (float *)phase;
    return threads == 256 ? spectrum_rows_typed<float, 256>(dev, lds, x, rows, cols, row_len, first_col, a, p, s)
         : threads == 512 ? spectrum_rows_typed<float, 512>(dev, lds, x, rows, cols, row_len, first_col, a, p, s)
                          : spectrum_rows_typed<float, 1024>(dev, lds, x, rows, cols, row_len, first_col, a, p, s);
  }
  const double *x = (const double *)src;
  double *a = (double *)amp, *p = (double *)phase;
  return threads == 256 ? spectrum_rows_typed<double, 256>(dev, lds, x, rows, cols, row_len, first_col, a, p, s)
       : threads == 512 ? spectrum_rows_typed<double, 512>(dev, lds, x, rows, cols, row_len, first_col, a, p, s)
                        : spectrum_rows_typed<double, 1024>(dev, lds, x, rows, cols, row_len, first_col, a, p, s);
}

}  // namespace pseg

namespace pseg {

hipError_t launch_hilbert(int dtype, int threads, const DevFft *dev, size_t lds, const void *src,
                          int32_t rows, int32_t cols, const int32_t *row_len, int32_t envelope,
                          void *dst, hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  if (dtype == 0) {
    const float *x = (const float *)src;
    float *d = (float *)dst;
    return threads == 256 ? hilbert_typed<float, 256>(dev, lds, x, rows, cols, row_len, envelope, d, s)
         : threads == 512 ? hilbert_typed<float, 512>(dev, lds, x, rows, cols, row_len, envelope, d, s)
                          : hilbert_typed<float, 1024>(dev, lds, x, rows, cols, row_len, envelope, d, s);
  }
  const double *x = (const double *)src;
  double *d = (double *)dst;
  return threads == 256 ? hilbert_typed<double, 256>(dev, lds, x, rows, cols, row_len, envelope, d, s)
       : threads == 512 ? hilbert_typed<double, 512>(dev, lds, x, rows, cols, row_len, envelope, d, s)
                        : hilbert_typed<double, 1024>(dev, lds, x, rows, cols, row_len, envelope, d, s);
}

hipError_t launch_analytic(int dtype, int threads, const DevFft *dev, size_t lds, const void *src,
                           int32_t rows, int32_t cols, const int32_t *row_len, double dt_ms,
                           double thr_scale, double *env, double *phase, double *hz,
                           hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  if (dtype == 0) {
    const float *x = (const float *)src;
    return threads == 256 ? analytic_typed<float, 256>(dev, lds, x, rows, cols, row_len, dt_ms, thr_scale, env, phase, hz, s)
         : threads == 512 ? analytic_typed<float, 512>(dev, lds, x, rows, cols, row_len, dt_ms, thr_scale, env, phase, hz, s)
                          : analytic_typed<float, 1024>(dev, lds, x, rows, cols, row_len, dt_ms, thr_scale, env, phase, hz, s);
  }
  const double *x = (const double *)src;
  return threads == 256 ? analytic_typed<double, 256>(dev, lds, x, rows, cols, row_len, dt_ms, thr_scale, env, phase, hz, s)
       : threads == 512 ? analytic_typed<double, 512>(dev, lds, x, rows, cols, row_len, dt_ms, thr_scale, env, phase, hz, s)
                        : analytic_typed<double, 1024>(dev, lds, x, rows, cols, row_len, dt_ms, thr_scale, env, phase, hz, s);
}

template <typename T, int NTC>
hipError_t sgram_avg_typed(const DevSpec *dev, size_t lds, int32_t bins, const T *src,
                           int32_t rows, int32_t cols, const int32_t *row_len, int32_t hop,
                           int32_t frames, int32_t groups, int32_t *status, double *parts,
                           double *used, double *amp, double *used_out, hipStream_t stream) {
  static const hipError_t a2 = hipFuncSetAttribute(
      (const void *)sgram_avg_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (a2 != hipSuccess) return a2;
  hipLaunchKernelGGL((sgram_status_kernel<T, NTC>), dim3(rows), dim3(NTC), 0, stream, src, cols,
                     row_len, status);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t acc_off = (lds + 15) / 16 * 16;
  const size_t total = acc_off + (size_t)bins * sizeof(double);
  for (int check = 1; check >= 0; check--) {
    hipLaunchKernelGGL((sgram_avg_kernel<T, NTC>), dim3(frames, groups), dim3(NTC), total, stream,
                       dev, src, rows, cols, row_len, status, hop, frames, check, acc_off, parts,
                       used);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int64_t n = (int64_t)frames * bins;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256));
  hipLaunchKernelGGL(sgram_finish_kernel, dim3(grid), dim3(256), 0, stream, parts, used, groups,
                     n, amp, used_out);
  return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256));
  hipLaunchKernelGGL(zero_f64_kernel, dim3(grid), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_spectrogram_average(int dtype, int threads, const DevSpec *dev, size_t lds,
                                      int32_t bins, const void *src, int32_t rows, int32_t cols,
                                      const int32_t *row_len, int32_t hop, int32_t frames,
                                      int32_t groups, int32_t *status, double *parts,
                                      double *used, double *amp, double *used_out,
                                      hipStream_t s) {
  if (dtype == 0) {
    const float *x = (const float *)src;
    return threads == 64 ? sgram_avg_typed<float, 64>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s)
         : threads == 256 ? sgram_avg_typed<float, 256>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s)
         : threads == 512 ? sgram_avg_typed<float, 512>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s)
                          : sgram_avg_typed<float, 1024>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s);
  }
  const double *x = (const double *)src;
  return threads == 64 ? sgram_avg_typed<double, 64>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s)
       : threads == 256 ? sgram_avg_typed<double, 256>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s)
       : threads == 512 ? sgram_avg_typed<double, 512>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s)
                        : sgram_avg_typed<double, 1024>(dev, lds, bins, x, rows, cols, row_len, hop, frames, groups, status, parts, used, amp, used_out, s);
}

}  // namespace pseg

namespace pseg {

hipError_t launch_spectrogram(int dtype, int threads, const DevSpec *dev, size_t lds,
                              const void *x, int32_t len, int32_t hop, int32_t frames,
                              double *amp, hipStream_t s) {
  if (frames == 0) return hipSuccess;
  if (dtype == 0) {
    const float *v = (const float *)x;
    return threads == 256 ? spectrogram_typed<float, 256>(dev, lds, v, len, hop, frames, amp, s)
         : threads == 512 ? spectrogram_typed<float, 512>(dev, lds, v, len, hop, frames, amp, s)
                          : spectrogram_typed<float, 1024>(dev, lds, v, len, hop, frames, amp, s);
  }
  const double *v = (const double *)x;
  return threads == 256 ? spectrogram_typed<double, 256>(dev, lds, v, len, hop, frames, amp, s)
       : threads == 512 ? spectrogram_typed<double, 512>(dev, lds, v, len, hop, frames, amp, s)
                        : spectrogram_typed<double, 1024>(dev, lds, v, len, hop, frames, amp, s);
}

}  // namespace pseg

// ------------------------------------------- the allocation-free spectrum

namespace {

// The partial sums of pseg_spectrum_average_noalloc: kSpecPartials float64
// (16 MiB) of static device memory, one copy per GPU the library runs on,
// held until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) double scratch[pseg::kSpecPartials];

constexpr int kMaxScratchDevices = 64;

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
hipError_t scratch_address(double **out) {
  static std::mutex mu;
  static double *addr[kMaxScratchDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = (double *)p;
  }
  *out = addr[dev];
  return hipSuccess;
}

}  // namespace

namespace pseg {

hipError_t launch_spectrum_noalloc(int dtype, int threads, const DevSpec *dev, size_t lds,
                                   const void *src, int32_t rows, int32_t cols,
                                   const int32_t *row_len, int32_t first_col, double *amp,
                                   double *used, int32_t bins, int32_t nfft, hipStream_t s) {
  double *parts = nullptr;
  hipError_t e = scratch_address(&parts);
  if (e != hipSuccess) return e;
  // every workgroup writes its whole partial before the finish kernel reads
  // it: nothing in the scratch needs clearing
  return launch_spectrum(dtype, threads, dev, lds, src, rows, cols, row_len, first_col, parts, amp,
                         used, bins, nfft, s);
}

}  // namespace pseg

// ------------------------------------------------------------------- f-k
//
// fkAmplitudeInto (fft/fk.nim:438-504) in two launches.  The time pass: one
// workgroup per trace that enters, its window tapered on both axes (the
// sample times the time weight, then the trace weight, in T), zero-padded to
// nTime and through the real FFT, the half spectrum X[k] (T pairs) into the
// workspace, trace-major.  The spatial pass: one workgroup per frequency
// bin, as the reference: the real parts across traces and the imaginary
// parts, each zero-padded to nK, through the real FFT; the complex
// transform of (re + i im) read off the two half spectra with the sign flip
// that puts later-on-later events at positive k (fft/fk.nim:400-410), the
// modulus in float64, sqrt(re^2 + im^2), or hypot when a float64 sample of
// the window leaves [1e-100, 1e140] (flag set by the time pass).

namespace pseg {
namespace {

// The half spectrum of the time pass: X[k] of trace ti (the split gives
// 2 X[k] between 0 Hz and Nyquist).
template <typename T> struct FkSpecOut {
  V2<T> *spec;
  int m;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    spec[k] = (k != 0 && k != m) ? mk<T>(x.x * (T)0.5, x.y * (T)0.5) : x;
    return x;
  }
};

// The first spatial transform (the real parts): its half spectrum into LDS.
template <typename T> struct FkKeepA {
  V2<T> *a;
  int m;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    a[k] = (k != 0 && k != m) ? mk<T>(x.x * (T)0.5, x.y * (T)0.5) : x;
    return x;
  }
};

// The second (the imaginary parts): bin k of both half spectra gives the
// outputs whose spatial bin is k (ik = nK/2 - k) and nK - k (ik = nK/2 + k,
// fft/fk.nim:400-410).
template <typename T> struct FkOut {
  const V2<T> *a;
  double *amp;
  int m, n_freq, jf;
  bool extreme;
  __device__ double modulus(double re, double im) const {
    return extreme ? hypot(re, im) : sqrt(re * re + im * im);
  }
  __device__ V2<T> operator()(V2<T> x, int k) const {
    const V2<T> b = (k != 0 && k != m) ? mk<T>(x.x * (T)0.5, x.y * (T)0.5) : x;
    const V2<T> av = a[k];
    const double ar = (double)av.x, ai = (double)av.y, br = (double)b.x, bi = (double)b.y;
    const int half = m;  // nK / 2
    // sk = k <= half: re = ar - bi, im = ai + br
    amp[(int64_t)(half - k) * n_freq + jf] = modulus(ar - bi, ai + br);
    if (k != 0 && k != half) {
      // sk = nK - k > half: re = ar + bi, im = br - ai
      amp[(int64_t)(half + k) * n_freq + jf] = modulus(ar + bi, br - ai);
    }
    return x;
  }
};

// leavesFloat64Range (fft/fk.nim:164-181) of one sample.
__device__ __forceinline__ bool fk_leaves(double x) {
  const double a = fabs(x);
  return !(a <= 1e140) || (a < 1e-100 && a != 0);
}

template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void fk_time_kernel(const DevFk *__restrict__ plan,
                                                      const T *__restrict__ src, int32_t cols,
                                                      const int32_t *__restrict__ row_len,
                                                      const int32_t *__restrict__ live_rows,
                                                      int32_t first_col, V2<T> *__restrict__ spec,
                                                      uint32_t *__restrict__ flag) {
  extern __shared__ __align__(16) char lds[];
  const DevFk &P = *plan;
  const DevFft &f = P.time;
  const int ti = blockIdx.x;
  const int64_t r = live_rows[ti];
  const int live = min(row_len[r], cols);
  const int n = max(min(live, first_col + P.samples) - first_col, 0);
  const T *s = src + r * (int64_t)cols + first_col;
  gptr<T> tw = as_global<T>(P.time_w);
  const T trw = as_global<T>(P.trace_w)[ti];
  T *buf = (T *)lds;
  bool bad = false;
  for (int j = threadIdx.x; j < f.n; j += NTC) {
    T y = (T)0;
    if (j < n) {
      const T x = s[j];
      if (sizeof(T) == 8) bad |= fk_leaves((double)x);
      y = (x * tw[j]) * trw;
    }
    buf[j] = y;
  }
  if (sizeof(T) == 8 && bad) atomicOr(flag, 1u);
  __syncthreads();
  real_forward<T, NTC>(f, (V2<T> *)lds, FkSpecOut<T>{spec + (int64_t)ti * P.n_freq, f.m});
}

template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void fk_space_kernel(const DevFk *__restrict__ plan,
                                                       const V2<T> *__restrict__ spec,
                                                       const uint32_t *__restrict__ flag,
                                                       double *__restrict__ amp) {
  extern __shared__ __align__(16) char lds[];
  const DevFk &P = *plan;
  const DevFft &f = P.space;
  const int traces = P.traces, nf = P.n_freq;
  // Workgroups go round robin over the 8 XCDs (workgroup b on XCD b mod 8):
  // each XCD takes a contiguous range of frequencies, so the 128-byte lines
  // of the time pass's spectra (consecutive frequencies of one trace) that
  // its workgroups read are shared in its own L2.  The grid is 8 x ceil(nf /
  // 8); workgroups past the last frequency leave at once.
  const int per = (nf + 7) / 8;
  const int jf = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (jf >= nf) return;
  V2<T> *A = (V2<T> *)lds;
  T *buf = (T *)lds;
  V2<T> *keep = (V2<T> *)(lds + P.space_lds);
  T *im = P.im_lds ? (T *)(lds + P.im_lds) : nullptr;
  const bool extreme = sizeof(T) == 8 && *flag != 0;
  // the real parts across traces, zero-padded to nK (the imaginary parts
  // kept beside them when there is room)
  for (int t = threadIdx.x; t < f.n; t += NTC) {
    T re = (T)0;
    if (t < traces) {
      const V2<T> v = spec[(int64_t)t * nf + jf];
      re = v.x;
      if (im) im[t] = v.y;
    }
    buf[t] = re;
  }
  __syncthreads();
  real_forward<T, NTC>(f, A, FkKeepA<T>{keep, f.m});
  __syncthreads();
  // the imaginary parts
  for (int t = threadIdx.x; t < f.n; t += NTC)
    buf[t] = t < traces ? (im ? im[t] : spec[(int64_t)t * nf + jf].y) : (T)0;
  __syncthreads();
  real_forward<T, NTC>(f, A, FkOut<T>{keep, amp, f.m, nf, jf, extreme});
}

template <typename T, int NTC>
hipError_t fk_typed(const DevFk *dev, size_t lds_time, size_t lds_space, int traces, int n_freq,
                    const T *src, int32_t cols, const int32_t *row_len, const int32_t *live_rows,
                    int32_t first_col, V2<T> *spec, uint32_t *flag, double *amp,
                    hipStream_t stream) {
  static const hipError_t a1 = hipFuncSetAttribute(
      (const void *)fk_time_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  static const hipError_t a2 = hipFuncSetAttribute(
      (const void *)fk_space_kernel<T, NTC>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  if (a1 != hipSuccess) return a1;
  if (a2 != hipSuccess) return a2;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, stream, flag, 1);
  hipLaunchKernelGGL((fk_time_kernel<T, NTC>), dim3(traces), dim3(NTC), lds_time, stream, dev, src,
                     cols, row_len, live_rows, first_col, spec, flag);
  hipLaunchKernelGGL((fk_space_kernel<T, NTC>), dim3(8 * ((n_freq + 7) / 8)), dim3(NTC), lds_space, stream, dev,
                     (const V2<T> *)spec, (const uint32_t *)flag, amp);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_fk(int dtype, int threads, const DevFk *dev, size_t lds_time, size_t lds_space,
                     int traces, int n_freq, const void *src, int32_t cols,
                     const int32_t *row_len, const int32_t *live_rows, int32_t first_col,
                     void *spec, uint32_t *flag, double *amp, hipStream_t s) {
  if (dtype == 0) {
    const float *x = (const float *)src;
    float2 *sp = (float2 *)spec;
    return threads == 256 ? fk_typed<float, 256>(dev, lds_time, lds_space, traces, n_freq, x, cols, row_len, live_rows, first_col, sp, flag, amp, s)
         : threads == 512 ? fk_typed<float, 512>(dev, lds_time, lds_space, traces, n_freq, x, cols, row_len, live_rows, first_col, sp, flag, amp, s)
                          : fk_typed<float, 1024>(dev, lds_time, lds_space, traces, n_freq, x, cols, row_len, live_rows, first_col, sp, flag, amp, s);
  }
  const double *x = (const double *)src;
  double2 *sp = (double2 *)spec;
  return threads == 256 ? fk_typed<double, 256>(dev, lds_time, lds_space, traces, n_freq, x, cols, row_len, live_rows, first_col, sp, flag, amp, s)
       : threads == 512 ? fk_typed<double, 512>(dev, lds_time, lds_space, traces, n_freq, x, cols, row_len, live_rows, first_col, sp, flag, amp, s)
                        : fk_typed<double, 1024>(dev, lds_time, lds_space, traces, n_freq, x, cols, row_len, live_rows, first_col, sp, flag, amp, s);
}

}  // namespace pseg
